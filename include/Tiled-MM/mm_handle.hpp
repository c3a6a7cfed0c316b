// Context of the out-of-core GEMM: the reference's gpu::mm_handle<T> and make_context
// (Tiled-MM src/Tiled-MM/mm_handle.hpp:11-76), backed by the C ABI of libtiledmm.so.
// Instantiated in the library for float, double, std::complex<float>, std::complex<double>.
#pragma once
#include <memory>
#include <tuple>

#include "../tiledmm.h"
#include "device_buffer.hpp"
#include "device_stream.hpp"
#include "device_vector.hpp"
#include "gpu_context.hpp"

namespace gpu {

template <typename Scalar>
class mm_handle {
 public:
  // Creates the streams on the current device; no device memory is reserved yet.
  mm_handle(int streams, int max_tile_m, int max_tile_n, int max_tile_k);
  ~mm_handle();

  mm_handle(mm_handle&&) = delete;
  mm_handle(const mm_handle&) = delete;
  mm_handle& operator=(const mm_handle&& other) = delete;

  void set_num_streams(int streams);
  int get_num_streams();

  gpu_context& get_gpu_context();

  void set_tile_sizes(int tile_size_m, int tile_size_n, int tile_size_k);
  void set_tile_sizes(int tile_size);
  void set_full_sizes(int m, int n, int k);
  // tile sizes the reference policy picks for a problem of size m, n, k
  std::tuple<int, int, int> optimal_tile_sizes(int m, int n, int k);
  std::tuple<int, int, int> get_max_tile_sizes();

  void set_streams_and_tiles(int streams, int tile_size_m, int tile_size_n, int tile_size_k);

  device_buffer<Scalar>& get_device_buffer_a();
  device_buffer<Scalar>& get_device_buffer_b();
  device_buffer<Scalar>& get_device_buffer_c();
  // resident C (column-major, ld = m) after a gemm call with copy_c_back = false
  device_vector<Scalar>& get_full_device_buffer_c();

  // The C handle underneath (extension).
  tmm_handle* native_handle() { return h_; }

 private:
  void sync_buffers();
  tmm_handle* h_ = nullptr;
  std::unique_ptr<gpu_context> ctx_;
  device_buffer<Scalar> a_buff, b_buff, c_buff;
  device_vector<Scalar> full_c_view;
};

template <typename Scalar>
std::unique_ptr<mm_handle<Scalar>> make_context(int streams, int max_tile_m, int max_tile_n, int max_tile_k) {
  return std::make_unique<mm_handle<Scalar>>(streams, max_tile_m, max_tile_n, max_tile_k);
}

template <typename Scalar>
std::unique_ptr<mm_handle<Scalar>> make_context() {
  return std::make_unique<mm_handle<Scalar>>(2, 5000, 5000, 5000);
}

}  // namespace gpu

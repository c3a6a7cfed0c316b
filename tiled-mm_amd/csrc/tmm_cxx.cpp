// C++ drop-in API (include/Tiled-MM/*.hpp) over the C ABI: gpu::mm_handle<T>, gpu::gemm<T>
// and their explicit instantiations for the reference's four element types
// (Tiled-MM src/Tiled-MM/tiled_mm.cpp:626-668, mm_handle.cpp:167-170).
#include <complex>
#include <iostream>
#include <stdexcept>
#include <string>

#include "../../include/Tiled-MM/tiled_mm.hpp"

namespace gpu {

namespace {

template <typename T> struct dtype_of;
template <> struct dtype_of<float> { static constexpr int value = TMM_F32; };
template <> struct dtype_of<double> { static constexpr int value = TMM_F64; };
template <> struct dtype_of<std::complex<float>> { static constexpr int value = TMM_C32; };
template <> struct dtype_of<std::complex<double>> { static constexpr int value = TMM_C64; };

// Invalid arguments raise std::runtime_error with the message; device failures print the
// message and raise std::runtime_error("GPU ERROR") like the reference (util.hpp:13-36).
void raise(int status, const char* msg) {
  if (status == TMM_SUCCESS) return;
  if (status == TMM_ERR_HIP || status == TMM_ERR_ALLOC) {
    std::cerr << "error: GPU API call : " << msg << std::endl;
    throw std::runtime_error("GPU ERROR");
  }
  throw std::runtime_error(msg);
}

int call_gemm(tmm_handle* h, float, char ta, char tb, int m, int n, int k, const float* al, const float* a, int lda,
              const float* b, int ldb, const float* be, float* c, int ldc, int pin, int cb) {
  return tmm_sgemm(h, ta, tb, m, n, k, al, a, lda, b, ldb, be, c, ldc, pin, cb);
}
int call_gemm(tmm_handle* h, double, char ta, char tb, int m, int n, int k, const double* al, const double* a,
              int lda, const double* b, int ldb, const double* be, double* c, int ldc, int pin, int cb) {
  return tmm_dgemm(h, ta, tb, m, n, k, al, a, lda, b, ldb, be, c, ldc, pin, cb);
}
int call_gemm(tmm_handle* h, std::complex<float>, char ta, char tb, int m, int n, int k, const std::complex<float>* al,
              const std::complex<float>* a, int lda, const std::complex<float>* b, int ldb,
              const std::complex<float>* be, std::complex<float>* c, int ldc, int pin, int cb) {
  return tmm_cgemm(h, ta, tb, m, n, k, (const float*)al, (const float*)a, lda, (const float*)b, ldb,
                   (const float*)be, (float*)c, ldc, pin, cb);
}
int call_gemm(tmm_handle* h, std::complex<double>, char ta, char tb, int m, int n, int k,
              const std::complex<double>* al, const std::complex<double>* a, int lda, const std::complex<double>* b,
              int ldb, const std::complex<double>* be, std::complex<double>* c, int ldc, int pin, int cb) {
  return tmm_zgemm(h, ta, tb, m, n, k, (const double*)al, (const double*)a, lda, (const double*)b, ldb,
                   (const double*)be, (double*)c, ldc, pin, cb);
}

}  // namespace

template <typename Scalar>
mm_handle<Scalar>::mm_handle(int streams, int max_tile_m, int max_tile_n, int max_tile_k) {
  int st = tmm_create(&h_, dtype_of<Scalar>::value, streams, max_tile_m, max_tile_n, max_tile_k);
  raise(st, tmm_last_error(nullptr));
  ctx_.reset(new gpu_context(h_));
  a_buff = device_buffer<Scalar>(streams);
  b_buff = device_buffer<Scalar>(streams);
  c_buff = device_buffer<Scalar>(streams);
}

template <typename Scalar>
mm_handle<Scalar>::~mm_handle() {
  ctx_.reset();
  tmm_destroy(h_);
}

template <typename Scalar>
void mm_handle<Scalar>::set_num_streams(int streams) {
  ctx_->set_num_streams(streams);
  a_buff.set_num_streams(streams);
  b_buff.set_num_streams(streams);
  c_buff.set_num_streams(streams);
}

template <typename Scalar>
int mm_handle<Scalar>::get_num_streams() {
  int n = 0;
  tmm_get_num_streams(h_, &n);
  return n;
}

template <typename Scalar>
gpu_context& mm_handle<Scalar>::get_gpu_context() {
  return *ctx_;
}

template <typename Scalar>
void mm_handle<Scalar>::sync_buffers() {
  int tm = 0, tn = 0, tk = 0;
  tmm_get_tile_sizes(h_, &tm, &tn, &tk);
  a_buff.set_tile_sizes({tm, tk});
  b_buff.set_tile_sizes({tk, tn});
  c_buff.set_tile_sizes({tm, tn});
}

template <typename Scalar>
void mm_handle<Scalar>::set_tile_sizes(int tile_m, int tile_n, int tile_k) {
  raise(tmm_set_tile_sizes(h_, tile_m, tile_n, tile_k), tmm_last_error(h_));
  sync_buffers();
}

template <typename Scalar>
void mm_handle<Scalar>::set_tile_sizes(int tile_size) {
  set_tile_sizes(tile_size, tile_size, tile_size);
}

template <typename Scalar>
void mm_handle<Scalar>::set_full_sizes(int m, int n, int k) {
  raise(tmm_set_full_sizes(h_, m, n, k), tmm_last_error(h_));
}

template <typename Scalar>
std::tuple<int, int, int> mm_handle<Scalar>::optimal_tile_sizes(int m, int n, int k) {
  int tm = 0, tn = 0, tk = 0;
  raise(tmm_optimal_tile_sizes(h_, m, n, k, &tm, &tn, &tk), tmm_last_error(h_));
  return std::tuple<int, int, int>{tm, tn, tk};
}

template <typename Scalar>
std::tuple<int, int, int> mm_handle<Scalar>::get_max_tile_sizes() {
  int tm = 0, tn = 0, tk = 0;
  tmm_get_max_tile_sizes(h_, &tm, &tn, &tk);
  return std::tuple<int, int, int>{tm, tn, tk};
}

template <typename Scalar>
void mm_handle<Scalar>::set_streams_and_tiles(int streams, int tile_m, int tile_n, int tile_k) {
  raise(tmm_set_tile_sizes(h_, tile_m, tile_n, tile_k), tmm_last_error(h_));
  set_num_streams(streams);
  sync_buffers();
}

template <typename Scalar>
device_buffer<Scalar>& mm_handle<Scalar>::get_device_buffer_a() {
  return a_buff;
}
template <typename Scalar>
device_buffer<Scalar>& mm_handle<Scalar>::get_device_buffer_b() {
  return b_buff;
}
template <typename Scalar>
device_buffer<Scalar>& mm_handle<Scalar>::get_device_buffer_c() {
  return c_buff;
}

template <typename Scalar>
device_vector<Scalar>& mm_handle<Scalar>::get_full_device_buffer_c() {
  void* p = nullptr;
  size_t n = 0;
  tmm_full_c_device_ptr(h_, &p, &n);
  full_c_view = device_vector<Scalar>::view((Scalar*)p, n);
  return full_c_view;
}

template <typename Scalar>
void gemm(mm_handle<Scalar>& handle, char trans_a, char trans_b, int m, int n, int k, Scalar alpha, Scalar* a,
          int ld_a, Scalar* b, int ld_b, Scalar beta, Scalar* c, int ld_c, bool pin_host_buffers, bool copy_c_back) {
  tmm_handle* h = handle.native_handle();
  const int st = call_gemm(h, Scalar{}, trans_a, trans_b, m, n, k, &alpha, a, ld_a, b, ld_b, &beta, c, ld_c,
                           pin_host_buffers ? 1 : 0, copy_c_back ? 1 : 0);
  raise(st, tmm_last_error(h));
}

template class mm_handle<float>;
template class mm_handle<double>;
template class mm_handle<std::complex<float>>;
template class mm_handle<std::complex<double>>;

template void gemm<float>(mm_handle<float>&, char, char, int, int, int, float, float*, int, float*, int, float,
                          float*, int, bool, bool);
template void gemm<double>(mm_handle<double>&, char, char, int, int, int, double, double*, int, double*, int, double,
                           double*, int, bool, bool);
template void gemm<std::complex<float>>(mm_handle<std::complex<float>>&, char, char, int, int, int,
                                        std::complex<float>, std::complex<float>*, int, std::complex<float>*, int,
                                        std::complex<float>, std::complex<float>*, int, bool, bool);
template void gemm<std::complex<double>>(mm_handle<std::complex<double>>&, char, char, int, int, int,
                                         std::complex<double>, std::complex<double>*, int, std::complex<double>*, int,
                                         std::complex<double>, std::complex<double>*, int, bool, bool);

}  // namespace gpu

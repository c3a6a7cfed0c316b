// Streams of a handle (reference: Tiled-MM src/Tiled-MM/gpu_context.hpp:13-40): n_streams
// compute streams plus one result stream used for device-to-host copies. The vendor-BLAS
// handles of the reference (get_blas_handle) have no counterpart: the tile GEMM is this
// library's own kernel.
#pragma once
#include <memory>
#include <stdexcept>
#include <vector>

#include "../tiledmm.h"
#include "device_event.hpp"
#include "device_stream.hpp"

namespace gpu {

class gpu_context {
 public:
  explicit gpu_context(tmm_handle* h) : h_(h) { refresh(); }
  gpu_context(gpu_context&&) = delete;
  gpu_context(gpu_context&) = delete;

  hipStream_t get_stream(int stream_id) const {
    check(stream_id);
    return streams_[stream_id].stream();
  }
  device_stream& get_device_stream(int stream_id) {
    check(stream_id);
    return streams_[stream_id];
  }
  device_event enqueue_event(int stream_id) const {
    check(stream_id);
    return streams_[stream_id].enqueue_event();
  }
  int get_num_streams() const { return (int)streams_.size(); }
  void set_num_streams(int streams) {
    if (tmm_set_num_streams(h_, streams) != TMM_SUCCESS) throw std::runtime_error(tmm_last_error(h_));
    refresh();
  }
  device_stream& get_result_stream() { return *result_; }

  // Re-read the handle's streams (after the stream count changed).
  void refresh() {
    int n = 0;
    tmm_get_num_streams(h_, &n);
    streams_.clear();
    for (int i = 0; i < n; ++i) {
      void* s = nullptr;
      tmm_get_stream(h_, i, &s);
      streams_.emplace_back((hipStream_t)s);
    }
    void* r = nullptr;
    tmm_get_stream(h_, -1, &r);
    result_.reset(new device_stream((hipStream_t)r));
  }

 private:
  void check(int stream_id) const {
    if (stream_id < 0 || stream_id >= (int)streams_.size())
      throw std::runtime_error("index of gpu stream has to be in the range [0, n_streams)");
  }
  tmm_handle* h_;
  std::vector<device_stream> streams_;
  std::unique_ptr<device_stream> result_;
};

}  // namespace gpu

// Per-stream tile slots: n_streams x tile.size() elements, stream s owning the slice at
// s*tile.size() (reference: Tiled-MM src/Tiled-MM/device_buffer.hpp:13-83). The device
// memory is allocated when a slot is first requested, not when tile sizes are set: the
// out-of-core scheduler of this library keeps its own work pool and does not use these slots.
#pragma once
#include <stdexcept>

#include "device_vector.hpp"
#include "tile_dim.hpp"

namespace gpu {

template <typename T>
class device_buffer {
 public:
  device_buffer() = default;
  explicit device_buffer(int streams) : n_streams_(streams) {}

  T* stream_buffer(int stream_id) {
    // the reference builds this error without throwing it (device_buffer.hpp:46-48)
    if (stream_id < 0 || stream_id >= n_streams_)
      throw std::runtime_error("stream id in device buffer has to be in the range [0, n_streams)");
    ensure();
    return d_vec_.data() + (std::size_t)stream_id * tile_.size();
  }
  T* data() {
    ensure();
    return d_vec_.data();
  }
  void set_num_streams(int streams) { n_streams_ = streams; }
  void set_tile_sizes(tile_dim tile) { tile_ = tile; }
  tile_dim get_tile_sizes() { return tile_; }
  void set_streams_and_tiles(int streams, tile_dim tile) {
    n_streams_ = streams;
    tile_ = tile;
  }

 private:
  void ensure() {
    const std::size_t need = (std::size_t)n_streams_ * (std::size_t)tile_.size();
    if (need > 0) d_vec_.resize(need);
  }
  int n_streams_ = 0;
  tile_dim tile_;
  device_vector<T> d_vec_;
};

}  // namespace gpu

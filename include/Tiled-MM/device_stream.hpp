// HIP stream wrapper (reference: Tiled-MM src/Tiled-MM/device_stream.hpp:9-86). A stream
// made by the default constructor is owned (non-blocking, default priority); one wrapping a
// handle's stream is a non-owning view.
#pragma once
#include "device_event.hpp"
#include "util.hpp"

namespace gpu {

class device_stream {
 public:
  device_stream() : owning_(true) { check_runtime_status(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking)); }
  explicit device_stream(hipStream_t s) : stream_(s), owning_(false) {}
  ~device_stream() {
    if (owning_ && stream_) (void)hipStreamDestroy(stream_);
  }
  device_stream(device_stream&& other) noexcept : stream_(other.stream_), owning_(other.owning_) {
    other.owning_ = false;
    other.stream_ = nullptr;
  }
  device_stream& operator=(device_stream&& other) {
    if (this != &other) {
      if (owning_ && stream_) check_runtime_status(hipStreamDestroy(stream_));
      stream_ = other.stream_;
      owning_ = other.owning_;
      other.owning_ = false;
      other.stream_ = nullptr;
    }
    return *this;
  }
  device_stream(device_stream&) = delete;
  device_stream& operator=(device_stream&) = delete;

  hipStream_t stream() const { return stream_; }
  device_event enqueue_event() const {
    device_event e;
    check_runtime_status(hipEventRecord(e.get(), stream_));
    return e;
  }
  void wait_on_event(device_event& e) const { check_runtime_status(hipStreamWaitEvent(stream_, e.get(), 0)); }

 private:
  hipStream_t stream_ = nullptr;
  bool owning_ = false;
};

}  // namespace gpu

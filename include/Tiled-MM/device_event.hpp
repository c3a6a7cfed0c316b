// RAII HIP event (reference: Tiled-MM src/Tiled-MM/device_event.hpp:9-77). Events are created
// with timing enabled, so time_since() works (the reference disables timing).
#pragma once
#include "util.hpp"

namespace gpu {

class device_event {
 public:
  device_event() {
    check_runtime_status(hipEventCreate(&event_));
    valid_ = true;
  }
  ~device_event() {
    if (valid_) (void)hipEventDestroy(event_);
  }
  device_event(device_event&& other) noexcept : valid_(other.valid_), event_(other.event_) { other.valid_ = false; }
  device_event& operator=(device_event&& other) {
    if (this != &other) {
      if (valid_) check_runtime_status(hipEventDestroy(event_));
      event_ = other.event_;
      valid_ = other.valid_;
      other.valid_ = false;
    }
    return *this;
  }
  device_event(device_event&) = delete;
  device_event& operator=(device_event&) = delete;

  hipEvent_t& get() { return event_; }
  void wait() { check_runtime_status(hipEventSynchronize(event_)); }
  // seconds between other and this event
  double time_since(device_event& other) {
    float ms = 0.0f;
    check_runtime_status(hipEventElapsedTime(&ms, other.get(), event_));
    return double(ms / 1.e3);
  }

 private:
  bool valid_ = false;
  hipEvent_t event_ = nullptr;
};

}  // namespace gpu

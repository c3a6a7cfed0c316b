// Grow-only device array with the reference's interface (Tiled-MM
// src/Tiled-MM/device_vector.hpp:11-107): 1.2x capacity slack, contents not preserved on
// growth, move-only. A vector returned by mm_handle::get_full_device_buffer_c() is a view of
// memory the handle owns and is never freed through it.
#pragma once
#include <cmath>
#include <cstddef>

#include "util.hpp"

namespace gpu {

template <typename T>
class device_vector {
 public:
  device_vector() = default;
  explicit device_vector(std::size_t n) : size_(n), capacity_((std::size_t)std::ceil(1.2 * n)) {
    data_ = capacity_ ? malloc_device<T>(capacity_) : nullptr;
  }
  device_vector(device_vector& other) = delete;
  device_vector& operator=(device_vector&) = delete;
  device_vector(device_vector&& other) noexcept { take(other); }
  device_vector& operator=(device_vector&& other) {
    if (this != &other) {
      free_data();
      take(other);
    }
    return *this;
  }
  ~device_vector() { free_data(); }

  T* data() { return data_; }
  std::size_t size() { return size_; }
  std::size_t capacity() { return capacity_; }

  void resize(std::size_t size) {
    if (size == 0) return;
    if (size > capacity_) {
      if (!owning_) throw std::runtime_error("device_vector: cannot grow a view of handle-owned memory");
      free_data();
      capacity_ = (std::size_t)std::ceil(1.2 * size);
      data_ = malloc_device<T>(capacity_);
    }
    size_ = size;
  }

  // Non-owning view (used by mm_handle for the resident C).
  static device_vector view(T* p, std::size_t n) {
    device_vector v;
    v.data_ = p;
    v.size_ = n;
    v.capacity_ = n;
    v.owning_ = false;
    return v;
  }

 private:
  void free_data() {
    if (owning_ && data_ && capacity_ > 0) check_runtime_status(hipFree(data_));
    data_ = nullptr;
    size_ = capacity_ = 0;
  }
  void take(device_vector& o) {
    data_ = o.data_;
    size_ = o.size_;
    capacity_ = o.capacity_;
    owning_ = o.owning_;
    o.data_ = nullptr;
    o.size_ = o.capacity_ = 0;
  }
  T* data_ = nullptr;
  std::size_t size_ = 0;
  std::size_t capacity_ = 0;
  bool owning_ = true;
};

}  // namespace gpu

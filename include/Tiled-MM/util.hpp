// Error checks and memory helpers with the reference's names and behaviour
// (Tiled-MM src/Tiled-MM/util.hpp:13-118), written for HIP on gfx950.
#pragma once
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstddef>
#include <iostream>
#include <stdexcept>

namespace gpu {

// Prints the HIP error and throws std::runtime_error("GPU ERROR") (util.hpp:13-19).
inline void check_runtime_status(hipError_t status) {
  if (status != hipSuccess) {
    std::cerr << "error: GPU API call : " << hipGetErrorString(status) << std::endl;
    throw std::runtime_error("GPU ERROR");
  }
}

inline void check_last_device_kernel(const char* errstr) {
  hipError_t status = hipGetLastError();
  if (status != hipSuccess) {
    std::cerr << "error: GPU kernel launch : " << errstr << " : " << hipGetErrorString(status) << std::endl;
    throw std::runtime_error("GPU ERROR");
  }
}

// Bytes of device memory in use (util.hpp:41-50).
inline std::size_t gpu_allocated_memory() {
  check_runtime_status(hipDeviceSynchronize());
  std::size_t free = 0, total = 0;
  return hipMemGetInfo(&free, &total) == hipSuccess ? total - free : (std::size_t)-1;
}

// Device allocation of n elements (util.hpp:56-62).
template <typename T>
T* malloc_device(std::size_t n) {
  void* p = nullptr;
  check_runtime_status(hipMalloc(&p, n * sizeof(T)));
  return (T*)p;
}

// Pinned host allocation of N elements filled with value (util.hpp:64-72).
template <typename T>
T* malloc_pinned(std::size_t N, T value = T()) {
  T* ptr = nullptr;
  check_runtime_status(hipHostMalloc((void**)&ptr, N * sizeof(T), hipHostMallocDefault));
  std::fill(ptr, ptr + N, value);
  return ptr;
}

template <typename T>
void free_pinned(T* ptr) {
  check_runtime_status(hipHostFree(ptr));
}

// Synchronous copies of n elements (util.hpp:78-88).
template <typename T>
void copy_to_device(const T* from, T* to, std::size_t n) {
  check_runtime_status(hipMemcpy(to, from, n * sizeof(T), hipMemcpyHostToDevice));
}

template <typename T>
void copy_to_host(const T* from, T* to, std::size_t n) {
  check_runtime_status(hipMemcpy(to, from, n * sizeof(T), hipMemcpyDeviceToHost));
}

// Asynchronous copies on a stream (util.hpp:90-118).
template <typename T>
void copy_to_device_async(const T* from, T* to, std::size_t n, hipStream_t stream = nullptr) {
  check_runtime_status(hipMemcpyAsync(to, from, n * sizeof(T), hipMemcpyHostToDevice, stream));
}

template <typename T>
void copy_to_host_async(const T* from, T* to, std::size_t n, hipStream_t stream = nullptr) {
  check_runtime_status(hipMemcpyAsync(to, from, n * sizeof(T), hipMemcpyDeviceToHost, stream));
}

}  // namespace gpu

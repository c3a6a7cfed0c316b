// Out-of-core GEMM entry point with the reference's signature (Tiled-MM
// src/Tiled-MM/tiled_mm.hpp:62-80): C = alpha*op(A)*op(B) + beta*C for column-major host
// matrices, op in {'N','T','C'}. Synchronous. With copy_c_back = false the result stays in
// handle.get_full_device_buffer_c() (ld = m) and C is only read (beta != 0).
//
// Deviations from the reference, all documented in the library's design notes:
//   * pin_host_buffers does not page-lock the caller's matrices; pageable memory is streamed
//     through pinned staging buffers, pinned memory is used directly;
//   * ld < rows and tile sizes above the maxima raise std::runtime_error instead of being
//     accepted silently.
#pragma once
#include "mm_handle.hpp"

namespace gpu {

template <typename Scalar>
void gemm(mm_handle<Scalar>& handle, char trans_a, char trans_b, int m, int n, int k, Scalar alpha, Scalar* a,
          int ld_a, Scalar* b, int ld_b, Scalar beta, Scalar* c, int ld_c, bool pin_host_buffers = true,
          bool copy_c_back = true);

}  // namespace gpu

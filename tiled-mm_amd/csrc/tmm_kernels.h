// Internal interface between the out-of-core scheduler and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tmm {

// C = alpha*op(A)*op(B) + (accumulate ? C : 0) + beta*C_in, all column-major on the device.
// op in {'N','T','C'}; for complex types a, b, c, cin point at interleaved (re, im) pairs and
// the *_im fields carry the imaginary parts of alpha and beta. beta == 0 means C_in is not read.
template <typename R>
struct GemmParams {
  char opa, opb;
  int m, n, k;
  R alpha, alpha_im;
  const R* a;
  long lda;
  const R* b;
  long ldb;
  R beta, beta_im;
  const R* cin;
  long ldcin;
  int accumulate;
  R* c;
  long ldc;
  int variant;  // FP64 and complex FP64 tile kernels: 0 = register-staged LDS double buffer, 1 = direct-to-LDS ring
  // FP64 only: when gate_a is set, the launch does its work only if *gate_a | *gate_b is non-zero
  // (device flags written by earlier kernels of the stream; the Ozaki paths hand a tile whose
  // operands hold non-finite or huge entries to the native kernel this way)
  const int* gate_a = nullptr;
  const int* gate_b = nullptr;
};

hipError_t launch_gemm_d(const GemmParams<double>& p, hipStream_t s);

hipError_t launch_gemm_s(const GemmParams<float>& p, hipStream_t s);
hipError_t launch_gemm_z(const GemmParams<double>& p, hipStream_t s);
hipError_t launch_gemm_c(const GemmParams<float>& p, hipStream_t s);

// Up to kDgemmGroupMax independent products of one element type with the same op flags, alpha
// and k, each groupable (dgemm_groupable, zgemm_groupable, sgemm_groupable, cgemm_groupable), in
// one launch of the direct-to-LDS kernel (per-product m, n, pointers, leading dimensions, beta and
// accumulate): one partially filled last wave for the group instead of one per product.
constexpr int kDgemmGroupMax = 16;
template <typename R>
struct TileGroup {
  GemmParams<R> base;  // op flags, alpha, k; per-product fields below
  int count;
  int wg_start[kDgemmGroupMax + 1];  // first workgroup of each product (multiples of 8)
  int m[kDgemmGroupMax], n[kDgemmGroupMax], accumulate[kDgemmGroupMax];
  const R* a[kDgemmGroupMax];
  const R* b[kDgemmGroupMax];
  R* c[kDgemmGroupMax];
  const R* cin[kDgemmGroupMax];
  long lda[kDgemmGroupMax], ldb[kDgemmGroupMax], ldc[kDgemmGroupMax], ldcin[kDgemmGroupMax];
  R beta[kDgemmGroupMax], beta_im[kDgemmGroupMax];
};
using DgemmGroup = TileGroup<double>;
using SgemmGroup = TileGroup<float>;
bool dgemm_groupable(const GemmParams<double>& p);
hipError_t launch_gemm_d_group(const GemmParams<double>* tiles, int count, hipStream_t s);
bool zgemm_groupable(const GemmParams<double>& p);
hipError_t launch_gemm_z_group(const GemmParams<double>* tiles, int count, hipStream_t s);
bool sgemm_groupable(const GemmParams<float>& p);
hipError_t launch_gemm_s_group(const GemmParams<float>* tiles, int count, hipStream_t s);
bool cgemm_groupable(const GemmParams<float>& p);
hipError_t launch_gemm_c_group(const GemmParams<float>* tiles, int count, hipStream_t s);

// Persistent batched FP64 GEMM over device-resident problems (all with the same op flags):
// items = n_items triples (problem index, tile row, tile column) of 128x128 C tiles; counter is
// one device int, zeroed by a kernel at the start of the launch sequence (graph-capturable).
// body 0: register-staged LDS double buffer; 1: direct-to-LDS ring, which needs every operand in
// a device arena with at least 16 readable bytes after its last column.
hipError_t launch_gemm_d_batched(char opa, char opb, const GemmParams<double>* probs, const int* items, int n_items,
                                 int* counter, int n_workgroups, hipStream_t s, int body);

}  // namespace tmm

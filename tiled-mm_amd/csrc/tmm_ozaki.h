// Ozaki split-precision DGEMM on INT8 matrix cores: internal interface.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace tmm {

constexpr int kOzakiMaxProducts = 16;

// One launch: C = alpha * 2^(ea[i]+eb[j]+shift) * sum_t (A_slice[pa[t]] * B_slice[pb[t]]^T)_ij
//                 + (accumulate ? C : 0) + beta * C_in,
// slices int8 [slice][row][kpad] (k contiguous, rows padded to 256), k range [k_begin, k_end).
struct OzakiGemmParams {
  int m, n, kpad, k_begin, k_end;
  const int8_t* a;
  long a_slice_stride;
  const int8_t* b;
  long b_slice_stride;
  int nprod;
  int pa[kOzakiMaxProducts], pb[kOzakiMaxProducts];
  int shift;
  const int* ea;
  const int* eb;
  double alpha, beta;
  const double* cin;
  long ldcin;
  int accumulate;
  double* c;
  long ldc;
};

hipError_t ozaki_exponents(const double* p, long ld, bool kc, int rows, int k, int* e, hipStream_t s);
hipError_t ozaki_split(const double* p, long ld, bool kc, int rows, int k, const int* e, int nslices, int rows_pad,
                       int kpad, int8_t* out, hipStream_t s);
hipError_t ozaki_gemm(const OzakiGemmParams& p, hipStream_t s);

// Whole emulated DGEMM on device data: C = alpha*op(A)*op(B) + (accumulate ? C : 0) + beta*C_in.
// `work` must hold ozaki_workspace_bytes(m, n, k, nslices) bytes. Returns the number of INT8
// slice products computed through *n_products (s(s+1)/2 for s slices).
size_t ozaki_workspace_bytes(int m, int n, int k, int nslices);
hipError_t ozaki_dgemm(char opa, char opb, int m, int n, int k, double alpha, const double* a, long lda,
                       const double* b, long ldb, double beta, const double* cin, long ldcin, bool accumulate,
                       double* c, long ldc, int nslices, void* work, hipStream_t s, int* n_products);

}  // namespace tmm

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
  // residue mode (res != nullptr): instead of the FP64 epilogue, store the exact int32 sum
  // reduced modulo `modulus` as uint8 in [0, modulus) at res[i + j*ldres] (nprod must be 1;
  // ldres a multiple of 16, rows m..ldres-1 receive padding);
  // magic = ceil(2^32 / modulus). The operands are then in k-step tiled layout (one contiguous
  // 16 KB block per 256 rows and 64 k, blocks of a row block in k order, 64-byte rows inside),
  // not row-major with stride kpad.
  uint8_t* res;
  long ldres;
  int modulus;
  unsigned magic;
  // when fa is set, the launch does nothing if *fa | *fb is non-zero (the operands' fallback
  // flags, see ozaki_exponents)
  const int* fa;
  const int* fb;
};

// Entries with |x| >= 2^kFallbackExp, and Inf and NaN, are not emulated: a product whose operands
// hold one runs on the native FP64 kernel instead (vendor DGEMM semantics: Inf and NaN propagate,
// overflow gives Inf), decided on the device from the operands' fallback flags. Below the limit
// every product and sum of the emulation stays far from the FP64 range (2^(2*480+17) < 2^1023).
constexpr int kFallbackExp = 480;

// e[r] (r < rows): exponent with max_k |x(r, k)| < 2^e[r]; *flag (when not null) becomes 1 if an
// entry is non-finite or at least 2^kFallbackExp in magnitude, else 0.
hipError_t ozaki_exponents(const double* p, long ld, bool kc, int rows, int k, int* e, int* flag, hipStream_t s);
hipError_t ozaki_split(const double* p, long ld, bool kc, int rows, int k, const int* e, int nslices, int rows_pad,
                       int kpad, int8_t* out, hipStream_t s);
hipError_t ozaki_gemm(const OzakiGemmParams& p, hipStream_t s);

// Whole emulated DGEMM on device data: C = alpha*op(A)*op(B) + (accumulate ? C : 0) + beta*C_in.
// `work` must hold ozaki_workspace_bytes(m, n, k, nslices) bytes. Returns the number of INT8
// slice products computed through *n_products (s(s+1)/2 for s slices). Operands with non-finite
// or huge entries (kFallbackExp) are multiplied by the native FP64 kernel instead.
size_t ozaki_workspace_bytes(int m, int n, int k, int nslices);

// Ozaki scheme II (CRT): the entries of op(A) / op(B) are scaled per row / column by powers of
// two and rounded to integers of at most `bits` bits; their exact integer product is recovered
// from its residues modulo nmod pairwise coprime moduli <= 256 (one INT8 product per modulus)
// by the CRT sum, evaluated exactly in 26-bit FP64 limbs.
constexpr int kCrtMaxModuli = 20, kCrtMinModuli = 6;
constexpr int kCrtMaxK = 65536;  // k per reconstruction pass: |sum| <= k * 128^2 < 2^30
size_t ozaki2_workspace_bytes(int m, int n, int k, int nmod);
// bits per operand for nmod moduli and inner dimension k (the exact product then stays inside
// the balanced range of the moduli's product)
int ozaki2_bits(int nmod, int k);

// The same in two steps, so that an operand used by several products is prepared once.
// A prepared operand of `rows` rows (rows of op(A), or columns of op(B)) and inner dimension
// k <= kCrtMaxK holds the per-row exponents and the nmod residue planes (k-step tiled layout):
// ozaki2_operand_bytes(rows, k, nmod) bytes. kc: entry (r, kk) at p[kk + r*ld], else p[r + kk*ld]
// (A: kc = opa != 'N'; B: kc = opb == 'N').
size_t ozaki2_operand_bytes(int rows, int k, int nmod);
hipError_t ozaki2_prepare(const double* p, long ld, bool kc, int rows, int k, int nmod, void* out, hipStream_t s);
// C = alpha * A'B' (reconstructed) + (accumulate ? C : 0) + beta * C_in from two operands prepared
// with the same k and nmod; `work` holds ozaki2_product_bytes(m, n, nmod) bytes (the residue
// planes of the product).
size_t ozaki2_product_bytes(int m, int n, int nmod);
hipError_t ozaki2_product(int m, int n, int k, double alpha, const void* pa, const void* pb, double beta,
                          const double* cin, long ldcin, bool accumulate, double* c, long ldc, int nmod, void* work,
                          hipStream_t s);
// Its two halves, for a caller that must wait for C_in in between: the nmod INT8 products into
// `work`, then the reconstruction from `work` (same arguments).
hipError_t ozaki2_products(int m, int n, int k, const void* pa, const void* pb, int nmod, void* work, hipStream_t s);
hipError_t ozaki2_reconstruct(int m, int n, int k, double alpha, const void* pa, const void* pb, double beta,
                              const double* cin, long ldcin, bool accumulate, double* c, long ldc, int nmod,
                              const void* work, hipStream_t s);
// The native FP64 product of the original operands (a: op(A) stored as opa says, b likewise), run
// on the device only if the prepared operands pa / pb carry a fallback flag (an Inf, a NaN or an
// entry of at least 2^kFallbackExp); the products and the reconstruction of such a pair do
// nothing. Enqueue it after ozaki2_reconstruct with the same C arguments.
hipError_t ozaki2_fallback(char opa, char opb, int m, int n, int k, double alpha, const double* a, long lda,
                           const double* b, long ldb, double beta, const double* cin, long ldcin, bool accumulate,
                           double* c, long ldc, const void* pa, const void* pb, hipStream_t s);

// Fixed per-device scratch for callers that cannot allocate (tmm_device_gemm_noalloc): a static
// device array of kOzakiScratchBytes that every GPU the library runs on holds until the process
// ends. ozaki_scratch_resolve() looks up its address on the current device (hipGetSymbolAddress,
// once per device; call it at set-up with that device current); ozaki_scratch(device) returns the
// address resolved for that device, or null.
constexpr size_t kOzakiScratchBytes = (size_t)256 << 20;
hipError_t ozaki_scratch_resolve();
void* ozaki_scratch(int device);

hipError_t ozaki2_dgemm(char opa, char opb, int m, int n, int k, double alpha, const double* a, long lda,
                        const double* b, long ldb, double beta, const double* cin, long ldcin, bool accumulate,
                        double* c, long ldc, int nmod, void* work, hipStream_t s, int* n_products);
hipError_t ozaki_dgemm(char opa, char opb, int m, int n, int k, double alpha, const double* a, long lda,
                       const double* b, long ldb, double beta, const double* cin, long ldcin, bool accumulate,
                       double* c, long ldc, int nslices, void* work, hipStream_t s, int* n_products);

}  // namespace tmm

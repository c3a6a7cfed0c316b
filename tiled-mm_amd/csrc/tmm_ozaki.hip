// Ozaki-style split-precision DGEMM on INT8 matrix cores (gfx950).
//
// Scheme (Ozaki, Ogita, Oishi, Rump 2012, "Error-free transformations of matrix multiplication
// by using fast routines of matrix multiplication and its applications", in the integer-MMU form
// of Ootomo, Ozaki, Yokota 2024):
//   * every row i of op(A) is scaled by 2^-ea[i] (ea[i] = exponent with max_k |a_ik| < 2^ea[i]),
//     every column j of op(B) by 2^-eb[j]; the scaled values lie in (-1, 1);
//   * each scaled value is cut into s slices of 7 bits: x = sum_p S_p 2^(-7p) + r, S_p in
//     [-127, 127] (int8), |r| < 2^(-7s); the cut is exact in FP64 (scaling by powers of two and
//     removing integer parts);
//   * C = alpha * 2^(ea[i]+eb[j]) * sum_{p+q <= s+1} 2^(-7(p+q)) (S^A_p S^B_q)_ij + beta * C,
//     where every slice product runs on v_mfma_i32_16x16x64_i8 with exact int32 accumulation:
//     one product is bounded by k*127^2, so a launch sums only as many products of one diagonal
//     d = p+q as keep d_count*k*127^2 < 2^31;
//   * each launch adds its exact int32 sum, converted to FP64 and scaled by the power of two,
//     into C (the alpha/beta epilogue is fused into the first launch of a C tile).
// FP64 is used only for the exponent extraction, the slicing, and the recombination/epilogue.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "tmm_ozaki.h"

namespace tmm {
namespace {

typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
constexpr int kExpNone = -100000;  // row without a nonzero entry

__device__ __forceinline__ int frexp_exp_of(double v) {
  // exponent e with |v| < 2^e (frexp convention), kExpNone for 0; subnormals get -1022
  const uint64_t bits = __double_as_longlong(v) & 0x7FFFFFFFFFFFFFFFull;
  if (bits == 0) return kExpNone;
  const int be = (int)(bits >> 52);
  return be == 0 ? -1022 : be - 1022;
}

// ---------------------------------------------------------------------------------------------
// Exponents: e[r] = max over k of frexp-exponent(|x(r, k)|). x(r, k) at p[k + r*ld] (KC) or
// p[r + k*ld] (RC). Grid: (row blocks, k splits); partial maxima meet in atomicMax.
// ---------------------------------------------------------------------------------------------
__global__ void exp_init_kernel(int* e, int rows) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < rows) e[r] = kExpNone;
}

__global__ __launch_bounds__(256) void exp_rows_rc_kernel(const double* __restrict__ p, long ld, int rows, int k,
                                                          int kper, int* __restrict__ e) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= rows) return;
  const int k0 = blockIdx.y * kper, k1 = min(k, k0 + kper);
  int m = kExpNone;
  for (int kk = k0; kk < k1; ++kk) m = max(m, frexp_exp_of(p[(long)r + (long)kk * ld]));
  if (m != kExpNone) atomicMax(&e[r], m);
}

__global__ __launch_bounds__(256) void exp_rows_kc_kernel(const double* __restrict__ p, long ld, int rows, int k,
                                                          int kper, int* __restrict__ e) {
  // one wave per (row, k range)
  const int wave = (blockIdx.x * 256 + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (wave >= rows) return;
  const int k0 = blockIdx.y * kper, k1 = min(k, k0 + kper);
  int m = kExpNone;
  const double* row = p + (long)wave * ld;
  for (int kk = k0 + lane; kk < k1; kk += 64) m = max(m, frexp_exp_of(row[kk]));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
  if (lane == 0 && m != kExpNone) atomicMax(&e[wave], m);
}

// ---------------------------------------------------------------------------------------------
// Slicing: S[p][r][k] (int8, row stride kpad, slice stride rows_pad*kpad) for r < rows_pad,
// k < kpad; zero outside the matrix. Each thread cuts 4 consecutive k of one row.
// ---------------------------------------------------------------------------------------------
template <bool KC>
__global__ __launch_bounds__(256) void split_kernel(const double* __restrict__ p, long ld, int rows, int k,
                                                    const int* __restrict__ e, int nslices, int rows_pad, int kpad,
                                                    int8_t* __restrict__ s) {
  const long tid = (long)blockIdx.x * 256 + threadIdx.x;
  const int k4 = kpad / 4;
  int r, kq;
  if (KC) {  // consecutive threads walk k of one row
    r = (int)(tid / k4);
    kq = (int)(tid % k4);
  } else {  // consecutive threads walk rows at one k
    r = (int)(tid % rows_pad);
    kq = (int)(tid / rows_pad);
  }
  if (r >= rows_pad || kq >= k4) return;
  double x[4];
  int sg[4];
  const int er = r < rows ? e[r] : kExpNone;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int kk = kq * 4 + t;
    double v = 0.0;
    if (r < rows && kk < k) v = KC ? p[(long)kk + (long)r * ld] : p[(long)r + (long)kk * ld];
    sg[t] = v < 0.0 ? -1 : 1;
    // |v| * 2^-er in [0, 1): exact (power-of-two scaling)
    x[t] = er == kExpNone ? 0.0 : ldexp(fabs(v), -er);
  }
  const long plane = (long)rows_pad * kpad;
  int8_t* out = s + (long)r * kpad + (long)kq * 4;
  for (int q = 0; q < nslices; ++q) {
    uint32_t word = 0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const double y = x[t] * 128.0;  // exact
      const double d = floor(y);       // 0..127
      x[t] = y - d;                    // exact
      const int v = sg[t] * (int)d;
      word |= ((uint32_t)(uint8_t)(int8_t)v) << (8 * t);
    }
    *(uint32_t*)(out + q * plane) = word;
  }
}

// ---------------------------------------------------------------------------------------------
// INT8 slice-product GEMM with FP64 scaled accumulation.
// WG tile 256x256, 4 waves (2x2) of 128x128 = 8x8 blocks of 16x16, k-step 64 per stage.
// ---------------------------------------------------------------------------------------------
constexpr int OZ_BM = 256, OZ_BN = 256, OZ_BK = 64;

struct OzTile {
  // one 256-row x 64-byte stage: 1024 chunks of 16 bytes, 4 per thread; LDS layout [g][row][16]
  i32x4 regs[4];
  __device__ __forceinline__ void load(const int8_t* __restrict__ base, long row_stride, int row0, int k0) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int c = threadIdx.x + 256 * u;
      const int row = c >> 2, g = c & 3;
      regs[u] = *(const i32x4*)(base + (long)(row0 + row) * row_stride + k0 + 16 * g);
    }
  }
  __device__ __forceinline__ void store(int8_t* lds) const {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int c = threadIdx.x + 256 * u;
      const int row = c >> 2, g = c & 3;
      *(i32x4*)(lds + ((long)g * 256 + row) * 16) = regs[u];
    }
  }
};

__global__ __launch_bounds__(256, 1) void ozaki_gemm_kernel(OzakiGemmParams prm) {
  __shared__ __attribute__((aligned(16))) int8_t lds[2][2][4 * 256 * 16];  // [buf][A/B][g][row][16]

  const int nbm = (prm.m + OZ_BM - 1) / OZ_BM, nbn = (prm.n + OZ_BN - 1) / OZ_BN;
  // grouped tile order: 8 tile-rows per group, column by column
  const int w = blockIdx.x;
  const int GROUP_M = 8, in_group = GROUP_M * nbn, group = w / in_group, first_m = group * GROUP_M;
  const int gsz = min(nbm - first_m, GROUP_M), rr = w - group * in_group;
  const int bm = first_m + rr % gsz, bn = rr / gsz;
  const int m0 = bm * OZ_BM, n0 = bn * OZ_BN;

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int li = lane & 15, lg = lane >> 4;

  i32x4 acc[8][8];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[i][j] = i32x4{0, 0, 0, 0};

  const int kt0 = prm.k_begin / OZ_BK;
  const int nkt = (prm.k_end - prm.k_begin) / OZ_BK;
  const int total = nkt * prm.nprod;
  OzTile ta, tb;
  auto slice_a = [&](int it) { return prm.a + (long)prm.pa[it / nkt] * prm.a_slice_stride; };
  auto slice_b = [&](int it) { return prm.b + (long)prm.pb[it / nkt] * prm.b_slice_stride; };
  if (total > 0) {
    ta.load(slice_a(0), prm.kpad, m0, kt0 * OZ_BK);
    tb.load(slice_b(0), prm.kpad, n0, kt0 * OZ_BK);
    ta.store(&lds[0][0][0]);
    tb.store(&lds[0][1][0]);
  }
  __syncthreads();
  for (int it = 0; it < total; ++it) {
    const int cur = it & 1;
    const bool more = it + 1 < total;
    if (more) {
      const int k0 = (kt0 + (it + 1) % nkt) * OZ_BK;
      ta.load(slice_a(it + 1), prm.kpad, m0, k0);
      tb.load(slice_b(it + 1), prm.kpad, n0, k0);
    }
    const int8_t* As = &lds[cur][0][0];
    const int8_t* Bs = &lds[cur][1][0];
    i32x4 af[8], bf[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) af[i] = *(const i32x4*)(As + ((long)lg * 256 + wm * 128 + i * 16 + li) * 16);
#pragma unroll
    for (int j = 0; j < 8; ++j) bf[j] = *(const i32x4*)(Bs + ((long)lg * 256 + wn * 128 + j * 16 + li) * 16);
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(bf[j], af[i], acc[i][j], 0, 0, 0);
    if (more) {
      ta.store(&lds[cur ^ 1][0][0]);
      tb.store(&lds[cur ^ 1][1][0]);
    }
    __syncthreads();
  }

  // epilogue: C = alpha * 2^(ea+eb+shift) * acc + (accumulate ? C : 0) + beta * C_in
  // lane holds C(i = m0 + wm*128 + bi*16 + li, j = n0 + wn*128 + bj*16 + 4*lg + r)
#pragma unroll
  for (int bi = 0; bi < 8; ++bi) {
    const int gi = m0 + wm * 128 + bi * 16 + li;
    if (gi >= prm.m) continue;
    const int ei = prm.ea[gi];
#pragma unroll
    for (int bj = 0; bj < 8; ++bj) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gj = n0 + wn * 128 + bj * 16 + 4 * lg + r;
        if (gj >= prm.n) continue;
        const int ej = prm.eb[gj];
        double v = 0.0;
        if (ei != kExpNone && ej != kExpNone) v = ldexp(prm.alpha * (double)acc[bi][bj][r], ei + ej + prm.shift);
        double* dst = prm.c + (long)gi + (long)gj * prm.ldc;
        if (prm.accumulate) v += *dst;
        if (prm.beta != 0.0) v += prm.beta * prm.cin[(long)gi + (long)gj * prm.ldcin];
        *dst = v;
      }
    }
  }
}

}  // namespace

hipError_t ozaki_exponents(const double* p, long ld, bool kc, int rows, int k, int* e, hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(exp_init_kernel, dim3((rows + 255) / 256), dim3(256), 0, s, e, rows);
  if (k <= 0) return hipGetLastError();
  if (kc) {
    const int ksplit = k >= 8192 ? 8 : 1, kper = (k + ksplit - 1) / ksplit;
    hipLaunchKernelGGL(exp_rows_kc_kernel, dim3((rows + 3) / 4, ksplit), dim3(256), 0, s, p, ld, rows, k, kper, e);
  } else {
    const int ksplit = std::min(64, std::max(1, k / 256)), kper = (k + ksplit - 1) / ksplit;
    hipLaunchKernelGGL(exp_rows_rc_kernel, dim3((rows + 255) / 256, ksplit), dim3(256), 0, s, p, ld, rows, k, kper, e);
  }
  return hipGetLastError();
}

hipError_t ozaki_split(const double* p, long ld, bool kc, int rows, int k, const int* e, int nslices, int rows_pad,
                       int kpad, int8_t* out, hipStream_t s) {
  const long threads = (long)rows_pad * (kpad / 4);
  if (threads == 0) return hipSuccess;
  const dim3 grid((unsigned)((threads + 255) / 256));
  if (kc) hipLaunchKernelGGL(split_kernel<true>, grid, dim3(256), 0, s, p, ld, rows, k, e, nslices, rows_pad, kpad, out);
  else hipLaunchKernelGGL(split_kernel<false>, grid, dim3(256), 0, s, p, ld, rows, k, e, nslices, rows_pad, kpad, out);
  return hipGetLastError();
}

hipError_t ozaki_gemm(const OzakiGemmParams& p, hipStream_t s) {
  if (p.m <= 0 || p.n <= 0) return hipSuccess;
  const int nbm = (p.m + OZ_BM - 1) / OZ_BM, nbn = (p.n + OZ_BN - 1) / OZ_BN;
  hipLaunchKernelGGL(ozaki_gemm_kernel, dim3(nbm * nbn), dim3(256), 0, s, p);
  return hipGetLastError();
}

namespace {
inline long pad_to(long x, long q) { return (x + q - 1) / q * q; }
inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
}  // namespace

size_t ozaki_workspace_bytes(int m, int n, int k, int nslices) {
  const long kpad = pad_to(std::max(k, 1), OZ_BK), mpad = pad_to(m, OZ_BM), npad = pad_to(n, OZ_BN);
  return align256((size_t)m * 4) + align256((size_t)n * 4) + align256((size_t)nslices * mpad * kpad) +
         align256((size_t)nslices * npad * kpad);
}

hipError_t ozaki_dgemm(char opa, char opb, int m, int n, int k, double alpha, const double* a, long lda,
                       const double* b, long ldb, double beta, const double* cin, long ldcin, bool accumulate,
                       double* c, long ldc, int nslices, void* work, hipStream_t s, int* n_products) {
  if (n_products) *n_products = 0;
  if (m <= 0 || n <= 0) return hipSuccess;
  if (nslices < 1 || nslices > 12) return hipErrorInvalidValue;
  const long kpad = pad_to(std::max(k, 1), OZ_BK), mpad = pad_to(m, OZ_BM), npad = pad_to(n, OZ_BN);
  char* w = (char*)work;
  int* ea = (int*)w;
  w += align256((size_t)m * 4);
  int* eb = (int*)w;
  w += align256((size_t)n * 4);
  int8_t* sa = (int8_t*)w;
  w += align256((size_t)nslices * mpad * kpad);
  int8_t* sb = (int8_t*)w;
  // op(A) rows: k-contiguous when A is transposed; op(B) columns: k-contiguous when B is not.
  const bool a_kc = opa != 'N', b_kc = opb == 'N';
  hipError_t e;
  if ((e = ozaki_exponents(a, lda, a_kc, m, k, ea, s)) != hipSuccess) return e;
  if ((e = ozaki_exponents(b, ldb, b_kc, n, k, eb, s)) != hipSuccess) return e;
  if ((e = ozaki_split(a, lda, a_kc, m, k, ea, nslices, (int)mpad, (int)kpad, sa, s)) != hipSuccess) return e;
  if ((e = ozaki_split(b, ldb, b_kc, n, k, eb, nslices, (int)npad, (int)kpad, sb, s)) != hipSuccess) return e;

  // products of one diagonal d = p+q share the power-of-two scale; a launch sums at most
  // gmax of them so the int32 accumulation stays exact: gmax * k * 127^2 < 2^31.
  const long per_product = (long)std::max(k, 1) * 127 * 127;
  const int gmax = (int)std::max<long>(1, std::min<long>(kOzakiMaxProducts, 2147483647L / per_product));
  // with k > 2^31/127^2 a single product could overflow: process k in ranges
  const long krange = per_product > 2147483647L ? pad_to(2147483647L / (127 * 127) - OZ_BK, OZ_BK) : kpad;
  OzakiGemmParams P{};
  P.m = m;
  P.n = n;
  P.kpad = (int)kpad;
  P.a = sa;
  P.a_slice_stride = mpad * kpad;
  P.b = sb;
  P.b_slice_stride = npad * kpad;
  P.ea = ea;
  P.eb = eb;
  P.alpha = alpha;
  P.c = c;
  P.ldc = ldc;
  bool first = true;
  int count = 0;
  for (long k0 = 0; k0 < kpad; k0 += krange) {
    P.k_begin = (int)k0;
    P.k_end = (int)std::min(kpad, k0 + krange);
    for (int d = 2; d <= nslices + 1; ++d) {
      int np = 0;
      for (int p = 1; p <= nslices; ++p) {
        const int q = d - p;
        if (q < 1 || q > nslices) continue;
        P.pa[np] = p - 1;
        P.pb[np] = q - 1;
        ++np;
        if (np == gmax || p == nslices || d - p - 1 < 1) {
          P.nprod = np;
          P.shift = -7 * d;
          P.accumulate = first ? (accumulate ? 1 : 0) : 1;
          P.beta = first ? beta : 0.0;
          P.cin = cin;
          P.ldcin = ldcin;
          if ((e = ozaki_gemm(P, s)) != hipSuccess) return e;
          first = false;
          count += np;
          np = 0;
        }
      }
    }
  }
  if (n_products) *n_products = count;
  return hipSuccess;
}

}  // namespace tmm

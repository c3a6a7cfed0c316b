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
#include <cmath>
#include <mutex>

#include "tmm_crt_loop.inc"
#include "tmm_kernels.h"
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

// p mod m for p < 2^32 and 2 <= m <= 256, magic = ceil(2^32 / m): the estimate q = p*magic/2^32
// is floor(p/m) or one more, so a single correction gives the remainder.
__device__ __forceinline__ unsigned mod_small(unsigned p, unsigned m, unsigned magic) {
  const unsigned q = __umulhi(p, magic);
  const int r = (int)(p - q * m);
  return (unsigned)(r < 0 ? r + (int)m : r);
}

// ---------------------------------------------------------------------------------------------
// Exponents: e[r] = max over k of frexp-exponent(|x(r, k)|). x(r, k) at p[k + r*ld] (KC) or
// p[r + k*ld] (RC). Grid: (row blocks, k splits); partial maxima meet in atomicMax. Inf and NaN
// have frexp-exponent 1025 here, so a partial maximum above kFallbackExp marks the operand for
// the native fallback (*flag = 1; every writer stores the same value).
// ---------------------------------------------------------------------------------------------
__global__ void exp_init_kernel(int* e, int rows, int* flag) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < rows) e[r] = kExpNone;
  if (flag && r == 0) *flag = 0;
}

__global__ __launch_bounds__(256) void exp_rows_rc_kernel(const double* __restrict__ p, long ld, int rows, int k,
                                                          int kper, int* __restrict__ e, int* __restrict__ flag) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= rows) return;
  const int k0 = blockIdx.y * kper, k1 = min(k, k0 + kper);
  int m = kExpNone;
  for (int kk = k0; kk < k1; ++kk) m = max(m, frexp_exp_of(p[(long)r + (long)kk * ld]));
  if (m != kExpNone) atomicMax(&e[r], m);
  if (m > kFallbackExp && flag) *flag = 1;
}

__global__ __launch_bounds__(256) void exp_rows_kc_kernel(const double* __restrict__ p, long ld, int rows, int k,
                                                          int kper, int* __restrict__ e, int* __restrict__ flag) {
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
  if (lane == 0 && m > kFallbackExp && flag) *flag = 1;
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
// WG tile 256x256, 4 waves (2x2) of 128x128 = 8x8 blocks of 16x16, k-step 64 per stage, one
// workgroup per CU. Operands go global -> LDS with global_load_lds_dwordx4 into an OZ_S-stage
// ring (OZ_S - 1 stages in flight while the MFMAs of one run): with one wave per SIMD nothing
// else hides the load latency, so a register-staged double buffer left the MFMAs idle most of
// each stage.
//
// Stage layout per operand: [row][slot][16 bytes], 64 bytes per row. The slices are k-contiguous
// with rows padded to 256 and k to 64, so every load is in bounds and unguarded. A wave
// instruction fills 16 consecutive rows (1 KB, lane l at byte 16*l): lanes 4j..4j+3 read the
// 64 contiguous bytes of row j, lane l landing in slot l % 4, which receives the k-chunk
// (l % 4) ^ oz_swz(row). The MFMA fragment read of one k-chunk by 16 consecutive rows then hits
// 8 distinct 16-byte bank groups per 8 lanes (rows r, r+1 share a swizzle and lie 64 bytes
// apart): no bank conflicts.
// ---------------------------------------------------------------------------------------------
constexpr int OZ_BM = 256, OZ_BN = 256, OZ_BK = 64;
constexpr int OZ_S = 4;                         // LDS ring stages
constexpr int OZ_TILE = OZ_BM * OZ_BK;          // bytes of one operand stage (16 KB)
constexpr int OZ_STAGE = 2 * OZ_TILE;           // A and B
constexpr int OZ_LOADS = OZ_TILE / (16 * 256);  // 16-byte loads per thread per operand and stage
static_assert(OZ_BM == OZ_BN && OZ_LOADS == 4, "load mapping assumes square 256x256 tiles, k-step 64");

__device__ __forceinline__ int oz_swz(int r) { return (r >> 1) & 3; }

// s_waitcnt immediate: vmcnt(vm) with lgkmcnt(0), expcnt not waited (gfx9 encoding: vmcnt bits
// [3:0] and [15:14], expcnt [6:4], lgkmcnt [11:8]).
template <int VM>
__device__ __forceinline__ void wait_vm_lgkm0() {
  static_assert(VM >= 0 && VM < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((VM & 15) | ((VM >> 4) << 14) | (7 << 4));
}

// Issue this wave's loads of one operand stage: OZ_LOADS instructions of 16 rows x 64 bytes.
__device__ __forceinline__ void oz_issue(const int8_t* __restrict__ base, long row_stride, int row0, int k0,
                                         int8_t* stage, int wave, int lane) {
#pragma unroll
  for (int t = 0; t < OZ_LOADS; ++t) {
    const int inst = wave * OZ_LOADS + t;  // 0..15
    const int local = inst * 16 + (lane >> 2);
    const int q = (lane & 3) ^ oz_swz(local);
    const int8_t* src = base + (long)(row0 + local) * row_stride + k0 + 16 * q;
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                     (__attribute__((address_space(3))) void*)(stage + inst * 1024), 16, 0, 0);
  }
}

__device__ __forceinline__ i32x4 oz_frag(const int8_t* stage, int row, int g) {
  return *(const i32x4*)(stage + row * 64 + ((g ^ oz_swz(row)) << 4));
}

// Output tile of this workgroup. XCD-aware order: workgroup w runs on XCD w % 8, so each XCD
// gets a contiguous range of tile ids, walked in groups of 4 tile-rows column by column (the ~32
// tiles an XCD runs at once then share A and B strips in its L2).
__device__ __forceinline__ void oz_tile(int m, int n, int& m0, int& n0) {
  const int nbm = (m + OZ_BM - 1) / OZ_BM, nbn = (n + OZ_BN - 1) / OZ_BN;
  int t = blockIdx.x;
  const int ntiles = nbm * nbn;
  if (ntiles >= 32) {
    const int xcd = t % 8, idx = t / 8, per = ntiles / 8, rem = ntiles % 8;
    t = xcd * per + min(xcd, rem) + idx;
  }
  const int GROUP_M = 4, in_group = GROUP_M * nbn, group = t / in_group, first_m = group * GROUP_M;
  const int gsz = min(nbm - first_m, GROUP_M), rr = t - group * in_group;
  m0 = (first_m + rr % gsz) * OZ_BM;
  n0 = (rr / gsz) * OZ_BN;
}

// Residue epilogue: res[i + j*ldres] = acc mod modulus for the workgroup's 256x256 tile. The lane
// holds C(i = m0 + wm*128 + bi*16 + li, j = n0 + wn*128 + bj*16 + 4*lg + r); |acc| <= k * 128^2 <
// 2^30 and modulus * 2^23 >= 2^30, so the biased sum is a non-negative 32-bit value congruent to
// acc. The bytes go through the (then idle) LDS ring as a column-major tile with a 272-byte column
// stride, then out as 16 consecutive rows of a column per lane (16 lanes write one column's 256
// bytes): 16 vector stores per lane instead of 256 byte stores. ldres is a multiple of 16 and
// every row m0..m0+255 below ldres may be written (rows >= m hold padding).
constexpr int OZ_RES_STRIDE = 272;  // 16-byte aligned, and columns 4 apart start 16 banks apart

__device__ __forceinline__ void oz_store_residues(const OzakiGemmParams& prm, const i32x4 (&acc)[8][8], int m0,
                                                  int n0, int wm, int wn, int li, int lg, int8_t* lds) {
  const unsigned mod = (unsigned)prm.modulus, bias = mod << 23;
  uint8_t* T = (uint8_t*)lds;
  __syncthreads();  // every wave is done with the ring
#pragma unroll
  for (int bi = 0; bi < 8; ++bi)
#pragma unroll
    for (int bj = 0; bj < 8; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        T[(wn * 128 + bj * 16 + 4 * lg + r) * OZ_RES_STRIDE + wm * 128 + bi * 16 + li] =
            (uint8_t)mod_small((unsigned)acc[bi][bj][r] + bias, mod, prm.magic);
  __syncthreads();
  const long rows_left = prm.ldres - m0;  // rows of the plane from m0 on (a multiple of 16)
#pragma unroll 4
  for (int q = 0; q < 16; ++q) {
    const int c = (int)threadIdx.x + 256 * q, col = c >> 4, ch = c & 15;
    if (n0 + col >= prm.n || 16 * ch >= rows_left) continue;
    const uint4 v = *(const uint4*)(T + col * OZ_RES_STRIDE + 16 * ch);
    *(uint4*)(prm.res + (long)(m0 + 16 * ch) + (long)(n0 + col) * prm.ldres) = v;
  }
}

__global__ __launch_bounds__(256, 1) void ozaki_gemm_kernel(OzakiGemmParams prm) {
  __shared__ __attribute__((aligned(16))) int8_t lds[OZ_S * OZ_STAGE];  // 128 KB
  if (prm.fa && (*prm.fa | *prm.fb)) return;  // the native fallback computes this product

  int m0, n0;
  oz_tile(prm.m, prm.n, m0, n0);

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int li = lane & 15, lg = lane >> 4;

  i32x4 acc[8][8];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[i][j] = i32x4{0, 0, 0, 0};

  // slice indices packed into 4-bit fields of two scalars, so the loop reads no kernel argument
  // memory (a scalar load in flight would make the compiler wait for every LDS read at once)
  unsigned long long pa_bits = 0, pb_bits = 0;
  for (int q = 0; q < prm.nprod; ++q) {
    pa_bits |= (unsigned long long)prm.pa[q] << (4 * q);
    pb_bits |= (unsigned long long)prm.pb[q] << (4 * q);
  }
  const int kt0 = prm.k_begin / OZ_BK;
  const int nkt = (prm.k_end - prm.k_begin) / OZ_BK;
  const int total = nkt * prm.nprod;
  // stage `it` = slice pair it / nkt at k-tile kt0 + it % nkt, into ring slot `slot`
  auto issue = [&](int it, int slot) {
    int8_t* st = lds + slot * OZ_STAGE;
    const int pr = it / nkt, k0 = (kt0 + it % nkt) * OZ_BK;
    oz_issue(prm.a + (long)((pa_bits >> (4 * pr)) & 15) * prm.a_slice_stride, prm.kpad, m0, k0, st, wave, lane);
    oz_issue(prm.b + (long)((pb_bits >> (4 * pr)) & 15) * prm.b_slice_stride, prm.kpad, n0, k0, st + OZ_TILE, wave,
             lane);
  };

  // wait until this wave's loads of stage s have landed, given that `later` further stages are
  // also issued (2 * OZ_LOADS loads each), then the barrier
  auto wait_stage = [&](int later) {
    if (later >= 3) wait_vm_lgkm0<6 * OZ_LOADS>();
    else if (later == 2) wait_vm_lgkm0<4 * OZ_LOADS>();
    else if (later == 1) wait_vm_lgkm0<2 * OZ_LOADS>();
    else wait_vm_lgkm0<0>();
    __builtin_amdgcn_s_barrier();
  };

  auto frag_a = [&](int it, int i) {
    return oz_frag(lds + (it % OZ_S) * OZ_STAGE, wm * 128 + i * 16 + li, lg);
  };
  auto frag_b = [&](int it, int j) {
    return oz_frag(lds + (it % OZ_S) * OZ_STAGE + OZ_TILE, wn * 128 + j * 16 + li, lg);
  };

  // One barrier per stage. Every iteration issues exactly one stage of loads: past the end it
  // reloads the last stage into the slot just freed (never read again), so (OZ_S - 2) stages are
  // always in flight behind the awaited one and the waits need no tail cases.
  // Measured alternatives, on 8192^3 with 4 slice products (MI355X): this loop 1.74 POPS; reading
  // the next stage's fragments across a mid-stage barrier 1.54 (the register allocator then
  // shuffles the 256 accumulator registers); two 256x128 workgroups per CU 1.54 (LDS-bound); the
  // register-staged double buffer 1.27.
  if (total > 0) {
#pragma unroll
    for (int s = 0; s < OZ_S - 1; ++s) issue(min(s, total - 1), s);
  }
  for (int it = 0; it < total; ++it) {
    // stage it has landed for every wave and every wave's reads of stage it-1 are done: that
    // slot takes stage it+OZ_S-1
    wait_stage(OZ_S - 2);
    issue(min(it + OZ_S - 1, total - 1), (it + OZ_S - 1) % OZ_S);
    // the fragment reads first, in the order the MFMAs consume them
    i32x4 af[8], bf[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) af[i] = frag_a(it, i);
#pragma unroll
    for (int j = 0; j < 8; ++j) bf[j] = frag_b(it, j);
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int i = 0; i < 8; ++i) acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(bf[j], af[i], acc[i][j], 0, 0, 0);
    __builtin_amdgcn_sched_group_barrier(0x100, 16, 0);  // DS reads
    __builtin_amdgcn_sched_group_barrier(0x008, 64, 0);  // then the MFMAs
  }
  // no LDS load may still be landing when the workgroup ends (its LDS is handed to the next one)
  wait_vm_lgkm0<0>();

  // lane holds C(i = m0 + wm*128 + bi*16 + li, j = n0 + wn*128 + bj*16 + 4*lg + r)
  // epilogue: C = alpha * 2^(ea+eb+shift) * acc + (accumulate ? C : 0) + beta * C_in
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

// Residue product of one modulus (nprod == 1, residue mode) with the main loop scheduled by hand:
// the same tile, ring slot layout, load mapping and fragment layout as ozaki_gemm_kernel, operands
// in k-step tiled layout, a 4-stage ring (128 KB of LDS) and the k-step schedule of
// gen_crt_loop.py (fragments double-buffered in registers, loads and LDS reads interleaved with
// the MFMAs, one barrier half-way through each k-step). Measured on MI355X, 8192^3: 2.47 int8
// POP/s against 1.68 for ozaki_gemm_kernel on row-major operands; a 5-stage ring (160 KB) gave
// the same rate, the limit being the L2 -> LDS data path (texture data unit busy 83% of cycles).
template <int S>
__global__ __launch_bounds__(256, 1) void crt_product_kernel(OzakiGemmParams prm) {
  static_assert(S == 4, "the generated loop exists for a 4-stage ring");
  __shared__ __attribute__((aligned(16))) int8_t lds[S * OZ_STAGE];
  if (prm.fa && (*prm.fa | *prm.fb)) return;  // the native fallback computes this product

  int m0, n0;
  oz_tile(prm.m, prm.n, m0, n0);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int li = lane & 15, lg = lane >> 4;

  i32x4 acc[8][8];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[i][j] = i32x4{0, 0, 0, 0};

  const int nkt = (prm.k_end - prm.k_begin) / OZ_BK;
  if (nkt > 0) {
    const unsigned lds0 = (unsigned)(uintptr_t)((__attribute__((address_space(3))) int8_t*)lds);
    // load t of a k-step fills LDS rows (4*wave + t)*16 .. +15 of each operand (the oz_issue mapping)
    const unsigned ldsw = __builtin_amdgcn_readfirstlane(lds0 + (unsigned)wave * 4 * 1024);
    // k-step tiled operands: row `local` of a 16 KB k-step block at byte local*64; load t covers
    // rows (4*wave + t)*16 + lane/4, i.e. the bytes at vo[t] (the swizzle depends on lane/4)
    unsigned vo[4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
      vo[t] = ((unsigned)(wave * 4 + t) * 16 + (lane >> 2)) * 64 + 16u * (unsigned)((lane & 3) ^ oz_swz(lane >> 2));
    // fragment reads (oz_frag): row wm*128 + i*16 + li at byte row*64 + 16*(lg ^ swz(row)); the
    // swizzle depends on li only, so fragment i sits at a fixed offset i*1024 from the lane base
    const unsigned lane_off = (unsigned)li * 64 + ((unsigned)(lg ^ oz_swz(li)) << 4);
    unsigned ra[S], rb[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      ra[s] = lds0 + s * OZ_STAGE + wm * 8192 + lane_off;
      rb[s] = lds0 + s * OZ_STAGE + OZ_TILE + wn * 8192 + lane_off;
    }
    const long nks = prm.kpad / OZ_BK, ks0 = prm.k_begin / OZ_BK;
    const int8_t* gA = prm.a + (long)prm.pa[0] * prm.a_slice_stride + ((m0 / OZ_BM) * nks + ks0) * OZ_TILE;
    const int8_t* gB = prm.b + (long)prm.pb[0] * prm.b_slice_stride + ((n0 / OZ_BN) * nks + ks0) * OZ_TILE;
    unsigned cnt = (unsigned)nkt, adv = (unsigned)nkt - 1, tmp, tmp2;
    i32x4 fa[8], fb[8], ga[8], gb[8];
    // M0 is written by the loop (LDS DMA destination); nothing else in this kernel uses it
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
    asm volatile(TMM_CRT_LOOP_S4 : TMM_CRT_OUTPUTS : TMM_CRT_INPUTS_S4 : "memory", "m0", "scc");
#pragma clang diagnostic pop
  }
  oz_store_residues(prm, acc, m0, n0, wm, wn, li, lg, lds);
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// Scratch of the no-allocation entry point. In host code `scratch` names a host-side stand-in,
// not device memory: only its per-device address from hipGetSymbolAddress is ever handed to a
// kernel or a copy. Nothing in it has to start at zero (every kernel of the emulation writes what
// it later reads), and calls that overlap on one device share it.
// ---------------------------------------------------------------------------------------------
static __device__ __align__(256) char scratch[kOzakiScratchBytes];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchDevices] = {};
}  // namespace

hipError_t ozaki_scratch_resolve() {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return hipSuccess;
  void* p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_ptr[dev] = p;
  return hipSuccess;
}

void* ozaki_scratch(int device) {
  if (device < 0 || device >= kScratchDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}

hipError_t ozaki_exponents(const double* p, long ld, bool kc, int rows, int k, int* e, int* flag, hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(exp_init_kernel, dim3((rows + 255) / 256), dim3(256), 0, s, e, rows, flag);
  if (k <= 0) return hipGetLastError();
  if (kc) {
    const int ksplit = k >= 8192 ? 8 : 1, kper = (k + ksplit - 1) / ksplit;
    hipLaunchKernelGGL(exp_rows_kc_kernel, dim3((rows + 3) / 4, ksplit), dim3(256), 0, s, p, ld, rows, k, kper, e,
                       flag);
  } else {
    const int ksplit = std::min(64, std::max(1, k / 256)), kper = (k + ksplit - 1) / ksplit;
    hipLaunchKernelGGL(exp_rows_rc_kernel, dim3((rows + 255) / 256, ksplit), dim3(256), 0, s, p, ld, rows, k, kper, e,
                       flag);
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
  if (p.res) {
    if (p.nprod != 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(crt_product_kernel<4>, dim3(nbm * nbn), dim3(256), 0, s, p);
  } else {
    hipLaunchKernelGGL(ozaki_gemm_kernel, dim3(nbm * nbn), dim3(256), 0, s, p);
  }
  return hipGetLastError();
}

namespace {
inline long pad_to(long x, long q) { return (x + q - 1) / q * q; }
inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
}  // namespace

// exponents of `rows` rows followed by the operand's fallback flag
static inline size_t exp_flag_bytes(int rows) { return align256((size_t)std::max(rows, 1) * 4 + 4); }

size_t ozaki_workspace_bytes(int m, int n, int k, int nslices) {
  const long kpad = pad_to(std::max(k, 1), OZ_BK), mpad = pad_to(m, OZ_BM), npad = pad_to(n, OZ_BN);
  return exp_flag_bytes(m) + exp_flag_bytes(n) + align256((size_t)nslices * mpad * kpad) +
         align256((size_t)nslices * npad * kpad);
}

namespace {
// The native FP64 kernel on the original operands, doing its work only when fa or fb is set.
hipError_t native_fallback(char opa, char opb, int m, int n, int k, double alpha, const double* a, long lda,
                           const double* b, long ldb, double beta, const double* cin, long ldcin, bool accumulate,
                           double* c, long ldc, const int* fa, const int* fb, hipStream_t s) {
  GemmParams<double> p;
  p.opa = opa;
  p.opb = opb;
  p.m = m;
  p.n = n;
  p.k = k;
  p.alpha = alpha;
  p.alpha_im = 0.0;
  p.a = a;
  p.lda = lda;
  p.b = b;
  p.ldb = ldb;
  p.beta = beta;
  p.beta_im = 0.0;
  p.cin = cin;
  p.ldcin = ldcin;
  p.accumulate = accumulate ? 1 : 0;
  p.c = c;
  p.ldc = ldc;
  p.variant = 1;
  p.gate_a = fa;
  p.gate_b = fb;
  return launch_gemm_d(p, s);
}
}  // namespace

hipError_t ozaki_dgemm(char opa, char opb, int m, int n, int k, double alpha, const double* a, long lda,
                       const double* b, long ldb, double beta, const double* cin, long ldcin, bool accumulate,
                       double* c, long ldc, int nslices, void* work, hipStream_t s, int* n_products) {
  if (n_products) *n_products = 0;
  if (m <= 0 || n <= 0) return hipSuccess;
  if (nslices < 1 || nslices > 12) return hipErrorInvalidValue;
  const long kpad = pad_to(std::max(k, 1), OZ_BK), mpad = pad_to(m, OZ_BM), npad = pad_to(n, OZ_BN);
  char* w = (char*)work;
  int* ea = (int*)w;
  w += exp_flag_bytes(m);
  int* eb = (int*)w;
  w += exp_flag_bytes(n);
  int* fa = ea + m;  // fallback flags
  int* fb = eb + n;
  int8_t* sa = (int8_t*)w;
  w += align256((size_t)nslices * mpad * kpad);
  int8_t* sb = (int8_t*)w;
  // op(A) rows: k-contiguous when A is transposed; op(B) columns: k-contiguous when B is not.
  const bool a_kc = opa != 'N', b_kc = opb == 'N';
  hipError_t e;
  if ((e = ozaki_exponents(a, lda, a_kc, m, k, ea, fa, s)) != hipSuccess) return e;
  if ((e = ozaki_exponents(b, ldb, b_kc, n, k, eb, fb, s)) != hipSuccess) return e;
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
  P.fa = fa;
  P.fb = fb;
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
  if ((e = native_fallback(opa, opb, m, n, k, alpha, a, lda, b, ldb, beta, cin, ldcin, accumulate, c, ldc, fa, fb,
                           s)) != hipSuccess)
    return e;
  if (n_products) *n_products = count;
  return hipSuccess;
}


// =============================================================================================
// Ozaki scheme II: integer products modulo pairwise coprime moduli, CRT reconstruction.
//
// (Ozaki, Uchino, Imamura 2025, "Ozaki Scheme II: A GEMM-oriented emulation of floating-point
// matrix multiplication using an integer modular technique".)
//   * a'(i,k) = round(|a(i,k)| * 2^(bits - ea[i])) * sign, |a'| <= 2^bits, likewise b'(k,j) with
//     eb[j]; op(A)op(B) = 2^(ea[i] + eb[j] - 2 bits) * (A'B')(i,j) up to the rounding of the inputs
//     (at most 2^-bits of the row / column maximum per entry);
//   * |(A'B')(i,j)| <= k 2^(2 bits) < M/4 with M the product of the moduli, so A'B' is fixed by its
//     residues r_t = A'B' mod m_t; each residue comes from one INT8 product of the balanced
//     residues of A' and B' (entries in [-128, 127], exact int32 sums for k <= 65536);
//   * CRT sum: A'B' = sum_t r_t W_t - q M with W_t = (M/m_t) y_t, y_t = (M/m_t)^-1 mod m_t, and
//     q = rint(sum_t r_t y_t / m_t) (in FP64 within 2^-35 of sum_t r_t W_t / M, whose distance to
//     the integer q is |A'B'| / M < 1/4, so the rounding is never in doubt); the sum is taken
//     exactly in 26-bit FP64 limbs (every partial sum an integer below 2^40), carried into
//     non-negative limbs and added from the top in double-double: exact when A'B' is
//     representable, otherwise within one unit in the last place.
// =============================================================================================
namespace {

constexpr int kCrtModuli[kCrtMaxModuli] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227,
                                           223, 211, 199, 197, 193, 191, 181, 179, 173, 167};

struct CrtParams {
  int n;
  unsigned m[kCrtMaxModuli];
  unsigned magic[kCrtMaxModuli];  // ceil(2^32 / m)
  unsigned c21[kCrtMaxModuli];    // 2^21 mod m
  unsigned c42[kCrtMaxModuli];    // 2^42 mod m
};

constexpr long pow_mod(long b, long e, long m) {
  long r = 1 % m;
  b %= m;
  while (e > 0) {
    if (e & 1) r = r * b % m;
    b = b * b % m;
    e >>= 1;
  }
  return r;
}

constexpr long inv_mod(long a, long m) {  // extended Euclid; a and m coprime
  long g = m, x = 0, x1 = 1, a1 = a % m;
  while (a1 != 0) {
    const long q = g / a1, t = g - q * a1;
    g = a1;
    a1 = t;
    const long tx = x - q * x1;
    x = x1;
    x1 = tx;
  }
  return ((x % m) + m) % m;
}

// Constants of the CRT sum for the first N moduli, fixed by the moduli list and known at compile
// time (the unrolled loops take them as scalar operands): f[t] = y_t / m_t rounded to FP64, the
// 26-bit limbs of W_t = (M/m_t) y_t and of M (M <= 256^N needs L = ceil(8N/26) limbs).
constexpr int kLimbBits = 26;
template <int N>
struct CrtSum {
  static constexpr int L = (8 * N + kLimbBits - 1) / kLimbBits;
  double f[N];
  double w[N][L];
  double mm[L];
};

// x *= s for a little-endian number of L 26-bit limbs (s <= 256; the product fits the L limbs)
template <int L>
constexpr void limbs_mul_small(unsigned long long (&x)[L], unsigned long long s) {
  unsigned long long carry = 0;
  for (int i = 0; i < L; ++i) {
    const unsigned long long v = x[i] * s + carry;
    x[i] = v & ((1ull << kLimbBits) - 1);
    carry = v >> kLimbBits;
  }
}

template <int N>
constexpr CrtSum<N> make_crt_sum() {
  constexpr int L = CrtSum<N>::L;
  CrtSum<N> d{};
  unsigned long long M[L] = {};
  M[0] = 1;
  for (int t = 0; t < N; ++t) limbs_mul_small<L>(M, (unsigned long long)kCrtModuli[t]);
  for (int i = 0; i < L; ++i) d.mm[i] = (double)M[i];
  for (int t = 0; t < N; ++t) {
    const long m = kCrtModuli[t];
    long mt_mod = 1;  // (M / m_t) mod m_t
    for (int s = 0; s < N; ++s)
      if (s != t) mt_mod = mt_mod * kCrtModuli[s] % m;
    const long y = inv_mod(mt_mod, m);
    unsigned long long W[L] = {};
    W[0] = (unsigned long long)y;
    for (int s = 0; s < N; ++s)
      if (s != t) limbs_mul_small<L>(W, (unsigned long long)kCrtModuli[s]);
    for (int i = 0; i < L; ++i) d.w[t][i] = (double)W[i];
    d.f[t] = (double)y / (double)m;
  }
  return d;
}

template <int N>
struct CrtSumTable {
  static constexpr CrtSum<N> v = make_crt_sum<N>();
};

CrtParams make_crt(int n) {
  CrtParams c{};
  c.n = n;
  for (int t = 0; t < n; ++t) {
    const unsigned m = (unsigned)kCrtModuli[t];
    c.m[t] = m;
    c.magic[t] = (unsigned)((((unsigned long long)1 << 32) + m - 1) / m);
    c.c21[t] = (unsigned)pow_mod(2, 21, m);
    c.c42[t] = (unsigned)pow_mod(2, 42, m);
  }
  return c;
}

// Scaled, rounded entry x = round(|v| 2^(bits - er)) <= 2^62, split exactly into 21-bit digits
// x = x2 2^42 + x1 2^21 + x0, so that x mod m = (x2 c42 + x1 c21 + x0) mod m with one reduction
// of a value below 2^30.
struct Split21 {
  unsigned x0, x1, x2;
  bool neg;
};
__device__ __forceinline__ Split21 split21(double v, int er, int bits) {
  Split21 s;
  s.neg = v < 0.0;
  const double x = er == kExpNone ? 0.0 : rint(ldexp(fabs(v), bits - er));
  const double h2 = floor(x * 0x1p-42);
  const double rem = x - h2 * 0x1p42;
  const double h1 = floor(rem * 0x1p-21);
  s.x2 = (unsigned)h2;
  s.x1 = (unsigned)h1;
  s.x0 = (unsigned)(rem - h1 * 0x1p21);
  return s;
}

// Balanced residue (int8) of a split entry modulo cp.m[q].
__device__ __forceinline__ uint32_t residue8(const Split21& s, const CrtParams& cp, int q) {
  const unsigned m = cp.m[q];
  unsigned r = mod_small(s.x2 * cp.c42[q] + s.x1 * cp.c21[q] + s.x0, m, cp.magic[q]);
  if (s.neg && r != 0) r = m - r;
  const int bal = 2 * r >= m ? (int)r - (int)m : (int)r;
  return (uint32_t)(uint8_t)(int8_t)bal;
}

// Byte offset of entry (r, k) in a residue plane of rows_pad x kpad in k-step tiled layout: one
// contiguous 16 KB block per (256-row block, 64-k step), blocks of a row block in k order, rows of
// 64 bytes inside a block. A k-step of a 256-row product tile is then one contiguous 16 KB read.
__device__ __forceinline__ long crt_tiled_offset(int r, int k, int kpad) {
  return ((long)(r >> 8) * (kpad >> 6) + (k >> 6)) * 16384 + (r & 255) * 64 + (k & 63);
}

// Balanced residues of the scaled, rounded entries: one int8 plane per modulus, rows_pad x kpad
// in k-step tiled layout (zero outside the matrix); k-contiguous source: each thread handles 4
// consecutive k of one row and writes one 4-byte word (consecutive threads, consecutive words).
__global__ __launch_bounds__(256) void crt_residue_kc_kernel(const double* __restrict__ p, long ld, int rows, int k,
                                                             const int* __restrict__ e, int bits, CrtParams cp,
                                                             int rows_pad, int kpad, int8_t* __restrict__ out) {
  // thread tid writes bytes 4*tid .. 4*tid+3 of each plane (k-step tiled layout, crt_tiled_offset)
  const long tid = (long)blockIdx.x * 256 + threadIdx.x;
  if (tid >= (long)rows_pad * (kpad / 4)) return;
  const long blk = tid >> 12;  // 16 KB block = (256-row block, 64-k step)
  const int nks = kpad / 64;
  const int r = (int)(blk / nks) * 256 + (int)((tid >> 4) & 255);
  const int kq = (int)(blk % nks) * 16 + (int)(tid & 15);
  const int er = r < rows ? e[r] : kExpNone;
  Split21 s[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int kk = kq * 4 + t;
    s[t] = split21(r < rows && kk < k ? p[(long)kk + (long)r * ld] : 0.0, er, bits);
  }
  const long plane = (long)rows_pad * kpad;
  int8_t* o = out + tid * 4;
  for (int q = 0; q < cp.n; ++q) {
    uint32_t word = 0;
#pragma unroll
    for (int t = 0; t < 4; ++t) word |= residue8(s[t], cp, q) << (8 * t);
    *(uint32_t*)(o + q * plane) = word;
  }
}

// Same for a row-contiguous source: a block transposes a 64-row x 64-k tile through LDS (reads
// of 64 consecutive rows per k), then each thread takes 16 consecutive k of one row and writes
// 16 bytes per plane (the block's 64 rows fill 4 KB of one 16 KB k-step block). Grid
// (rows_pad / 64, kpad / 64).
__global__ __launch_bounds__(256) void crt_residue_rc_kernel(const double* __restrict__ p, long ld, int rows, int k,
                                                             const int* __restrict__ e, int bits, CrtParams cp,
                                                             int kpad, long plane, int8_t* __restrict__ out) {
  __shared__ double tile[64][64];  // [k][row]
  const int r0 = blockIdx.x * 64, k0 = blockIdx.y * 64, t = threadIdx.x;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int kl = (t >> 6) + 4 * j, rl = t & 63, r = r0 + rl, kk = k0 + kl;
    tile[kl][rl] = r < rows && kk < k ? p[(long)r + (long)kk * ld] : 0.0;
  }
  __syncthreads();
  const int rl = t & 63, g = t >> 6, r = r0 + rl;
  const int er = r < rows ? e[r] : kExpNone;
  Split21 s[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) s[i] = split21(tile[16 * g + i][rl], er, bits);
  int8_t* o = out + crt_tiled_offset(r, k0 + 16 * g, kpad);
  for (int q = 0; q < cp.n; ++q) {
    uint32_t w[4] = {0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < 16; ++i) w[i >> 2] |= residue8(s[i], cp, q) << (8 * (i & 3));
    *(uint4*)(o + q * plane) = uint4{w[0], w[1], w[2], w[3]};
  }
}

// S_i -> [0, 2^26) for every limb but the top, the carries moving up (each step exact: the limbs
// are integers below 2^40 and floor of a power-of-two scaling is exact).
template <int L>
__device__ __forceinline__ void crt_carry(double* S) {
#pragma unroll
  for (int i = 0; i < L - 1; ++i) {
    const double c = floor(S[i] * 0x1p-26);
    S[i] = fma(-c, 0x1p26, S[i]);
    S[i + 1] = S[i + 1] + c;
  }
}

// FP64 value of Y = sum_i S_i 2^(26 i) (exact integer limbs, |Y| < M/4): carried, made
// non-negative (the sign is the top limb's once the others lie in [0, 2^26)), paired into limbs
// below 2^52 and added from the top in double-double. Exact when Y is representable, as every
// top-down partial sum then is; otherwise within one unit in the last place. Contraction is off
// so that each line is one IEEE operation (tests/crt_model.py mirrors them).
template <int L>
__device__ __forceinline__ double crt_limbs_value(double* S) {
#pragma clang fp contract(off)
  crt_carry<L>(S);
  const double sgn = S[L - 1] < 0.0 ? -1.0 : 1.0;
#pragma unroll
  for (int i = 0; i < L; ++i) S[i] = sgn * S[i];
  crt_carry<L>(S);
  constexpr int NP = (L + 1) / 2;
  double P[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) P[j] = 2 * j + 1 < L ? fma(S[2 * j + 1], 0x1p26, S[2 * j]) : S[2 * j];
  double h = ldexp(P[NP - 1], 52 * (NP - 1)), lo = 0.0;
#pragma unroll
  for (int j = NP - 2; j >= 0; --j) {
    const double T = ldexp(P[j], 52 * j);
    const double s = h + T;  // TwoSum(h, T) = s + e exactly
    const double bb = s - h;
    const double e = (h - (s - bb)) + (T - bb);
    h = s;
    lo = lo + e;
  }
  return sgn * (h + lo);
}

// C(i,j) = alpha * 2^(ea[i] + eb[j] - 2 bits) * CRT(residues) + (accumulate ? C : 0) + beta*C_in.
// A thread takes 4 consecutive rows of one column: the residues of the 4 sit in one 32-bit word
// per plane (ldres is a multiple of 16), and each constant of the sum serves the 4 rows in turn.
// Grid: x over row quads, y over columns (strided).
template <int N>
__global__ __launch_bounds__(256) void crt_reconstruct_kernel(const uint8_t* __restrict__ res, long plane, long ldres,
                                                              int m, int n, const int* __restrict__ ea,
                                                              const int* __restrict__ eb, int bits,
                                                              double alpha, double beta, const double* __restrict__ cin,
                                                              long ldcin, int accumulate, double* __restrict__ c,
                                                              long ldc, const int* __restrict__ fa,
                                                              const int* __restrict__ fb) {
  constexpr int L = CrtSum<N>::L;
  constexpr CrtSum<N> D = CrtSumTable<N>::v;
  if (*fa | *fb) return;  // the native fallback writes this C
  const int i0 = 4 * (blockIdx.x * 256 + threadIdx.x);
  if (i0 >= m) return;
  int ei[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) ei[u] = i0 + u < m ? ea[i0 + u] : kExpNone;
  for (int j = blockIdx.y; j < n; j += gridDim.y) {
    const long off = (long)i0 + (long)j * ldres;
    uint32_t w[N];
#pragma unroll
    for (int t = 0; t < N; ++t) w[t] = *(const uint32_t*)(res + off + (long)t * plane);
    const int ej = eb[j];
    // per row: the q sum sum_t r_t f_t and the limb sums sum_t r_t W_ti (exact integers)
    double qs[4] = {0.0, 0.0, 0.0, 0.0}, S[4][L];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int i = 0; i < L; ++i) S[u][i] = 0.0;
#pragma unroll
    for (int t = 0; t < N; ++t) {
      double r[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) r[u] = (double)((w[t] >> (8 * u)) & 255u);
#pragma unroll
      for (int u = 0; u < 4; ++u) qs[u] = fma(r[u], D.f[t], qs[u]);
#pragma unroll
      for (int i = 0; i < L; ++i)
#pragma unroll
        for (int u = 0; u < 4; ++u) S[u][i] = fma(r[u], D.w[t][i], S[u][i]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (i0 + u >= m) break;
      const double q = rint(qs[u]);
#pragma unroll
      for (int i = 0; i < L; ++i) S[u][i] = fma(-q, D.mm[i], S[u][i]);
      const double y = crt_limbs_value<L>(S[u]);
      double out = 0.0;
      if (ei[u] != kExpNone && ej != kExpNone) out = alpha * ldexp(y, ei[u] + ej - 2 * bits);
      const int i = i0 + u;
      double* dst = c + (long)i + (long)j * ldc;
      if (accumulate) out += *dst;
      if (beta != 0.0) out += beta * cin[(long)i + (long)j * ldcin];
      *dst = out;
    }
  }
}

template <int N>
void launch_reconstruct(dim3 grid, hipStream_t s, const uint8_t* res, long plane, long ldres, int m, int n,
                        const int* ea, const int* eb, int bits, double alpha, double beta,
                        const double* cin, long ldcin, int accumulate, double* c, long ldc) {
  // the fallback flags follow the exponents of each prepared operand
  hipLaunchKernelGGL(crt_reconstruct_kernel<N>, grid, dim3(256), 0, s, res, plane, ldres, m, n, ea, eb, bits, alpha,
                     beta, cin, ldcin, accumulate, c, ldc, ea + m, eb + n);
}

}  // namespace

int ozaki2_bits(int nmod, int k) {
  double l = 0;
  for (int t = 0; t < nmod && t < kCrtMaxModuli; ++t) l += std::log2((double)kCrtModuli[t]);
  const int lk = (int)std::ceil(std::log2((double)std::max(k, 1)));
  // k * 2^(2 bits) <= 2^(l - 2) < M / 4
  const int bits = (int)std::floor((l - 2.0 - lk) / 2.0);
  return std::max(1, std::min(62, bits));
}

size_t ozaki2_operand_bytes(int rows, int k, int nmod) {
  const long kr = pad_to(std::max(k, 1), OZ_BK), rows_pad = pad_to(std::max(rows, 1), OZ_BM);
  return exp_flag_bytes(rows) + align256((size_t)nmod * rows_pad * kr);
}

size_t ozaki2_product_bytes(int m, int n, int nmod) {
  return align256((size_t)nmod * pad_to(std::max(m, 1), 16) * std::max(n, 1));
}

size_t ozaki2_workspace_bytes(int m, int n, int k, int nmod) {
  const int kr = std::min(std::max(k, 1), kCrtMaxK);
  return ozaki2_operand_bytes(m, kr, nmod) + ozaki2_operand_bytes(n, kr, nmod) + ozaki2_product_bytes(m, n, nmod);
}

hipError_t ozaki2_prepare(const double* p, long ld, bool kc, int rows, int k, int nmod, void* out, hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  if (nmod < kCrtMinModuli || nmod > kCrtMaxModuli || k > kCrtMaxK) return hipErrorInvalidValue;
  const int kp = (int)pad_to(std::max(k, 1), OZ_BK);
  const long rows_pad = pad_to(rows, OZ_BM);
  int* ex = (int*)out;
  int8_t* res = (int8_t*)out + exp_flag_bytes(rows);
  const CrtParams cp = make_crt(nmod);
  const int bits = ozaki2_bits(nmod, k);
  hipError_t e;
  if ((e = ozaki_exponents(p, ld, kc, rows, k, ex, ex + rows, s)) != hipSuccess) return e;
  if (kc)
    hipLaunchKernelGGL(crt_residue_kc_kernel, dim3((unsigned)((rows_pad * (kp / 4) + 255) / 256)), dim3(256), 0, s, p,
                       ld, rows, k, ex, bits, cp, (int)rows_pad, kp, res);
  else
    hipLaunchKernelGGL(crt_residue_rc_kernel, dim3((unsigned)(rows_pad / 64), (unsigned)(kp / 64)), dim3(256), 0, s, p,
                       ld, rows, k, ex, bits, cp, kp, rows_pad * kp, res);
  return hipGetLastError();
}

hipError_t ozaki2_products(int m, int n, int k, const void* pa, const void* pb, int nmod, void* work, hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  if (nmod < kCrtMinModuli || nmod > kCrtMaxModuli || k > kCrtMaxK) return hipErrorInvalidValue;
  const int kp = (int)pad_to(std::max(k, 1), OZ_BK);
  const long mpad = pad_to(m, OZ_BM), npad = pad_to(n, OZ_BN);
  const int8_t* ra = (const int8_t*)pa + exp_flag_bytes(m);
  const int8_t* rb = (const int8_t*)pb + exp_flag_bytes(n);
  uint8_t* rc = (uint8_t*)work;
  const CrtParams cp = make_crt(nmod);
  hipError_t e;
  OzakiGemmParams P{};
  P.m = m;
  P.n = n;
  P.kpad = kp;
  P.k_begin = 0;
  P.k_end = kp;
  P.nprod = 1;
  P.fa = (const int*)pa + m;  // operands with a fallback flag skip the products
  P.fb = (const int*)pb + n;
  const long ldres = pad_to(m, 16), plane = ldres * n;
  P.ldres = ldres;
  for (int t = 0; t < nmod; ++t) {
    P.a = ra + (long)t * mpad * kp;
    P.b = rb + (long)t * npad * kp;
    P.res = rc + (long)t * plane;
    P.modulus = (int)cp.m[t];
    P.magic = cp.magic[t];
    if ((e = ozaki_gemm(P, s)) != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t ozaki2_reconstruct(int m, int n, int k, double alpha, const void* pa, const void* pb, double beta,
                              const double* cin, long ldcin, bool accumulate, double* c, long ldc, int nmod,
                              const void* work, hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  if (nmod < kCrtMinModuli || nmod > kCrtMaxModuli || k > kCrtMaxK) return hipErrorInvalidValue;
  const int* ea = (const int*)pa;
  const int* eb = (const int*)pb;
  const uint8_t* rc = (const uint8_t*)work;
  const int bits = ozaki2_bits(nmod, k);
  const long ldres = pad_to(m, 16), plane = ldres * n;
  const dim3 grid((unsigned)((m + 1023) / 1024), (unsigned)std::min(n, 32768));
  const int acc = accumulate ? 1 : 0;
  switch (nmod) {
#define TMM_CRT_CASE(NN) \
  case NN:                                                                                                     \
    launch_reconstruct<NN>(grid, s, rc, plane, ldres, m, n, ea, eb, bits, alpha, beta, cin, ldcin, acc, c, ldc); \
    break;
    TMM_CRT_CASE(6) TMM_CRT_CASE(7) TMM_CRT_CASE(8) TMM_CRT_CASE(9) TMM_CRT_CASE(10) TMM_CRT_CASE(11)
    TMM_CRT_CASE(12) TMM_CRT_CASE(13) TMM_CRT_CASE(14) TMM_CRT_CASE(15) TMM_CRT_CASE(16) TMM_CRT_CASE(17)
    TMM_CRT_CASE(18) TMM_CRT_CASE(19) TMM_CRT_CASE(20)
#undef TMM_CRT_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t ozaki2_fallback(char opa, char opb, int m, int n, int k, double alpha, const double* a, long lda,
                           const double* b, long ldb, double beta, const double* cin, long ldcin, bool accumulate,
                           double* c, long ldc, const void* pa, const void* pb, hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  return native_fallback(opa, opb, m, n, k, alpha, a, lda, b, ldb, beta, cin, ldcin, accumulate, c, ldc,
                         (const int*)pa + m, (const int*)pb + n, s);
}

hipError_t ozaki2_product(int m, int n, int k, double alpha, const void* pa, const void* pb, double beta,
                          const double* cin, long ldcin, bool accumulate, double* c, long ldc, int nmod, void* work,
                          hipStream_t s) {
  const hipError_t e = ozaki2_products(m, n, k, pa, pb, nmod, work, s);
  if (e != hipSuccess) return e;
  return ozaki2_reconstruct(m, n, k, alpha, pa, pb, beta, cin, ldcin, accumulate, c, ldc, nmod, work, s);
}

hipError_t ozaki2_dgemm(char opa, char opb, int m, int n, int k, double alpha, const double* a, long lda,
                        const double* b, long ldb, double beta, const double* cin, long ldcin, bool accumulate,
                        double* c, long ldc, int nmod, void* work, hipStream_t s, int* n_products) {
  if (n_products) *n_products = 0;
  if (m <= 0 || n <= 0) return hipSuccess;
  if (nmod < kCrtMinModuli || nmod > kCrtMaxModuli) return hipErrorInvalidValue;
  const int kr = std::min(std::max(k, 1), kCrtMaxK);
  char* pa = (char*)work;
  char* pb = pa + ozaki2_operand_bytes(m, kr, nmod);
  char* rc = pb + ozaki2_operand_bytes(n, kr, nmod);
  const bool a_kc = opa != 'N', b_kc = opb == 'N';
  hipError_t e;
  int count = 0;
  const int nranges = (std::max(k, 1) + kCrtMaxK - 1) / kCrtMaxK;
  for (int q = 0; q < nranges; ++q) {
    // one pass per k range of at most kCrtMaxK: exponents and residues of both operands, nmod
    // INT8 products and a reconstruction that adds into C (alpha/beta/accumulate on the first
    // pass only)
    const int k0 = q * kCrtMaxK, kk = std::min(k, k0 + kCrtMaxK) - k0;
    const double* ap = a + (a_kc ? (long)k0 : (long)k0 * lda);
    const double* bp = b + (b_kc ? (long)k0 : (long)k0 * ldb);
    if ((e = ozaki2_prepare(ap, lda, a_kc, m, kk, nmod, pa, s)) != hipSuccess) return e;
    if ((e = ozaki2_prepare(bp, ldb, b_kc, n, kk, nmod, pb, s)) != hipSuccess) return e;
    if ((e = ozaki2_product(m, n, kk, alpha, pa, pb, q == 0 ? beta : 0.0, cin, ldcin, q == 0 ? accumulate : true, c,
                            ldc, nmod, rc, s)) != hipSuccess)
      return e;
    if ((e = ozaki2_fallback(opa, opb, m, n, kk, alpha, ap, lda, bp, ldb, q == 0 ? beta : 0.0, cin, ldcin,
                             q == 0 ? accumulate : true, c, ldc, pa, pb, s)) != hipSuccess)
      return e;
    count += nmod;
  }
  if (n_products) *n_products = count;
  return hipSuccess;
}

}  // namespace tmm

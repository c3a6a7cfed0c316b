// Per-tile GEMM kernels for gfx950 (MI355X): C_out = alpha*op(A)*op(B) + gamma*C_out + beta*C_in.
//
// This is the device arithmetic that the reference hands to the vendor BLAS for every tile
// (Tiled-MM src/Tiled-MM/tiled_mm.cpp:181-268, cublas_gemm_wrapper -> {s,d,c,z}gemm). Here it
// is written directly for CDNA4 matrix cores:
//   * double  : v_mfma_f64_16x16x4_f64, one 16x16 block per instruction;
//   * float   : v_mfma_f32_16x16x4_f32;
//   * complex : four real MFMAs per block (re*re, -im*im, re*im, im*re), conjugation of
//               op='C' folded into the load path.
// The epilogue is fused: alpha scaling, accumulation into the existing tile (gamma in {0,1}),
// and an optional beta*C_in term that lets the scheduler add the caller's C on the last
// k-chunk instead of the first one.
//
// Layout decisions (all matrices column-major):
//   * every operand tile is staged in LDS "k-paired": chunk q holds KV consecutive k values of
//     one row (16 bytes), stored [q][row][KV]; a wave's 16 lanes of one k-group read 16
//     consecutive 16-byte chunks, which is conflict-free for ds_read_b128;
//   * each lane of the MFMA takes 4 consecutive k values of its row (k = 16p + 4g + s, with
//     g = lane/16 and s the MFMA step), the same permutation for A and B, so the sum over
//     k is unchanged;
//   * operands are fed to the MFMA swapped (B fragment as the "A" operand), so that a lane's
//     output row index is lane%16 and 16 lanes store 16 consecutive rows of one column.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include "tmm_kernels.h"

namespace tmm {
namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

// 16-byte vector of the real type, aligned only to the element (column pitches are arbitrary).
typedef double v2d __attribute__((ext_vector_type(2)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename R> struct Vec16;
template <> struct Vec16<double> { typedef v2d type; };
template <> struct Vec16<float> { typedef v4f type; };

template <typename R> struct Acc;
template <> struct Acc<double> { typedef dbl4 type; };
template <> struct Acc<float> { typedef flt4 type; };

__device__ __forceinline__ dbl4 mfma(double a, double b, dbl4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ flt4 mfma(float a, float b, flt4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Output column (within a 16x16 block) held by lane l in accumulator register r.
// f64 16x16x4: register r of lane l is row (l/16)+4r of the instruction's output;
// f32 16x16x4: row 4(l/16)+r. With the swapped operands the instruction's rows are C columns.
template <typename R> __device__ __forceinline__ int out_col(int lane, int r);
template <> __device__ __forceinline__ int out_col<double>(int lane, int r) { return (lane >> 4) + 4 * r; }
template <> __device__ __forceinline__ int out_col<float>(int lane, int r) { return 4 * (lane >> 4) + r; }

template <typename V>
__device__ __forceinline__ V load16(const void* p) {
  V v;
  __builtin_memcpy(&v, p, sizeof(V));
  return v;
}

// Map a linear workgroup id to a (tile_m, tile_n) pair.
// 1) Workgroups are dispatched round-robin over the 8 XCDs (id % 8); give every XCD a
//    contiguous range of tile ids so that the tiles it runs at the same time share its L2.
// 2) Inside a range, walk tiles in groups of GROUP_M tile-rows, column by column, so that
//    the ~64 concurrently resident tiles of one XCD form a compact block.
__device__ __forceinline__ void tile_coords(int w, int nbm, int nbn, int& bm, int& bn) {
  const int ntiles = nbm * nbn;
  const int NXCD = 8;
  int t = w;
  if (ntiles >= NXCD * 4) {
    const int xcd = w % NXCD, idx = w / NXCD;
    const int per = ntiles / NXCD, rem = ntiles % NXCD;
    t = xcd * per + (xcd < rem ? xcd : rem) + idx;
  }
  const int GROUP_M = 8;
  const int in_group = GROUP_M * nbn;
  const int group = t / in_group;
  const int first_m = group * GROUP_M;
  const int gsz = (nbm - first_m) < GROUP_M ? (nbm - first_m) : GROUP_M;
  const int r = t - group * in_group;
  bm = first_m + r % gsz;
  bn = r / gsz;
}

// ---------------------------------------------------------------------------------------------
// Global -> register staging of one operand tile (NROWS x BK), then register -> LDS.
//
// KC ("k-contiguous"): element (r, k) at p[k + r*ld]  (A with op T/C, B with op N)
// RC ("row-contiguous"): element (r, k) at p[r + k*ld] (A with op N, B with op T/C)
// CPLX: elements are (re, im) pairs of R; LDS holds separate re and im planes.
// NEG_IM: negate the imaginary part (conjugate transpose).
// ---------------------------------------------------------------------------------------------
// LDS position of chunk q of tile row `row` inside plane q. Operands loaded k-contiguous (KC)
// are written by 8 consecutive lanes holding chunks q = 0..7 of one row; with planes a multiple
// of 128 bytes apart these 8 writes would all fall into the same banks. XOR-ing the row with
// q & 7 spreads them over 8 distinct 16-byte bank groups, and since it only permutes rows inside
// aligned groups of 8, the fragment reads (16 consecutive rows of one plane) stay conflict-free.
template <bool SWZ>
__device__ __forceinline__ int lds_row(int q, int row) {
  return SWZ ? (row ^ (q & 7)) : row;
}

template <typename R, bool CPLX, bool KC, bool NEG_IM, int NROWS, int BK>
struct TileLoader {
  static constexpr int E = CPLX ? 2 * (int)sizeof(R) : (int)sizeof(R);  // bytes per element
  static constexpr int KV = 16 / (int)sizeof(R);                          // k values per LDS chunk
  static constexpr int NQ = BK / KV;                                       // LDS chunks per row
  static constexpr int PLANE = NQ * NROWS * KV;                            // reals per LDS plane
  // 16-byte global loads per thread per stage: the NROWS x BK tile spread over 256 threads.
  static constexpr int LOADS = NROWS * BK * E / (16 * 256);
  static_assert(NROWS * BK * E == LOADS * 16 * 256, "tile must split into whole 16-byte loads per thread");
  typedef typename Vec16<R>::type V;
  V regs[LOADS];

  // KC real: a "unit" is one 16-byte global chunk = one LDS chunk.
  // KC complex: a unit is 32 bytes (two loads) = one re chunk + one im chunk.
  // RC: a unit is a block of RB rows x KV k-columns, RB = 16/E rows per 16-byte load.
  static constexpr int RB = 16 / E;
  static constexpr int UNIT_LOADS = KC ? (CPLX ? 2 : 1) : KV;
  static constexpr int UNITS = LOADS / UNIT_LOADS;  // per thread
  static_assert(UNITS * UNIT_LOADS == LOADS && UNITS >= 1, "loads must form whole units");

  __device__ __forceinline__ void unit_pos(int u, int& row, int& q) const {
    const int c = (int)threadIdx.x + 256 * u;
    if (KC) {
      row = c / NQ;
      q = c % NQ;
    } else {
      const int nrb = NROWS / RB;
      row = (c % nrb) * RB;
      q = c / nrb;
    }
  }

  // full: the whole NROWS x BK tile is inside the matrix (wave-uniform), so no element needs a
  // bounds check.
  __device__ __forceinline__ void load(const R* __restrict__ p, long ld, int row0, int nrows, int k0, int nk,
                                       bool full) {
    if (full) {
#pragma unroll
      for (int u = 0; u < UNITS; ++u) {
        int row, q;
        unit_pos(u, row, q);
        const int gr = row0 + row, gk = k0 + q * KV;
        if (KC) {
          const char* base = (const char*)p + ((long)gk + (long)gr * ld) * E;
#pragma unroll
          for (int h = 0; h < UNIT_LOADS; ++h) regs[u * UNIT_LOADS + h] = load16<V>(base + 16 * h);
        } else {
#pragma unroll
          for (int kk = 0; kk < KV; ++kk)
            regs[u * UNIT_LOADS + kk] = load16<V>((const char*)p + ((long)gr + (long)(gk + kk) * ld) * E);
        }
      }
      return;
    }
#pragma unroll
    for (int u = 0; u < UNITS; ++u) {
      int row, q;
      unit_pos(u, row, q);
      const int gr = row0 + row;
      const int gk = k0 + q * KV;
      if (KC) {
        // KV consecutive elements of row gr, starting at column gk.
        const char* base = (const char*)p + ((long)gk + (long)gr * ld) * E;
        if (gr < nrows && gk + KV <= nk) {
#pragma unroll
          for (int h = 0; h < UNIT_LOADS; ++h) regs[u * UNIT_LOADS + h] = load16<V>(base + 16 * h);
        } else {
#pragma unroll
          for (int h = 0; h < UNIT_LOADS; ++h) {
            V v;
#pragma unroll
            for (int e = 0; e < KV; ++e) {
              // real index e of this 16-byte piece -> element (e*sizeof(R))/E of the piece
              const int el = (16 * h + e * (int)sizeof(R)) / E;
              const bool ok = gr < nrows && (gk + el) < nk;
              v[e] = ok ? ((const R*)base)[h * KV + e] : R(0);
            }
            regs[u * UNIT_LOADS + h] = v;
          }
        }
      } else {
        // RB consecutive rows at each of KV columns gk..gk+KV-1.
#pragma unroll
        for (int kk = 0; kk < KV; ++kk) {
          const char* base = (const char*)p + ((long)gr + (long)(gk + kk) * ld) * E;
          if (gr + RB <= nrows && gk + kk < nk) {
            regs[u * UNIT_LOADS + kk] = load16<V>(base);
          } else {
            V v;
#pragma unroll
            for (int e = 0; e < KV; ++e) {
              const int el = (e * (int)sizeof(R)) / E;
              const bool ok = (gr + el) < nrows && (gk + kk) < nk;
              v[e] = ok ? ((const R*)base)[e] : R(0);
            }
            regs[u * UNIT_LOADS + kk] = v;
          }
        }
      }
    }
  }

  // Write the staged registers to LDS. lds_re / lds_im: planes of layout [NQ][NROWS][KV].
  __device__ __forceinline__ void store(R* lds_re, R* lds_im) const {
#pragma unroll
    for (int u = 0; u < UNITS; ++u) {
      int row, q;
      unit_pos(u, row, q);
      if (KC) {
        if (!CPLX) {
          *(V*)(lds_re + ((long)q * NROWS + lds_row<true>(q, row)) * KV) = regs[u];
        } else {
          // 2 loads = KV complex values (re0, im0, re1, im1, ...) -> split into planes.
          V re, im;
#pragma unroll
          for (int e = 0; e < KV; ++e) {
            const int h = (2 * e) / KV, o = (2 * e) % KV;
            re[e] = regs[u * 2 + h][o];
            im[e] = regs[u * 2 + h][o + 1];
            if (NEG_IM) im[e] = -im[e];
          }
          *(V*)(lds_re + ((long)q * NROWS + lds_row<true>(q, row)) * KV) = re;
          *(V*)(lds_im + ((long)q * NROWS + lds_row<true>(q, row)) * KV) = im;
        }
      } else {
        // regs[kk] holds RB rows at column kk; LDS chunk of row r gathers all kk.
        if (!CPLX) {
#pragma unroll
          for (int r = 0; r < RB; ++r) {
            V v;
#pragma unroll
            for (int kk = 0; kk < KV; ++kk) v[kk] = regs[u * UNIT_LOADS + kk][r];
            *(V*)(lds_re + ((long)q * NROWS + row + r) * KV) = v;
          }
        } else {
#pragma unroll
          for (int r = 0; r < RB; ++r) {
            V re, im;
#pragma unroll
            for (int kk = 0; kk < KV; ++kk) {
              re[kk] = regs[u * UNIT_LOADS + kk][2 * r];
              im[kk] = regs[u * UNIT_LOADS + kk][2 * r + 1];
              if (NEG_IM) im[kk] = -im[kk];
            }
            *(V*)(lds_re + ((long)q * NROWS + row + r) * KV) = re;
            *(V*)(lds_im + ((long)q * NROWS + row + r) * KV) = im;
          }
        }
      }
    }
  }
};

// Fragment of consecutive k values of one row, read from LDS. The lane's k values are
// k = 16p + 4g + s (s = 0..3); "half" h selects s in {2h, 2h+1} for 8-byte reals (one 16-byte
// chunk) and all four s for 4-byte reals (h must be 0).
template <typename R, int NROWS, bool SWZ>
__device__ __forceinline__ typename Vec16<R>::type read_chunk(const R* lds, int p, int g, int h, int row) {
  constexpr int KV = 16 / (int)sizeof(R);
  typedef typename Vec16<R>::type V;
  const int q = KV == 2 ? 8 * p + 2 * g + h : 4 * p + g;
  return *(const V*)(lds + ((long)q * NROWS + lds_row<SWZ>(q, row)) * KV);
}

// ---------------------------------------------------------------------------------------------
// Real GEMM: 256 threads, 2x2 waves, wave tile 64x64 (4x4 MFMA blocks), WG tile 128x128.
// ---------------------------------------------------------------------------------------------
template <typename R> struct RealCfg { static constexpr int BK = sizeof(R) == 8 ? 16 : 32, OCC = 2; };

// One 128x128 C tile at (m0, n0) of the problem in prm.
template <typename R, bool A_KC, bool B_KC, int BK>
__device__ __forceinline__ void real_tile(const GemmParams<R>& prm, int m0, int n0, R* lds) {
  constexpr int BM = 128, BN = 128;
  constexpr int WB = 4;  // 16x16 blocks per wave along m and along n
  typedef TileLoader<R, false, A_KC, false, BM, BK> LA;
  typedef TileLoader<R, false, B_KC, false, BN, BK> LB;
  typedef typename Acc<R>::type AccT;
  constexpr int STAGE = LA::PLANE + LB::PLANE;

  const int M = prm.m, N = prm.n, K = prm.k;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int li = lane & 15, lg = lane >> 4;

  AccT acc[WB][WB];
#pragma unroll
  for (int i = 0; i < WB; ++i)
#pragma unroll
    for (int j = 0; j < WB; ++j) acc[i][j] = AccT{0, 0, 0, 0};

  LA la;
  LB lb;
  const int nkt = (K + BK - 1) / BK;
  const bool a_rows_full = m0 + BM <= M, b_rows_full = n0 + BN <= N;
  if (nkt > 0) {
    la.load(prm.a, prm.lda, m0, M, 0, K, a_rows_full && BK <= K);
    lb.load(prm.b, prm.ldb, n0, N, 0, K, b_rows_full && BK <= K);
    la.store(lds, nullptr);
    lb.store(lds + LA::PLANE, nullptr);
  }
  __syncthreads();

  for (int kt = 0; kt < nkt; ++kt) {
    const int cur = kt & 1;
    const bool more = kt + 1 < nkt;
    if (more) {
      const bool kfull = (kt + 2) * BK <= K;
      la.load(prm.a, prm.lda, m0, M, (kt + 1) * BK, K, a_rows_full && kfull);
      lb.load(prm.b, prm.ldb, n0, N, (kt + 1) * BK, K, b_rows_full && kfull);
    }
    const R* As = lds + cur * STAGE;
    const R* Bs = As + LA::PLANE;
    constexpr int KV = 16 / (int)sizeof(R);
    constexpr int NH = 4 / KV;  // chunk reads per 4 k values: 2 for double, 1 for float
#pragma unroll
    for (int p = 0; p < BK / 16; ++p) {
#pragma unroll
      for (int h = 0; h < NH; ++h) {
        typename Vec16<R>::type af[WB], bf[WB];
#pragma unroll
        for (int i = 0; i < WB; ++i) af[i] = read_chunk<R, BM, A_KC>(As, p, lg, h, wm * 64 + i * 16 + li);
#pragma unroll
        for (int j = 0; j < WB; ++j) bf[j] = read_chunk<R, BN, B_KC>(Bs, p, lg, h, wn * 64 + j * 16 + li);
#pragma unroll
        for (int s = 0; s < KV; ++s)
#pragma unroll
          for (int i = 0; i < WB; ++i)
#pragma unroll
            for (int j = 0; j < WB; ++j) acc[i][j] = mfma(bf[j][s], af[i][s], acc[i][j]);
      }
    }
    if (more) {
      R* Ln = lds + (cur ^ 1) * STAGE;
      la.store(Ln, nullptr);
      lb.store(Ln + LA::PLANE, nullptr);
    }
    __syncthreads();
  }

  // Epilogue: out = alpha*acc + gamma*C_out + beta*C_in.
  const R alpha = prm.alpha, beta = prm.beta;
  const bool acc_prev = prm.accumulate != 0;
  const bool use_cin = beta != R(0);
#pragma unroll
  for (int i = 0; i < WB; ++i) {
    const int gi = m0 + wm * 64 + i * 16 + li;
    if (gi >= M) continue;
#pragma unroll
    for (int j = 0; j < WB; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gj = n0 + wn * 64 + j * 16 + out_col<R>(lane, r);
        if (gj >= N) continue;
        R v = alpha * acc[i][j][r];
        R* dst = prm.c + (long)gi + (long)gj * prm.ldc;
        if (acc_prev) v += *dst;
        if (use_cin) v += beta * prm.cin[(long)gi + (long)gj * prm.ldcin];
        *dst = v;
      }
    }
  }
}

template <typename R, int BK>
constexpr int real_lds_elems() {
  return 2 * (TileLoader<R, false, true, false, 128, BK>::PLANE + TileLoader<R, false, true, false, 128, BK>::PLANE);
}

// Real GEMM: 256 threads, 2x2 waves, wave tile 64x64 (4x4 MFMA blocks), WG tile 128x128.
template <typename R, bool A_KC, bool B_KC, int BK = RealCfg<R>::BK, int OCC = RealCfg<R>::OCC>
__global__ __launch_bounds__(256, OCC) void gemm_real_kernel(GemmParams<R> prm) {
  __shared__ __attribute__((aligned(16))) R lds[real_lds_elems<R, BK>()];
  if (prm.gate_a && (*prm.gate_a | *prm.gate_b) == 0) return;  // uniform: before any barrier
  const int nbm = (prm.m + 127) / 128, nbn = (prm.n + 127) / 128;
  int bm, bn;
  tile_coords(blockIdx.x, nbm, nbn, bm, bn);
  real_tile<R, A_KC, B_KC, BK>(prm, bm * 128, bn * 128, lds);
}

// Persistent batched GEMM: a fixed grid pulls work items (problem, tile m, tile n) from a
// list in order of decreasing cost through an atomic counter; every workgroup exits when the
// counter passes the end of the list. counter[0] must be 0 at launch.
template <typename R, bool A_KC, bool B_KC>
__global__ __launch_bounds__(256, 2) void gemm_real_batched_kernel(const GemmParams<R>* __restrict__ probs,
                                                                  const int* __restrict__ items, int n_items,
                                                                  int* __restrict__ counter) {
  constexpr int BK = RealCfg<R>::BK;
  __shared__ __attribute__((aligned(16))) R lds[real_lds_elems<R, BK>()];
  __shared__ int next;
  for (;;) {
    if (threadIdx.x == 0) next = atomicAdd(counter, 1);
    __syncthreads();
    const int it = next;
    __syncthreads();  // everyone has read `next` before thread 0 overwrites it
    if (it >= n_items) break;
    const int pidx = items[3 * it], bm = items[3 * it + 1], bn = items[3 * it + 2];
    real_tile<R, A_KC, B_KC, BK>(probs[pidx], bm * 128, bn * 128, lds);
    __syncthreads();  // LDS is reused by the next item
  }
}

__global__ void zero_counter_kernel(int* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) c[0] = 0;
}

// ---------------------------------------------------------------------------------------------
// Complex GEMM: 256 threads, 2x2 waves, wave tile 32x32 complex (2x2 blocks, re and im
// accumulators), WG tile 64x64 complex. Each block update is four real MFMAs.
// ---------------------------------------------------------------------------------------------
template <typename R, bool A_KC, bool B_KC, bool A_CONJ, bool B_CONJ>
__global__ __launch_bounds__(256, 2) void gemm_cplx_kernel(GemmParams<R> prm) {
  constexpr int BM = 64, BN = 64;
  constexpr int BK = sizeof(R) == 8 ? 16 : 32;
  constexpr int WB = 2;
  typedef TileLoader<R, true, A_KC, A_CONJ, BM, BK> LA;
  typedef TileLoader<R, true, B_KC, B_CONJ, BN, BK> LB;
  typedef typename Acc<R>::type AccT;

  // planes: A re, A im, B re, B im
  __shared__ __attribute__((aligned(16))) R lds[2][2 * LA::PLANE + 2 * LB::PLANE];

  const int M = prm.m, N = prm.n, K = prm.k;
  const int nbm = (M + BM - 1) / BM, nbn = (N + BN - 1) / BN;
  int bm, bn;
  tile_coords(blockIdx.x, nbm, nbn, bm, bn);
  const int m0 = bm * BM, n0 = bn * BN;

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int li = lane & 15, lg = lane >> 4;

  AccT cre[WB][WB], cim[WB][WB];
#pragma unroll
  for (int i = 0; i < WB; ++i)
#pragma unroll
    for (int j = 0; j < WB; ++j) { cre[i][j] = AccT{0, 0, 0, 0}; cim[i][j] = AccT{0, 0, 0, 0}; }

  LA la;
  LB lb;
  const int nkt = (K + BK - 1) / BK;
  const bool a_rows_full = m0 + BM <= M, b_rows_full = n0 + BN <= N;
  if (nkt > 0) {
    la.load(prm.a, prm.lda, m0, M, 0, K, a_rows_full && BK <= K);
    lb.load(prm.b, prm.ldb, n0, N, 0, K, b_rows_full && BK <= K);
    R* L = &lds[0][0];
    la.store(L, L + LA::PLANE);
    lb.store(L + 2 * LA::PLANE, L + 2 * LA::PLANE + LB::PLANE);
  }
  __syncthreads();

  for (int kt = 0; kt < nkt; ++kt) {
    const int cur = kt & 1;
    const bool more = kt + 1 < nkt;
    if (more) {
      const bool kfull = (kt + 2) * BK <= K;
      la.load(prm.a, prm.lda, m0, M, (kt + 1) * BK, K, a_rows_full && kfull);
      lb.load(prm.b, prm.ldb, n0, N, (kt + 1) * BK, K, b_rows_full && kfull);
    }
    const R* L = &lds[cur][0];
    constexpr int KV = 16 / (int)sizeof(R);
    constexpr int NH = 4 / KV;
#pragma unroll
    for (int p = 0; p < BK / 16; ++p) {
#pragma unroll
      for (int h = 0; h < NH; ++h) {
        typename Vec16<R>::type ar[WB], ai[WB], br[WB], bi[WB];
#pragma unroll
        for (int i = 0; i < WB; ++i) {
          const int row = wm * 32 + i * 16 + li;
          ar[i] = read_chunk<R, BM, A_KC>(L, p, lg, h, row);
          ai[i] = read_chunk<R, BM, A_KC>(L + LA::PLANE, p, lg, h, row);
        }
#pragma unroll
        for (int j = 0; j < WB; ++j) {
          const int row = wn * 32 + j * 16 + li;
          br[j] = read_chunk<R, BN, B_KC>(L + 2 * LA::PLANE, p, lg, h, row);
          bi[j] = read_chunk<R, BN, B_KC>(L + 2 * LA::PLANE + LB::PLANE, p, lg, h, row);
        }
#pragma unroll
        for (int s = 0; s < KV; ++s)
#pragma unroll
          for (int i = 0; i < WB; ++i) {
            const R nai = -ai[i][s];
#pragma unroll
            for (int j = 0; j < WB; ++j) {
              cre[i][j] = mfma(br[j][s], ar[i][s], cre[i][j]);
              cre[i][j] = mfma(bi[j][s], nai, cre[i][j]);
              cim[i][j] = mfma(bi[j][s], ar[i][s], cim[i][j]);
              cim[i][j] = mfma(br[j][s], ai[i][s], cim[i][j]);
            }
          }
      }
    }
    if (more) {
      R* Ln = &lds[cur ^ 1][0];
      la.store(Ln, Ln + LA::PLANE);
      lb.store(Ln + 2 * LA::PLANE, Ln + 2 * LA::PLANE + LB::PLANE);
    }
    __syncthreads();
  }

  const R alr = prm.alpha, ali = prm.alpha_im;
  const R ber = prm.beta, bei = prm.beta_im;
  const bool acc_prev = prm.accumulate != 0;
  const bool use_cin = ber != R(0) || bei != R(0);
#pragma unroll
  for (int i = 0; i < WB; ++i) {
    const int gi = m0 + wm * 32 + i * 16 + li;
    if (gi >= M) continue;
#pragma unroll
    for (int j = 0; j < WB; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gj = n0 + wn * 32 + j * 16 + out_col<R>(lane, r);
        if (gj >= N) continue;
        const R xr = cre[i][j][r], xi = cim[i][j][r];
        R vr = alr * xr - ali * xi;
        R vi = alr * xi + ali * xr;
        R* dst = prm.c + 2 * ((long)gi + (long)gj * prm.ldc);
        if (acc_prev) { vr += dst[0]; vi += dst[1]; }
        if (use_cin) {
          const R* ci = prm.cin + 2 * ((long)gi + (long)gj * prm.ldcin);
          const R cr = ci[0], cm = ci[1];
          vr += ber * cr - bei * cm;
          vi += ber * cm + bei * cr;
        }
        dst[0] = vr;
        dst[1] = vi;
      }
    }
  }
}


// ---------------------------------------------------------------------------------------------
// FP64 GEMM with direct global->LDS loads (global_load_lds_dwordx4) into an S-stage LDS ring.
// WG tile 128x128, 2x2 waves of 64x64, k-step 8 per stage, S stages of 16 KB (S = 3, three
// workgroups per CU).
// The loads of stage kt+S-1 are issued right after the barrier of stage kt, so S-1 stages are in
// flight while the MFMAs run; no registers or ds_write instructions carry the operands.
//   k-contiguous operand (A for 'T', B for 'N'): LDS [row][slot][2], a 16-byte chunk = 2
//     consecutive k of one row, the 4 chunks of a row XOR-permuted over its 4 slots; a wave
//     instruction fills 16 consecutive rows, reading 64 contiguous bytes of each.
//   row-contiguous operand (A for 'N', B for 'T'): LDS [k][row], a 16-byte chunk = 2 consecutive
//     rows at one k; a wave instruction fills all 128 rows of one k (coalesced in global memory).
// Rows beyond the matrix load a clamped in-bounds row: they only feed outputs that are not
// stored. K must be a multiple of 8 (the launcher hands a remainder to gemm_real_kernel).
// ---------------------------------------------------------------------------------------------
template <bool KC>
struct DldsOperand {
  // k-contiguous stage layout: [row][slot][2], 64 bytes per row, slot s of row r holding the k
  // pair (2q, 2q+1) with q = s ^ kc_swz(r). The fragment read of one k pair by 16 consecutive rows
  // then touches 8 distinct 16-byte bank groups per 8 lanes (rows r, r+1 share a swizzle and sit
  // 64 bytes apart), i.e. it is free of bank conflicts.
  __device__ __forceinline__ static int kc_swz(int r) { return (r >> 1) & 3; }
  // issue this wave's 2 loads of one stage (8 instructions per workgroup). ARENA: the operand lies
  // in a device arena with at least 16 more bytes after its last column (the batched path), so a
  // row-contiguous operand of odd row count may read its last 2-row chunk one row past the end
  // (that row only feeds unstored outputs) instead of needing an even row count.
  template <bool ARENA = false>
  __device__ __forceinline__ static void issue(const double* __restrict__ p, long ld, int r0, int nrows, int k0,
                                               double* stage, int wave, int lane) {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int inst = wave * 2 + t;  // 0..7
      const double* src;
      double* dst;
      if (KC) {
        // 16 rows x 64 bytes per instruction: lanes 4j..4j+3 read the 4 chunks of row j, so the
        // global access is contiguous per row; lane l lands in slot l % 4 of its row, which holds
        // chunk (l % 4) ^ swz(row) (see kc_slot)
        const int local = inst * 16 + (lane >> 2);
        const int q = (lane & 3) ^ kc_swz(local);
        const int row = min(r0 + local, nrows - 1);
        src = p + (long)(2 * q + k0) + (long)row * ld;
        dst = stage + inst * 128;
      } else {
        const int kk = inst;
        const int row = min(r0 + 2 * lane, ARENA ? ((nrows + 1) & ~1) - 2 : nrows - 2);
        src = p + (long)row + (long)(k0 + kk) * ld;
        dst = stage + kk * 128;
      }
      __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                       (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
    }
  }
  // values at k = 2g, 2g+1 of row `row` of the stage
  __device__ __forceinline__ static void frag(const double* stage, int g, int row, double& v0, double& v1) {
    if (KC) {
      const v2d c = *(const v2d*)(stage + row * 8 + (g ^ kc_swz(row)) * 2);
      v0 = c[0];
      v1 = c[1];
    } else {
      v0 = stage[(2 * g) * 128 + row];
      v1 = stage[(2 * g + 1) * 128 + row];
    }
  }
};

// One 128x128 C tile at (m0, n0). ARENA (the batched path): operands in the batch's device arena
// (see DldsOperand::issue), any m, n >= 1, and K need not be a multiple of 8: the last K % 8
// values are added in the epilogue from global memory.
template <bool A_KC, bool B_KC, int S, bool ARENA>
__device__ __forceinline__ void dlds_tile_at(const GemmParams<double>& prm, int m0, int n0) {
  constexpr int STAGE = 2 * 1024;  // A and B, 1024 doubles each
  __shared__ __attribute__((aligned(16))) double lds[S * STAGE];
  const int M = prm.m, N = prm.n, K = prm.k;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int li = lane & 15, lg = lane >> 4;
  const int nkt = K / 8;

  dbl4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = dbl4{0, 0, 0, 0};

  auto issue = [&](int kt) {
    double* st = lds + (kt % S) * STAGE;
    DldsOperand<A_KC>::template issue<ARENA>(prm.a, prm.lda, m0, M, kt * 8, st, wave, lane);
    DldsOperand<B_KC>::template issue<ARENA>(prm.b, prm.ldb, n0, N, kt * 8, st + 1024, wave, lane);
  };
#pragma unroll
  for (int s = 0; s < S - 1; ++s)
    if (s < nkt) issue(s);
  for (int kt = 0; kt < nkt; ++kt) {
    // this wave's loads of stage kt have landed once at most 4 per later stage are pending; its
    // LDS reads of stage kt-1 are done (lgkmcnt 0); the barrier extends both to the workgroup
    const int later = min(S - 2, nkt - 1 - kt);
    if (later >= 3) __builtin_amdgcn_s_waitcnt(0x070 | 12);
    else if (later == 2) __builtin_amdgcn_s_waitcnt(0x070 | 8);
    else if (later == 1) __builtin_amdgcn_s_waitcnt(0x070 | 4);
    else __builtin_amdgcn_s_waitcnt(0x070);
    __builtin_amdgcn_s_barrier();
    if (kt + S - 1 < nkt) issue(kt + S - 1);  // into the slot of stage kt-1, free after the barrier
    const double* st = lds + (kt % S) * STAGE;
    double a0[4], a1[4], b0[4], b1[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) DldsOperand<A_KC>::frag(st, lg, wm * 64 + i * 16 + li, a0[i], a1[i]);
#pragma unroll
    for (int j = 0; j < 4; ++j) DldsOperand<B_KC>::frag(st + 1024, lg, wn * 64 + j * 16 + li, b0[j], b1[j]);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = mfma(b0[j], a0[i], acc[i][j]);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = mfma(b1[j], a1[i], acc[i][j]);
  }

  const double alpha = prm.alpha, beta = prm.beta;
  const bool acc_prev = prm.accumulate != 0, use_cin = beta != 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int gi = m0 + wm * 64 + i * 16 + li;
    if (gi >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gj = n0 + wn * 64 + j * 16 + out_col<double>(lane, r);
        if (gj >= N) continue;
        double sum = acc[i][j][r];
        if constexpr (ARENA) {
          for (int kk = nkt * 8; kk < K; ++kk) {
            const double av = A_KC ? prm.a[(long)kk + (long)gi * prm.lda] : prm.a[(long)gi + (long)kk * prm.lda];
            const double bv = B_KC ? prm.b[(long)kk + (long)gj * prm.ldb] : prm.b[(long)gj + (long)kk * prm.ldb];
            sum = fma(av, bv, sum);
          }
        }
        double v = alpha * sum;
        double* dst = prm.c + (long)gi + (long)gj * prm.ldc;
        if (acc_prev) v += *dst;
        if (use_cin) v += beta * prm.cin[(long)gi + (long)gj * prm.ldcin];
        *dst = v;
      }
    }
  }
}

template <bool A_KC, bool B_KC, int S>
__device__ __forceinline__ void dlds_tile(const GemmParams<double>& prm, int block) {
  const int nbm = (prm.m + 127) / 128, nbn = (prm.n + 127) / 128;
  int bm, bn;
  tile_coords(block, nbm, nbn, bm, bn);
  dlds_tile_at<A_KC, B_KC, S, false>(prm, bm * 128, bn * 128);
}

template <bool A_KC, bool B_KC, int S, int OCC = 2>
__global__ __launch_bounds__(256, OCC) void gemm_d_dlds_kernel(GemmParams<double> prm) {
  if (prm.gate_a && (*prm.gate_a | *prm.gate_b) == 0) return;  // uniform: before any barrier
  dlds_tile<A_KC, B_KC, S>(prm, blockIdx.x);
}

// ---------------------------------------------------------------------------------------------
// FP32 GEMM with direct global->LDS loads: the FP64 kernel's ring with 16-byte chunks of 4 floats,
// so a stage holds k-step 16 (64 bytes per operand row), 3 stages; workgroup tiles of 256 x 256
// (8 waves of 64 x 128, 32 KB per stage, one workgroup per CU), 256 x 128 (8 waves of 64 x 64, 24
// KB per stage, two per CU) or 128 x 128 (4 waves, 16 KB per stage, three per CU). A lane's
// 16-byte fragment holds k = 4g..4g+3 of its row (g = lane / 16) and
// feeds four v_mfma_f32_16x16x4f32, the same k permutation for A and B.
//   k-contiguous operand: LDS [row][slot][4], chunk q (k = 4q..4q+3) of row r in slot q ^ swz(r),
//     a wave instruction fills 16 rows, 64 contiguous bytes of each (as in the FP64 kernel).
//   row-contiguous operand: LDS [k][row], a 16-byte chunk = 4 consecutive rows at one k; a wave
//     instruction fills 256 rows at one k, or 128 rows at two k (lanes 0-31 the first, 32-63 the
//     second). Needs a row count that is a multiple of 4 (the launcher sends other shapes to
//     gemm_real_kernel).
// Rows beyond the matrix read a clamped in-bounds row (they feed only unstored outputs). K must be
// a multiple of 16 (the launcher hands a remainder to gemm_real_kernel).
// ---------------------------------------------------------------------------------------------
template <bool KC, int ROWS, int NW>
struct SdldsOperand {
  static_assert(ROWS == 128 || ROWS == 256, "operand tile of 128 or 256 rows");
  static constexpr int NINST = ROWS / 16;  // wave instructions per stage
  static_assert(NINST % NW == 0, "the stage's instructions spread evenly over the waves");
  __device__ __forceinline__ static int kc_swz(int r) { return (r >> 1) & 3; }
  // this wave's loads of one stage
  __device__ __forceinline__ static void issue(const float* __restrict__ p, long ld, int r0, int nrows, int k0,
                                               float* stage, int wave, int lane) {
#pragma unroll
    for (int t = 0; t < NINST / NW; ++t) {
      const int inst = wave * (NINST / NW) + t;
      const float* src;
      if (KC) {  // 16 rows x 64 bytes
        const int local = inst * 16 + (lane >> 2);
        const int q = (lane & 3) ^ kc_swz(local);
        const int row = min(r0 + local, nrows - 1);
        src = p + (long)(4 * q + k0) + (long)row * ld;
      } else if (ROWS == 128) {  // 128 rows at two k
        const int kk = 2 * inst + (lane >> 5);
        const int row = min(r0 + 4 * (lane & 31), nrows - 4);
        src = p + (long)row + (long)(k0 + kk) * ld;
      } else {  // 256 rows at one k
        const int row = min(r0 + 4 * lane, nrows - 4);
        src = p + (long)row + (long)(k0 + inst) * ld;
      }
      __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                       (__attribute__((address_space(3))) void*)(stage + inst * 256), 16, 0, 0);
    }
  }
  // values at k = 4g..4g+3 of row `row` of the stage
  __device__ __forceinline__ static v4f frag(const float* stage, int g, int row) {
    if (KC) return *(const v4f*)(stage + row * 16 + (g ^ kc_swz(row)) * 4);
    v4f v;
#pragma unroll
    for (int t = 0; t < 4; ++t) v[t] = stage[(4 * g + t) * ROWS + row];
    return v;
  }
};

// One BM x BN C tile at (m0, n0): WM x 2 waves of 64 x (BN / 2).
template <bool A_KC, bool B_KC, int S, int BM, int BN = 128>
__device__ __forceinline__ void sdlds_tile_at(const GemmParams<float>& prm, int m0, int n0) {
  constexpr int WM = BM / 64, NW = 2 * WM, TN = BN / 2, TJ = TN / 16;
  constexpr int A_ELEMS = BM * 16, STAGE = A_ELEMS + BN * 16;
  constexpr int PER = (BM / 16 + BN / 16) / NW;  // load instructions per wave and stage
  using OpA = SdldsOperand<A_KC, BM, NW>;
  using OpB = SdldsOperand<B_KC, BN, NW>;
  __shared__ __attribute__((aligned(16))) float lds[S * STAGE];
  const int M = prm.m, N = prm.n, K = prm.k;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int li = lane & 15, lg = lane >> 4;
  const int nkt = K / 16;

  flt4 acc[4][TJ];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = flt4{0, 0, 0, 0};

  auto issue = [&](int kt) {
    float* st = lds + (kt % S) * STAGE;
    OpA::issue(prm.a, prm.lda, m0, M, kt * 16, st, wave, lane);
    OpB::issue(prm.b, prm.ldb, n0, N, kt * 16, st + A_ELEMS, wave, lane);
  };
#pragma unroll
  for (int s = 0; s < S - 1; ++s)
    if (s < nkt) issue(s);
  for (int kt = 0; kt < nkt; ++kt) {
    // this wave's loads of stage kt have landed once at most PER per later stage are pending; its
    // LDS reads of stage kt-1 are done (lgkmcnt 0); the barrier extends both to the workgroup
    const int later = min(S - 2, nkt - 1 - kt);
    if (later >= 2) __builtin_amdgcn_s_waitcnt(0x070 | (2 * PER));
    else if (later == 1) __builtin_amdgcn_s_waitcnt(0x070 | PER);
    else __builtin_amdgcn_s_waitcnt(0x070);
    __builtin_amdgcn_s_barrier();
    if (kt + S - 1 < nkt) issue(kt + S - 1);  // into the slot of stage kt-1, free after the barrier
    const float* st = lds + (kt % S) * STAGE;
    v4f a[4], b[TJ];
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = OpA::frag(st, lg, wm * 64 + i * 16 + li);
#pragma unroll
    for (int j = 0; j < TJ; ++j) b[j] = OpB::frag(st + A_ELEMS, lg, wn * TN + j * 16 + li);
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) acc[i][j] = mfma(b[j][t], a[i][t], acc[i][j]);
  }

  const float alpha = prm.alpha, beta = prm.beta;
  const bool acc_prev = prm.accumulate != 0, use_cin = beta != 0.0f;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int gi = m0 + wm * 64 + i * 16 + li;
    if (gi >= M) continue;
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gj = n0 + wn * TN + j * 16 + out_col<float>(lane, r);
        if (gj >= N) continue;
        float v = alpha * acc[i][j][r];
        float* dst = prm.c + (long)gi + (long)gj * prm.ldc;
        if (acc_prev) v += *dst;
        if (use_cin) v += beta * prm.cin[(long)gi + (long)gj * prm.ldcin];
        *dst = v;
      }
    }
  }
}

// The FP32 tile software-pipelined as dbig_tile_at (FP64): the fragments of stage kt+1 are read
// into a second register set while the 16 x TJ MFMAs of stage kt run, the loads of stage kt+3 go
// between its first MFMAs, scheduling-group barriers pin that order, and the loads always run
// (clamped to the last stage) so that every wait count is a constant. Per output block the MFMA
// sequence is sdlds_tile_at's: the results are the same bits.
template <bool A_KC, bool B_KC, int BM, int BN>
__device__ __forceinline__ void spipe_tile_at(const GemmParams<float>& prm, int m0, int n0) {
  constexpr int S = 3;
  constexpr int WM = BM / 64, NW = 2 * WM, TN = BN / 2, TJ = TN / 16;
  constexpr int A_ELEMS = BM * 16, STAGE = A_ELEMS + BN * 16;
  constexpr int PER = (BM / 16 + BN / 16) / NW;  // load instructions per wave and stage
  constexpr int MF = 16 * TJ;                    // MFMAs per wave and stage
  constexpr int NDS = 4 * (A_KC ? 1 : 4) + TJ * (B_KC ? 1 : 4);  // LDS reads per wave and stage
  constexpr int DSG = PER + 2 * NDS <= MF ? 2 : 1;
  constexpr int REST = MF - PER - DSG * NDS;
  static_assert(REST >= 0, "stage schedule exceeds the stage's MFMAs");
  using OpA = SdldsOperand<A_KC, BM, NW>;
  using OpB = SdldsOperand<B_KC, BN, NW>;
  __shared__ __attribute__((aligned(16))) float lds[S * STAGE];
  const int M = prm.m, N = prm.n, K = prm.k;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int li = lane & 15, lg = lane >> 4;
  const int nkt = K / 16;

  flt4 acc[4][TJ];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = flt4{0, 0, 0, 0};
  auto issue = [&](int kt, int so) {
    OpA::issue(prm.a, prm.lda, m0, M, kt * 16, lds + so, wave, lane);
    OpB::issue(prm.b, prm.ldb, n0, N, kt * 16, lds + so + A_ELEMS, wave, lane);
  };
  struct Frags {
    v4f a[4], b[TJ];
  };
  auto read = [&](Frags& f, int so) {
    const float* st = lds + so;
#pragma unroll
    for (int i = 0; i < 4; ++i) f.a[i] = OpA::frag(st, lg, wm * 64 + i * 16 + li);
#pragma unroll
    for (int j = 0; j < TJ; ++j) f.b[j] = OpB::frag(st + A_ELEMS, lg, wn * TN + j * 16 + li);
  };
  if (nkt > 0) {
#pragma unroll
    for (int s = 0; s < S; ++s) issue(min(s, nkt - 1), s * STAGE);
    __builtin_amdgcn_s_waitcnt(0x070 | ((S - 1) * PER));
    __builtin_amdgcn_s_barrier();
    Frags F0, F1;
    read(F0, 0);
    int so = 0;
    auto step = [&](Frags& cur, Frags& nxt, int kt) {
      __builtin_amdgcn_s_waitcnt(0x070 | ((S - 2) * PER));
      __builtin_amdgcn_s_barrier();
      __builtin_amdgcn_sched_barrier(0);
      issue(min(kt + S, nkt - 1), so);
      const int sn = so + STAGE == S * STAGE ? 0 : so + STAGE;
      read(nxt, sn);
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < TJ; ++j) acc[i][j] = mfma(cur.b[j][t], cur.a[i][t], acc[i][j]);
#pragma unroll
      for (int r = 0; r < PER; ++r) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
        __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // VMEM read (LDS-DMA load)
      }
#pragma unroll
      for (int r = 0; r < NDS; ++r) {
        __builtin_amdgcn_sched_group_barrier(0x008, DSG, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // DS read
      }
      __builtin_amdgcn_sched_group_barrier(0x008, REST, 0);
      __builtin_amdgcn_sched_barrier(0);
      so = sn;
    };
    int kt = 0;
    for (; kt + 1 < nkt; kt += 2) {
      step(F0, F1, kt);
      step(F1, F0, kt + 1);
    }
    if (kt < nkt) step(F0, F1, kt);
    // the last stages' loads (clamped) land before the workgroup's LDS is released
    __builtin_amdgcn_s_waitcnt(0x070);
  }

  const float alpha = prm.alpha, beta = prm.beta;
  const bool acc_prev = prm.accumulate != 0, use_cin = beta != 0.0f;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int gi = m0 + wm * 64 + i * 16 + li;
    if (gi >= M) continue;
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gj = n0 + wn * TN + j * 16 + out_col<float>(lane, r);
        if (gj >= N) continue;
        float v = alpha * acc[i][j][r];
        float* dst = prm.c + (long)gi + (long)gj * prm.ldc;
        if (acc_prev) v += *dst;
        if (use_cin) v += beta * prm.cin[(long)gi + (long)gj * prm.ldcin];
        *dst = v;
      }
    }
  }
}

// 128 x 128 tiles, two workgroups per CU
template <bool A_KC, bool B_KC>
__global__ __launch_bounds__(256, 2) void gemm_s_pipe_kernel(GemmParams<float> prm) {
  const int nbm = (prm.m + 127) / 128, nbn = (prm.n + 127) / 128;
  int bm, bn;
  tile_coords(blockIdx.x, nbm, nbn, bm, bn);
  spipe_tile_at<A_KC, B_KC, 128, 128>(prm, bm * 128, bn * 128);
}

template <bool A_KC, bool B_KC, int S, int OCC, int BM, int BN = 128>
__global__ __launch_bounds__(BM * 2, OCC) void gemm_s_dlds_kernel(GemmParams<float> prm) {
  const int nbm = (prm.m + BM - 1) / BM, nbn = (prm.n + BN - 1) / BN;
  int bm, bn;
  tile_coords(blockIdx.x, nbm, nbn, bm, bn);
  sdlds_tile_at<A_KC, B_KC, S, BM, BN>(prm, bm * BM, bn * BN);
}

// Persistent batched FP64 GEMM on the direct-to-LDS tile body: the same item protocol as
// gemm_real_batched_kernel, operands in the batch's device arena (dlds_tile_at with ARENA: any
// m, n >= 1 and any k), three workgroups per CU like gemm_d_dlds_kernel.
template <bool A_KC, bool B_KC, int S>
__global__ __launch_bounds__(256, 3) void gemm_d_dlds_batched_kernel(const GemmParams<double>* __restrict__ probs,
                                                                    const int* __restrict__ items, int n_items,
                                                                    int* __restrict__ counter) {
  __shared__ int next;
  for (;;) {
    if (threadIdx.x == 0) next = atomicAdd(counter, 1);
    __syncthreads();
    const int it = next;
    __syncthreads();  // everyone has read `next` before thread 0 overwrites it
    if (it >= n_items) break;
    const GemmParams<double> prm = probs[items[3 * it]];
    dlds_tile_at<A_KC, B_KC, S, true>(prm, items[3 * it + 1] * 128, items[3 * it + 2] * 128);
    __syncthreads();  // every wave is done with the LDS ring before the next item refills it
  }
}

// Several independent FP64 products of one shape class (same ops, alpha, k) in one launch: the
// grid is the concatenation of the products' workgroup ranges, each padded to a multiple of 8 so
// that a workgroup's XCD (blockIdx.x % 8) is the one tile_coords assumes. Padding workgroups exit
// before any barrier.
template <bool A_KC, bool B_KC, int S, int OCC = 2>
__global__ __launch_bounds__(256, OCC) void gemm_d_dlds_group_kernel(DgemmGroup g) {
  const int w = blockIdx.x;
  int t = 0;
  while (t + 1 < g.count && w >= g.wg_start[t + 1]) ++t;
  GemmParams<double> prm = g.base;
  prm.m = g.m[t];
  prm.n = g.n[t];
  prm.a = g.a[t];
  prm.lda = g.lda[t];
  prm.b = g.b[t];
  prm.ldb = g.ldb[t];
  prm.c = g.c[t];
  prm.ldc = g.ldc[t];
  prm.cin = g.cin[t];
  prm.ldcin = g.ldcin[t];
  prm.beta = g.beta[t];
  prm.accumulate = g.accumulate[t];
  const int local = w - g.wg_start[t];
  if (local >= ((prm.m + 127) / 128) * ((prm.n + 127) / 128)) return;
  dlds_tile<A_KC, B_KC, S>(prm, local);
}

// ---------------------------------------------------------------------------------------------
// Complex FP32 GEMM with direct global->LDS loads (CGEMM): the FP32 ring with 16-byte chunks of
// two complex floats, k-step 8 complex per stage (64 bytes per operand row, 16 KB per stage, 3
// stages). WG tile 128x128 complex, 2x2 waves of 64x64 (4x4 blocks, re and im accumulators, four
// real v_mfma_f32_16x16x4f32 per block and k). A lane's fragment holds k = 2g, 2g+1 of its row
// (g = lane / 16), the same k permutation for A and B.
//   k-contiguous operand (A for 'T'/'C', B for 'N'): LDS [row][slot], chunk q (k = 2q, 2q+1) of
//     row r in slot q ^ swz(r); a wave instruction fills 16 rows, 64 contiguous bytes of each.
//   row-contiguous operand (A for 'N', B for 'T'/'C'): LDS [k][row], a 16-byte chunk = 2
//     consecutive rows at one k; a wave instruction fills the 128 rows of one k. Needs an even row
//     count (the launcher sends other shapes to gemm_cplx_kernel).
// Conjugation ('C') negates the imaginary fragment in registers. Rows beyond the matrix read a
// clamped in-bounds row. K must be a multiple of 8 (the launcher hands a remainder to
// gemm_cplx_kernel).
// ---------------------------------------------------------------------------------------------
template <bool KC>
struct CdldsOperand {
  __device__ __forceinline__ static int kc_swz(int r) { return (r >> 1) & 3; }
  // this wave's 2 loads of one stage (8 instructions of 1 KB per workgroup and operand)
  __device__ __forceinline__ static void issue(const float* __restrict__ p, long ld, int r0, int nrows, int k0,
                                               float* stage, int wave, int lane) {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int inst = wave * 2 + t;  // 0..7
      const float* src;
      if (KC) {
        const int local = inst * 16 + (lane >> 2);
        const int q = (lane & 3) ^ kc_swz(local);
        const int row = min(r0 + local, nrows - 1);
        src = p + 2 * ((long)(2 * q + k0) + (long)row * ld);
      } else {
        const int row = min(r0 + 2 * lane, nrows - 2);
        src = p + 2 * ((long)row + (long)(k0 + inst) * ld);
      }
      __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                       (__attribute__((address_space(3))) void*)(stage + inst * 256), 16, 0, 0);
    }
  }
  // (re, im) of k = 2g and of k = 2g + 1 of row `row` of the stage
  __device__ __forceinline__ static v4f frag(const float* stage, int g, int row) {
    if (KC) return *(const v4f*)(stage + row * 16 + (g ^ kc_swz(row)) * 4);
    v4f v;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const float* e = stage + 2 * ((2 * g + t) * 128 + row);
      v[2 * t] = e[0];
      v[2 * t + 1] = e[1];
    }
    return v;
  }
};

template <bool A_KC, bool B_KC, bool A_CONJ, bool B_CONJ, int S>
__device__ __forceinline__ void cdlds_tile(const GemmParams<float>& prm, int block) {
  constexpr int TILE = 128 * 16;  // floats per operand and stage (8 KB)
  constexpr int STAGE = 2 * TILE;
  __shared__ __attribute__((aligned(16))) float lds[S * STAGE];
  const int M = prm.m, N = prm.n, K = prm.k;
  const int nbm = (M + 127) / 128, nbn = (N + 127) / 128;
  int bm, bn;
  tile_coords(block, nbm, nbn, bm, bn);
  const int m0 = bm * 128, n0 = bn * 128;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int li = lane & 15, lg = lane >> 4;
  const int nkt = K / 8;

  flt4 cre[4][4], cim[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      cre[i][j] = flt4{0, 0, 0, 0};
      cim[i][j] = flt4{0, 0, 0, 0};
    }

  auto issue = [&](int kt) {
    float* st = lds + (kt % S) * STAGE;
    CdldsOperand<A_KC>::issue(prm.a, prm.lda, m0, M, kt * 8, st, wave, lane);
    CdldsOperand<B_KC>::issue(prm.b, prm.ldb, n0, N, kt * 8, st + TILE, wave, lane);
  };
#pragma unroll
  for (int s = 0; s < S - 1; ++s)
    if (s < nkt) issue(s);
  for (int kt = 0; kt < nkt; ++kt) {
    // this wave's loads of stage kt have landed once at most 4 per later stage are pending, and
    // its LDS reads of stage kt-1 are done; the barrier extends both to the workgroup
    const int later = min(S - 2, nkt - 1 - kt);
    if (later >= 2) __builtin_amdgcn_s_waitcnt(0x070 | 8);
    else if (later == 1) __builtin_amdgcn_s_waitcnt(0x070 | 4);
    else __builtin_amdgcn_s_waitcnt(0x070);
    __builtin_amdgcn_s_barrier();
    if (kt + S - 1 < nkt) issue(kt + S - 1);
    const float* st = lds + (kt % S) * STAGE;
    v4f a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      a[i] = CdldsOperand<A_KC>::frag(st, lg, wm * 64 + i * 16 + li);
      if (A_CONJ) {
        a[i][1] = -a[i][1];
        a[i][3] = -a[i][3];
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      b[j] = CdldsOperand<B_KC>::frag(st + TILE, lg, wn * 64 + j * 16 + li);
      if (B_CONJ) {
        b[j][1] = -b[j][1];
        b[j][3] = -b[j][3];
      }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float ar = a[i][2 * t], ai = a[i][2 * t + 1], nai = -ai;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const float br = b[j][2 * t], bi = b[j][2 * t + 1];
          cre[i][j] = mfma(br, ar, cre[i][j]);
          cre[i][j] = mfma(bi, nai, cre[i][j]);
          cim[i][j] = mfma(bi, ar, cim[i][j]);
          cim[i][j] = mfma(br, ai, cim[i][j]);
        }
      }
  }

  const float alr = prm.alpha, ali = prm.alpha_im;
  const float ber = prm.beta, bei = prm.beta_im;
  const bool acc_prev = prm.accumulate != 0;
  const bool use_cin = ber != 0.0f || bei != 0.0f;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int gi = m0 + wm * 64 + i * 16 + li;
    if (gi >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gj = n0 + wn * 64 + j * 16 + out_col<float>(lane, r);
        if (gj >= N) continue;
        const float xr = cre[i][j][r], xi = cim[i][j][r];
        float vr = alr * xr - ali * xi;
        float vi = alr * xi + ali * xr;
        float* dst = prm.c + 2 * ((long)gi + (long)gj * prm.ldc);
        if (acc_prev) {
          vr += dst[0];
          vi += dst[1];
        }
        if (use_cin) {
          const float* ci = prm.cin + 2 * ((long)gi + (long)gj * prm.ldcin);
          const float cr = ci[0], cm = ci[1];
          vr += ber * cr - bei * cm;
          vi += ber * cm + bei * cr;
        }
        dst[0] = vr;
        dst[1] = vi;
      }
    }
  }
}

// The same tile software-pipelined as spipe_tile_at (two workgroups per CU): the fragments of
// stage kt+1 are read into a second register set (for a row-contiguous B, into the registers of
// each B fragment as it is used up) while the 128 MFMAs of stage kt run, the loads of
// stage kt+3 go between its first MFMAs, scheduling-group barriers pin that order, and the loads
// always run (clamped to the last stage) so that every wait count is a constant. Per output block
// the MFMA sequence is cdlds_tile's: the results are the same bits.
template <bool A_KC, bool B_KC, bool A_CONJ, bool B_CONJ>
__device__ __forceinline__ void cpipe_tile(const GemmParams<float>& prm, int block) {
  constexpr int TILE = 128 * 16, STAGE = 2 * TILE, S = 3;
  constexpr int LPS = 4;                                          // loads per wave and stage
  constexpr int NDS = 4 * (A_KC ? 1 : 2) + 4 * (B_KC ? 1 : 2);    // LDS reads per wave and stage
  constexpr int REST = 128 - LPS - 2 * NDS;
  __shared__ __attribute__((aligned(16))) float lds[S * STAGE];
  const int M = prm.m, N = prm.n, K = prm.k;
  const int nbm = (M + 127) / 128, nbn = (N + 127) / 128;
  int bm, bn;
  tile_coords(block, nbm, nbn, bm, bn);
  const int m0 = bm * 128, n0 = bn * 128;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int li = lane & 15, lg = lane >> 4;
  const int nkt = K / 8;

  flt4 cre[4][4], cim[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      cre[i][j] = flt4{0, 0, 0, 0};
      cim[i][j] = flt4{0, 0, 0, 0};
    }
  auto issue = [&](int kt, int so) {
    CdldsOperand<A_KC>::issue(prm.a, prm.lda, m0, M, kt * 8, lds + so, wave, lane);
    CdldsOperand<B_KC>::issue(prm.b, prm.ldb, n0, N, kt * 8, lds + so + TILE, wave, lane);
  };
  struct Frags {
    v4f a[4], b[4];
  };
  auto read = [&](Frags& f, int so) {
    const float* st = lds + so;
#pragma unroll
    for (int i = 0; i < 4; ++i) f.a[i] = CdldsOperand<A_KC>::frag(st, lg, wm * 64 + i * 16 + li);
#pragma unroll
    for (int j = 0; j < 4; ++j) f.b[j] = CdldsOperand<B_KC>::frag(st + TILE, lg, wn * 64 + j * 16 + li);
  };
  if constexpr (!B_KC) {
    // A row-contiguous B: two fragment sets of A but one of B, each B fragment read for stage kt+1
    // into its registers as soon as the 32 MFMAs of stage kt that use it are issued (two full sets
    // and 128 accumulators would exceed the register file). Per output block the MFMA sequence is
    // unchanged.
    constexpr int NDA = 4 * (A_KC ? 1 : 2);  // LDS reads of the A fragments per wave and stage
    struct AFrags {
      v4f a[4];
    };
    v4f b[4];
    auto read_a = [&](AFrags& f, int so) {
#pragma unroll
      for (int i = 0; i < 4; ++i) f.a[i] = CdldsOperand<A_KC>::frag(lds + so, lg, wm * 64 + i * 16 + li);
    };
    if (nkt > 0) {
#pragma unroll
      for (int s = 0; s < S; ++s) issue(min(s, nkt - 1), s * STAGE);
      __builtin_amdgcn_s_waitcnt(0x070 | ((S - 1) * LPS));
      __builtin_amdgcn_s_barrier();
      AFrags F0, F1;
      read_a(F0, 0);
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = CdldsOperand<false>::frag(lds + TILE, lg, wn * 64 + j * 16 + li);
      int so = 0;
      auto step = [&](AFrags& cur, AFrags& nxt, int kt) {
        // the waves' B reads of stage kt+1 issued last step are done (lgkmcnt 0) before the
        // barrier, which releases stage kt's buffer to the loads of stage kt+3
        __builtin_amdgcn_s_waitcnt(0x070 | ((S - 2) * LPS));
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_sched_barrier(0);
        issue(min(kt + S, nkt - 1), so);
        const int sn = so + STAGE == S * STAGE ? 0 : so + STAGE;
        read_a(nxt, sn);
        v4f a[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          a[i] = cur.a[i];
          if (A_CONJ) {
            a[i][1] = -a[i][1];
            a[i][3] = -a[i][3];
          }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          v4f bj = b[j];
          if (B_CONJ) {
            bj[1] = -bj[1];
            bj[3] = -bj[3];
          }
#pragma unroll
          for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const float ar = a[i][2 * t], ai = a[i][2 * t + 1], nai = -ai;
              const float br = bj[2 * t], bi = bj[2 * t + 1];
              cre[i][j] = mfma(br, ar, cre[i][j]);
              cre[i][j] = mfma(bi, nai, cre[i][j]);
              cim[i][j] = mfma(bi, ar, cim[i][j]);
              cim[i][j] = mfma(br, ai, cim[i][j]);
            }
          b[j] = CdldsOperand<false>::frag(lds + sn + TILE, lg, wn * 64 + j * 16 + li);
        }
        // 32 MFMAs per B fragment; the loads and the A reads go among the first 32, the two reads
        // of B fragment j after the 32 that use it
#pragma unroll
        for (int r = 0; r < LPS; ++r) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
          __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // VMEM read (LDS-DMA load)
        }
#pragma unroll
        for (int r = 0; r < NDA; ++r) {
          __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
          __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // DS read
        }
        __builtin_amdgcn_sched_group_barrier(0x008, 32 - LPS - 2 * NDA, 0);
#pragma unroll
        for (int j = 1; j < 4; ++j) {
          __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
          __builtin_amdgcn_sched_group_barrier(0x008, 32, 0);
        }
        __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
        __builtin_amdgcn_sched_barrier(0);
        so = sn;
      };
      int kt = 0;
      for (; kt + 1 < nkt; kt += 2) {
        step(F0, F1, kt);
        step(F1, F0, kt + 1);
      }
      if (kt < nkt) step(F0, F1, kt);
      __builtin_amdgcn_s_waitcnt(0x070);
    }
  } else if (nkt > 0) {
#pragma unroll
    for (int s = 0; s < S; ++s) issue(min(s, nkt - 1), s * STAGE);
    __builtin_amdgcn_s_waitcnt(0x070 | ((S - 1) * LPS));
    __builtin_amdgcn_s_barrier();
    Frags F0, F1;
    read(F0, 0);
    int so = 0;
    auto step = [&](Frags& cur, Frags& nxt, int kt) {
      __builtin_amdgcn_s_waitcnt(0x070 | ((S - 2) * LPS));
      __builtin_amdgcn_s_barrier();
      __builtin_amdgcn_sched_barrier(0);
      issue(min(kt + S, nkt - 1), so);
      const int sn = so + STAGE == S * STAGE ? 0 : so + STAGE;
      read(nxt, sn);
      v4f a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        a[i] = cur.a[i];
        if (A_CONJ) {
          a[i][1] = -a[i][1];
          a[i][3] = -a[i][3];
        }
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        b[j] = cur.b[j];
        if (B_CONJ) {
          b[j][1] = -b[j][1];
          b[j][3] = -b[j][3];
        }
      }
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float ar = a[i][2 * t], ai = a[i][2 * t + 1], nai = -ai;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const float br = b[j][2 * t], bi = b[j][2 * t + 1];
            cre[i][j] = mfma(br, ar, cre[i][j]);
            cre[i][j] = mfma(bi, nai, cre[i][j]);
            cim[i][j] = mfma(bi, ar, cim[i][j]);
            cim[i][j] = mfma(br, ai, cim[i][j]);
          }
        }
#pragma unroll
      for (int r = 0; r < LPS; ++r) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
        __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // VMEM read (LDS-DMA load)
      }
#pragma unroll
      for (int r = 0; r < NDS; ++r) {
        __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // DS read
      }
      __builtin_amdgcn_sched_group_barrier(0x008, REST, 0);
      __builtin_amdgcn_sched_barrier(0);
      so = sn;
    };
    int kt = 0;
    for (; kt + 1 < nkt; kt += 2) {
      step(F0, F1, kt);
      step(F1, F0, kt + 1);
    }
    if (kt < nkt) step(F0, F1, kt);
    // the last stages' loads (clamped) land before the workgroup's LDS is released
    __builtin_amdgcn_s_waitcnt(0x070);
  }

  const float alr = prm.alpha, ali = prm.alpha_im;
  const float ber = prm.beta, bei = prm.beta_im;
  const bool acc_prev = prm.accumulate != 0;
  const bool use_cin = ber != 0.0f || bei != 0.0f;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int gi = m0 + wm * 64 + i * 16 + li;
    if (gi >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gj = n0 + wn * 64 + j * 16 + out_col<float>(lane, r);
        if (gj >= N) continue;
        const float xr = cre[i][j][r], xi = cim[i][j][r];
        float vr = alr * xr - ali * xi;
        float vi = alr * xi + ali * xr;
        float* dst = prm.c + 2 * ((long)gi + (long)gj * prm.ldc);
        if (acc_prev) {
          vr += dst[0];
          vi += dst[1];
        }
        if (use_cin) {
          const float* ci = prm.cin + 2 * ((long)gi + (long)gj * prm.ldcin);
          const float cr = ci[0], cm = ci[1];
          vr += ber * cr - bei * cm;
          vi += ber * cm + bei * cr;
        }
        dst[0] = vr;
        dst[1] = vi;
      }
    }
  }
}

template <bool A_KC, bool B_KC, bool A_CONJ, bool B_CONJ>
__global__ __launch_bounds__(256, 2) void gemm_c_pipe_kernel(GemmParams<float> prm) {
  cpipe_tile<A_KC, B_KC, A_CONJ, B_CONJ>(prm, blockIdx.x);
}

template <bool A_KC, bool B_KC, bool A_CONJ, bool B_CONJ, int S, int OCC>
__global__ __launch_bounds__(256, OCC) void gemm_c_dlds_kernel(GemmParams<float> prm) {
  cdlds_tile<A_KC, B_KC, A_CONJ, B_CONJ, S>(prm, blockIdx.x);
}

// The product of a grouped launch that workgroup w works on, with its parameters; -1 for a padding
// workgroup. Products start on multiples of 8 workgroups, so that a workgroup's XCD
// (blockIdx.x % 8) is the one tile_coords assumes.
template <typename R>
__device__ __forceinline__ int group_member(const TileGroup<R>& g, int w, int tiles_m, int tiles_n,
                                            GemmParams<R>& prm, int& local) {
  int t = 0;
  while (t + 1 < g.count && w >= g.wg_start[t + 1]) ++t;
  prm = g.base;
  prm.m = g.m[t];
  prm.n = g.n[t];
  prm.a = g.a[t];
  prm.lda = g.lda[t];
  prm.b = g.b[t];
  prm.ldb = g.ldb[t];
  prm.c = g.c[t];
  prm.ldc = g.ldc[t];
  prm.cin = g.cin[t];
  prm.ldcin = g.ldcin[t];
  prm.beta = g.beta[t];
  prm.beta_im = g.beta_im[t];
  prm.accumulate = g.accumulate[t];
  local = w - g.wg_start[t];
  const int nbm = (prm.m + tiles_m - 1) / tiles_m, nbn = (prm.n + tiles_n - 1) / tiles_n;
  return local < nbm * nbn ? t : -1;
}

// ---------------------------------------------------------------------------------------------
// Software-pipelined FP64 GEMM, two tile widths BN: 128 x 256 workgroup tiles (2 x 2 waves of
// 64 x 128, one workgroup per CU, so each wave has a SIMD to itself and its 32 accumulator blocks,
// 256 registers, live in AGPRs) for the largest products, and 128 x 128 tiles (waves of 64 x 64,
// two workgroups per CU, accumulators in VGPRs) for the others. The stage layouts are those of
// gemm_d_dlds_kernel (k-step 8, a 3-stage direct-to-LDS ring; (128 + BN) x 64 bytes per stage).
// With one or two waves per SIMD little hides a stall, so the loop is pipelined by hand: the
// fragments of stage kt+1 are read into a second register set while the 2 x 4 x BN/32 MFMAs of
// stage kt run, the loads of stage kt+3 go between the first MFMAs, and scheduling-group barriers
// pin that order (one load after each of the first MFMAs, then one fragment read per two MFMAs);
// scheduling barriers at both ends of a stage keep the compiler from hoisting MFMAs above the
// workgroup barrier, which made it rotate accumulators between AGPRs on every iteration. The loads
// always run, the k index of stages past the end clamped to the last one (its slot is not read
// again), so every wait count is a constant. Each output block sees the MFMA sequence of
// gemm_d_dlds_kernel: the results are the same bits.
// TN 8192^3 / 16384^3 on HBM-resident operands: 73.3 / 73.4 TFLOP/s with 128 x 256 tiles and
// 73.2 / 72.0 with 128 x 128 tiles at two per CU, against 71.1 / 71.1 for gemm_d_dlds_kernel (128 x
// 128, three per CU, no software pipelining); 128 x 256 without the scheduling groups 69.9 / 70.1,
// with a 4-stage ring 73.0 / 73.1, with one load per 4 MFMAs 72.4 / 72.6. 40000^3: 72.5 with 128 x
// 256 tiles, 69.0 with 128 x 128 (twice the B re-reads); on the 8064 x 8192 x 8000 blocks of the
// host-to-host call 128 x 128 tiles win (72.3 against 69.3 with C read in the epilogue): with one
// workgroup per CU the epilogue idles the CU, with two the other workgroup's MFMAs cover it.
// ---------------------------------------------------------------------------------------------
template <bool KC, int ROWS>
struct BigOperand {
  static constexpr int PER_WAVE = ROWS / 64;  // load instructions per wave and stage (ROWS / 16 in all)
  const double* base[PER_WAVE];
  int dst[PER_WAVE];
  long kstride;  // elements per step of k: 1 (k-contiguous) or ld (row-contiguous)
  // k-contiguous: instruction i fills rows 16i..16i+15 (64 bytes each, slots XOR-permuted as in
  // DldsOperand); row-contiguous: instruction i fills the 128 rows (i % P) * 128.. at k = i / P,
  // P = ROWS / 128, 2 rows per lane (needs an even row count).
  __device__ __forceinline__ BigOperand(const double* p, long ld, int r0, int nrows, int wave, int lane) {
    kstride = KC ? 1 : ld;
#pragma unroll
    for (int t = 0; t < PER_WAVE; ++t) {
      const int inst = wave * PER_WAVE + t;
      if (KC) {
        const int local = inst * 16 + (lane >> 2);
        const int q = (lane & 3) ^ DldsOperand<true>::kc_swz(local);
        base[t] = p + 2 * q + (long)min(r0 + local, nrows - 1) * ld;
        dst[t] = inst * 128;
      } else {
        constexpr int P = ROWS / 128;
        const int kk = inst / P, part = inst % P;
        base[t] = p + min(r0 + part * 128 + 2 * lane, nrows - 2) + (long)kk * ld;
        dst[t] = kk * ROWS + part * 128;
      }
    }
  }
  __device__ __forceinline__ void issue(int k0, double* stage) const {
#pragma unroll
    for (int t = 0; t < PER_WAVE; ++t)
      __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(base[t] + (long)k0 * kstride),
                                       (__attribute__((address_space(3))) void*)(stage + dst[t]), 16, 0, 0);
  }
  // stage offset (doubles) of row `row`'s values at k = 2g, 2g+1 (row-contiguous: the second one
  // ROWS further)
  __device__ __forceinline__ static int frag_off(int row, int g) {
    return KC ? row * 8 + (g ^ DldsOperand<true>::kc_swz(row)) * 2 : 2 * g * ROWS + row;
  }
  __device__ __forceinline__ static v2d frag(const double* stage, int off) {
    if (KC) return *(const v2d*)(stage + off);
    v2d v;
    v[0] = stage[off];
    v[1] = stage[off + ROWS];
    return v;
  }
  static constexpr int READS = KC ? 1 : 2;  // LDS read instructions per fragment
};

// One 128 x BN C tile at (m0, n0); K a multiple of 8 (the launcher hands a remainder to
// gemm_real_kernel).
template <bool A_KC, bool B_KC, int BN>
__device__ __forceinline__ void dbig_tile_at(const GemmParams<double>& prm, int m0, int n0) {
  constexpr int BM = 128, S = 3, A_ST = BM * 8, STAGE = (BM + BN) * 8;
  constexpr int TI = 4, TJ = BN / 32;  // 16-row blocks of the wave tile (64 x BN/2)
  using OA = BigOperand<A_KC, BM>;
  using OB = BigOperand<B_KC, BN>;
  constexpr int LPS = OA::PER_WAVE + OB::PER_WAVE;         // loads per wave and stage
  constexpr int NDS = TI * OA::READS + TJ * OB::READS;     // fragment reads per wave and stage
  constexpr int DSG = NDS <= 12 ? 2 : 1;                   // MFMAs per fragment read
  constexpr int REST = 2 * TI * TJ - LPS - NDS * DSG;
  static_assert(REST >= 0, "stage schedule exceeds the stage's MFMAs");
  __shared__ __attribute__((aligned(16))) double lds[S * STAGE];
  const int M = prm.m, N = prm.n, K = prm.k;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int li = lane & 15, lg = lane >> 4;
  const int nkt = K / 8;
  const OA opa(prm.a, prm.lda, m0, M, wave, lane);
  const OB opb(prm.b, prm.ldb, n0, N, wave, lane);
  auto issue = [&](int kt, int so) {
    opa.issue(kt * 8, lds + so);
    opb.issue(kt * 8, lds + so + A_ST);
  };
  int offa[TI], offb[TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i) offa[i] = OA::frag_off(wm * 64 + i * 16 + li, lg);
#pragma unroll
  for (int j = 0; j < TJ; ++j) offb[j] = OB::frag_off(wn * (BN / 2) + j * 16 + li, lg);  // within B's part
  struct Frags {
    v2d a[TI], b[TJ];
  };
  auto read = [&](Frags& f, int so) {
    const double* st = lds + so;
#pragma unroll
    for (int i = 0; i < TI; ++i) f.a[i] = OA::frag(st, offa[i]);
#pragma unroll
    for (int j = 0; j < TJ; ++j) f.b[j] = OB::frag(st + A_ST, offb[j]);
  };
  dbl4 acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = dbl4{0, 0, 0, 0};

  if (nkt > 0) {
    // stages 0, 1, 2 in flight (clamped), stage 0 read into F0
#pragma unroll
    for (int s = 0; s < S; ++s) issue(min(s, nkt - 1), s * STAGE);
    __builtin_amdgcn_s_waitcnt(0x070 | ((S - 1) * LPS));
    __builtin_amdgcn_s_barrier();
    Frags F0, F1;
    read(F0, 0);
    int so = 0;  // slot of stage kt
    auto step = [&](Frags& cur, Frags& nxt, int kt) {
      // stage kt+1 has landed (stage kt+2's loads may still run) and this wave's reads of stage
      // kt are done; the barrier extends both to the workgroup
      __builtin_amdgcn_s_waitcnt(0x070 | ((S - 2) * LPS));
      __builtin_amdgcn_s_barrier();
      __builtin_amdgcn_sched_barrier(0);
      issue(min(kt + S, nkt - 1), so);  // into the slot of stage kt
      const int sn = so + STAGE == S * STAGE ? 0 : so + STAGE;
      read(nxt, sn);  // stage kt+1 (not used after the last stage)
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int i = 0; i < TI; ++i)
#pragma unroll
          for (int j = 0; j < TJ; ++j) acc[i][j] = mfma(cur.b[j][h], cur.a[i][h], acc[i][j]);
#pragma unroll
      for (int r = 0; r < LPS; ++r) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
        __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // VMEM read (LDS-DMA load)
      }
#pragma unroll
      for (int r = 0; r < NDS; ++r) {
        __builtin_amdgcn_sched_group_barrier(0x008, DSG, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // DS read
      }
      __builtin_amdgcn_sched_group_barrier(0x008, REST, 0);
      __builtin_amdgcn_sched_barrier(0);
      so = sn;
    };
    int kt = 0;
    for (; kt + 1 < nkt; kt += 2) {
      step(F0, F1, kt);
      step(F1, F0, kt + 1);
    }
    if (kt < nkt) step(F0, F1, kt);
  }

  // Epilogue: the accumulators leave the AGPRs only through stores into LDS and come back as
  // ordinary values for C = alpha * acc (+ C) (+ beta * Cin). Arithmetic on them directly made the
  // compiler emit the MFMAs in their VGPR form, where 256 accumulators next to the fragments and
  // addresses overflow into AGPR copies and scratch inside the loop. Per 16-row block each wave
  // stores its 32 values per lane into its own 16 KB of the ring (element-major: conflict-free) and
  // reads back its own slots after a workgroup barrier, whose fence keeps the compiler from
  // folding the round trip. The block's C (or beta * Cin) values are loaded before the round trip so that their
  // latency overlaps it: with one workgroup per CU nothing else hides an epilogue.
  const double alpha = prm.alpha, beta = prm.beta;
  const bool acc_prev = prm.accumulate != 0, use_cin = beta != 0.0, pre_any = acc_prev || use_cin;
  // The loop's last stages issued loads into the ring (clamped to the last stage) that nobody waits
  // for; they must land before the ring holds the epilogue's values: each wave waits for all of
  // its loads, then the barrier
  __builtin_amdgcn_s_waitcnt(0x070);
  __syncthreads();
  // A generic pointer to this wave's part of the ring whose LDS origin the compiler does not see:
  // its stores are FLAT stores, which take the accumulators straight from AGPRs (LDS stores made
  // the compiler move them to VGPRs as well).
  const unsigned long long ep_flat = (unsigned long long)(uintptr_t)(lds + wave * (TJ * 4 * 64));
  double* ep = (double*)(uintptr_t)(((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(ep_flat >> 32)) << 32) |
                                    __builtin_amdgcn_readfirstlane((unsigned)ep_flat));
#pragma unroll
  for (int i = 0; i < TI; ++i) {
    const int gi = m0 + wm * 64 + i * 16 + li;
    const bool row_ok = gi < M;
    double pre[TJ][4];  // C when accumulating, else beta * Cin (else unused)
#pragma unroll
    for (int j = 0; j < TJ; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gj = n0 + wn * (BN / 2) + j * 16 + out_col<double>(lane, r);
        pre[j][r] = 0.0;
        if (pre_any && row_ok && gj < N)
          pre[j][r] = acc_prev ? prm.c[(long)gi + (long)gj * prm.ldc] : beta * prm.cin[(long)gi + (long)gj * prm.ldcin];
      }
#pragma unroll
    for (int j = 0; j < TJ; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) ep[(j * 4 + r) * 64 + lane] = acc[i][j][r];
    __syncthreads();
    if (row_ok) {
#pragma unroll
      for (int j = 0; j < TJ; ++j) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gj = n0 + wn * (BN / 2) + j * 16 + out_col<double>(lane, r);
          if (gj >= N) continue;
          double v = alpha * ep[(j * 4 + r) * 64 + lane];
          if (pre_any) v += pre[j][r];
          if (acc_prev && use_cin) v += beta * prm.cin[(long)gi + (long)gj * prm.ldcin];
          prm.c[(long)gi + (long)gj * prm.ldc] = v;
        }
      }
    }
    __syncthreads();  // this block's reads before the next block's stores
  }
}

template <bool A_KC, bool B_KC, int BN>
__device__ __forceinline__ void dbig_tile(const GemmParams<double>& prm, int w) {
  const int nbm = (prm.m + 127) / 128, nbn = (prm.n + BN - 1) / BN;
  int bm, bn;
  tile_coords(w, nbm, nbn, bm, bn);
  dbig_tile_at<A_KC, B_KC, BN>(prm, bm * 128, bn * BN);
}

// 128 x 256 tiles, one workgroup per CU
template <bool A_KC, bool B_KC>
__global__ __launch_bounds__(256, 1) void gemm_d_big_kernel(GemmParams<double> prm) {
  if (prm.gate_a && (*prm.gate_a | *prm.gate_b) == 0) return;  // uniform: before any barrier
  dbig_tile<A_KC, B_KC, 256>(prm, blockIdx.x);
}

// 128 x 128 tiles, two workgroups per CU
template <bool A_KC, bool B_KC>
__global__ __launch_bounds__(256, 2) void gemm_d_pipe_kernel(GemmParams<double> prm) {
  if (prm.gate_a && (*prm.gate_a | *prm.gate_b) == 0) return;  // uniform: before any barrier
  dbig_tile<A_KC, B_KC, 128>(prm, blockIdx.x);
}

// Grouped FP64 tile GEMMs on the same kernels (128 x 256: every member has m, n > 128). Padding
// workgroups exit before any barrier.
template <bool A_KC, bool B_KC>
__global__ __launch_bounds__(256, 1) void gemm_d_big_group_kernel(DgemmGroup g) {
  GemmParams<double> prm;
  int local;
  if (group_member(g, blockIdx.x, 128, 256, prm, local) < 0) return;
  dbig_tile<A_KC, B_KC, 256>(prm, local);
}

template <bool A_KC, bool B_KC>
__global__ __launch_bounds__(256, 2) void gemm_d_pipe_group_kernel(DgemmGroup g) {
  GemmParams<double> prm;
  int local;
  if (group_member(g, blockIdx.x, 128, 128, prm, local) < 0) return;
  dbig_tile<A_KC, B_KC, 128>(prm, local);
}

// Grouped FP32 tile GEMMs on the pipelined 128 x 128 tiles. Padding workgroups exit before any
// barrier.
template <bool A_KC, bool B_KC>
__global__ __launch_bounds__(256, 2) void gemm_s_pipe_group_kernel(SgemmGroup g) {
  GemmParams<float> prm;
  int local;
  if (group_member(g, blockIdx.x, 128, 128, prm, local) < 0) return;
  const int nbm = (prm.m + 127) / 128, nbn = (prm.n + 127) / 128;
  int bm, bn;
  tile_coords(local, nbm, nbn, bm, bn);
  spipe_tile_at<A_KC, B_KC, 128, 128>(prm, bm * 128, bn * 128);
}

// Grouped FP32 tile GEMMs (256 x 256 workgroup tiles: every member has m, n > 128). Padding
// workgroups exit before any barrier.
template <bool A_KC, bool B_KC, int S, int OCC>
__global__ __launch_bounds__(512, OCC) void gemm_s_dlds_group_kernel(SgemmGroup g) {
  GemmParams<float> prm;
  int local;
  if (group_member(g, blockIdx.x, 256, 256, prm, local) < 0) return;
  const int nbm = (prm.m + 255) / 256, nbn = (prm.n + 255) / 256;
  int bm, bn;
  tile_coords(local, nbm, nbn, bm, bn);
  sdlds_tile_at<A_KC, B_KC, S, 256, 256>(prm, bm * 256, bn * 256);
}

// Grouped complex FP32 tile GEMMs (128 x 128 complex workgroup tiles).
template <bool A_KC, bool B_KC, bool A_CONJ, bool B_CONJ>
__global__ __launch_bounds__(256, 2) void gemm_c_pipe_group_kernel(SgemmGroup g) {
  GemmParams<float> prm;
  int local;
  if (group_member(g, blockIdx.x, 128, 128, prm, local) < 0) return;
  cpipe_tile<A_KC, B_KC, A_CONJ, B_CONJ>(prm, local);
}

template <bool A_KC, bool B_KC, bool A_CONJ, bool B_CONJ, int S, int OCC>
__global__ __launch_bounds__(256, OCC) void gemm_c_dlds_group_kernel(SgemmGroup g) {
  GemmParams<float> prm;
  int local;
  if (group_member(g, blockIdx.x, 128, 128, prm, local) < 0) return;
  cdlds_tile<A_KC, B_KC, A_CONJ, B_CONJ, S>(prm, local);
}

// ---------------------------------------------------------------------------------------------
// Complex FP64 GEMM with direct global->LDS loads (ZGEMM), the complex sibling of
// gemm_d_dlds_kernel. WG tile 64x64 complex, 2x2 waves of 32x32 (2x2 blocks, re and im
// accumulators, four real MFMAs per block update), k-step 8 complex per stage, S-stage ring of
// 16 KB stages (S = 3, three workgroups per CU). A 16-byte chunk is one complex element, so one
// ds_read_b128 yields both MFMA operands (re, im) of a lane.
//   k-contiguous operand (A for 'T'/'C', B for 'N'): LDS [row][slot], 128 bytes per row; a wave
//     instruction fills 8 rows, lanes 8j..8j+7 reading the 128 contiguous bytes of row j, lane l
//     landing in slot l % 8 which receives element (l % 8) ^ (row & 7): the fragment read of one
//     element by 8 consecutive rows then covers all 8 16-byte bank groups.
//   row-contiguous operand (A for 'N', B for 'T'/'C'): LDS [k][row], a wave instruction fills
//     the 64 rows of one k (1 KB contiguous in global memory).
// Conjugation ('C') negates the imaginary fragment in registers. Rows beyond the matrix read a
// clamped in-bounds row. K must be a multiple of 8 (the launcher hands a remainder to
// gemm_cplx_kernel).
// ---------------------------------------------------------------------------------------------
template <bool KC>
struct ZdldsOperand {
  // this wave's 2 loads of one stage (8 instructions of 1 KB per workgroup and operand)
  __device__ __forceinline__ static void issue(const double* __restrict__ p, long ld, int r0, int nrows, int k0,
                                               double* stage, int wave, int lane) {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int inst = wave * 2 + t;  // 0..7
      const double* src;
      if (KC) {
        const int local = inst * 8 + (lane >> 3);
        const int q = (lane & 7) ^ (local & 7);
        const int row = min(r0 + local, nrows - 1);
        src = p + 2 * ((long)(k0 + q) + (long)row * ld);
      } else {
        const int row = min(r0 + lane, nrows - 1);
        src = p + 2 * ((long)row + (long)(k0 + inst) * ld);
      }
      __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                       (__attribute__((address_space(3))) void*)(stage + inst * 128), 16, 0, 0);
    }
  }
  // complex element (row, k) of the stage, k in [0, 8)
  __device__ __forceinline__ static v2d elem(const double* stage, int row, int k) {
    if (KC) return *(const v2d*)(stage + 2 * (row * 8 + (k ^ (row & 7))));
    return *(const v2d*)(stage + 2 * (k * 64 + row));
  }
};

template <bool A_KC, bool B_KC, bool A_CONJ, bool B_CONJ, int S>
__device__ __forceinline__ void zdlds_tile(const GemmParams<double>& prm, int block) {
  constexpr int TILE = 64 * 8 * 2;  // doubles per operand and stage (8 KB)
  constexpr int STAGE = 2 * TILE;
  __shared__ __attribute__((aligned(16))) double lds[S * STAGE];
  const int M = prm.m, N = prm.n, K = prm.k;
  const int nbm = (M + 63) / 64, nbn = (N + 63) / 64;
  int bm, bn;
  tile_coords(block, nbm, nbn, bm, bn);
  const int m0 = bm * 64, n0 = bn * 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int li = lane & 15, lg = lane >> 4;
  const int nkt = K / 8;

  dbl4 cre[2][2], cim[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      cre[i][j] = dbl4{0, 0, 0, 0};
      cim[i][j] = dbl4{0, 0, 0, 0};
    }

  auto issue = [&](int kt) {
    double* st = lds + (kt % S) * STAGE;
    ZdldsOperand<A_KC>::issue(prm.a, prm.lda, m0, M, kt * 8, st, wave, lane);
    ZdldsOperand<B_KC>::issue(prm.b, prm.ldb, n0, N, kt * 8, st + TILE, wave, lane);
  };
#pragma unroll
  for (int s = 0; s < S - 1; ++s)
    if (s < nkt) issue(s);
  for (int kt = 0; kt < nkt; ++kt) {
    // this wave's loads of stage kt have landed once at most 4 per later stage are pending, and
    // its LDS reads of stage kt-1 are done; the barrier extends both to the workgroup
    const int later = min(S - 2, nkt - 1 - kt);
    if (later >= 3) __builtin_amdgcn_s_waitcnt(0x070 | 12);
    else if (later == 2) __builtin_amdgcn_s_waitcnt(0x070 | 8);
    else if (later == 1) __builtin_amdgcn_s_waitcnt(0x070 | 4);
    else __builtin_amdgcn_s_waitcnt(0x070);
    __builtin_amdgcn_s_barrier();
    if (kt + S - 1 < nkt) issue(kt + S - 1);
    const double* st = lds + (kt % S) * STAGE;
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
      v2d a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        a[i] = ZdldsOperand<A_KC>::elem(st, wm * 32 + i * 16 + li, 4 * kk + lg);
        if (A_CONJ) a[i][1] = -a[i][1];
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        b[j] = ZdldsOperand<B_KC>::elem(st + TILE, wn * 32 + j * 16 + li, 4 * kk + lg);
        if (B_CONJ) b[j][1] = -b[j][1];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const double nai = -a[i][1];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          cre[i][j] = mfma(b[j][0], a[i][0], cre[i][j]);
          cre[i][j] = mfma(b[j][1], nai, cre[i][j]);
          cim[i][j] = mfma(b[j][1], a[i][0], cim[i][j]);
          cim[i][j] = mfma(b[j][0], a[i][1], cim[i][j]);
        }
      }
    }
  }

  const double alr = prm.alpha, ali = prm.alpha_im;
  const double ber = prm.beta, bei = prm.beta_im;
  const bool acc_prev = prm.accumulate != 0;
  const bool use_cin = ber != 0.0 || bei != 0.0;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int gi = m0 + wm * 32 + i * 16 + li;
    if (gi >= M) continue;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gj = n0 + wn * 32 + j * 16 + out_col<double>(lane, r);
        if (gj >= N) continue;
        const double xr = cre[i][j][r], xi = cim[i][j][r];
        double vr = alr * xr - ali * xi;
        double vi = alr * xi + ali * xr;
        double* dst = prm.c + 2 * ((long)gi + (long)gj * prm.ldc);
        if (acc_prev) {
          vr += dst[0];
          vi += dst[1];
        }
        if (use_cin) {
          const double* ci = prm.cin + 2 * ((long)gi + (long)gj * prm.ldcin);
          const double cr = ci[0], cm = ci[1];
          vr += ber * cr - bei * cm;
          vi += ber * cm + bei * cr;
        }
        dst[0] = vr;
        dst[1] = vi;
      }
    }
  }
}

// The same tile software-pipelined as dbig_tile_at (two workgroups per CU): the fragments of stage
// kt+1 are read into a second register set while the 32 MFMAs of stage kt run, the 4 loads of
// stage kt+3 go between its first MFMAs, scheduling-group barriers pin that order, and the loads
// always run (clamped to the last stage) so that every wait count is a constant. Per output block
// the MFMA sequence is zdlds_tile's: the results are the same bits.
template <bool A_KC, bool B_KC, bool A_CONJ, bool B_CONJ>
__device__ __forceinline__ void zpipe_tile(const GemmParams<double>& prm, int block) {
  constexpr int TILE = 64 * 8 * 2, STAGE = 2 * TILE, S = 3;
  constexpr int LPS = 4;  // loads per wave and stage (2 per operand)
  constexpr int NDS = 8;  // fragment reads per wave and stage
  __shared__ __attribute__((aligned(16))) double lds[S * STAGE];
  const int M = prm.m, N = prm.n, K = prm.k;
  const int nbm = (M + 63) / 64, nbn = (N + 63) / 64;
  int bm, bn;
  tile_coords(block, nbm, nbn, bm, bn);
  const int m0 = bm * 64, n0 = bn * 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int li = lane & 15, lg = lane >> 4;
  const int nkt = K / 8;

  dbl4 cre[2][2], cim[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      cre[i][j] = dbl4{0, 0, 0, 0};
      cim[i][j] = dbl4{0, 0, 0, 0};
    }
  auto issue = [&](int kt, int so) {
    ZdldsOperand<A_KC>::issue(prm.a, prm.lda, m0, M, kt * 8, lds + so, wave, lane);
    ZdldsOperand<B_KC>::issue(prm.b, prm.ldb, n0, N, kt * 8, lds + so + TILE, wave, lane);
  };
  struct Frags {
    v2d a[2][2], b[2][2];  // [kk][block]
  };
  auto read = [&](Frags& f, int so) {
    const double* st = lds + so;
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
#pragma unroll
      for (int i = 0; i < 2; ++i) f.a[kk][i] = ZdldsOperand<A_KC>::elem(st, wm * 32 + i * 16 + li, 4 * kk + lg);
#pragma unroll
      for (int j = 0; j < 2; ++j) f.b[kk][j] = ZdldsOperand<B_KC>::elem(st + TILE, wn * 32 + j * 16 + li, 4 * kk + lg);
    }
  };
  if (nkt > 0) {
#pragma unroll
    for (int s = 0; s < S; ++s) issue(min(s, nkt - 1), s * STAGE);
    __builtin_amdgcn_s_waitcnt(0x070 | ((S - 1) * LPS));
    __builtin_amdgcn_s_barrier();
    Frags F0, F1;
    read(F0, 0);
    int so = 0;
    auto step = [&](Frags& cur, Frags& nxt, int kt) {
      __builtin_amdgcn_s_waitcnt(0x070 | ((S - 2) * LPS));
      __builtin_amdgcn_s_barrier();
      __builtin_amdgcn_sched_barrier(0);
      issue(min(kt + S, nkt - 1), so);
      const int sn = so + STAGE == S * STAGE ? 0 : so + STAGE;
      read(nxt, sn);
#pragma unroll
      for (int kk = 0; kk < 2; ++kk) {
        v2d a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          a[i] = cur.a[kk][i];
          if (A_CONJ) a[i][1] = -a[i][1];
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          b[j] = cur.b[kk][j];
          if (B_CONJ) b[j][1] = -b[j][1];
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          const double nai = -a[i][1];
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            cre[i][j] = mfma(b[j][0], a[i][0], cre[i][j]);
            cre[i][j] = mfma(b[j][1], nai, cre[i][j]);
            cim[i][j] = mfma(b[j][1], a[i][0], cim[i][j]);
            cim[i][j] = mfma(b[j][0], a[i][1], cim[i][j]);
          }
        }
      }
#pragma unroll
      for (int r = 0; r < LPS; ++r) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
        __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // VMEM read (LDS-DMA load)
      }
#pragma unroll
      for (int r = 0; r < NDS; ++r) {
        __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // DS read
      }
      __builtin_amdgcn_sched_group_barrier(0x008, 32 - LPS - 2 * NDS, 0);
      __builtin_amdgcn_sched_barrier(0);
      so = sn;
    };
    int kt = 0;
    for (; kt + 1 < nkt; kt += 2) {
      step(F0, F1, kt);
      step(F1, F0, kt + 1);
    }
    if (kt < nkt) step(F0, F1, kt);
    // the last stages' loads (clamped) land before the workgroup's LDS is released
    __builtin_amdgcn_s_waitcnt(0x070);
  }

  const double alr = prm.alpha, ali = prm.alpha_im;
  const double ber = prm.beta, bei = prm.beta_im;
  const bool acc_prev = prm.accumulate != 0;
  const bool use_cin = ber != 0.0 || bei != 0.0;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int gi = m0 + wm * 32 + i * 16 + li;
    if (gi >= M) continue;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gj = n0 + wn * 32 + j * 16 + out_col<double>(lane, r);
        if (gj >= N) continue;
        const double xr = cre[i][j][r], xi = cim[i][j][r];
        double vr = alr * xr - ali * xi;
        double vi = alr * xi + ali * xr;
        double* dst = prm.c + 2 * ((long)gi + (long)gj * prm.ldc);
        if (acc_prev) {
          vr += dst[0];
          vi += dst[1];
        }
        if (use_cin) {
          const double* ci = prm.cin + 2 * ((long)gi + (long)gj * prm.ldcin);
          const double cr = ci[0], cm = ci[1];
          vr += ber * cr - bei * cm;
          vi += ber * cm + bei * cr;
        }
        dst[0] = vr;
        dst[1] = vi;
      }
    }
  }
}

template <bool A_KC, bool B_KC, bool A_CONJ, bool B_CONJ>
__global__ __launch_bounds__(256, 2) void gemm_z_pipe_kernel(GemmParams<double> prm) {
  zpipe_tile<A_KC, B_KC, A_CONJ, B_CONJ>(prm, blockIdx.x);
}

template <bool A_KC, bool B_KC, bool A_CONJ, bool B_CONJ>
__global__ __launch_bounds__(256, 2) void gemm_z_pipe_group_kernel(DgemmGroup g) {
  GemmParams<double> prm;
  int local;
  if (group_member(g, blockIdx.x, 64, 64, prm, local) < 0) return;
  zpipe_tile<A_KC, B_KC, A_CONJ, B_CONJ>(prm, local);
}

template <bool A_KC, bool B_KC, bool A_CONJ, bool B_CONJ, int S, int OCC = 2>
__global__ __launch_bounds__(256, OCC) void gemm_z_dlds_kernel(GemmParams<double> prm) {
  zdlds_tile<A_KC, B_KC, A_CONJ, B_CONJ, S>(prm, blockIdx.x);
}

// The complex sibling of gemm_d_dlds_group_kernel (64x64 complex workgroup tiles).
template <bool A_KC, bool B_KC, bool A_CONJ, bool B_CONJ, int S, int OCC = 2>
__global__ __launch_bounds__(256, OCC) void gemm_z_dlds_group_kernel(DgemmGroup g) {
  const int w = blockIdx.x;
  int t = 0;
  while (t + 1 < g.count && w >= g.wg_start[t + 1]) ++t;
  GemmParams<double> prm = g.base;
  prm.m = g.m[t];
  prm.n = g.n[t];
  prm.a = g.a[t];
  prm.lda = g.lda[t];
  prm.b = g.b[t];
  prm.ldb = g.ldb[t];
  prm.c = g.c[t];
  prm.ldc = g.ldc[t];
  prm.cin = g.cin[t];
  prm.ldcin = g.ldcin[t];
  prm.beta = g.beta[t];
  prm.beta_im = g.beta_im[t];
  prm.accumulate = g.accumulate[t];
  const int local = w - g.wg_start[t];
  if (local >= ((prm.m + 63) / 64) * ((prm.n + 63) / 64)) return;
  zdlds_tile<A_KC, B_KC, A_CONJ, B_CONJ, S>(prm, local);
}

inline bool is_kc_a(char op) { return op != 'N'; }
inline bool is_kc_b(char op) { return op == 'N'; }

template <typename R>
hipError_t launch_real(const GemmParams<R>& p, hipStream_t s) {
  const int nbm = (p.m + 127) / 128, nbn = (p.n + 127) / 128;
  dim3 grid(nbm * nbn), block(256);
  const bool akc = is_kc_a(p.opa), bkc = is_kc_b(p.opb);
  if (akc && bkc) hipLaunchKernelGGL((gemm_real_kernel<R, true, true>), grid, block, 0, s, p);
  else if (akc && !bkc) hipLaunchKernelGGL((gemm_real_kernel<R, true, false>), grid, block, 0, s, p);
  else if (!akc && bkc) hipLaunchKernelGGL((gemm_real_kernel<R, false, true>), grid, block, 0, s, p);
  else hipLaunchKernelGGL((gemm_real_kernel<R, false, false>), grid, block, 0, s, p);
  return hipGetLastError();
}

template <typename R, bool AKC, bool BKC>
hipError_t launch_cplx_conj(const GemmParams<R>& p, hipStream_t s, dim3 grid, dim3 block) {
  const bool ca = p.opa == 'C', cb = p.opb == 'C';
  if (ca && cb) hipLaunchKernelGGL((gemm_cplx_kernel<R, AKC, BKC, true, true>), grid, block, 0, s, p);
  else if (ca) hipLaunchKernelGGL((gemm_cplx_kernel<R, AKC, BKC, true, false>), grid, block, 0, s, p);
  else if (cb) hipLaunchKernelGGL((gemm_cplx_kernel<R, AKC, BKC, false, true>), grid, block, 0, s, p);
  else hipLaunchKernelGGL((gemm_cplx_kernel<R, AKC, BKC, false, false>), grid, block, 0, s, p);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_cplx(const GemmParams<R>& p, hipStream_t s) {
  const int nbm = (p.m + 63) / 64, nbn = (p.n + 63) / 64;
  dim3 grid(nbm * nbn), block(256);
  const bool akc = is_kc_a(p.opa), bkc = is_kc_b(p.opb);
  if (akc && bkc) return launch_cplx_conj<R, true, true>(p, s, grid, block);
  if (akc && !bkc) return launch_cplx_conj<R, true, false>(p, s, grid, block);
  if (!akc && bkc) return launch_cplx_conj<R, false, true>(p, s, grid, block);
  return launch_cplx_conj<R, false, false>(p, s, grid, block);
}

// Which FP64 kernel a product of m x n runs on (for a group: its members together), by what the
// measurements favour: the 128 x 256 tiles for products of at least 2^27 elements of C (their
// halved B re-reads pay; 40000^2: 72.5 against 69.0 TFLOP/s), the pipelined 128 x 128 tiles at
// two per CU below that (8064 x 8192 blocks: 72.3 against 69.3 with C read in the epilogue), and
// the 128 x 128 kernel at three per CU when neither grid has two workgroups per CU to fill the
// device with.
int device_cus() {
  static const int cus = [] {
    int dev = 0, c = 256;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      c = 256;
    return c;
  }();
  return cus;
}
long dgemm_tiles(int m, int n, int bn) { return (long)((m + 127) / 128) * ((n + bn - 1) / bn); }
enum DgemmKind { kDgemmRing3 = 0, kDgemmPipe = 1, kDgemmBig = 2 };
// largest: the largest member's m x n; big_ok: every member has m, n > 128; tiles128 / tiles256:
// the workgroups of all members on either grid
// FP32: the pipelined 128 x 128 tiles when the product (a group: its members together) has two of
// them per CU
bool sgemm_pipe(long tiles128) { return tiles128 >= 2L * device_cus(); }
bool sgemm_pipe(int m, int n) { return sgemm_pipe(dgemm_tiles(m, n, 128)); }
int dgemm_kind(long largest, bool big_ok, long tiles128, long tiles256) {
  if (big_ok && largest >= (1L << 27) && tiles256 >= 2L * device_cus()) return kDgemmBig;
  if (tiles128 >= 2L * device_cus()) return kDgemmPipe;
  return kDgemmRing3;
}

template <bool AKC, bool BKC>
void launch_dlds(const GemmParams<double>& p, hipStream_t s) {
  const int kind = dgemm_kind((long)p.m * p.n, p.m > 128 && p.n > 128, dgemm_tiles(p.m, p.n, 128),
                              dgemm_tiles(p.m, p.n, 256));
  if (kind == kDgemmBig) {
    hipLaunchKernelGGL((gemm_d_big_kernel<AKC, BKC>), dim3(dgemm_tiles(p.m, p.n, 256)), dim3(256), 0, s, p);
    return;
  }
  if (kind == kDgemmPipe) {
    hipLaunchKernelGGL((gemm_d_pipe_kernel<AKC, BKC>), dim3(dgemm_tiles(p.m, p.n, 128)), dim3(256), 0, s, p);
    return;
  }
  const int nbm = (p.m + 127) / 128, nbn = (p.n + 127) / 128;
  // 3-stage ring, 3 workgroups per CU (48 KB of LDS each): 70.9 TFLOP/s on TN 8192^3 against
  // 68.8 for 4 stages at 2 per CU, 68.7 for 5 at 2, 43.5 for 6 at 1
  hipLaunchKernelGGL((gemm_d_dlds_kernel<AKC, BKC, 3, 3>), dim3(nbm * nbn), dim3(256), 0, s, p);
}

// FP64 through the direct-LDS kernel for the largest multiple of 8 in K; a remainder of K goes to
// gemm_real_kernel, accumulating (beta only on the first part).
hipError_t launch_gemm_d_direct(const GemmParams<double>& p, hipStream_t s) {
  const bool akc = is_kc_a(p.opa), bkc = is_kc_b(p.opb);
  const int k8 = p.k / 8 * 8;
  // a row-contiguous operand with an odd row count would need a 2-row chunk that straddles the
  // last row: such shapes keep the register-staged kernel
  const bool pairs_ok = (akc || p.m % 2 == 0) && (bkc || p.n % 2 == 0);
  if (k8 == 0 || p.m < 2 || p.n < 2 || !pairs_ok) return launch_real<double>(p, s);
  GemmParams<double> q = p;
  q.k = k8;
  if (akc && bkc) launch_dlds<true, true>(q, s);
  else if (akc) launch_dlds<true, false>(q, s);
  else if (bkc) launch_dlds<false, true>(q, s);
  else launch_dlds<false, false>(q, s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || k8 == p.k) return e;
  GemmParams<double> r = p;
  r.k = p.k - k8;
  r.a = p.a + (akc ? (long)k8 : (long)k8 * p.lda);
  r.b = p.b + (bkc ? (long)k8 : (long)k8 * p.ldb);
  r.accumulate = 1;
  r.beta = 0.0;
  return launch_real<double>(r, s);
}

}  // namespace

bool dgemm_groupable(const GemmParams<double>& p) {
  const bool akc = is_kc_a(p.opa), bkc = is_kc_b(p.opb);
  return p.variant == 1 && p.k >= 8 && p.k % 8 == 0 && p.m >= 2 && p.n >= 2 && (akc || p.m % 2 == 0) &&
         (bkc || p.n % 2 == 0);
}

bool zgemm_groupable(const GemmParams<double>& p) { return p.variant == 1 && p.k >= 8 && p.k % 8 == 0 && p.m >= 1 && p.n >= 1; }
// FP32 groups run the 256 x 256 tiles: m, n > 128, K a multiple of 16, row-contiguous operands in
// whole 4-row chunks (what launch_gemm_s_direct sends to that kernel without a remainder launch)
bool sgemm_groupable(const GemmParams<float>& p) {
  const bool akc = is_kc_a(p.opa), bkc = is_kc_b(p.opb);
  return p.variant == 1 && p.k >= 16 && p.k % 16 == 0 && p.m > 128 && p.n > 128 && (akc || p.m % 4 == 0) &&
         (bkc || p.n % 4 == 0);
}
bool cgemm_groupable(const GemmParams<float>& p) {
  const bool akc = is_kc_a(p.opa), bkc = is_kc_b(p.opb);
  return p.variant == 1 && p.k >= 8 && p.k % 8 == 0 && p.m >= 1 && p.n >= 1 && (akc || p.m % 2 == 0) &&
         (bkc || p.n % 2 == 0);
}

namespace {
// The group descriptor of `count` products of one shape class; false if they do not form one.
// edge_m, edge_n: rows and columns of C per workgroup of the group's kernel.
template <typename R, typename Pred>
bool make_group(const GemmParams<R>* tiles, int count, int edge_m, int edge_n, Pred groupable, TileGroup<R>& g,
                int& wg) {
  if (count <= 0 || count > kDgemmGroupMax) return false;
  g = TileGroup<R>{};
  g.base = tiles[0];
  g.count = count;
  wg = 0;
  for (int t = 0; t < count; ++t) {
    const GemmParams<R>& p = tiles[t];
    if (!groupable(p) || p.k != g.base.k || p.opa != g.base.opa || p.opb != g.base.opb ||
        p.alpha != g.base.alpha || p.alpha_im != g.base.alpha_im)
      return false;
    g.wg_start[t] = wg;
    wg += (((p.m + edge_m - 1) / edge_m) * ((p.n + edge_n - 1) / edge_n) + 7) / 8 * 8;
    g.m[t] = p.m;
    g.n[t] = p.n;
    g.a[t] = p.a;
    g.lda[t] = p.lda;
    g.b[t] = p.b;
    g.ldb[t] = p.ldb;
    g.c[t] = p.c;
    g.ldc[t] = p.ldc;
    g.cin[t] = p.cin;
    g.ldcin[t] = p.ldcin;
    g.beta[t] = p.beta;
    g.beta_im[t] = p.beta_im;
    g.accumulate[t] = p.accumulate;
  }
  g.wg_start[count] = wg;
  return true;
}
}  // namespace

hipError_t launch_gemm_z_group(const GemmParams<double>* tiles, int count, hipStream_t s) {
  DgemmGroup g;
  int wg = 0;
  if (count <= 0) return hipSuccess;
  if (!make_group(tiles, count, 64, 64, zgemm_groupable, g, wg)) return hipErrorInvalidValue;
  const bool akc = is_kc_a(g.base.opa), bkc = is_kc_b(g.base.opb);
  const bool ca = g.base.opa == 'C', cb = g.base.opb == 'C';
  // the pipelined tile at two workgroups per CU when the group has two tiles per CU to fill it
  const bool pipe = wg >= 2L * device_cus();
#define TMM_ZG(AK, BK, CA, CB)                                                                        \
  do {                                                                                               \
    if (pipe) hipLaunchKernelGGL((gemm_z_pipe_group_kernel<AK, BK, CA, CB>), dim3(wg), dim3(256), 0, s, g); \
    else hipLaunchKernelGGL((gemm_z_dlds_group_kernel<AK, BK, CA, CB, 3, 3>), dim3(wg), dim3(256), 0, s, g); \
  } while (0)
  if (akc && bkc) {
    if (ca) TMM_ZG(true, true, true, false);
    else TMM_ZG(true, true, false, false);
  } else if (akc) {
    if (ca && cb) TMM_ZG(true, false, true, true);
    else if (ca) TMM_ZG(true, false, true, false);
    else if (cb) TMM_ZG(true, false, false, true);
    else TMM_ZG(true, false, false, false);
  } else if (bkc) {
    TMM_ZG(false, true, false, false);
  } else {
    if (cb) TMM_ZG(false, false, false, true);
    else TMM_ZG(false, false, false, false);
  }
#undef TMM_ZG
  return hipGetLastError();
}

hipError_t launch_gemm_d_group(const GemmParams<double>* tiles, int count, hipStream_t s) {
  DgemmGroup g;
  int wg = 0;
  if (count <= 0) return hipSuccess;
  bool big_ok = true;
  long largest = 0, t128 = 0, t256 = 0;
  for (int t = 0; t < count; ++t) {
    big_ok = big_ok && tiles[t].m > 128 && tiles[t].n > 128;
    largest = std::max(largest, (long)tiles[t].m * tiles[t].n);
    t128 += dgemm_tiles(tiles[t].m, tiles[t].n, 128);
    t256 += dgemm_tiles(tiles[t].m, tiles[t].n, 256);
  }
  const int kind = dgemm_kind(largest, big_ok, t128, t256);
  if (!make_group(tiles, count, 128, kind == kDgemmBig ? 256 : 128, dgemm_groupable, g, wg))
    return hipErrorInvalidValue;
  const bool akc = is_kc_a(g.base.opa), bkc = is_kc_b(g.base.opb);
  if (kind == kDgemmBig) {
    if (akc && bkc) hipLaunchKernelGGL((gemm_d_big_group_kernel<true, true>), dim3(wg), dim3(256), 0, s, g);
    else if (akc) hipLaunchKernelGGL((gemm_d_big_group_kernel<true, false>), dim3(wg), dim3(256), 0, s, g);
    else if (bkc) hipLaunchKernelGGL((gemm_d_big_group_kernel<false, true>), dim3(wg), dim3(256), 0, s, g);
    else hipLaunchKernelGGL((gemm_d_big_group_kernel<false, false>), dim3(wg), dim3(256), 0, s, g);
    return hipGetLastError();
  }
  if (kind == kDgemmPipe) {
    if (akc && bkc) hipLaunchKernelGGL((gemm_d_pipe_group_kernel<true, true>), dim3(wg), dim3(256), 0, s, g);
    else if (akc) hipLaunchKernelGGL((gemm_d_pipe_group_kernel<true, false>), dim3(wg), dim3(256), 0, s, g);
    else if (bkc) hipLaunchKernelGGL((gemm_d_pipe_group_kernel<false, true>), dim3(wg), dim3(256), 0, s, g);
    else hipLaunchKernelGGL((gemm_d_pipe_group_kernel<false, false>), dim3(wg), dim3(256), 0, s, g);
    return hipGetLastError();
  }
  if (akc && bkc) hipLaunchKernelGGL((gemm_d_dlds_group_kernel<true, true, 3, 3>), dim3(wg), dim3(256), 0, s, g);
  else if (akc) hipLaunchKernelGGL((gemm_d_dlds_group_kernel<true, false, 3, 3>), dim3(wg), dim3(256), 0, s, g);
  else if (bkc) hipLaunchKernelGGL((gemm_d_dlds_group_kernel<false, true, 3, 3>), dim3(wg), dim3(256), 0, s, g);
  else hipLaunchKernelGGL((gemm_d_dlds_group_kernel<false, false, 3, 3>), dim3(wg), dim3(256), 0, s, g);
  return hipGetLastError();
}

hipError_t launch_gemm_s_group(const GemmParams<float>* tiles, int count, hipStream_t s) {
  SgemmGroup g;
  int wg = 0;
  if (count <= 0) return hipSuccess;
  long t128 = 0;
  for (int t = 0; t < count; ++t) t128 += dgemm_tiles(tiles[t].m, tiles[t].n, 128);
  const bool pipe = sgemm_pipe(t128);
  if (!make_group(tiles, count, pipe ? 128 : 256, pipe ? 128 : 256, sgemm_groupable, g, wg))
    return hipErrorInvalidValue;
  const bool akc = is_kc_a(g.base.opa), bkc = is_kc_b(g.base.opb);
  if (pipe) {
    if (akc && bkc) hipLaunchKernelGGL((gemm_s_pipe_group_kernel<true, true>), dim3(wg), dim3(256), 0, s, g);
    else if (akc) hipLaunchKernelGGL((gemm_s_pipe_group_kernel<true, false>), dim3(wg), dim3(256), 0, s, g);
    else if (bkc) hipLaunchKernelGGL((gemm_s_pipe_group_kernel<false, true>), dim3(wg), dim3(256), 0, s, g);
    else hipLaunchKernelGGL((gemm_s_pipe_group_kernel<false, false>), dim3(wg), dim3(256), 0, s, g);
    return hipGetLastError();
  }
  if (akc && bkc) hipLaunchKernelGGL((gemm_s_dlds_group_kernel<true, true, 3, 1>), dim3(wg), dim3(512), 0, s, g);
  else if (akc) hipLaunchKernelGGL((gemm_s_dlds_group_kernel<true, false, 3, 1>), dim3(wg), dim3(512), 0, s, g);
  else if (bkc) hipLaunchKernelGGL((gemm_s_dlds_group_kernel<false, true, 3, 1>), dim3(wg), dim3(512), 0, s, g);
  else hipLaunchKernelGGL((gemm_s_dlds_group_kernel<false, false, 3, 1>), dim3(wg), dim3(512), 0, s, g);
  return hipGetLastError();
}

template <bool AK, bool BK, bool CA, bool CB>
void launch_cg(bool pipe, int wg, hipStream_t s, const SgemmGroup& g) {
  if (pipe) hipLaunchKernelGGL((gemm_c_pipe_group_kernel<AK, BK, CA, CB>), dim3(wg), dim3(256), 0, s, g);
  else hipLaunchKernelGGL((gemm_c_dlds_group_kernel<AK, BK, CA, CB, 3, 2>), dim3(wg), dim3(256), 0, s, g);
}
template <bool AK, bool BK, bool CA, bool CB>
void launch_c1(bool pipe, dim3 grid, hipStream_t s, const GemmParams<float>& q) {
  if (pipe) hipLaunchKernelGGL((gemm_c_pipe_kernel<AK, BK, CA, CB>), grid, dim3(256), 0, s, q);
  else hipLaunchKernelGGL((gemm_c_dlds_kernel<AK, BK, CA, CB, 3, 2>), grid, dim3(256), 0, s, q);
}

hipError_t launch_gemm_c_group(const GemmParams<float>* tiles, int count, hipStream_t s) {
  SgemmGroup g;
  int wg = 0;
  if (count <= 0) return hipSuccess;
  if (!make_group(tiles, count, 128, 128, cgemm_groupable, g, wg)) return hipErrorInvalidValue;
  const bool akc = is_kc_a(g.base.opa), bkc = is_kc_b(g.base.opb);
  const bool ca = g.base.opa == 'C', cb = g.base.opb == 'C';
  // the pipelined tile when the group has two tiles per CU to fill it
  const bool pipe = sgemm_pipe(wg);
#define TMM_CG(AK, BK, CA, CB) launch_cg<AK, BK, CA, CB>(pipe, wg, s, g)
  if (akc && bkc) {
    if (ca) TMM_CG(true, true, true, false);
    else TMM_CG(true, true, false, false);
  } else if (akc) {
    if (ca && cb) TMM_CG(true, false, true, true);
    else if (ca) TMM_CG(true, false, true, false);
    else if (cb) TMM_CG(true, false, false, true);
    else TMM_CG(true, false, false, false);
  } else if (bkc) {
    TMM_CG(false, true, false, false);
  } else {
    if (cb) TMM_CG(false, false, false, true);
    else TMM_CG(false, false, false, false);
  }
#undef TMM_CG
  return hipGetLastError();
}

hipError_t launch_gemm_d_batched(char opa, char opb, const GemmParams<double>* probs, const int* items, int n_items,
                                 int* counter, int n_workgroups, hipStream_t s, int body) {
  hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(64), 0, s, counter);
  if (n_items <= 0) return hipGetLastError();
  dim3 grid(n_workgroups), block(256);
  const bool akc = is_kc_a(opa), bkc = is_kc_b(opb);
  if (body == 1) {
    if (akc && bkc) hipLaunchKernelGGL((gemm_d_dlds_batched_kernel<true, true, 3>), grid, block, 0, s, probs, items, n_items, counter);
    else if (akc) hipLaunchKernelGGL((gemm_d_dlds_batched_kernel<true, false, 3>), grid, block, 0, s, probs, items, n_items, counter);
    else if (bkc) hipLaunchKernelGGL((gemm_d_dlds_batched_kernel<false, true, 3>), grid, block, 0, s, probs, items, n_items, counter);
    else hipLaunchKernelGGL((gemm_d_dlds_batched_kernel<false, false, 3>), grid, block, 0, s, probs, items, n_items, counter);
    return hipGetLastError();
  }
  if (akc && bkc) hipLaunchKernelGGL((gemm_real_batched_kernel<double, true, true>), grid, block, 0, s, probs, items, n_items, counter);
  else if (akc) hipLaunchKernelGGL((gemm_real_batched_kernel<double, true, false>), grid, block, 0, s, probs, items, n_items, counter);
  else if (bkc) hipLaunchKernelGGL((gemm_real_batched_kernel<double, false, true>), grid, block, 0, s, probs, items, n_items, counter);
  else hipLaunchKernelGGL((gemm_real_batched_kernel<double, false, false>), grid, block, 0, s, probs, items, n_items, counter);
  return hipGetLastError();
}

hipError_t launch_gemm_d(const GemmParams<double>& p, hipStream_t s) {
  if (p.m <= 0 || p.n <= 0) return hipSuccess;
  if (p.variant == 1) return launch_gemm_d_direct(p, s);
  return launch_real<double>(p, s);
}
// FP32 through the direct-LDS kernel for the largest multiple of 16 in K; a remainder of K goes to
// gemm_real_kernel, accumulating (beta only on the first part).
hipError_t launch_gemm_s_direct(const GemmParams<float>& p, hipStream_t s) {
  const bool akc = is_kc_a(p.opa), bkc = is_kc_b(p.opb);
  const int k16 = p.k / 16 * 16;
  // a row-contiguous operand needs whole 4-row chunks
  const bool rows_ok = (akc || p.m % 4 == 0) && (bkc || p.n % 4 == 0);
  if (k16 == 0 || !rows_ok) return launch_real<float>(p, s);
  GemmParams<float> q = p;
  q.k = k16;
  // 256 x 256 workgroup tiles (8 waves of 64 x 128, one workgroup per CU) once m and n allow:
  // SGEMM TN 16384^3 142.3, 40000^3 140.0, 8064^2 x 8000 134.0 TFLOP/s against 135.8, 134.5, 127.8
  // for 256 x 128 tiles (8 waves of 64 x 64, two per CU; 32 against 48 bytes of LDS writes per
  // MFMA); 256 x 128 in turn ran NN 138.2 and TN 135.1 against 134.8 and 134.3 for 128 x 128 tiles
  // at three per CU
  if (sgemm_pipe(p.m, p.n)) {
    const dim3 grid(dgemm_tiles(p.m, p.n, 128)), block(256);
    if (akc && bkc) hipLaunchKernelGGL((gemm_s_pipe_kernel<true, true>), grid, block, 0, s, q);
    else if (akc) hipLaunchKernelGGL((gemm_s_pipe_kernel<true, false>), grid, block, 0, s, q);
    else if (bkc) hipLaunchKernelGGL((gemm_s_pipe_kernel<false, true>), grid, block, 0, s, q);
    else hipLaunchKernelGGL((gemm_s_pipe_kernel<false, false>), grid, block, 0, s, q);
  } else if (p.m > 128 && p.n > 128) {
    const dim3 grid(((p.m + 255) / 256) * ((p.n + 255) / 256)), block(512);
    if (akc && bkc) hipLaunchKernelGGL((gemm_s_dlds_kernel<true, true, 3, 1, 256, 256>), grid, block, 0, s, q);
    else if (akc) hipLaunchKernelGGL((gemm_s_dlds_kernel<true, false, 3, 1, 256, 256>), grid, block, 0, s, q);
    else if (bkc) hipLaunchKernelGGL((gemm_s_dlds_kernel<false, true, 3, 1, 256, 256>), grid, block, 0, s, q);
    else hipLaunchKernelGGL((gemm_s_dlds_kernel<false, false, 3, 1, 256, 256>), grid, block, 0, s, q);
  } else if (p.m > 128) {
    const dim3 grid(((p.m + 255) / 256) * ((p.n + 127) / 128)), block(512);
    if (akc && bkc) hipLaunchKernelGGL((gemm_s_dlds_kernel<true, true, 3, 2, 256>), grid, block, 0, s, q);
    else if (akc) hipLaunchKernelGGL((gemm_s_dlds_kernel<true, false, 3, 2, 256>), grid, block, 0, s, q);
    else if (bkc) hipLaunchKernelGGL((gemm_s_dlds_kernel<false, true, 3, 2, 256>), grid, block, 0, s, q);
    else hipLaunchKernelGGL((gemm_s_dlds_kernel<false, false, 3, 2, 256>), grid, block, 0, s, q);
  } else {
    const dim3 grid((p.n + 127) / 128), block(256);
    if (akc && bkc) hipLaunchKernelGGL((gemm_s_dlds_kernel<true, true, 3, 3, 128>), grid, block, 0, s, q);
    else if (akc) hipLaunchKernelGGL((gemm_s_dlds_kernel<true, false, 3, 3, 128>), grid, block, 0, s, q);
    else if (bkc) hipLaunchKernelGGL((gemm_s_dlds_kernel<false, true, 3, 3, 128>), grid, block, 0, s, q);
    else hipLaunchKernelGGL((gemm_s_dlds_kernel<false, false, 3, 3, 128>), grid, block, 0, s, q);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || k16 == p.k) return e;
  GemmParams<float> r = p;
  r.k = p.k - k16;
  r.a = p.a + (akc ? (long)k16 : (long)k16 * p.lda);
  r.b = p.b + (bkc ? (long)k16 : (long)k16 * p.ldb);
  r.accumulate = 1;
  r.beta = 0.0f;
  return launch_real<float>(r, s);
}

hipError_t launch_gemm_s(const GemmParams<float>& p, hipStream_t s) {
  if (p.m <= 0 || p.n <= 0) return hipSuccess;
  if (p.variant == 1) return launch_gemm_s_direct(p, s);
  return launch_real<float>(p, s);
}
hipError_t launch_gemm_z(const GemmParams<double>& p, hipStream_t s) {
  if (p.m <= 0 || p.n <= 0) return hipSuccess;
  if (p.variant == 0) return launch_cplx<double>(p, s);
  // direct-LDS kernel on the largest multiple of 8 in K, the remainder through gemm_cplx_kernel
  // accumulating (beta only on the first part)
  const int k8 = p.k / 8 * 8;
  if (k8 == 0) return launch_cplx<double>(p, s);
  const bool akc = is_kc_a(p.opa), bkc = is_kc_b(p.opb);
  GemmParams<double> q = p;
  q.k = k8;
  const dim3 grid(((p.m + 63) / 64) * ((p.n + 63) / 64)), block(256);
  const bool ca = p.opa == 'C', cb = p.opb == 'C';
  // 3-stage ring, 3 workgroups per CU: 69.7 TFLOP/s on CN 8192^3 against 67.9 for 4 stages at 2
  const bool pipe = (long)grid.x >= 2L * device_cus();  // as for the groups
#define TMM_ZDLDS(AK, BK, CA, CB)                                                              \
  do {                                                                                         \
    if (pipe) hipLaunchKernelGGL((gemm_z_pipe_kernel<AK, BK, CA, CB>), grid, block, 0, s, q);  \
    else hipLaunchKernelGGL((gemm_z_dlds_kernel<AK, BK, CA, CB, 3, 3>), grid, block, 0, s, q); \
  } while (0)
  if (akc && bkc) {
    if (ca) TMM_ZDLDS(true, true, true, false);
    else TMM_ZDLDS(true, true, false, false);
  } else if (akc) {
    if (ca && cb) TMM_ZDLDS(true, false, true, true);
    else if (ca) TMM_ZDLDS(true, false, true, false);
    else if (cb) TMM_ZDLDS(true, false, false, true);
    else TMM_ZDLDS(true, false, false, false);
  } else if (bkc) {
    TMM_ZDLDS(false, true, false, false);
  } else {
    if (cb) TMM_ZDLDS(false, false, false, true);
    else TMM_ZDLDS(false, false, false, false);
  }
#undef TMM_ZDLDS
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || k8 == p.k) return e;
  GemmParams<double> r = p;
  r.k = p.k - k8;
  r.a = p.a + 2 * (akc ? (long)k8 : (long)k8 * p.lda);
  r.b = p.b + 2 * (bkc ? (long)k8 : (long)k8 * p.ldb);
  r.accumulate = 1;
  r.beta = 0.0;
  r.beta_im = 0.0;
  return launch_cplx<double>(r, s);
}
hipError_t launch_gemm_c(const GemmParams<float>& p, hipStream_t s) {
  if (p.m <= 0 || p.n <= 0) return hipSuccess;
  const bool akc = is_kc_a(p.opa), bkc = is_kc_b(p.opb);
  const int k8 = p.k / 8 * 8;
  // a row-contiguous operand needs whole 2-row chunks
  const bool rows_ok = (akc || p.m % 2 == 0) && (bkc || p.n % 2 == 0);
  if (p.variant != 1 || k8 == 0 || !rows_ok) return launch_cplx<float>(p, s);
  // direct-LDS kernel on the largest multiple of 8 in K, the remainder through gemm_cplx_kernel
  // accumulating (beta only on the first part)
  GemmParams<float> q = p;
  q.k = k8;
  const dim3 grid(((p.m + 127) / 128) * ((p.n + 127) / 128));
  const bool ca = p.opa == 'C', cb = p.opb == 'C';
  const bool pipe = sgemm_pipe((long)grid.x);  // as for the groups
#define TMM_CDLDS(AK, BK, CA, CB) launch_c1<AK, BK, CA, CB>(pipe, grid, s, q)
  if (akc && bkc) {
    if (ca) TMM_CDLDS(true, true, true, false);
    else TMM_CDLDS(true, true, false, false);
  } else if (akc) {
    if (ca && cb) TMM_CDLDS(true, false, true, true);
    else if (ca) TMM_CDLDS(true, false, true, false);
    else if (cb) TMM_CDLDS(true, false, false, true);
    else TMM_CDLDS(true, false, false, false);
  } else if (bkc) {
    TMM_CDLDS(false, true, false, false);
  } else {
    if (cb) TMM_CDLDS(false, false, false, true);
    else TMM_CDLDS(false, false, false, false);
  }
#undef TMM_CDLDS
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || k8 == p.k) return e;
  GemmParams<float> r = p;
  r.k = p.k - k8;
  r.a = p.a + 2 * (akc ? (long)k8 : (long)k8 * p.lda);
  r.b = p.b + 2 * (bkc ? (long)k8 : (long)k8 * p.ldb);
  r.accumulate = 1;
  r.beta = 0.0f;
  r.beta_im = 0.0f;
  return launch_cplx<float>(r, s);
}

}  // namespace tmm

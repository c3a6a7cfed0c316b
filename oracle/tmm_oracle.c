/*
 * tmm_oracle.c - CPU restatement of Tiled-MM's host-to-host GEMM path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker that the tests, the smoke check and the
 * benchmark's CPU-baseline leg compare the library against; no product code links it.
 *
 * What is restated (paths relative to the Tiled-MM source tree):
 *   - tile-size policy ............ src/Tiled-MM/mm_handle.cpp:89-110 (optimal_tile_size),
 *                                   :117-133 (optimal_tile_sizes)
 *   - host tile geometry .......... src/Tiled-MM/tiled_matrix.cpp:7-90 (clamped tile dims, ceil
 *                                   tile counts, short edge tiles, tile offset), with 64-bit
 *                                   offsets instead of int
 *   - transposed tile coordinates . src/Tiled-MM/tiled_mm.cpp:32-42, 126-165
 *   - tile packing ................ src/Tiled-MM/tiled_mm.cpp:45-66 (device pitch = tile rows)
 *   - round-robin GEMM semantics .. src/Tiled-MM/tiled_mm.cpp:270-365 and 367-454: for every C
 *                                   tile, k-tiles in order, C_t = alpha*op(A_t)*op(B_t) + beta'*C_t
 *                                   with beta' = beta on k-tile 0 and 1 afterwards; the caller's
 *                                   C tile is read only when beta != 0 (:325, :423)
 *   - per-tile arithmetic ......... the vendor {s,d,c,z}gemm the reference calls
 *                                   (:181-268): BLAS semantics, summed in working precision in
 *                                   natural k order; op 'C' conjugates (get_blas_operation,
 *                                   :168-179), any letter other than 'T'/'C' means 'N'; alpha = 0
 *                                   gives C_t = beta'*C_t without reading the A and B tiles (the
 *                                   reference BLAS dgemm's quick path, which the vendor
 *                                   libraries keep).
 *
 * How it is pinned (tests/test_oracle.py): the policy and geometry against the reference's
 * own sources compiled into oracle/_ref and against golden tables; the GEMM against exact
 * integer known-answer tests (the reference test's data style, tests/test-multiply.cpp:58-98),
 * numpy's BLAS, brute force in exact rational arithmetic on tiny sizes, and algebraic
 * identities (identity matrices, transposes, alpha = 0, beta = 0 with NaN in C).
 */
#include <complex.h>
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

/* ---- tile policy (mm_handle.cpp:89-110) ---- */
int oracle_optimal_tile_size(int dim, int curr_tile_size, int max_tile_size) {
  if (curr_tile_size < max_tile_size && dim <= max_tile_size) return dim;
  int tile = 1;
  int limit = max_tile_size < dim ? max_tile_size : dim;
  for (int i = 1; i <= limit; ++i)
    if (dim % i == 0) tile = i;
  if (abs(max_tile_size - tile) <= max_tile_size / 2) return tile;
  return limit;
}

/* ---- geometry (tiled_matrix.cpp:7-90) ----
 * A rows x cols matrix split into tiles of (tile_rows x tile_cols), both clamped to the
 * matrix (tiled_matrix.cpp:13-14). Returns the extent of tile (ti, tj) and its element
 * offset in the host matrix of leading dimension ld. */
static int clamp_tile(int t, int len) { return t < len ? t : len; }

int oracle_num_tiles(int len, int tile) {
  tile = clamp_tile(tile, len);
  return tile > 0 ? (len + tile - 1) / tile : 0;
}

/* actual_size (tiled_matrix.cpp:74-80): last tile is short when len % tile != 0 */
int oracle_tile_extent(int len, int tile, int idx) {
  tile = clamp_tile(tile, len);
  int n = oracle_num_tiles(len, tile);
  int rem = len % tile;
  return (idx == n - 1 && rem > 0) ? rem : tile;
}

/* tile_offset (tiled_matrix.cpp:61-67), 64-bit */
long long oracle_tile_offset(int rows, int cols, long long ld, int tile_rows, int tile_cols, int ti, int tj) {
  tile_rows = clamp_tile(tile_rows, rows);
  tile_cols = clamp_tile(tile_cols, cols);
  return (long long)tj * tile_cols * ld + (long long)ti * tile_rows;
}

/* copy_tile_to_device_async (tiled_mm.cpp:45-66): tile (ti, tj) packed with pitch = its rows. */
void oracle_pack_tile_d(const double* host, int rows, int cols, long long ld, int tile_rows, int tile_cols, int ti,
                        int tj, double* out, int* out_rows, int* out_cols) {
  int r = oracle_tile_extent(rows, tile_rows, ti);
  int c = oracle_tile_extent(cols, tile_cols, tj);
  const double* src = host + oracle_tile_offset(rows, cols, ld, tile_rows, tile_cols, ti, tj);
  for (int j = 0; j < c; ++j)
    for (int i = 0; i < r; ++i) out[i + (size_t)j * r] = src[i + (size_t)j * ld];
  *out_rows = r;
  *out_cols = c;
}

/* ---- tiled GEMM (round_robin semantics) ---- */
static char norm_op(char t) {
  if (t >= 'a' && t <= 'z') t = (char)(t - 'a' + 'A');
  return (t == 'T' || t == 'C') ? t : 'N';
}

#define DEFINE_TILED_GEMM(NAME, T, CONJ)                                                                            \
  void NAME(char ta, char tb, int m, int n, int k, T alpha, const T* A, long long lda, const T* B, long long ldb,    \
            T beta, T* C, long long ldc, int tile_m, int tile_n, int tile_k) {                                         \
    ta = norm_op(ta);                                                                                                 \
    tb = norm_op(tb);                                                                                                 \
    if (m <= 0 || n <= 0) return;                                                                                     \
    /* tiles clamped to the problem, as tiled_matrix does */                                                          \
    int tm = clamp_tile(tile_m, m), tn = clamp_tile(tile_n, n), tk = k > 0 ? clamp_tile(tile_k, k) : 1;              \
    int nTm = (m + tm - 1) / tm, nTn = (n + tn - 1) / tn, nTk = k > 0 ? (k + tk - 1) / tk : 0;                       \
    int beta_nz = beta != (T)0;                                                                                       \
    T* ctile = (T*)malloc(sizeof(T) * (size_t)tm * tn);                                                               \
    for (int t = 0; t < nTm * nTn; ++t) {                                                                             \
      /* tile order of round_robin: m_tile = i / nTn, n_tile = i % nTn (tiled_mm.cpp:300-301) */                      \
      int it = t / nTn, jt = t % nTn;                                                                                 \
      int r0 = it * tm, c0 = jt * tn;                                                                                 \
      int rows = oracle_tile_extent(m, tm, it), cols = oracle_tile_extent(n, tn, jt);                                 \
      /* device C tile: caller's values when beta != 0, otherwise never read */                                        \
      for (int j = 0; j < cols; ++j)                                                                                  \
        for (int i = 0; i < rows; ++i) ctile[i + (size_t)j * rows] = beta_nz ? C[(r0 + i) + (size_t)(c0 + j) * ldc] : (T)0; \
      if (nTk == 0) { /* k == 0: C = beta*C (BLAS) */                                                                 \
        for (int j = 0; j < cols; ++j)                                                                                \
          for (int i = 0; i < rows; ++i) ctile[i + (size_t)j * rows] = beta_nz ? beta * ctile[i + (size_t)j * rows] : (T)0; \
      }                                                                                                               \
      for (int kt = 0; kt < nTk; ++kt) {                                                                              \
        int k0 = kt * tk, kc = oracle_tile_extent(k, tk, kt);                                                         \
        T bp = kt == 0 ? beta : (T)1;                                                                                 \
        for (int j = 0; j < cols; ++j)                                                                                \
          for (int i = 0; i < rows; ++i) {                                                                            \
            T s = 0;                                                                                                  \
            if (alpha == (T)0) { /* A and B not read */                                                               \
              T* cz = &ctile[i + (size_t)j * rows];                                                                   \
              *cz = (bp == (T)0) ? (T)0 : bp * *cz;                                                                   \
              continue;                                                                                               \
            }                                                                                                         \
            for (int kk = 0; kk < kc; ++kk) {                                                                         \
              long long gi = r0 + i, gj = c0 + j, gk = k0 + kk;                                                       \
              T a = ta == 'N' ? A[gi + gk * lda] : A[gk + gi * lda];                                                  \
              T b = tb == 'N' ? B[gk + gj * ldb] : B[gj + gk * ldb];                                                  \
              if (ta == 'C') a = CONJ(a);                                                                             \
              if (tb == 'C') b = CONJ(b);                                                                             \
              s += a * b;                                                                                             \
            }                                                                                                         \
            T* cp = &ctile[i + (size_t)j * rows];                                                                     \
            /* BLAS: beta' == 0 means C is not read */                                                                \
            *cp = (bp == (T)0) ? alpha * s : alpha * s + bp * *cp;                                                    \
          }                                                                                                           \
      }                                                                                                               \
      for (int j = 0; j < cols; ++j)                                                                                  \
        for (int i = 0; i < rows; ++i) C[(r0 + i) + (size_t)(c0 + j) * ldc] = ctile[i + (size_t)j * rows];          \
    }                                                                                                                 \
    free(ctile);                                                                                                      \
  }

#define REAL_CONJ(x) (x)
DEFINE_TILED_GEMM(oracle_sgemm, float, REAL_CONJ)
DEFINE_TILED_GEMM(oracle_dgemm, double, REAL_CONJ)
DEFINE_TILED_GEMM(oracle_cgemm_impl, float complex, conjf)
DEFINE_TILED_GEMM(oracle_zgemm_impl, double complex, conj)

/* complex entry points take alpha/beta as (re, im) pairs so they are callable through ctypes */
void oracle_cgemm(char ta, char tb, int m, int n, int k, const float* alpha, const float complex* A, long long lda,
                  const float complex* B, long long ldb, const float* beta, float complex* C, long long ldc,
                  int tile_m, int tile_n, int tile_k) {
  oracle_cgemm_impl(ta, tb, m, n, k, alpha[0] + alpha[1] * I, A, lda, B, ldb, beta[0] + beta[1] * I, C, ldc, tile_m,
                    tile_n, tile_k);
}

void oracle_zgemm(char ta, char tb, int m, int n, int k, const double* alpha, const double complex* A, long long lda,
                  const double complex* B, long long ldb, const double* beta, double complex* C, long long ldc,
                  int tile_m, int tile_n, int tile_k) {
  oracle_zgemm_impl(ta, tb, m, n, k, alpha[0] + alpha[1] * I, A, lda, B, ldb, beta[0] + beta[1] * I, C, ldc, tile_m,
                    tile_n, tile_k);
}

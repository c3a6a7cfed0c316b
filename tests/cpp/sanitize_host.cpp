// Host-side planner and CPU plan executor under AddressSanitizer + UndefinedBehaviorSanitizer.
// Built by tests/test_sanitizers.py with g++ from tiled-mm_amd/csrc/tmm_plan.cpp and
// tmm_cpu_exec.cpp (no HIP): every plan below is built, its operand-reuse plan and tile groups
// are derived, and it is executed on the CPU with several random queue interleavings; the result
// must equal a naive host GEMM exactly (small integer data). Prints "SANITIZE HOST OK".
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

extern "C" {
const char* tmmh_last_error();
void tmmh_set_defer_cin(int d);
int tmmh_plan_info(char opa, char opb, int m, int n, int k, int lda, int ldb, int ldc, int beta_nonzero, int copy_back,
                   int tm, int tn, int tk, int nq, long budget, int first_chunk, long* info);
long tmmh_operand_reuse(char opa, char opb, int m, int n, int k, int lda, int ldb, int ldc, int beta_nonzero,
                        int copy_back, int tm, int tn, int tk, int nq, long budget, int first_chunk, int capacity,
                        int max_k, long* out);
int tmmh_tile_groups(char opa, char opb, int m, int n, int k, int lda, int ldb, int ldc, int beta_nonzero,
                     int copy_back, int tm, int tn, int tk, int nq, long budget, int first_chunk, int max_group,
                     int unsafe, long* info);
int tmmh_simulate_gemm(int dtype, char opa, char opb, int m, int n, int k, const void* alpha, const void* A, int lda,
                       const void* B, int ldb, const void* beta, void* C, int ldc, int copy_back, void* fullc, int tm,
                       int tn, int tk, int nq, long budget, int first_chunk, unsigned seed);
}

namespace {

int failures = 0;

template <typename T> T conj_if(T v, bool) { return v; }
template <typename R> std::complex<R> conj_if(std::complex<R> v, bool c) { return c ? std::conj(v) : v; }

template <typename T>
T op_at(const std::vector<T>& M, int ld, char op, int i, int kk) {  // op(M)(i, kk)
  return op == 'N' ? M[i + (size_t)kk * ld] : conj_if(M[kk + (size_t)i * ld], op == 'C');
}

template <typename T>
void run_case(int dtype, char ta, char tb, int m, int n, int k, int pad, int beta_nz, int copy_back, int tile, int nq,
              long budget, int first, int defer) {
  const int lda = (ta == 'N' ? m : k) + pad, ldb = (tb == 'N' ? k : n) + pad, ldc = m + pad;
  const int acols = ta == 'N' ? k : m, bcols = tb == 'N' ? n : k;
  std::vector<T> A((size_t)lda * acols), B((size_t)ldb * bcols), C0((size_t)ldc * n);
  unsigned s = 12345u + m * 7 + n * 13 + k * 17;
  auto next = [&]() { s = s * 1103515245u + 12345u; return (int)((s >> 16) % 7) - 3; };
  for (auto& v : A) v = T(next());
  for (auto& v : B) v = T(next());
  for (auto& v : C0) v = T(next());
  const T alpha = T(2), beta = beta_nz ? T(-1) : T(0);
  std::vector<T> ref(C0);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      T acc = T(0);
      for (int kk = 0; kk < k; ++kk) acc += op_at(A, lda, ta, i, kk) * op_at(B, ldb, tb, kk, j);
      ref[i + (size_t)j * ldc] = alpha * acc + (beta_nz ? beta * C0[i + (size_t)j * ldc] : T(0));
    }
  tmmh_set_defer_cin(defer);
  long info[11];
  if (tmmh_plan_info(ta, tb, m, n, k, lda, ldb, ldc, beta_nz, copy_back, tile, tile, tile, nq, budget, first, info)) {
    std::printf("plan failed: %s\n", tmmh_last_error());
    ++failures;
    return;
  }
  const long nops = tmmh_operand_reuse(ta, tb, m, n, k, lda, ldb, ldc, beta_nz, copy_back, tile, tile, tile, nq,
                                       budget, first, 0, 0, nullptr);
  std::vector<long> reuse(8 * (size_t)(nops > 0 ? nops : 1));
  for (int cap : {1, 2, 3, 64})
    if (tmmh_operand_reuse(ta, tb, m, n, k, lda, ldb, ldc, beta_nz, copy_back, tile, tile, tile, nq, budget, first,
                           cap, 1 << 30, reuse.data()) != nops) {
      std::printf("operand reuse failed: %s\n", tmmh_last_error());
      ++failures;
    }
  long ginfo[3];
  if (tmmh_tile_groups(ta, tb, m, n, k, lda, ldb, ldc, beta_nz, copy_back, tile, tile, tile, nq, budget, first, 4, 0,
                       ginfo) || ginfo[2] != 1) {
    std::printf("tile groups failed or stalled\n");
    ++failures;
  }
  for (unsigned seed = 0; seed < 3; ++seed) {
    std::vector<T> C(C0), full((size_t)m * n, T(0));
    if (tmmh_simulate_gemm(dtype, ta, tb, m, n, k, &alpha, A.data(), lda, B.data(), ldb, &beta, C.data(), ldc,
                           copy_back, full.data(), tile, tile, tile, nq, budget, first, seed)) {
      std::printf("simulate failed: %s\n", tmmh_last_error());
      ++failures;
      return;
    }
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < m; ++i) {
        const T got = copy_back ? C[i + (size_t)j * ldc] : full[i + (size_t)j * m];
        if (got != ref[i + (size_t)j * ldc]) {
          std::printf("mismatch dtype %d %c%c %dx%dx%d tile %d nq %d copy_back %d defer %d seed %u at (%d,%d)\n", dtype,
                      ta, tb, m, n, k, tile, nq, copy_back, defer, seed, i, j);
          ++failures;
          return;
        }
      }
  }
}

}  // namespace

int main() {
  const char ops[] = {'N', 'T'};
  const int shapes[][3] = {{1, 1, 1}, {5, 2, 2}, {50, 200, 21}, {37, 29, 53}, {64, 64, 64}, {17, 90, 33}};
  for (auto& sh : shapes)
    for (char ta : ops)
      for (char tb : ops)
        for (int cb = 0; cb < 2; ++cb)
          for (int bz = 0; bz < 2; ++bz)
            for (int tile : {4, 16}) {
              const int defer = (sh[0] + tile) % 3;
              run_case<double>(1, ta, tb, sh[0], sh[1], sh[2], sh[0] % 3, bz, cb, tile, 2, 0, 0, defer);
            }
  // complex with conjugation, 1 and 3 compute queues, a first chunk, a small pool budget
  for (char ta : {'N', 'C'})
    for (char tb : {'T', 'C'}) {
      run_case<std::complex<double>>(3, ta, tb, 33, 21, 45, 1, 1, 1, 8, 3, 0, 0, 1);
      run_case<std::complex<double>>(3, ta, tb, 33, 21, 45, 0, 1, 0, 8, 1, 0, 4, 0);
    }
  run_case<double>(1, 'T', 'N', 40, 40, 40, 2, 1, 1, 8, 2, 3 * 8 * 8 * 4, 0, 1);
  run_case<float>(0, 'N', 'N', 30, 20, 70, 1, 1, 1, 16, 4, 0, 16, 2);
  tmmh_set_defer_cin(0);
  if (failures) {
    std::printf("%d failures\n", failures);
    return 1;
  }
  std::printf("SANITIZE HOST OK\n");
  return 0;
}

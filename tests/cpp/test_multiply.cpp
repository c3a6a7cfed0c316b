// C++ drop-in check: the reference's test program flow (Tiled-MM tests/test-multiply.cpp:116-351)
// against this library's gpu:: API, with the same command-line flags and the same pass/fail rule,
// but an independent host reference in place of the reference's device-BLAS call: the data are
// integers in [0, 9] (test-multiply.cpp:58-66), so every dot product is an integer below 2^53 and
// the reference, summed in double, is exact. Exit status 0 = CORRECT.
//
// One flag of its own: --check_cols N compares N evenly spaced columns of C (first and last
// included) instead of all of them, so that the reference's large cases (tests/CMakeLists.txt:12,14,
// up to 12345 x 23456 x 67891) are checked in seconds. The fill and the reference run on
// OMP_NUM_THREADS threads (default: all hardware threads).
#include <Tiled-MM/tiled_mm.hpp>
#include <Tiled-MM/util.hpp>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

namespace {

int arg_int(int argc, char** argv, const char* name, int def) {
  for (int i = 1; i + 1 < argc; ++i)
    if (!std::strcmp(argv[i], name)) return std::atoi(argv[i + 1]);
  return def;
}

double arg_double(int argc, char** argv, const char* name, double def) {
  for (int i = 1; i + 1 < argc; ++i)
    if (!std::strcmp(argv[i], name)) return std::atof(argv[i + 1]);
  return def;
}

std::string arg_str(int argc, char** argv, const char* name, const char* def) {
  for (int i = 1; i + 1 < argc; ++i)
    if (!std::strcmp(argv[i], name)) return argv[i + 1];
  return def;
}

int n_threads() {
  const char* e = std::getenv("OMP_NUM_THREADS");
  int t = e ? std::atoi(e) : 0;
  if (t <= 0) t = (int)std::thread::hardware_concurrency();
  return std::max(1, std::min(t, 64));
}

template <typename F>
void parallel_for(long n, F&& f) {
  const int nt = (int)std::min<long>(n_threads(), std::max(1L, n));
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&, t]() {
      for (long i = n * t / nt; i < n * (t + 1) / nt; ++i) f(i);
    });
  for (auto& x : th) x.join();
}

// Integers in [0, 9]: chunks of 2^20 entries, each drawn from its own mt19937 seeded from 42 and
// the chunk index, so the data do not depend on the thread count.
void fill_matrix(double* p, size_t n, unsigned salt) {
  const size_t chunk = 1u << 20;
  parallel_for((long)((n + chunk - 1) / chunk), [&](long c) {
    std::mt19937 rng(42u + 7919u * salt + 104729u * (unsigned)c);
    std::uniform_int_distribution<int> dist(0, 9);
    const size_t e = std::min(n, (size_t)(c + 1) * chunk);
    for (size_t i = (size_t)c * chunk; i < e; ++i) p[i] = dist(rng);
  });
}

// c(:, j) = beta*c(:, j) + alpha*op(A)*op(B)(:, j) for the listed columns j. The dot products are
// integers below 2^53 (entries in [0, 9], k < 2^31), so summing them in double is exact.
void reference(const double* a, const double* b, double* c, int m, int k, int lda, int ldb, int ldc, double alpha,
               double beta, char ta, char tb, const std::vector<int>& cols) {
  const int rb = 256;
  const int nblk = (m + rb - 1) / rb;
  const size_t nc = cols.size();
  parallel_for(nblk, [&](long blk) {
    const int i0 = (int)blk * rb, i1 = std::min(m, i0 + rb);
    std::vector<double> acc((size_t)(i1 - i0) * nc, 0.0);
    auto bval = [&](int p, int j) { return tb == 'N' ? b[p + (size_t)j * ldb] : b[j + (size_t)p * ldb]; };
    if (ta == 'N') {  // columns of A are contiguous in i: axpy form
      for (int p = 0; p < k; ++p) {
        const double* ap = a + (size_t)p * lda;
        for (size_t q = 0; q < nc; ++q) {
          const double bv = bval(p, cols[q]);
          if (bv == 0) continue;
          double* out = acc.data() + q * (i1 - i0);
          for (int i = i0; i < i1; ++i) out[i - i0] += ap[i] * bv;
        }
      }
    } else {  // rows of op(A) are contiguous in p: dot form
      std::vector<double> bcol(k);
      for (size_t q = 0; q < nc; ++q) {
        for (int p = 0; p < k; ++p) bcol[p] = bval(p, cols[q]);
        for (int i = i0; i < i1; ++i) {
          const double* ar = a + (size_t)i * lda;
          double s = 0;
          for (int p = 0; p < k; ++p) s += ar[p] * bcol[p];
          acc[q * (i1 - i0) + (i - i0)] = s;
        }
      }
    }
    for (size_t q = 0; q < nc; ++q)
      for (int i = i0; i < i1; ++i) {
        double& cv = c[i + (size_t)cols[q] * ldc];
        const long double s = acc[q * (i1 - i0) + (i - i0)];
        cv = (double)(alpha * s + (beta != 0 ? beta * (long double)cv : 0.0L));
      }
  });
}

}  // namespace

int main(int argc, char** argv) {
  const int m = arg_int(argc, argv, "-m", 1000), n = arg_int(argc, argv, "-n", 1000), k = arg_int(argc, argv, "-k", 1000);
  const int tile_m = arg_int(argc, argv, "--tile_m", 5000), tile_n = arg_int(argc, argv, "--tile_n", 5000),
            tile_k = arg_int(argc, argv, "--tile_k", 5000);
  const int n_streams = arg_int(argc, argv, "--n_streams", 2);
  std::string transpose = arg_str(argc, argv, "--transpose", "NN");
  std::transform(transpose.begin(), transpose.end(), transpose.begin(), [](char ch) { return (char)std::toupper(ch); });
  const double alpha = arg_double(argc, argv, "--alpha", 1.0), beta = arg_double(argc, argv, "--beta", 0.0);
  const char trans_a = transpose[0], trans_b = transpose[1];
  const int a_m = trans_a == 'N' ? m : k, a_n = trans_a == 'N' ? k : m;
  const int b_m = trans_b == 'N' ? k : n, b_n = trans_b == 'N' ? n : k;
  const int ld_a = std::max(a_m, arg_int(argc, argv, "--ld_a", 0));
  const int ld_b = std::max(b_m, arg_int(argc, argv, "--ld_b", 0));
  const int ld_c = std::max(m, arg_int(argc, argv, "--ld_c", 0));
  const int check_cols = arg_int(argc, argv, "--check_cols", 0);

  std::vector<int> cols;
  if (check_cols <= 0 || check_cols >= n) {
    for (int j = 0; j < n; ++j) cols.push_back(j);
  } else {
    for (int q = 0; q < check_cols; ++q) cols.push_back((int)((long)q * (n - 1) / (check_cols - 1 > 0 ? check_cols - 1 : 1)));
    cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
  }
  std::printf("test-multiply: m=%d n=%d k=%d op=%c%c tiles=%d,%d,%d streams=%d alpha=%g beta=%g, checking %zu columns\n",
              m, n, k, trans_a, trans_b, tile_m, tile_n, tile_k, n_streams, alpha, beta, cols.size());
  std::fflush(stdout);

  auto a_host = gpu::malloc_pinned<double>((size_t)ld_a * a_n, 1);
  auto b_host = gpu::malloc_pinned<double>((size_t)ld_b * b_n, 1);
  auto c_host = gpu::malloc_pinned<double>((size_t)ld_c * n, 0);
  auto c_host2 = gpu::malloc_pinned<double>((size_t)ld_c * n, 0);
  auto c_ref = gpu::malloc_pinned<double>((size_t)ld_c * n, 0);
  fill_matrix(a_host, (size_t)ld_a * a_n, 1);
  fill_matrix(b_host, (size_t)ld_b * b_n, 2);
  fill_matrix(c_host, (size_t)ld_c * n, 3);
  std::memcpy(c_host2, c_host, sizeof(double) * ld_c * n);
  std::memcpy(c_ref, c_host, sizeof(double) * ld_c * n);
  reference(a_host, b_host, c_ref, m, k, ld_a, ld_b, ld_c, alpha, beta, trans_a, trans_b, cols);

  auto ctx = gpu::make_context<double>(n_streams, tile_m, tile_n, tile_k);
  // with copying C back
  gpu::gemm(*ctx, trans_a, trans_b, m, n, k, alpha, a_host, ld_a, b_host, ld_b, beta, c_host, ld_c, false, true);
  bool correct = true;
  for (size_t q = 0; q < cols.size() && correct; ++q) {
    const int j = cols[q];
    for (int i = 0; i < m; ++i)
      if (std::abs(c_host[i + (size_t)j * ld_c] - c_ref[i + (size_t)j * ld_c]) > 1e-6) {
        std::printf("mismatch (copy back) at (%d, %d): %f vs %f\n", i, j, c_host[i + (size_t)j * ld_c],
                    c_ref[i + (size_t)j * ld_c]);
        correct = false;
        break;
      }
  }
  // without copying C back: the result stays in the resident device C (ld = m)
  gpu::gemm(*ctx, trans_a, trans_b, m, n, k, alpha, a_host, ld_a, b_host, ld_b, beta, c_host2, ld_c, false, false);
  auto& full = ctx->get_full_device_buffer_c();
  if (full.size() != (size_t)m * n) {
    std::printf("resident C has %zu elements, expected %zu\n", full.size(), (size_t)m * n);
    correct = false;
  }
  double* dev_c = gpu::malloc_pinned<double>((size_t)m * n, 0);
  gpu::copy_to_host(full.data(), dev_c, (size_t)m * n);
  for (size_t q = 0; q < cols.size() && correct; ++q) {
    const int j = cols[q];
    for (int i = 0; i < m; ++i)
      if (std::abs(dev_c[i + (size_t)j * m] - c_ref[i + (size_t)j * ld_c]) > 1e-6) {
        std::printf("mismatch (resident C) at (%d, %d)\n", i, j);
        correct = false;
        break;
      }
  }
  std::printf("The result is %s\n", correct ? "CORRECT" : "NOT CORRECT");
  gpu::free_pinned(a_host);
  gpu::free_pinned(b_host);
  gpu::free_pinned(c_host);
  gpu::free_pinned(c_host2);
  gpu::free_pinned(c_ref);
  gpu::free_pinned(dev_c);
  return correct ? 0 : 1;
}

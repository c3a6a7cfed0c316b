// C++ drop-in check: the reference's test program flow (Tiled-MM tests/test-multiply.cpp:116-351)
// against this library's gpu:: API, with the same command-line flags and the same pass/fail rule,
// but an independent host reference (exact integer arithmetic on the integer data in [0, 9]) in
// place of the reference's device-BLAS call. Exit status 0 = CORRECT.
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

// C = beta*C + alpha*op(A)*op(B) in long double: exact for the integer data used here.
void reference(const double* a, const double* b, double* c, int m, int n, int k, int lda, int ldb, int ldc, double alpha,
               double beta, char ta, char tb) {
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      long double s = 0;
      for (int p = 0; p < k; ++p) {
        const double av = ta == 'N' ? a[i + (size_t)p * lda] : a[p + (size_t)i * lda];
        const double bv = tb == 'N' ? b[p + (size_t)j * ldb] : b[j + (size_t)p * ldb];
        s += (long double)av * bv;
      }
      double& cv = c[i + (size_t)j * ldc];
      cv = (double)(alpha * s + (beta != 0 ? beta * (long double)cv : 0.0L));
    }
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

  auto a_host = gpu::malloc_pinned<double>((size_t)ld_a * a_n, 1);
  auto b_host = gpu::malloc_pinned<double>((size_t)ld_b * b_n, 1);
  auto c_host = gpu::malloc_pinned<double>((size_t)ld_c * n, 0);
  auto c_host2 = gpu::malloc_pinned<double>((size_t)ld_c * n, 0);
  auto c_ref = gpu::malloc_pinned<double>((size_t)ld_c * n, 0);
  std::mt19937 rng(42);
  std::uniform_int_distribution<int> dist(0, 9);
  for (size_t i = 0; i < (size_t)ld_a * a_n; ++i) a_host[i] = dist(rng);
  for (size_t i = 0; i < (size_t)ld_b * b_n; ++i) b_host[i] = dist(rng);
  for (size_t i = 0; i < (size_t)ld_c * n; ++i) c_host[i] = dist(rng);
  std::memcpy(c_host2, c_host, sizeof(double) * ld_c * n);
  std::memcpy(c_ref, c_host, sizeof(double) * ld_c * n);
  reference(a_host, b_host, c_ref, m, n, k, ld_a, ld_b, ld_c, alpha, beta, trans_a, trans_b);

  auto ctx = gpu::make_context<double>(n_streams, tile_m, tile_n, tile_k);
  // with copying C back
  gpu::gemm(*ctx, trans_a, trans_b, m, n, k, alpha, a_host, ld_a, b_host, ld_b, beta, c_host, ld_c, false, true);
  bool correct = true;
  for (int j = 0; j < n && correct; ++j)
    for (int i = 0; i < m; ++i)
      if (std::abs(c_host[i + (size_t)j * ld_c] - c_ref[i + (size_t)j * ld_c]) > 1e-6) {
        std::printf("mismatch (copy back) at (%d, %d): %f vs %f\n", i, j, c_host[i + (size_t)j * ld_c],
                    c_ref[i + (size_t)j * ld_c]);
        correct = false;
        break;
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
  for (int j = 0; j < n && correct; ++j)
    for (int i = 0; i < m; ++i)
      if (std::abs(dev_c[i + (size_t)j * m] - c_ref[i + (size_t)j * ld_c]) > 1e-6) {
        std::printf("mismatch (resident C) at (%d, %d)\n", i, j);
        correct = false;
        break;
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

// Host-side utilities for tests and the benchmark (libtmm_hostutil.so): multi-threaded,
// deterministic fills of large column-major matrices. Element (i, j) depends only on
// (seed, i, j), not on the leading dimension or on the thread count, so a sub-block can be
// regenerated on its own to check results.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// uniform in [0, 1) from the stream (seed, idx, lane)
inline double u01(uint64_t seed, uint64_t idx, uint64_t lane) {
  const uint64_t r = splitmix64(splitmix64(seed * 4 + lane) ^ idx);
  return (double)(r >> 11) * (1.0 / 9007199254740992.0);
}

template <typename F>
void parallel_cols(long cols, int threads, F&& f) {
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  if (threads > cols) threads = (int)std::max(1L, cols);
  std::vector<std::thread> th;
  for (int t = 0; t < threads; ++t)
    th.emplace_back([&, t]() {
      const long j0 = cols * t / threads, j1 = cols * (t + 1) / threads;
      for (long j = j0; j < j1; ++j) f(j);
    });
  for (auto& x : th) x.join();
}

}  // namespace

extern "C" {

// Element (i0+i, j0+j) of a conceptual matrix with `total_rows` rows, written at p[i + j*ld]:
// the block [i0, i0+rows) x [j0, j0+cols) of the full matrix. Real: lo + (hi-lo)*u.
void tmmu_fill_uniform_d(double* p, long rows, long cols, long ld, long i0, long j0, long total_rows, uint64_t seed,
                         double lo, double hi, int threads) {
  parallel_cols(cols, threads, [&](long j) {
    const uint64_t base = (uint64_t)(j0 + j) * (uint64_t)total_rows + (uint64_t)i0;
    double* col = p + j * ld;
    for (long i = 0; i < rows; ++i) col[i] = lo + (hi - lo) * u01(seed, base + i, 0);
  });
}

void tmmu_fill_uniform_s(float* p, long rows, long cols, long ld, long i0, long j0, long total_rows, uint64_t seed,
                         double lo, double hi, int threads) {
  parallel_cols(cols, threads, [&](long j) {
    const uint64_t base = (uint64_t)(j0 + j) * (uint64_t)total_rows + (uint64_t)i0;
    float* col = p + j * ld;
    for (long i = 0; i < rows; ++i) col[i] = (float)(lo + (hi - lo) * u01(seed, base + i, 0));
  });
}

// Integers lo, lo+1, ..., hi with equal probability (the reference's tests fill their matrices
// with integers in [0, 9], Tiled-MM tests/test-multiply.cpp:58-66), as doubles.
void tmmu_fill_int_d(double* p, long rows, long cols, long ld, long i0, long j0, long total_rows, uint64_t seed,
                     long lo, long hi, int threads) {
  const double span = (double)(hi - lo + 1);
  parallel_cols(cols, threads, [&](long j) {
    const uint64_t base = (uint64_t)(j0 + j) * (uint64_t)total_rows + (uint64_t)i0;
    double* col = p + j * ld;
    for (long i = 0; i < rows; ++i) col[i] = (double)lo + std::floor(span * u01(seed, base + i, 0));
  });
}

// Complex (interleaved): real part from lane 0, imaginary part from lane 1.
void tmmu_fill_uniform_z(double* p, long rows, long cols, long ld, long i0, long j0, long total_rows, uint64_t seed,
                         double lo, double hi, int threads) {
  parallel_cols(cols, threads, [&](long j) {
    const uint64_t base = (uint64_t)(j0 + j) * (uint64_t)total_rows + (uint64_t)i0;
    double* col = p + 2 * j * ld;
    for (long i = 0; i < rows; ++i) {
      col[2 * i] = lo + (hi - lo) * u01(seed, base + i, 0);
      col[2 * i + 1] = lo + (hi - lo) * u01(seed, base + i, 1);
    }
  });
}

// Values of wide dynamic range: (u - 0.5) * exp(phi * g), u uniform, g standard normal
// (Box-Muller on two further uniform streams).
void tmmu_fill_range_d(double* p, long rows, long cols, long ld, long i0, long j0, long total_rows, uint64_t seed,
                       double phi, int threads) {
  parallel_cols(cols, threads, [&](long j) {
    const uint64_t base = (uint64_t)(j0 + j) * (uint64_t)total_rows + (uint64_t)i0;
    double* col = p + j * ld;
    for (long i = 0; i < rows; ++i) {
      const double u = u01(seed, base + i, 0);
      const double u1 = 1.0 - u01(seed, base + i, 1), u2 = u01(seed, base + i, 2);
      const double g = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
      col[i] = (u - 0.5) * std::exp(phi * g);
    }
  });
}

// Parallel strided copy of a rows x cols column-major block (bytes per element es).
void tmmu_copy_block(void* dst, long ldd, const void* src, long lds, long rows, long cols, int es, int threads) {
  parallel_cols(cols, threads, [&](long j) {
    const char* s = (const char*)src + (size_t)j * lds * es;
    char* d = (char*)dst + (size_t)j * ldd * es;
    std::memcpy(d, s, (size_t)rows * es);
  });
}

}  // extern "C"

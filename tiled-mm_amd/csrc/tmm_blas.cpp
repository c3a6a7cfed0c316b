// BLAS-ABI interposer (libtmm_blas.so): Fortran-convention dgemm_/zgemm_/sgemm_/cgemm_ that run
// large products through the out-of-core GEMM (one handle per thread and element type) and
// hand small ones to the next BLAS in the symbol search order (dlsym(RTLD_NEXT)), so an
// unmodified BLAS caller offloads its large host GEMMs when this library is preloaded or linked
// first. This is the use Tiled-MM describes as "almost a drop-in replacement for cublasXt"
// (Tiled-MM README.md:14); the reference itself has no such entry point.
//
// Environment:
//   TMM_BLAS_MIN_FLOPS  products below this many flops go to the next BLAS (default 2e9);
//                       when no next BLAS exists, everything runs on the GPU
//   TMM_BLAS_STREAMS / TMM_BLAS_TILE   handle configuration (default 2 streams, 8192 tiles)
//   TMM_BLAS_POOL_LIMIT_MB             device work-pool budget of the handles (TMM_OPT_POOL_LIMIT_MB)
//
// Failures: the Fortran interface has no status. When the GPU path fails before anything was
// written to C (argument check, planning, allocation, no GPU), the call is re-run on the next BLAS;
// without one, or once C may be partly written, the message goes to stderr and the program stops
// (abort), as the reference BLAS stops in xerbla: a caller never continues with a wrong C.
#include <dlfcn.h>

#include <complex>
#include <cstdio>
#include <cstdlib>
#include <memory>

#include "../../include/tiledmm.h"

namespace {

double min_flops() {
  static const double v = [] {
    const char* e = std::getenv("TMM_BLAS_MIN_FLOPS");
    return e ? std::atof(e) : 2e9;
  }();
  return v;
}

int env_int(const char* name, int def) {
  const char* e = std::getenv(name);
  return e ? std::atoi(e) : def;
}

struct HandleDeleter {
  void operator()(tmm_handle* h) const { tmm_destroy(h); }
};

tmm_handle* thread_handle(int dtype) {
  thread_local std::unique_ptr<tmm_handle, HandleDeleter> handles[4];
  auto& slot = handles[dtype];
  if (!slot) {
    tmm_handle* h = nullptr;
    const int tile = env_int("TMM_BLAS_TILE", 8192);
    if (tmm_create(&h, dtype, env_int("TMM_BLAS_STREAMS", 2), tile, tile, tile) != TMM_SUCCESS) return nullptr;
    slot.reset(h);
    const int pool_mb = env_int("TMM_BLAS_POOL_LIMIT_MB", 0);
    if (pool_mb > 0) tmm_set_option(h, TMM_OPT_POOL_LIMIT_MB, pool_mb);
  }
  return slot.get();
}

template <typename T>
using gemm_fn = void (*)(const char*, const char*, const int*, const int*, const int*, const T*, const T*, const int*,
                         const T*, const int*, const T*, T*, const int*);

template <typename T>
gemm_fn<T> next_gemm(const char* name) {
  return (gemm_fn<T>)dlsym(RTLD_NEXT, name);
}

void report(const char* name, const char* msg) {
  std::fprintf(stderr, "%s (Tiled-MM for MI355X): %s\n", name, msg ? msg : "unknown error");
  std::fflush(stderr);
}

[[noreturn]] void stop(const char* name, const char* why) {
  report(name, why);
  std::abort();
}

template <typename T, typename R, typename F>
void dispatch(const char* name, F gpu, int dtype, double flop_per_mac, const char* ta, const char* tb, const int* m,
              const int* n, const int* k, const T* alpha, const T* a, const int* lda, const T* b, const int* ldb,
              const T* beta, T* c, const int* ldc) {
  static const gemm_fn<T> next = next_gemm<T>(name);
  const double flops = flop_per_mac * (double)*m * (double)*n * (double)*k;
  if (next && flops < min_flops()) {
    next(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
    return;
  }
  tmm_handle* h = thread_handle(dtype);
  if (!h) {
    if (!next) stop(name, "no GPU handle and no host BLAS to run the product on");
    report(name, "no GPU handle; running on the host BLAS");
    next(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
    return;
  }
  if (gpu(h, *ta, *tb, *m, *n, *k, (const R*)alpha, (const R*)a, *lda, (const R*)b, *ldb, (const R*)beta, (R*)c, *ldc,
          0, 1) == TMM_SUCCESS)
    return;
  report(name, tmm_last_error(h));
  double st[20] = {0};
  const bool c_written = tmm_last_call_stats(h, st, 20) != TMM_SUCCESS || st[19] != 0;
  if (c_written) stop(name, "C may be partly written by the failed GPU call: stopping");
  if (!next) stop(name, "C is unchanged and there is no host BLAS to run the product on: stopping");
  report(name, "C is unchanged; running the product on the host BLAS");
  next(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

}  // namespace

extern "C" {

void dgemm_(const char* ta, const char* tb, const int* m, const int* n, const int* k, const double* alpha,
            const double* a, const int* lda, const double* b, const int* ldb, const double* beta, double* c,
            const int* ldc) {
  dispatch<double, double>("dgemm_", tmm_dgemm, TMM_F64, 2.0, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

void sgemm_(const char* ta, const char* tb, const int* m, const int* n, const int* k, const float* alpha,
            const float* a, const int* lda, const float* b, const int* ldb, const float* beta, float* c,
            const int* ldc) {
  dispatch<float, float>("sgemm_", tmm_sgemm, TMM_F32, 2.0, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

void zgemm_(const char* ta, const char* tb, const int* m, const int* n, const int* k, const std::complex<double>* alpha,
            const std::complex<double>* a, const int* lda, const std::complex<double>* b, const int* ldb,
            const std::complex<double>* beta, std::complex<double>* c, const int* ldc) {
  dispatch<std::complex<double>, double>("zgemm_", tmm_zgemm, TMM_C64, 8.0, ta, tb, m, n, k, alpha, a, lda, b, ldb,
                                         beta, c, ldc);
}

void cgemm_(const char* ta, const char* tb, const int* m, const int* n, const int* k, const std::complex<float>* alpha,
            const std::complex<float>* a, const int* lda, const std::complex<float>* b, const int* ldb,
            const std::complex<float>* beta, std::complex<float>* c, const int* ldc) {
  dispatch<std::complex<float>, float>("cgemm_", tmm_cgemm, TMM_C32, 8.0, ta, tb, m, n, k, alpha, a, lda, b, ldb,
                                       beta, c, ldc);
}

}  // extern "C"

// C ABI of libtiledmm.so (declarations and reference citations in include/tiledmm.h).
#include <cstring>
#include <string>

#include "../../include/tiledmm.h"
#include "tmm_runtime.hpp"
#include "tmm_ozaki.h"

struct tmm_handle {
  tmm::Handle* impl;
};

namespace {

thread_local std::string g_last_error;

template <typename F>
int guard(tmm_handle* h, F&& f) {
  try {
    f();
    return TMM_SUCCESS;
  } catch (tmm::Error& e) {
    if (h) h->impl->last_error = e.what();
    g_last_error = e.what();
    return e.code;
  } catch (std::bad_alloc&) {
    if (h) h->impl->last_error = "host allocation failed";
    g_last_error = "host allocation failed";
    return TMM_ERR_ALLOC;
  } catch (std::exception& e) {
    if (h) h->impl->last_error = e.what();
    g_last_error = e.what();
    return TMM_ERR_INTERNAL;
  }
}

void need(bool ok, const char* msg) {
  if (!ok) throw tmm::Error(TMM_ERR_INVALID_ARG, msg);
}

int host_gemm(tmm_handle* h, int dtype, char ta, char tb, int m, int n, int k, const void* alpha, const void* a,
              int lda, const void* b, int ldb, const void* beta, void* c, int ldc, int /*pin*/, int copy_back) {
  if (!h) {
    g_last_error = "null handle";
    return TMM_ERR_INVALID_ARG;
  }
  return guard(h, [&] {
    if (h->impl->dtype != dtype) throw tmm::Error(TMM_ERR_DTYPE, "handle was created for another element type");
    need(alpha && beta, "alpha and beta must not be null");
    tmm::GemmCall call{ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, copy_back != 0};
    tmm::gemm_host(*h->impl, call);
  });
}

}  // namespace

extern "C" {

int tmm_create(tmm_handle** out, int dtype, int n_streams, int tm, int tn, int tk) {
  if (!out) {
    g_last_error = "null output pointer";
    return TMM_ERR_INVALID_ARG;
  }
  *out = nullptr;
  return guard(nullptr, [&] {
    need(dtype >= 0 && dtype <= 3, "dtype must be TMM_F32, TMM_F64, TMM_C32 or TMM_C64");
    need(n_streams > 0, "number of streams must be positive");
    need(tm > 0 && tn > 0 && tk > 0, "maximum tile sizes must be positive");
    auto* h = new tmm_handle{nullptr};
    try {
      h->impl = new tmm::Handle(dtype, n_streams, tm, tn, tk);
      // set-up of tmm_device_gemm_noalloc: the address of this device's fixed scratch
      tmm::hip_check(tmm::ozaki_scratch_resolve(), "hipGetSymbolAddress (device scratch)");
    } catch (...) {
      delete h->impl;
      delete h;
      throw;
    }
    *out = h;
  });
}

int tmm_destroy(tmm_handle* h) {
  if (!h) return TMM_SUCCESS;
  delete h->impl;
  delete h;
  return TMM_SUCCESS;
}

int tmm_get_dtype(tmm_handle* h, int* dtype) {
  if (!h || !dtype) return TMM_ERR_INVALID_ARG;
  *dtype = h->impl->dtype;
  return TMM_SUCCESS;
}

int tmm_set_num_streams(tmm_handle* h, int n_streams) {
  if (!h) return TMM_ERR_INVALID_ARG;
  return guard(h, [&] { h->impl->set_num_streams(n_streams); });
}

int tmm_get_num_streams(tmm_handle* h, int* n_streams) {
  if (!h || !n_streams) return TMM_ERR_INVALID_ARG;
  *n_streams = h->impl->n_streams;
  return TMM_SUCCESS;
}

int tmm_set_tile_sizes(tmm_handle* h, int tm, int tn, int tk) {
  if (!h) return TMM_ERR_INVALID_ARG;
  return guard(h, [&] {
    auto& H = *h->impl;
    need(tm > 0 && tn > 0 && tk > 0, "tile sizes must be positive");
    // the reference asserts these bounds (mm_handle.cpp:59-61)
    need(tm <= H.max_tm && tn <= H.max_tn && tk <= H.max_tk, "tile sizes must not exceed the maximum tile sizes");
    H.cur_tm = tm;
    H.cur_tn = tn;
    H.cur_tk = tk;
    H.tilebuf_elems[0] = H.tilebuf_elems[1] = H.tilebuf_elems[2] = 0;
  });
}

int tmm_get_tile_sizes(tmm_handle* h, int* tm, int* tn, int* tk) {
  if (!h || !tm || !tn || !tk) return TMM_ERR_INVALID_ARG;
  *tm = h->impl->cur_tm;
  *tn = h->impl->cur_tn;
  *tk = h->impl->cur_tk;
  return TMM_SUCCESS;
}

int tmm_get_max_tile_sizes(tmm_handle* h, int* tm, int* tn, int* tk) {
  if (!h || !tm || !tn || !tk) return TMM_ERR_INVALID_ARG;
  *tm = h->impl->max_tm;
  *tn = h->impl->max_tn;
  *tk = h->impl->max_tk;
  return TMM_SUCCESS;
}

int tmm_set_streams_and_tiles(tmm_handle* h, int n_streams, int tm, int tn, int tk) {
  if (!h) return TMM_ERR_INVALID_ARG;
  int st = tmm_set_tile_sizes(h, tm, tn, tk);
  if (st != TMM_SUCCESS) return st;
  return tmm_set_num_streams(h, n_streams);
}

int tmm_set_full_sizes(tmm_handle* h, int m, int n, int k) {
  if (!h) return TMM_ERR_INVALID_ARG;
  return guard(h, [&] {
    need(m > 0 && n > 0 && k > 0, "full sizes must be positive");
    auto& H = *h->impl;
    const size_t need_elems = (size_t)m * n;
    H.fullc.reserve(need_elems * H.elem_bytes);
    H.fullc_elems = need_elems;
  });
}

int tmm_optimal_tile_size(int dim, int curr, int max) { return tmm::optimal_tile_size(dim, curr, max); }

int tmm_optimal_tile_sizes(tmm_handle* h, int m, int n, int k, int* tm, int* tn, int* tk) {
  if (!h || !tm || !tn || !tk) return TMM_ERR_INVALID_ARG;
  auto& H = *h->impl;
  *tm = tmm::optimal_tile_size(m, H.cur_tm, H.max_tm);
  *tn = tmm::optimal_tile_size(n, H.cur_tn, H.max_tn);
  *tk = tmm::optimal_tile_size(k, H.cur_tk, H.max_tk);
  return TMM_SUCCESS;
}

int tmm_sgemm(tmm_handle* h, char ta, char tb, int m, int n, int k, const float* alpha, const float* a, int lda,
              const float* b, int ldb, const float* beta, float* c, int ldc, int pin, int copy_back) {
  return host_gemm(h, TMM_F32, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, pin, copy_back);
}
int tmm_dgemm(tmm_handle* h, char ta, char tb, int m, int n, int k, const double* alpha, const double* a, int lda,
              const double* b, int ldb, const double* beta, double* c, int ldc, int pin, int copy_back) {
  return host_gemm(h, TMM_F64, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, pin, copy_back);
}
int tmm_cgemm(tmm_handle* h, char ta, char tb, int m, int n, int k, const float* alpha, const float* a, int lda,
              const float* b, int ldb, const float* beta, float* c, int ldc, int pin, int copy_back) {
  return host_gemm(h, TMM_C32, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, pin, copy_back);
}
int tmm_zgemm(tmm_handle* h, char ta, char tb, int m, int n, int k, const double* alpha, const double* a, int lda,
              const double* b, int ldb, const double* beta, double* c, int ldc, int pin, int copy_back) {
  return host_gemm(h, TMM_C64, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, pin, copy_back);
}

int tmm_dgemm_batched(tmm_handle* h, char ta, char tb, int count, const int* m, const int* n, const int* k,
                      const double* alpha, const double* const* a, const int* lda, const double* const* b,
                      const int* ldb, const double* beta, double* const* c, const int* ldc) {
  if (!h) return TMM_ERR_INVALID_ARG;
  return guard(h, [&] {
    for (double& s : h->impl->stats) s = 0;
    tmm::gemm_batched_host(*h->impl, ta, tb, count, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
  });
}

int tmm_batch_rates(tmm_handle* h, double* rates, int n) {
  if (!h || !rates || n < 0) return TMM_ERR_INVALID_ARG;
  const tmm::TuneRates& r = h->impl->rates;
  const double v[7] = {r.h2d_Bps, r.d2h_Bps, r.kernel_flops, r.group_overhead_s, r.kernel_flops_dlds, r.emu_flops,
                       r.emu_overhead_s};
  for (int i = 0; i < n && i < 7; ++i) rates[i] = v[i];
  return TMM_SUCCESS;
}

int tmm_batch_config(tmm_handle* h, int* groups, int* workgroups, int* graph, double* predicted_s, double* rates) {
  if (!h) return TMM_ERR_INVALID_ARG;
  auto& H = *h->impl;
  if (groups) *groups = H.last_batch.groups;
  if (workgroups) *workgroups = H.last_batch.workgroups;
  if (graph) *graph = H.last_batch.use_graph ? 1 : 0;
  if (predicted_s) *predicted_s = H.last_batch.predicted_s;
  if (rates) {
    rates[0] = H.rates.h2d_Bps;
    rates[1] = H.rates.d2h_Bps;
    rates[2] = H.rates.kernel_flops;
    rates[3] = H.rates.group_overhead_s;
  }
  return TMM_SUCCESS;
}

int tmm_full_c_device_ptr(tmm_handle* h, void** ptr, size_t* n) {
  if (!h || !ptr || !n) return TMM_ERR_INVALID_ARG;
  *ptr = h->impl->fullc.ptr;
  *n = h->impl->fullc_elems;
  return TMM_SUCCESS;
}

int tmm_tile_buffer(tmm_handle* h, int which, int stream_id, void** ptr, size_t* n) {
  if (!h || !ptr || !n) return TMM_ERR_INVALID_ARG;
  return guard(h, [&] {
    auto& H = *h->impl;
    need(which >= 0 && which <= 2, "which must be 0 (A), 1 (B) or 2 (C)");
    need(stream_id >= 0 && stream_id < H.n_streams, "stream id in device buffer has to be in the range [0, n_streams)");
    // tile shapes as in mm_handle::set_tile_sizes: A (tm x tk), B (tk x tn), C (tm x tn)
    const size_t tile = which == 0 ? (size_t)H.cur_tm * H.cur_tk
                      : which == 1 ? (size_t)H.cur_tk * H.cur_tn
                                   : (size_t)H.cur_tm * H.cur_tn;
    if (tile == 0) {
      *ptr = nullptr;
      *n = 0;
      return;
    }
    H.tilebuf[which].reserve(tile * H.n_streams * H.elem_bytes);
    H.tilebuf_elems[which] = tile * H.n_streams;
    *ptr = (char*)H.tilebuf[which].ptr + (size_t)stream_id * tile * H.elem_bytes;
    *n = tile;
  });
}

int tmm_get_stream(tmm_handle* h, int stream_id, void** stream) {
  if (!h || !stream) return TMM_ERR_INVALID_ARG;
  return guard(h, [&] {
    auto& H = *h->impl;
    if (stream_id == -1) {
      *stream = (void*)H.d2h;
      return;
    }
    need(stream_id >= 0 && stream_id < H.n_streams, "index of gpu stream has to be in the range [0, n_streams)");
    *stream = (void*)H.compute_stream(stream_id);
  });
}

const char* tmm_last_error(tmm_handle* h) {
  if (h) return h->impl->last_error.c_str();
  return g_last_error.c_str();
}

int tmm_device_gemm(tmm_handle* h, char ta, char tb, int m, int n, int k, const void* alpha, const void* a, int lda,
                    const void* b, int ldb, const void* beta, void* c, int ldc, void* stream) {
  if (!h) return TMM_ERR_INVALID_ARG;
  return guard(h, [&] {
    need(alpha && beta, "alpha and beta must not be null");
    tmm::GemmCall call{ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, true};
    tmm::gemm_device(*h->impl, call, (hipStream_t)stream);
  });
}

static_assert(TMM_NOALLOC_SCRATCH_BYTES == tmm::kOzakiScratchBytes, "scratch size in the header and the library");

int tmm_device_gemm_noalloc(tmm_handle* h, char ta, char tb, int m, int n, int k, const void* alpha, const void* a,
                            int lda, const void* b, int ldb, const void* beta, void* c, int ldc, void* stream) {
  if (!h) return TMM_ERR_INVALID_ARG;
  return guard(h, [&] {
    need(alpha && beta, "alpha and beta must not be null");
    tmm::GemmCall call{ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, true};
    tmm::gemm_device_noalloc(*h->impl, call, (hipStream_t)stream);
  });
}

int tmm_synchronize(tmm_handle* h) {
  if (!h) return TMM_ERR_INVALID_ARG;
  return guard(h, [&] {
    auto& H = *h->impl;
    H.sync_compute();
    tmm::hip_check(hipStreamSynchronize(H.h2d), "hipStreamSynchronize");
    tmm::hip_check(hipStreamSynchronize(H.d2h), "hipStreamSynchronize");
  });
}

int tmm_set_option(tmm_handle* h, int option, long value) {
  if (!h) return TMM_ERR_INVALID_ARG;
  return guard(h, [&] {
    auto& H = *h->impl;
    switch (option) {
      case TMM_OPT_COARSEN: H.coarsen = value != 0; break;
      case TMM_OPT_FIRST_CHUNK: need(value >= -1, "first chunk must be >= -1"); H.first_chunk = value; break;
      case TMM_OPT_HOST_THREADS: need(value >= 0 && value <= 256, "host threads in [0, 256]"); H.host_threads = (int)value; break;
      case TMM_OPT_POOL_LIMIT_MB: need(value >= 0, "pool limit must be >= 0"); H.pool_limit_mb = value; break;
      case TMM_OPT_PROFILE: H.profile = value != 0; break;
      case TMM_OPT_OZAKI_SLICES: need(value >= 0 && value <= 12, "Ozaki slices in [0, 12]"); H.ozaki_slices = (int)value; break;
      case TMM_OPT_OZAKI_MODULI:
        need(value == 0 || (value >= tmm::kCrtMinModuli && value <= tmm::kCrtMaxModuli), "Ozaki moduli 0 or in [6, 20]");
        H.ozaki_moduli = (int)value;
        break;
      case TMM_OPT_DGEMM_KERNEL: need(value == 0 || value == 1, "tile kernel variant 0 or 1"); H.dgemm_variant = (int)value; break;
      case TMM_OPT_OZAKI_REUSE: need(value == 0 || value == 1, "Ozaki reuse 0 or 1"); H.ozaki_reuse = value != 0; break;
      case TMM_OPT_BATCH_GRAPH:
        need(value >= -1 && value <= 1, "batch graph -1, 0 or 1");
        H.batch_graph = (int)value;
        break;
      case TMM_OPT_TILE_GROUPS: need(value >= 0 && value <= 16, "tile groups 0 to 16"); H.tile_groups = (int)value; break;
      case TMM_OPT_DEFER_CIN: need(value >= 0 && value <= 1024, "defer 0..1024 tiles"); H.defer_cin = (int)value; break;
      case TMM_OPT_BATCH_BODY: need(value >= -1 && value <= 1, "batch tile body -1, 0 or 1"); H.batch_body = (int)value; break;
      case TMM_OPT_BATCH_QUEUES: need(value >= 0 && value <= 2, "batch compute queues 0, 1 or 2"); H.batch_queues = (int)value; break;
      case TMM_OPT_BATCH_EMULATION:
        need(value >= -1 && value <= 1, "batch emulation -1, 0 or 1");
        H.batch_emulation = (int)value;
        break;
      default: need(false, "unknown option");
    }
  });
}

int tmm_get_option(tmm_handle* h, int option, long* value) {
  if (!h || !value) return TMM_ERR_INVALID_ARG;
  return guard(h, [&] {
    auto& H = *h->impl;
    switch (option) {
      case TMM_OPT_COARSEN: *value = H.coarsen ? 1 : 0; break;
      case TMM_OPT_FIRST_CHUNK: *value = H.first_chunk; break;
      case TMM_OPT_HOST_THREADS: *value = H.host_threads; break;
      case TMM_OPT_POOL_LIMIT_MB: *value = H.pool_limit_mb; break;
      case TMM_OPT_PROFILE: *value = H.profile ? 1 : 0; break;
      case TMM_OPT_OZAKI_SLICES: *value = H.ozaki_slices; break;
      case TMM_OPT_OZAKI_MODULI: *value = H.ozaki_moduli; break;
      case TMM_OPT_DGEMM_KERNEL: *value = H.dgemm_variant; break;
      case TMM_OPT_OZAKI_REUSE: *value = H.ozaki_reuse ? 1 : 0; break;
      case TMM_OPT_BATCH_GRAPH: *value = H.batch_graph; break;
      case TMM_OPT_TILE_GROUPS: *value = H.tile_groups; break;
      case TMM_OPT_DEFER_CIN: *value = H.defer_cin; break;
      case TMM_OPT_BATCH_BODY: *value = H.batch_body; break;
      case TMM_OPT_BATCH_QUEUES: *value = H.batch_queues; break;
      case TMM_OPT_BATCH_EMULATION: *value = H.batch_emulation; break;
      default: need(false, "unknown option");
    }
  });
}

int tmm_last_call_stats(tmm_handle* h, double* stats, int n) {
  if (!h || !stats || n < 0) return TMM_ERR_INVALID_ARG;
  for (int i = 0; i < n && i < 24; ++i) stats[i] = h->impl->stats[i];
  return TMM_SUCCESS;
}

}  // extern "C"

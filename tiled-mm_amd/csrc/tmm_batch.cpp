// Batched host-to-host FP64 GEMM: many independent problems of uneven size in one call.
//
// No counterpart in the reference (a caller there issues one gpu::gemm per problem,
// Tiled-MM src/Tiled-MM/tiled_mm.cpp:492-624, each with its own tiling, copies and a
// device-wide synchronisation). Here the whole batch is one pipeline:
//   * the problems are packed into one device arena and cut into G groups of about equal
//     host-link bytes; group g's H2D copies (spread over two copy queues), one persistent
//     batched-kernel launch over all 128x128 C tiles of the group (largest k first), and its D2H
//     copies run on their own queues, so H2D of group g+1 overlaps the kernel of group g and the
//     D2H of group g-1;
//   * with pinned host matrices the whole sequence is captured once into a hipGraph and
//     launched as one graph (no per-problem launch overhead);
//   * G, the persistent grid, the tile body (register-staged or direct-to-LDS), the compute
//     queue count and, on a handle with Ozaki moduli set, which problems run emulated come from
//     an autotuner (tmm_autotune.cpp) that models the pipeline with rates measured on the device.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "tmm_kernels.h"
#include "tmm_ozaki.h"
#include "tmm_runtime.hpp"

namespace tmm {

namespace {

struct Prob {
  int m, n, k;
  long lda, ldb, ldc;
  const double* a;
  const double* b;
  double* c;
  double alpha, beta;
  size_t offA, offB, offC;  // arena offsets (elements)
  long rowsA, colsA, rowsB, colsB;
};

}  // namespace

BatchConfig choose_batch_config(Handle& h, const BatchShape& shape);  // tmm_autotune.cpp

void gemm_batched_host(Handle& h, char opa, char opb, int count, const int* m, const int* n, const int* k,
                       const double* alpha, const double* const* a, const int* lda, const double* const* b,
                       const int* ldb, const double* beta, double* const* c, const int* ldc) {
  if (h.dtype != 1) throw Error(4, "batched GEMM is available for double handles");
  if (count < 0) throw Error(1, "negative batch count");
  if (count == 0) return;
  if (!m || !n || !k || !alpha || !a || !lda || !b || !ldb || !beta || !c || !ldc)
    throw Error(1, "null batch array");
  opa = (char)std::toupper((unsigned char)opa);
  opb = (char)std::toupper((unsigned char)opb);
  if (opa != 'T' && opa != 'C') opa = 'N';
  if (opb != 'T' && opb != 'C') opb = 'N';
  if (opa == 'C') opa = 'T';
  if (opb == 'C') opb = 'T';

  std::vector<Prob> P(count);
  size_t arena = 0;
  BatchShape shape;
  for (int i = 0; i < count; ++i) {
    Prob& p = P[i];
    p.m = m[i];
    p.n = n[i];
    p.k = k[i];
    if (p.m < 0 || p.n < 0 || p.k < 0) throw Error(1, "matrix dimensions must be non-negative");
    p.lda = lda[i];
    p.ldb = ldb[i];
    p.ldc = ldc[i];
    p.rowsA = opa == 'N' ? p.m : p.k;
    p.colsA = opa == 'N' ? p.k : p.m;
    p.rowsB = opb == 'N' ? p.k : p.n;
    p.colsB = opb == 'N' ? p.n : p.k;
    if (p.lda < std::max<long>(1, p.rowsA)) throw Error(1, "Leading dimensions for matrix A must be >= the number of rows.");
    if (p.ldb < std::max<long>(1, p.rowsB)) throw Error(1, "Leading dimensions for matrix B must be >= the number of rows.");
    if (p.ldc < std::max(1, p.m)) throw Error(1, "Leading dimensions for matrix C must be >= the number of rows.");
    p.a = a[i];
    p.b = b[i];
    p.c = c[i];
    p.alpha = alpha[i];
    p.beta = beta[i];
    p.offA = arena;
    arena += (size_t)p.rowsA * p.colsA;
    p.offB = arena;
    arena += (size_t)p.rowsB * p.colsB;
    p.offC = arena;
    arena += (size_t)p.m * p.n;
    arena = (arena + 31) & ~(size_t)31;  // 256-byte aligned matrices
    shape.flops += 2.0 * p.m * p.n * p.k;
    shape.h2d_bytes += 8.0 * ((double)p.rowsA * p.colsA + (double)p.rowsB * p.colsB + (p.beta != 0.0 ? (double)p.m * p.n : 0.0));
    shape.d2h_bytes += 8.0 * p.m * p.n;
    shape.count += 1;
    shape.m.push_back(p.m);
    shape.n.push_back(p.n);
    shape.k.push_back(p.k);
  }

  // pinned detection for every matrix: the graph path needs all of them pinned
  bool all_pinned = true;
  for (const Prob& p : P) {
    auto span = [](long rows, long cols, long ld) -> size_t {
      return rows == 0 || cols == 0 ? 0 : (size_t)((cols - 1) * ld + rows) * 8;
    };
    all_pinned = all_pinned && is_pinned_range(p.a, span(p.rowsA, p.colsA, p.lda)) &&
                 is_pinned_range(p.b, span(p.rowsB, p.colsB, p.ldb)) && is_pinned_range(p.c, span(p.m, p.n, p.ldc));
    if (!all_pinned) break;
  }
  shape.pinned = all_pinned;
  BatchConfig cfg = choose_batch_config(h, shape);
  cfg.queues = std::max(1, std::min(cfg.queues, (int)h.compute.size()));
  // emulated problems: one workspace per compute queue (groups alternate over the queues),
  // reserved before any capture
  size_t emu_ws = 0;
  for (int i = 0; i < count; ++i)
    if (cfg.emulate[i]) emu_ws = std::max(emu_ws, ozaki2_workspace_bytes(P[i].m, P[i].n, P[i].k, h.ozaki_moduli));
  std::vector<void*> emu_work(cfg.queues, nullptr);
  if (emu_ws)
    for (int q = 0; q < cfg.queues; ++q) emu_work[q] = h.ozwork_acquire(q, emu_ws, nullptr);

  // groups of about equal host-link bytes, in the caller's order
  const int G = std::max(1, std::min(cfg.groups, count));
  std::vector<int> gstart(G + 1, count);
  {
    std::vector<double> bytes(count);
    for (int i = 0; i < count; ++i)
      bytes[i] = 8.0 * ((double)P[i].rowsA * P[i].colsA + (double)P[i].rowsB * P[i].colsB + 2.0 * P[i].m * P[i].n);
    const double total = std::accumulate(bytes.begin(), bytes.end(), 0.0);
    double acc = 0;
    int g = 0;
    gstart[0] = 0;
    for (int i = 0; i < count; ++i) {
      if (g + 1 < G && acc >= total * (g + 1) / G) gstart[++g] = i;
      acc += bytes[i];
    }
    for (int gg = g + 1; gg <= G; ++gg) gstart[gg] = count;
  }

  // device arena, problem descriptors and work items
  h.pool.reserve(arena * 8);
  double* dev = (double*)h.pool.ptr;
  std::vector<GemmParams<double>> desc(count);
  std::vector<std::vector<int>> items(G);
  for (int g = 0; g < G; ++g) {
    std::vector<int> order;
    for (int i = gstart[g]; i < gstart[g + 1]; ++i) order.push_back(i);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return P[x].k > P[y].k; });
    for (int i : order)
      if (!cfg.emulate[i])
      for (int bm = 0; bm < (P[i].m + 127) / 128; ++bm)
        for (int bn = 0; bn < (P[i].n + 127) / 128; ++bn) {
          items[g].push_back(i);
          items[g].push_back(bm);
          items[g].push_back(bn);
        }
  }
  for (int i = 0; i < count; ++i) {
    const Prob& p = P[i];
    GemmParams<double>& d = desc[i];
    d.opa = opa;
    d.opb = opb;
    d.m = p.m;
    d.n = p.n;
    d.k = p.k;
    d.alpha = p.alpha;
    d.alpha_im = 0;
    d.a = dev + p.offA;
    d.lda = std::max<long>(1, p.rowsA);
    d.b = dev + p.offB;
    d.ldb = std::max<long>(1, p.rowsB);
    d.beta = p.beta;
    d.beta_im = 0;
    d.cin = dev + p.offC;  // in place: the beta term is read before the same thread writes
    d.ldcin = std::max(1, p.m);
    d.accumulate = 0;
    d.c = dev + p.offC;
    d.ldc = std::max(1, p.m);
  }
  size_t meta_bytes = sizeof(GemmParams<double>) * count + 256;
  std::vector<size_t> item_off(G + 1, 0);
  for (int g = 0; g < G; ++g) item_off[g + 1] = item_off[g] + items[g].size();
  meta_bytes += item_off[G] * sizeof(int) + 256 + 64 * G;
  h.batch_meta.reserve(meta_bytes);
  char* meta = (char*)h.batch_meta.ptr;
  GemmParams<double>* d_desc = (GemmParams<double>*)meta;
  int* d_items = (int*)(meta + ((sizeof(GemmParams<double>) * count + 255) & ~(size_t)255));
  int* d_counters = d_items + ((item_off[G] + 63) & ~(size_t)63);
  hip_check(hipMemcpy(d_desc, desc.data(), sizeof(GemmParams<double>) * count, hipMemcpyHostToDevice), "hipMemcpy");
  {
    std::vector<int> flat(item_off[G]);
    for (int g = 0; g < G; ++g) std::copy(items[g].begin(), items[g].end(), flat.begin() + item_off[g]);
    if (!flat.empty())
      hip_check(hipMemcpy(d_items, flat.data(), flat.size() * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
  }

  hipStream_t sh = h.h2d, sd = h.d2h;
  std::vector<hipStream_t> scs(cfg.queues);
  for (int q = 0; q < cfg.queues; ++q) scs[q] = h.compute_stream(q);
  hipStream_t sc = scs[0];
  // the group's kernels: the persistent launch over its native items, then its emulated problems
  auto compute_group = [&](int g, hipStream_t s) {
    hip_check(launch_gemm_d_batched(opa, opb, d_desc, d_items + item_off[g], (int)(items[g].size() / 3),
                                    d_counters + 16 * g, cfg.workgroups, s, cfg.body),
              "batched GEMM launch");
    for (int i = gstart[g]; i < gstart[g + 1]; ++i) {
      if (!cfg.emulate[i]) continue;
      const GemmParams<double>& d = desc[i];
      int np = 0;
      hip_check(ozaki2_dgemm(opa, opb, d.m, d.n, d.k, d.alpha, d.a, d.lda, d.b, d.ldb, d.beta, d.cin, d.ldcin, false,
                             d.c, d.ldc, h.ozaki_moduli, emu_work[g % cfg.queues], s, &np),
                "batched Ozaki (CRT) product");
    }
  };
  // H2D copies go to two copy queues, each copy to the one with fewer bytes issued so far in this
  // call: a DMA command carries a fixed cost of several microseconds, and with two queues one
  // queue's transfer covers the other's. The 512-problem batch (m, n, k log-uniform in [64, 4096],
  // 1536 H2D copies of 4 KB - 130 MB, 64 groups) ran 232-234 ms per call with one H2D queue, 213 ms
  // with two balanced by bytes; a third queue, or a second D2H queue, ran slower.
  hipStream_t sh2 = h.h2d_aux_stream();
  std::vector<hipEvent_t> ev_in(G), ev_in2(G), ev_k(G);
  for (int g = 0; g < G; ++g) {
    ev_in[g] = h.event(3 * g);
    ev_in2[g] = h.event(3 * g + 1);
    ev_k[g] = h.event(3 * g + 2);
  }
  const size_t ev0 = 3 * (size_t)G;  // then: fork, D2H join, second H2D queue join, compute joins
  hipEvent_t ev_fork = h.event(ev0), ev_join_d = h.event(ev0 + 1), ev_join_h2 = h.event(ev0 + 2);
  for (size_t q = 0; q < scs.size(); ++q) (void)h.event(ev0 + 3 + q);  // created before any capture

  // contiguous matrices (leading dimension = rows) as one linear copy
  auto copy = [&](void* dst, size_t dpitch, const void* src, size_t spitch, size_t width, size_t cols,
                  hipMemcpyKind kind, hipStream_t s) {
    if (dpitch == width && spitch == width)
      hip_check(hipMemcpyAsync(dst, src, width * cols, kind, s), "hipMemcpyAsync");
    else
      hip_check(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, cols, kind, s), "hipMemcpy2DAsync");
  };
  double issued[2] = {0, 0};
  auto h2d_copy = [&](double* dst, const double* src, long rows, long cols, long ld) {
    if (!rows || !cols) return;
    const int q = issued[1] < issued[0] ? 1 : 0;
    issued[q] += (double)rows * cols;
    copy(dst, (size_t)rows * 8, src, (size_t)ld * 8, (size_t)rows * 8, (size_t)cols, hipMemcpyHostToDevice,
         q ? sh2 : sh);
  };
  auto h2d = [&](const Prob& p) {
    h2d_copy(dev + p.offA, p.a, p.rowsA, p.colsA, p.lda);
    h2d_copy(dev + p.offB, p.b, p.rowsB, p.colsB, p.ldb);
    if (p.beta != 0.0) h2d_copy(dev + p.offC, p.c, p.m, p.n, p.ldc);
  };
  auto d2h = [&](const Prob& p, hipStream_t s) {
    if (p.m && p.n)
      copy(p.c, (size_t)p.ldc * 8, dev + p.offC, (size_t)p.m * 8, (size_t)p.m * 8, (size_t)p.n, hipMemcpyDeviceToHost,
           s);
  };
  auto issue = [&]() {
    // fork the other queues from the H2D queue (needed under capture)
    hip_check(hipEventRecord(ev_fork, sh), "hipEventRecord");
    for (hipStream_t q : scs) hip_check(hipStreamWaitEvent(q, ev_fork, 0), "hipStreamWaitEvent");
    hip_check(hipStreamWaitEvent(sd, ev_fork, 0), "hipStreamWaitEvent");
    hip_check(hipStreamWaitEvent(sh2, ev_fork, 0), "hipStreamWaitEvent");
    issued[0] = issued[1] = 0;
    for (int g = 0; g < G; ++g) {
      hipStream_t sg = scs[g % cfg.queues];
      for (int i = gstart[g]; i < gstart[g + 1]; ++i) h2d(P[i]);
      hip_check(hipEventRecord(ev_in[g], sh), "hipEventRecord");
      hip_check(hipEventRecord(ev_in2[g], sh2), "hipEventRecord");
      hip_check(hipStreamWaitEvent(sg, ev_in[g], 0), "hipStreamWaitEvent");
      hip_check(hipStreamWaitEvent(sg, ev_in2[g], 0), "hipStreamWaitEvent");
      compute_group(g, sg);
      hip_check(hipEventRecord(ev_k[g], sg), "hipEventRecord");
      hip_check(hipStreamWaitEvent(sd, ev_k[g], 0), "hipStreamWaitEvent");
      for (int i = gstart[g]; i < gstart[g + 1]; ++i) d2h(P[i], sd);
    }
    // join back into the H2D queue
    for (size_t q = 0; q < scs.size(); ++q) {
      hip_check(hipEventRecord(h.event(ev0 + 3 + q), scs[q]), "hipEventRecord");
      hip_check(hipStreamWaitEvent(sh, h.event(ev0 + 3 + q), 0), "hipStreamWaitEvent");
    }
    hip_check(hipEventRecord(ev_join_d, sd), "hipEventRecord");
    hip_check(hipStreamWaitEvent(sh, ev_join_d, 0), "hipStreamWaitEvent");
    hip_check(hipEventRecord(ev_join_h2, sh2), "hipEventRecord");
    hip_check(hipStreamWaitEvent(sh, ev_join_h2, 0), "hipStreamWaitEvent");
  };

  if (all_pinned && cfg.use_graph) {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hip_check(hipStreamBeginCapture(sh, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
      issue();
    } catch (...) {
      hipGraph_t g2 = nullptr;
      (void)hipStreamEndCapture(sh, &g2);
      if (g2) (void)hipGraphDestroy(g2);
      throw;
    }
    hip_check(hipStreamEndCapture(sh, &graph), "hipStreamEndCapture");
    hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(graph);
      hip_check(e, "hipGraphInstantiate");
    }
    e = hipGraphLaunch(exec, sh);
    if (e == hipSuccess) e = hipStreamSynchronize(sh);
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    hip_check(e, "hipGraphLaunch");
    h.stats[14] = 1;
  } else if (all_pinned) {
    issue();
    hip_check(hipStreamSynchronize(sh), "hipStreamSynchronize");
    h.stats[14] = 0;
  } else {
    // pageable host memory: stage each matrix through one pinned bounce buffer per direction
    // (copies cannot be captured), then the same kernels
    auto stage_in = [&](double* dst, const double* src, long rows, long cols, long ld) {
      if (!rows || !cols) return;
      h.stage_in.ensure(1, std::max<size_t>(h.stage_in.slot_bytes, (size_t)rows * cols * 8));
      char* buf = (char*)h.stage_in.slot[0];
      for (long j = 0; j < cols; ++j) std::memcpy(buf + j * rows * 8, src + j * ld, rows * 8);
      hip_check(hipMemcpy(dst, buf, (size_t)rows * cols * 8, hipMemcpyHostToDevice), "hipMemcpy");
    };
    for (int g = 0; g < G; ++g) {
      for (int i = gstart[g]; i < gstart[g + 1]; ++i) {
        const Prob& p = P[i];
        stage_in(dev + p.offA, p.a, p.rowsA, p.colsA, p.lda);
        stage_in(dev + p.offB, p.b, p.rowsB, p.colsB, p.ldb);
        if (p.beta != 0.0) stage_in(dev + p.offC, p.c, p.m, p.n, p.ldc);
      }
      compute_group(g, sc);
    }
    hip_check(hipStreamSynchronize(sc), "hipStreamSynchronize");
    for (const Prob& p : P) {
      if (!p.m || !p.n) continue;
      h.stage_out.ensure(1, std::max<size_t>(h.stage_out.slot_bytes, (size_t)p.m * p.n * 8));
      char* buf = (char*)h.stage_out.slot[0];
      hip_check(hipMemcpy(buf, dev + p.offC, (size_t)p.m * p.n * 8, hipMemcpyDeviceToHost), "hipMemcpy");
      for (long j = 0; j < p.n; ++j) std::memcpy(p.c + j * p.ldc, buf + j * p.m * 8, (size_t)p.m * 8);
    }
    h.stats[14] = 0;
  }
  h.stats[15] = G;
  h.stats[20] = cfg.body;
  h.stats[21] = cfg.queues;
  h.stats[22] = cfg.emulated;
}

}  // namespace tmm

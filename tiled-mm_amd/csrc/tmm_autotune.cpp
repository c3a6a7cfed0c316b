// Autotuner of the batched pipeline.
//
// The batch runs as G groups through three kinds of queues (H2D copies, the compute queues, D2H
// copies). Per group, with stage times h = H2D/G / BW_h2d, c = compute/G and d = D2H/G /
// BW_d2h and a fixed per-group cost o (launch, events), a three-stage pipeline over G equal
// groups takes
//     T(G) = h + c + d + (G - 1) * max(h, c, d) + G * o.
// What the autotuner chooses, all from rates measured on the device (calibrate_rates: a pinned
// 64 MiB copy each way; both persistent tile bodies on 6 * CUs tiles of 1024^3 problems; with a handle that
// has Ozaki moduli set, one emulated 4096^3 and one 128^3 call), optionally cached in the file
// named by TMM_TUNE_FILE:
//   * the persistent tile body: the register-staged LDS double buffer or the direct-to-LDS ring
//     (whichever ran faster);
//   * per problem, native FP64 or the Ozaki scheme II emulation (only on a handle whose caller
//     opted into the emulation's accuracy with TMM_OPT_OZAKI_MODULI): a problem is emulated when
//     its modelled emulated time, 2mnk / (R_emu * min(1, tiles / CUs)) + o_emu with tiles its
//     256x256 product tiles, beats 2mnk / R_native;
//   * the compute queue count (1 or 2): with 2, consecutive groups' kernels may overlap, which
//     hides the last partial wave of each group (modelled as one 128x128 tile of the group's
//     largest k per group);
//   * G in {1, 2, 4, ..., 64}, and one captured hipGraph vs the queues (pinned memory only): this
//     ROCm's graph executor runs copy nodes one after the other, so with the graph H2D and D2H
//     add up: T = max(H + D + c/G, C + h + d) without per-group host cost.
// The persistent grid fills the CUs at the chosen body's occupancy (2 or 3 workgroups per CU).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "tmm_kernels.h"
#include "tmm_ozaki.h"
#include "tmm_runtime.hpp"

namespace tmm {

namespace {

bool load_rates(TuneRates& r) {
  const char* f = std::getenv("TMM_TUNE_FILE");
  if (!f) return false;
  FILE* fp = std::fopen(f, "r");
  if (!fp) return false;
  TuneRates t;
  const int got = std::fscanf(fp, "h2d_Bps %lf d2h_Bps %lf kernel_flops %lf group_overhead_s %lf kernel_flops_dlds %lf",
                              &t.h2d_Bps, &t.d2h_Bps, &t.kernel_flops, &t.group_overhead_s, &t.kernel_flops_dlds);
  std::fclose(fp);
  if (got != 5 || t.h2d_Bps <= 0 || t.d2h_Bps <= 0 || t.kernel_flops <= 0 || t.kernel_flops_dlds <= 0) return false;
  r = t;
  return true;
}

void store_rates(const TuneRates& r) {
  const char* f = std::getenv("TMM_TUNE_FILE");
  if (!f) return;
  FILE* fp = std::fopen(f, "w");
  if (!fp) return;
  std::fprintf(fp, "h2d_Bps %.6e d2h_Bps %.6e kernel_flops %.6e group_overhead_s %.6e kernel_flops_dlds %.6e\n",
               r.h2d_Bps, r.d2h_Bps, r.kernel_flops, r.group_overhead_s, r.kernel_flops_dlds);
  std::fclose(fp);
}

}  // namespace

void calibrate_rates(Handle& h) {
  if (h.rates.h2d_Bps > 0) return;
  if (load_rates(h.rates)) return;
  hipStream_t s = h.compute[0];
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "hipEventCreate");
  hip_check(hipEventCreate(&e1), "hipEventCreate");
  auto elapsed = [&]() {
    float ms = 0;
    hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
    hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    return ms * 1e-3;
  };
  // link rates
  const size_t bytes = 64u << 20;
  void* host = nullptr;
  DeviceBuffer dev;
  dev.reserve(bytes);
  hip_check(hipHostMalloc(&host, bytes, hipHostMallocDefault), "hipHostMalloc");
  hip_check(hipMemcpyAsync(dev.ptr, host, bytes, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
  hip_check(hipEventRecord(e0, s), "hipEventRecord");
  hip_check(hipMemcpyAsync(dev.ptr, host, bytes, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
  hip_check(hipEventRecord(e1, s), "hipEventRecord");
  h.rates.h2d_Bps = bytes / elapsed();
  hip_check(hipMemcpyAsync(host, dev.ptr, bytes, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  hip_check(hipEventRecord(e0, s), "hipEventRecord");
  hip_check(hipMemcpyAsync(host, dev.ptr, bytes, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  hip_check(hipEventRecord(e1, s), "hipEventRecord");
  h.rates.d2h_Bps = bytes / elapsed();
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  hip_check(hipHostFree(host), "hipHostFree");
  dev.release();

  // persistent kernel rates on problems of 1024^3 (contents irrelevant: zero-filled); 64 tiles
  // each, 6 * CUs tiles in all, so that both grids (2 and 3 workgroups per CU) run whole rounds
  int dev_id = 0, cus = 256;
  hip_check(hipGetDevice(&dev_id), "hipGetDevice");
  hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev_id), "hipDeviceGetAttribute");
  const int sz = 1024, np = std::max(1, 6 * cus / 64);
  DeviceBuffer mats, meta;
  mats.reserve((size_t)3 * np * sz * sz * 8);
  hip_check(hipMemsetAsync(mats.ptr, 0, (size_t)3 * np * sz * sz * 8, s), "hipMemsetAsync");
  std::vector<GemmParams<double>> desc(np);
  std::vector<int> items;
  for (int i = 0; i < np; ++i) {
    double* base = (double*)mats.ptr + (size_t)3 * i * sz * sz;
    desc[i] = GemmParams<double>{'N', 'N', sz, sz, sz, 1.0, 0.0, base, sz, base + (size_t)sz * sz, sz, 0.0, 0.0,
                                 nullptr, sz, 0, base + (size_t)2 * sz * sz, sz};
    for (int bm = 0; bm < sz / 128; ++bm)
      for (int bn = 0; bn < sz / 128; ++bn) {
        items.push_back(i);
        items.push_back(bm);
        items.push_back(bn);
      }
  }
  meta.reserve(sizeof(GemmParams<double>) * np + items.size() * sizeof(int) + 1024);
  GemmParams<double>* d_desc = (GemmParams<double>*)meta.ptr;
  int* d_items = (int*)((char*)meta.ptr + ((sizeof(GemmParams<double>) * np + 255) & ~(size_t)255));
  int* d_counter = d_items + ((items.size() + 63) & ~(size_t)63);
  hip_check(hipMemcpy(d_desc, desc.data(), sizeof(GemmParams<double>) * np, hipMemcpyHostToDevice), "hipMemcpy");
  hip_check(hipMemcpy(d_items, items.data(), items.size() * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
  const int n_items = (int)(items.size() / 3);
  for (int body = 0; body < 2; ++body) {
    const int grid = (body ? 3 : 2) * cus;  // the body's workgroups per CU
    hip_check(launch_gemm_d_batched('N', 'N', d_desc, d_items, n_items, d_counter, grid, s, body), "batched launch");
    hip_check(hipEventRecord(e0, s), "hipEventRecord");
    hip_check(launch_gemm_d_batched('N', 'N', d_desc, d_items, n_items, d_counter, grid, s, body), "batched launch");
    hip_check(hipEventRecord(e1, s), "hipEventRecord");
    const double t = elapsed();
    (body == 0 ? h.rates.kernel_flops : h.rates.kernel_flops_dlds) = 2.0 * np * (double)sz * sz * sz / t;
  }
  // per-group fixed cost: an empty item list still runs the counter kernel and the grid
  hip_check(hipEventRecord(e0, s), "hipEventRecord");
  for (int r = 0; r < 8; ++r)
    hip_check(launch_gemm_d_batched('N', 'N', d_desc, d_items, 0, d_counter, 3 * cus, s, 1), "batched launch");
  hip_check(hipEventRecord(e1, s), "hipEventRecord");
  h.rates.group_overhead_s = std::max(10e-6, elapsed() / 8 + 10e-6);  // + events and copy-queue switches
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  store_rates(h.rates);
}

namespace {

// Ozaki scheme II rates at the handle's modulus count: one emulated 4096^3 call (FP64-equivalent
// flop rate with 256 product tiles of 256x256, one per CU) and one 128^3 call (the fixed cost of
// a product), on zero data.
void calibrate_emulation(Handle& h) {
  if (h.ozaki_moduli <= 0 || h.batch_emulation == 0 || h.rates.emu_moduli == h.ozaki_moduli) return;
  hipStream_t s = h.compute[0];
  const int big = 4096, small = 128;
  const size_t ws = ozaki2_workspace_bytes(big, big, big, h.ozaki_moduli);
  DeviceBuffer mats, work;
  mats.reserve((size_t)3 * big * big * 8);
  work.reserve(ws);
  hip_check(hipMemsetAsync(mats.ptr, 0, (size_t)3 * big * big * 8, s), "hipMemsetAsync");
  double* A = (double*)mats.ptr;
  double* B = A + (size_t)big * big;
  double* C = B + (size_t)big * big;
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "hipEventCreate");
  hip_check(hipEventCreate(&e1), "hipEventCreate");
  auto timed = [&](int sz) {
    int np = 0;
    hip_check(ozaki2_dgemm('N', 'N', sz, sz, sz, 1.0, A, sz, B, sz, 0.0, C, sz, false, C, sz, h.ozaki_moduli, work.ptr, s,
                           &np), "Ozaki calibration");
    hip_check(hipEventRecord(e0, s), "hipEventRecord");
    hip_check(ozaki2_dgemm('N', 'N', sz, sz, sz, 1.0, A, sz, B, sz, 0.0, C, sz, false, C, sz, h.ozaki_moduli, work.ptr, s,
                           &np), "Ozaki calibration");
    hip_check(hipEventRecord(e1, s), "hipEventRecord");
    hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    return ms * 1e-3;
  };
  h.rates.emu_overhead_s = timed(small);
  // per CU: the product kernel runs one 256x256 tile per CU at a time
  int dev_id = 0, cus = 256;
  hip_check(hipGetDevice(&dev_id), "hipGetDevice");
  hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev_id), "hipDeviceGetAttribute");
  const double tiles = (double)(big / 256) * (big / 256);
  h.rates.emu_flops = 2.0 * big * (double)big * big / std::max(1e-9, timed(big) - h.rates.emu_overhead_s) *
                      std::max(1.0, cus / tiles);
  h.rates.emu_moduli = h.ozaki_moduli;
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

}  // namespace

BatchConfig choose_batch_config(Handle& h, const BatchShape& shape) {
  calibrate_rates(h);
  calibrate_emulation(h);
  int dev_id = 0, cus = 256;
  hip_check(hipGetDevice(&dev_id), "hipGetDevice");
  hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev_id), "hipDeviceGetAttribute");
  const TuneRates& r = h.rates;
  BatchConfig best;
  best.use_graph = false;
  best.predicted_s = 1e300;
  // tile body: the faster persistent kernel, with as many workgroups as fit the CUs
  best.body = h.batch_body >= 0 ? h.batch_body : (r.kernel_flops_dlds >= r.kernel_flops ? 1 : 0);
  best.workgroups = (best.body ? 3 : 2) * cus;
  const double rate = best.body ? r.kernel_flops_dlds : r.kernel_flops;
  // native or emulated per problem; the compute total and the largest k of the native items
  const int cnt = (int)shape.m.size();
  best.emulate.assign(cnt, 0);
  double csum = 0;
  int kmax = 1;
  for (int i = 0; i < cnt; ++i) {
    const double f = 2.0 * shape.m[i] * (double)shape.n[i] * shape.k[i];
    const double tn = f / rate;
    if (h.ozaki_moduli > 0 && r.emu_flops > 0 && shape.k[i] > 0 && h.batch_emulation != 0) {
      // the INT8 products run one 256x256 tile per CU: a product of fewer tiles than CUs leaves
      // the rest idle
      const double tiles = (double)((shape.m[i] + 255) / 256) * ((shape.n[i] + 255) / 256);
      const double te = f / (r.emu_flops * std::min(1.0, tiles / cus)) + r.emu_overhead_s;
      if (te < tn || h.batch_emulation == 1) {
        best.emulate[i] = 1;
        best.emulated += 1;
        csum += te;
        continue;
      }
    }
    csum += tn;
    kmax = std::max(kmax, shape.k[i]);
  }
  // a group's last partial wave: about one 128x128 tile of the largest k on one CU's share
  const double tail = 2.0 * 128 * 128 * (double)kmax / (rate / (3.0 * cus));
  const double hsum = shape.h2d_bytes / r.h2d_Bps, dsum = shape.d2h_bytes / r.d2h_Bps;
  const bool graph_ok = shape.pinned && h.batch_graph != 0, graph_only = graph_ok && h.batch_graph == 1;
  for (int G = 1; G <= 64; G *= 2) {
    if (G > std::max(1, shape.count)) break;
    for (int q = 1; q <= 2; ++q) {
      if (h.batch_queues > 0 && q != h.batch_queues) continue;
      const double hh = hsum / G, dd = dsum / G, cc = csum / G + tail / q;
      const double T = hh + cc + dd + (G - 1) * std::max(hh, std::max(cc, dd)) + G * r.group_overhead_s;
      if (!graph_only && T < best.predicted_s) {
        best.predicted_s = T;
        best.groups = G;
        best.queues = q;
        best.use_graph = false;
      }
      if (graph_ok) {
        const double Tg = std::max(hsum + dsum + cc, csum + G * tail / q + hh + dd);
        if (Tg < best.predicted_s) {
          best.predicted_s = Tg;
          best.groups = G;
          best.queues = q;
          best.use_graph = true;
        }
      }
    }
  }
  h.last_batch = best;
  return best;
}

}  // namespace tmm

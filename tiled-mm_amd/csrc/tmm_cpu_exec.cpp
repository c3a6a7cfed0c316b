// CPU executor of out-of-core plans, for host-side tests only (built into libtmm_host.so,
// not into libtiledmm.so). It models each queue as a FIFO and runs queue heads in a random
// order that respects only the plan's explicit dependencies, so a missing dependency between
// queues (a slot overwritten while a GEMM still needs it, a copy-back before the last k-chunk,
// ...) produces a wrong result for some seeds. Device memory is poisoned with NaN.
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "tmm_plan.hpp"

using namespace tmm;

namespace {

thread_local std::string g_err;

template <typename T> T conj_if(T v, bool) { return v; }
template <typename R> std::complex<R> conj_if(std::complex<R> v, bool c) { return c ? std::conj(v) : v; }
template <typename T> T nan_value() { return std::numeric_limits<T>::quiet_NaN(); }
template <> std::complex<float> nan_value<std::complex<float>>() { return {nan_value<float>(), nan_value<float>()}; }
template <> std::complex<double> nan_value<std::complex<double>>() { return {nan_value<double>(), nan_value<double>()}; }

template <typename T>
void run(const Plan& P, char opa, char opb, int m, int n, T alpha, const T* A, const T* B, T beta, T* C, T* fullc,
         unsigned seed) {
  std::vector<T> pool(P.pool_elems, nan_value<T>());
  auto dev = [&](const DevRef& r) -> T* {
    return r.region == Region::POOL ? pool.data() + r.off : fullc + r.off;
  };
  const T* host_in[3] = {A, B, C};
  const size_t nops = P.ops.size();
  for (size_t i = 0; i < nops; ++i)
    for (int d : P.ops[i].deps)
      if (d < 0 || (size_t)d >= nops || (size_t)d == i || ((size_t)d > i && d != P.ops[i].late_dep))
        throw std::logic_error("dependency does not point to an earlier op (or the GEMM's late C_in copy)");

  std::vector<std::vector<int>> queues(P.n_queues);
  for (size_t i = 0; i < nops; ++i) queues.at(P.ops[i].queue).push_back((int)i);
  std::vector<size_t> head(P.n_queues, 0);
  std::vector<char> done(nops, 0);
  std::mt19937 rng(seed);
  size_t remaining = nops;
  while (remaining > 0) {
    std::vector<int> ready;
    for (int q = 0; q < P.n_queues; ++q) {
      if (head[q] >= queues[q].size()) continue;
      const Op& op = P.ops[queues[q][head[q]]];
      bool ok = true;
      for (int d : op.deps) ok = ok && done[d];
      if (ok) ready.push_back(q);
    }
    if (ready.empty()) throw std::logic_error("plan deadlocks");
    const int q = ready[rng() % ready.size()];
    const int oi = queues[q][head[q]++];
    const Op& op = P.ops[oi];
    if (op.kind == OpKind::H2D) {
      const T* src = host_in[(int)op.mat] + op.host_off;
      T* dst = dev(op.dev);
      for (int c = 0; c < op.cols; ++c)
        for (int r = 0; r < op.rows; ++r) dst[r + (size_t)c * op.dev.ld] = src[r + (size_t)c * op.host_ld];
    } else if (op.kind == OpKind::D2H) {
      const T* src = dev(op.dev);
      T* dst = C + op.host_off;
      for (int c = 0; c < op.cols; ++c)
        for (int r = 0; r < op.rows; ++r) dst[r + (size_t)c * op.host_ld] = src[r + (size_t)c * op.dev.ld];
    } else {
      const T* a = dev(op.a);
      const T* b = dev(op.b);
      T* cc = dev(op.c);
      const T* ci = op.use_cin ? dev(op.cin) : nullptr;
      for (int j = 0; j < op.n; ++j)
        for (int i = 0; i < op.m; ++i) {
          T s = T(0);
          for (int kk = 0; kk < op.k; ++kk) {
            const T av = opa == 'N' ? a[i + (size_t)kk * op.a.ld] : conj_if(a[kk + (size_t)i * op.a.ld], opa == 'C');
            const T bv = opb == 'N' ? b[kk + (size_t)j * op.b.ld] : conj_if(b[j + (size_t)kk * op.b.ld], opb == 'C');
            s += av * bv;
          }
          T v = alpha * s;
          T& out = cc[i + (size_t)j * op.c.ld];
          if (op.accumulate) v += out;
          if (op.use_cin) v += beta * ci[i + (size_t)j * op.cin.ld];
          out = v;
        }
    }
    done[oi] = 1;
    --remaining;
  }
}

int g_defer_cin = 0;  // PlanRequest::defer_cin of the plans built here (tmmh_set_defer_cin)
int g_pair_max = 0;   // plan_tile_groups pair_max in tmmh_tile_groups (tmmh_set_pair_max)

PlanRequest request(char opa, char opb, int m, int n, int k, int lda, int ldb, int ldc, int beta_nonzero, int copy_back,
                    int tm, int tn, int tk, int nq, long budget, int first_chunk) {
  PlanRequest rq;
  rq.opa = opa;
  rq.opb = opb;
  rq.m = m;
  rq.n = n;
  rq.k = k;
  rq.lda = lda;
  rq.ldb = ldb;
  rq.ldc = ldc;
  rq.beta_nonzero = beta_nonzero != 0;
  rq.copy_back = copy_back != 0;
  rq.tile_m = tm;
  rq.tile_n = tn;
  rq.tile_k = tk;
  rq.n_compute_queues = nq;
  rq.pool_budget_elems = budget > 0 ? (size_t)budget : (size_t)-1;
  rq.first_chunk_k = first_chunk;
  rq.defer_cin = g_defer_cin;
  return rq;
}

}  // namespace

extern "C" {

const char* tmmh_last_error() { return g_err.c_str(); }

// PlanRequest::defer_cin for the plans the functions below build (0 = off, the default).
void tmmh_set_defer_cin(int d) { g_defer_cin = d; }

// pair_max of plan_tile_groups in tmmh_tile_groups (0 = off, the default).
void tmmh_set_pair_max(int v) { g_pair_max = v; }

int tmmh_optimal_tile_size(int dim, int curr, int max) { return optimal_tile_size(dim, curr, max); }

int tmmh_shell_order(int nTm, int nTn, int* out) {
  auto v = shell_order(nTm, nTn);
  for (size_t i = 0; i < v.size(); ++i) {
    out[2 * i] = v[i].first;
    out[2 * i + 1] = v[i].second;
  }
  return (int)v.size();
}

// info: [0] ops, [1] GEMM ops, [2] h2d elems, [3] d2h elems, [4] pool elems, [5] nTm, [6] nTn,
//       [7] nTk, [8] H2D ops, [9] D2H ops
int tmmh_plan_info(char opa, char opb, int m, int n, int k, int lda, int ldb, int ldc, int beta_nonzero, int copy_back,
                   int tm, int tn, int tk, int nq, long budget, int first_chunk, long* info) {
  try {
    Plan P = make_plan(request(opa, opb, m, n, k, lda, ldb, ldc, beta_nonzero, copy_back, tm, tn, tk, nq, budget,
                               first_chunk));
    long g = 0, h = 0, d = 0, late = 0;
    for (auto& op : P.ops) {
      g += op.kind == OpKind::GEMM;
      h += op.kind == OpKind::H2D;
      d += op.kind == OpKind::D2H;
      late += op.late_dep >= 0;
    }
    info[0] = (long)P.ops.size();
    info[1] = g;
    info[2] = (long)P.h2d_elems;
    info[3] = (long)P.d2h_elems;
    info[4] = (long)P.pool_elems;
    info[5] = P.nTm;
    info[6] = P.nTn;
    info[7] = P.nTk;
    info[8] = h;
    info[9] = d;
    info[10] = late;
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// Operand reuse plan of a plan (plan_operand_reuse): per op, out[8*i + ...] = kind, queue,
// A producer, slot_a, prep_a, B producer, slot_b, prep_b. Returns the op count (call with
// out == nullptr to size the buffer).
long tmmh_operand_reuse(char opa, char opb, int m, int n, int k, int lda, int ldb, int ldc, int beta_nonzero,
                        int copy_back, int tm, int tn, int tk, int nq, long budget, int first_chunk, int capacity,
                        int max_k, long* out) {
  try {
    Plan P = make_plan(request(opa, opb, m, n, k, lda, ldb, ldc, beta_nonzero, copy_back, tm, tn, tk, nq, budget,
                               first_chunk));
    if (!out) return (long)P.ops.size();
    const OperandReuse R = plan_operand_reuse(P, capacity, max_k);
    for (size_t i = 0; i < P.ops.size(); ++i) {
      const Op& op = P.ops[i];
      const bool g = op.kind == OpKind::GEMM;
      long* o = out + 8 * i;
      o[0] = (long)op.kind;
      o[1] = op.queue;
      o[2] = g ? operand_producer(P, i, Mat::A) : -1;
      o[3] = R.slot_a[i];
      o[4] = R.prep_a[i];
      o[5] = g ? operand_producer(P, i, Mat::B) : -1;
      o[6] = R.slot_b[i];
      o[7] = R.prep_b[i];
    }
    return (long)P.ops.size();
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// Members of the grouped tile launches of a plan (plan_tile_groups, every GEMM groupable, runs of
// at most max_group in steady chunks and pair_max in the others): out[9*i + ...] = group, tile
// row, tile column, k-chunk, m, n, k, accumulate, use_cin. Returns the member count (call with
// out == nullptr to size the buffer), -1 on a planning error.
long tmmh_group_members(char opa, char opb, int m, int n, int k, int lda, int ldb, int ldc, int beta_nonzero,
                        int copy_back, int tm, int tn, int tk, int nq, long budget, int first_chunk, int max_group,
                        int pair_max, long* out) {
  try {
    Plan P = make_plan(request(opa, opb, m, n, k, lda, ldb, ldc, beta_nonzero, copy_back, tm, tn, tk, nq, budget,
                               first_chunk));
    const auto groups = plan_tile_groups(P, max_group, [](const Op&) { return true; }, pair_max);
    long cnt = 0;
    for (size_t g = 0; g < groups.size(); ++g)
      for (size_t oi : groups[g]) {
        if (out) {
          const Op& op = P.ops[oi];
          long* o = out + 9 * cnt;
          o[0] = (long)g;
          o[1] = op.ti;
          o[2] = op.tj;
          o[3] = op.tc;
          o[4] = op.m;
          o[5] = op.n;
          o[6] = op.k;
          o[7] = op.accumulate ? 1 : 0;
          o[8] = op.use_cin ? 1 : 0;
        }
        ++cnt;
      }
    return cnt;
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// Grouped tile launches of a plan (plan_tile_groups, every GEMM groupable, groups of at most
// max_group) and a progress check of the executor's three threads with them: the compute thread
// takes its ops in plan order and starts a group once every dependency of every member is
// enqueued; each copy thread takes its ops in plan order and enqueues one once its
// dependencies are enqueued. info: [0] groups, [1] grouped GEMMs, [2] 1 if every thread
// finishes (0: some thread waits forever). Returns 0 or 1 on a planning error.
// unsafe = 1 groups runs without the dependency rule (for the test that shows why it is needed).
int tmmh_tile_groups(char opa, char opb, int m, int n, int k, int lda, int ldb, int ldc, int beta_nonzero,
                     int copy_back, int tm, int tn, int tk, int nq, long budget, int first_chunk, int max_group,
                     int unsafe, long* info) {
  try {
    Plan P = make_plan(request(opa, opb, m, n, k, lda, ldb, ldc, beta_nonzero, copy_back, tm, tn, tk, nq, budget,
                               first_chunk));
    std::vector<std::vector<size_t>> groups;
    if (!unsafe) {
      groups = plan_tile_groups(P, max_group, [](const Op&) { return true; }, g_pair_max);
    } else {
      for (int q = 0; q < P.n_queues; ++q) {
        std::vector<size_t> cur;
        for (size_t oi = 0; oi < P.ops.size(); ++oi) {
          const Op& op = P.ops[oi];
          if (op.kind != OpKind::GEMM || op.queue - 2 != q) continue;
          const bool same = !cur.empty() && P.ops[cur[0]].tc == op.tc && (int)cur.size() < max_group;
          if (!same || op.tc == P.nTk - 1) {
            if (cur.size() >= 2) groups.push_back(cur);
            cur.clear();
          }
          if (op.tc != P.nTk - 1) cur.push_back(oi);
        }
        if (cur.size() >= 2) groups.push_back(cur);
      }
    }
    const size_t nops = P.ops.size();
    std::vector<int> group_of(nops, -1);
    long grouped = 0;
    for (size_t g = 0; g < groups.size(); ++g)
      for (size_t i : groups[g]) {
        group_of[i] = (int)g;
        ++grouped;
      }
    std::vector<char> done(nops, 0);
    // threads: 0 = H2D, 1 = D2H, 2 = compute (all compute queues share the one thread)
    std::vector<std::vector<size_t>> seq(3);
    for (size_t i = 0; i < nops; ++i) seq[P.ops[i].queue >= 2 ? 2 : P.ops[i].queue].push_back(i);
    size_t pos[3] = {0, 0, 0};
    bool progress = true;
    while (progress) {
      progress = false;
      for (int th = 0; th < 3; ++th) {
        while (pos[th] < seq[th].size()) {
          const size_t i = seq[th][pos[th]];
          if (done[i]) {  // a later member of a group already started
            ++pos[th];
            continue;
          }
          if (th == 2 && group_of[i] >= 0) {
            const auto& g = groups[group_of[i]];
            bool ok = true;
            for (size_t gi : g)
              for (int d : P.ops[gi].deps)
                if (d >= 0 && P.ops[d].queue != P.ops[gi].queue && !done[d]) ok = false;
            if (!ok) break;
            for (size_t gi : g) done[gi] = 1;
          } else {
            // same-queue dependencies are met by stream order once enqueued
            bool ok = true;
            for (int d : P.ops[i].deps)
              if (d >= 0 && !done[d] && !(P.ops[d].queue == P.ops[i].queue)) ok = false;
            if (!ok) break;
            done[i] = 1;
          }
          ++pos[th];
          progress = true;
        }
      }
    }
    bool all = true;
    for (size_t i = 0; i < nops; ++i) all = all && done[i];
    info[0] = (long)groups.size();
    info[1] = grouped;
    info[2] = all ? 1 : 0;
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int tmmh_simulate_gemm(int dtype, char opa, char opb, int m, int n, int k, const void* alpha, const void* A, int lda,
                       const void* B, int ldb, const void* beta, void* C, int ldc, int copy_back, void* fullc, int tm,
                       int tn, int tk, int nq, long budget, int first_chunk, unsigned seed) {
  try {
    bool beta_nz;
    switch (dtype) {
      case 0: beta_nz = *(const float*)beta != 0; break;
      case 1: beta_nz = *(const double*)beta != 0; break;
      case 2: beta_nz = *(const std::complex<float>*)beta != std::complex<float>(0); break;
      default: beta_nz = *(const std::complex<double>*)beta != std::complex<double>(0); break;
    }
    Plan P = make_plan(request(opa, opb, m, n, k, lda, ldb, ldc, beta_nz, copy_back, tm, tn, tk, nq, budget,
                               first_chunk));
    switch (dtype) {
      case 0:
        run<float>(P, opa, opb, m, n, *(const float*)alpha, (const float*)A, (const float*)B, *(const float*)beta,
                   (float*)C, (float*)fullc, seed);
        break;
      case 1:
        run<double>(P, opa, opb, m, n, *(const double*)alpha, (const double*)A, (const double*)B, *(const double*)beta,
                    (double*)C, (double*)fullc, seed);
        break;
      case 2: {
        using Z = std::complex<float>;
        run<Z>(P, opa, opb, m, n, *(const Z*)alpha, (const Z*)A, (const Z*)B, *(const Z*)beta, (Z*)C, (Z*)fullc, seed);
        break;
      }
      default: {
        using Z = std::complex<double>;
        run<Z>(P, opa, opb, m, n, *(const Z*)alpha, (const Z*)A, (const Z*)B, *(const Z*)beta, (Z*)C, (Z*)fullc, seed);
        break;
      }
    }
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

}  // extern "C"

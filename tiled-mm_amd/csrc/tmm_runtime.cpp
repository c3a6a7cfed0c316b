// Host runtime: handle, staging pipeline and GPU executor of the out-of-core plan.
//
// The reference drives one gemm call from the caller's thread by enqueueing, per C tile and
// k-tile, H2D copies of A/B(/C) into per-stream slots, a vendor GEMM and a D2H copy
// (Tiled-MM src/Tiled-MM/tiled_mm.cpp:270-365), after page-locking the caller's matrices
// with HostRegister (tiled_mm.cpp:529-554). Here:
//   * the plan (tmm_plan.cpp) moves every A and B block over PCIe once;
//   * one H2D queue, one D2H queue and n_streams compute queues run it, ordered only by the
//     plan's explicit dependencies (events between queues);
//   * pinned caller memory is copied with hipMemcpy2DAsync directly; pageable caller memory
//     is streamed through a ring of pinned staging slots filled by a host thread pool (the
//     caller's pages are never registered), and results come back through a second ring that
//     a worker thread scatters into the caller's C.
#include "tmm_runtime.hpp"

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cstring>
#include <memory>
#include <sstream>

#include "tmm_kernels.h"
#include "tmm_ozaki.h"

namespace tmm {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    (void)hipGetLastError();
    std::ostringstream os;
    os << what << ": " << hipGetErrorString(e);
    throw Error(e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation ? 3 : 2, os.str());
  }
}

// ------------------------------------------------------------------------------------------
// ThreadPool
// ------------------------------------------------------------------------------------------
ThreadPool::ThreadPool(int n) {
  for (int i = 0; i < n; ++i)
    workers_.emplace_back([this]() {
      for (;;) {
        std::shared_ptr<Job> j;
        {
          std::unique_lock<std::mutex> lk(mu_);
          cv_.wait(lk, [&] { return stop_ || !jobs_.empty(); });
          if (stop_ && jobs_.empty()) return;
          j = jobs_.front();
          // every item of the front job is claimed once a worker sees next >= n
          if (j->next.load() >= j->n) {
            jobs_.pop_front();
            continue;
          }
        }
        run_job(*j);
      }
    });
}

ThreadPool::~ThreadPool() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
    jobs_.clear();
  }
  cv_.notify_all();
  for (auto& t : workers_) t.join();
}

void ThreadPool::run_job(Job& j) {
  for (;;) {
    const int i = j.next.fetch_add(1);
    if (i >= j.n) break;
    (*j.fn)(i);
    if (j.done.fetch_add(1) + 1 == j.n) {
      std::lock_guard<std::mutex> lk(mu_);
      done_cv_.notify_all();
    }
  }
}

void ThreadPool::parallel_for(int n, const std::function<void(int)>& fn) {
  if (n <= 0) return;
  if (n == 1 || workers_.empty()) {
    for (int i = 0; i < n; ++i) fn(i);
    return;
  }
  auto j = std::make_shared<Job>();
  j->fn = &fn;
  j->n = n;
  {
    std::lock_guard<std::mutex> lk(mu_);
    jobs_.push_back(j);
  }
  cv_.notify_all();
  run_job(*j);
  std::unique_lock<std::mutex> lk(mu_);
  done_cv_.wait(lk, [&] { return j->done.load() >= j->n; });
  auto it = std::find(jobs_.begin(), jobs_.end(), j);
  if (it != jobs_.end()) jobs_.erase(it);
}

// ------------------------------------------------------------------------------------------
// Memory
// ------------------------------------------------------------------------------------------
void DeviceBuffer::reserve(size_t b) {
  if (b <= bytes) return;
  release();
  void* p = nullptr;
  hip_check(hipMalloc(&p, b), "hipMalloc");
  ptr = p;
  bytes = b;
}

void DeviceBuffer::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

void StagingRing::ensure(int n, size_t bytes) {
  if ((int)slot.size() >= n && slot_bytes >= bytes) return;
  release();
  slot.assign(n, nullptr);
  ev.assign(n, nullptr);
  ev_pending.assign(n, false);
  slot_bytes = bytes;
  for (int i = 0; i < n; ++i) {
    hip_check(hipHostMalloc(&slot[i], bytes, hipHostMallocDefault), "hipHostMalloc (staging)");
    hip_check(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming), "hipEventCreate");
  }
  next = 0;
}

void StagingRing::release() {
  for (size_t i = 0; i < slot.size(); ++i) {
    if (ev[i]) {
      if (ev_pending[i]) (void)hipEventSynchronize(ev[i]);
      (void)hipEventDestroy(ev[i]);
    }
    if (slot[i]) (void)hipHostFree(slot[i]);
  }
  slot.clear();
  ev.clear();
  ev_pending.clear();
  slot_bytes = 0;
  next = 0;
}

bool is_pinned_range(const void* p, size_t bytes) {
  if (!p || bytes == 0) return true;
  auto pinned_at = [](const void* q, void** devp) {
    hipPointerAttribute_t at;
    hipError_t e = hipPointerGetAttributes(&at, q);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    if (at.type != hipMemoryTypeHost) return false;
    *devp = at.devicePointer;
    return true;
  };
  void *d0 = nullptr, *d1 = nullptr;
  const char* last = (const char*)p + bytes - 1;
  if (!pinned_at(p, &d0) || !pinned_at(last, &d1)) return false;
  // both ends must map into one allocation with the same offset
  return (const char*)d1 - (const char*)d0 == (ptrdiff_t)(bytes - 1);
}

// ------------------------------------------------------------------------------------------
// Handle
// ------------------------------------------------------------------------------------------
Handle::Handle(int dt, int streams, int tm, int tn, int tk)
    : dtype(dt), n_streams(streams), max_tm(tm), max_tn(tn), max_tk(tk) {
  elem_bytes = dt == 0 ? 4 : dt == 1 ? 8 : dt == 2 ? 8 : 16;
  hip_check(hipGetDevice(&device), "hipGetDevice");
  // The copy streams first, then one compute stream; further compute streams only when a plan
  // runs tile GEMMs on several queues (compute_stream). HIP streams share the device's few
  // hardware queues once they outnumber them, and a copy stream sharing one with the busy
  // compute stream would execute its copies in order with the GEMMs; creating only the streams
  // in use, the three busy ones back to back, keeps them on distinct queues.
  hip_check(hipStreamCreateWithFlags(&h2d, hipStreamNonBlocking), "hipStreamCreate");
  hip_check(hipStreamCreateWithFlags(&d2h, hipStreamNonBlocking), "hipStreamCreate");
  set_num_streams(streams);
}

Handle::~Handle() {
  for (auto s : compute)
    if (s) (void)hipStreamSynchronize(s);
  for (hipStream_t s : {h2d, d2h, h2d_aux})
    if (s) (void)hipStreamSynchronize(s);
  for (auto s : compute)
    if (s) (void)hipStreamDestroy(s);
  for (hipStream_t s : {h2d, d2h, h2d_aux})
    if (s) (void)hipStreamDestroy(s);
  for (size_t i = 0; i < ozwork_ev.size(); ++i)
    if (ozwork_ev[i]) {
      if (ozwork_pending[i]) (void)hipEventSynchronize(ozwork_ev[i]);
      (void)hipEventDestroy(ozwork_ev[i]);
    }
  for (auto e : events) (void)hipEventDestroy(e);
  for (auto e : prof_events) (void)hipEventDestroy(e);
  delete threads;
}

void* Handle::ozwork_acquire(size_t i, size_t bytes, hipStream_t s) {
  if (ozwork.size() <= i) ozwork.resize(i + 1);
  if (ozwork_ev.size() <= i) {
    ozwork_ev.resize(i + 1, nullptr);
    ozwork_pending.resize(i + 1, 0);
  }
  if (ozwork_pending[i]) {
    if (!s || bytes > ozwork[i].bytes) {
      // host caller, or the buffer is about to be freed and reallocated: wait for its last user
      hip_check(hipEventSynchronize(ozwork_ev[i]), "hipEventSynchronize (Ozaki workspace)");
      ozwork_pending[i] = 0;
    } else {
      hip_check(hipStreamWaitEvent(s, ozwork_ev[i], 0), "hipStreamWaitEvent (Ozaki workspace)");
    }
  }
  ozwork[i].reserve(bytes);
  return ozwork[i].ptr;
}

void Handle::ozwork_release(size_t i, hipStream_t s) {
  if (!ozwork_ev[i]) hip_check(hipEventCreateWithFlags(&ozwork_ev[i], hipEventDisableTiming), "hipEventCreate");
  hip_check(hipEventRecord(ozwork_ev[i], s), "hipEventRecord (Ozaki workspace)");
  ozwork_pending[i] = 1;
}

void Handle::ozwork_quiesce() {
  for (size_t i = 0; i < ozwork_pending.size(); ++i)
    if (ozwork_pending[i]) {
      hip_check(hipEventSynchronize(ozwork_ev[i]), "hipEventSynchronize (Ozaki workspace)");
      ozwork_pending[i] = 0;
    }
}

void Handle::set_num_streams(int s) {
  if (s <= 0) throw Error(1, "number of streams must be positive");
  for (auto st : compute) {
    if (!st) continue;
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    hip_check(hipStreamDestroy(st), "hipStreamDestroy");
  }
  compute.assign(s, nullptr);
  hip_check(hipStreamCreateWithFlags(&compute[0], hipStreamNonBlocking), "hipStreamCreate");
  n_streams = s;
  tilebuf_elems[0] = tilebuf_elems[1] = tilebuf_elems[2] = 0;
}

hipStream_t Handle::compute_stream(int i) {
  if (!compute[i]) hip_check(hipStreamCreateWithFlags(&compute[i], hipStreamNonBlocking), "hipStreamCreate");
  return compute[i];
}

hipStream_t Handle::h2d_aux_stream() {
  if (!h2d_aux) hip_check(hipStreamCreateWithFlags(&h2d_aux, hipStreamNonBlocking), "hipStreamCreate");
  return h2d_aux;
}

void Handle::sync_compute() {
  for (auto s : compute)
    if (s) hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

ThreadPool& Handle::pool_threads() {
  int want = host_threads > 0 ? host_threads : (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
  if (!threads || threads->size() != want - 1) {
    delete threads;
    threads = new ThreadPool(want - 1);  // the calling thread works too
  }
  return *threads;
}

hipEvent_t Handle::event(size_t i) {
  while (events.size() <= i) {
    hipEvent_t e;
    hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    events.push_back(e);
  }
  return events[i];
}

// ------------------------------------------------------------------------------------------
// Kernel dispatch
// ------------------------------------------------------------------------------------------
namespace {

template <typename R>
GemmParams<R> make_params(char opa, char opb, int m, int n, int k, const R* alpha, bool cplx, const void* a, long lda,
                          const void* b, long ldb, const R* beta, bool use_beta, const void* cin, long ldcin,
                          bool accumulate, void* c, long ldc) {
  GemmParams<R> p;
  p.opa = opa;
  p.opb = opb;
  p.m = m;
  p.n = n;
  p.k = k;
  p.alpha = alpha[0];
  p.alpha_im = cplx ? alpha[1] : R(0);
  p.a = (const R*)a;
  p.lda = lda;
  p.b = (const R*)b;
  p.ldb = ldb;
  p.beta = use_beta ? beta[0] : R(0);
  p.beta_im = use_beta && cplx ? beta[1] : R(0);
  p.cin = (const R*)cin;
  p.ldcin = ldcin;
  p.accumulate = accumulate ? 1 : 0;
  p.c = (R*)c;
  p.ldc = ldc;
  p.variant = 0;
  return p;
}

void launch(int dtype, char opa, char opb, int m, int n, int k, const void* alpha, const void* a, long lda,
            const void* b, long ldb, const void* beta, bool use_beta, const void* cin, long ldcin, bool accumulate,
            void* c, long ldc, hipStream_t s, int variant = 0) {
  hipError_t e = hipSuccess;
  switch (dtype) {
    case 0: {
      GemmParams<float> p = make_params<float>(opa, opb, m, n, k, (const float*)alpha, false, a, lda, b, ldb,
                                               (const float*)beta, use_beta, cin, ldcin, accumulate, c, ldc);
      p.variant = variant;
      e = launch_gemm_s(p, s);
      break;
    }
    case 1: {
      GemmParams<double> p = make_params<double>(opa, opb, m, n, k, (const double*)alpha, false, a, lda, b, ldb,
                                                 (const double*)beta, use_beta, cin, ldcin, accumulate, c, ldc);
      p.variant = variant;
      e = launch_gemm_d(p, s);
      break;
    }
    case 2: {
      GemmParams<float> p = make_params<float>(opa, opb, m, n, k, (const float*)alpha, true, a, lda, b, ldb,
                                               (const float*)beta, use_beta, cin, ldcin, accumulate, c, ldc);
      p.variant = variant;
      e = launch_gemm_c(p, s);
      break;
    }
    case 3: {
      GemmParams<double> p = make_params<double>(opa, opb, m, n, k, (const double*)alpha, true, a, lda, b, ldb,
                                                 (const double*)beta, use_beta, cin, ldcin, accumulate, c, ldc);
      p.variant = variant;
      e = launch_gemm_z(p, s);
      break;
    }
    default:
      throw Error(4, "unknown element type");
  }
  hip_check(e, "tile GEMM launch");
}

// Tile GEMM through the handle's compute mode: native MFMA kernels, or for FP64 with
// ozaki_slices > 0 the split-precision INT8 path (workspace `ws` of the launching queue).
// Returns the INT8 slice products computed (0 for native).
bool ozaki_active(const Handle& h) { return h.dtype == 1 && (h.ozaki_moduli > 0 || h.ozaki_slices > 0); }

size_t ozaki_ws_bytes(const Handle& h, int m, int n, int k) {
  if (h.dtype != 1) return 0;
  if (h.ozaki_moduli > 0) return ozaki2_workspace_bytes(m, n, k, h.ozaki_moduli);
  if (h.ozaki_slices > 0) return ozaki_workspace_bytes(m, n, k, h.ozaki_slices);
  return 0;
}

// fixed_work: the emulation's workspace is this buffer (at least ozaki_ws_bytes) instead of the
// handle's workspace ws, with no ordering against other users.
int launch_mode(Handle& h, int ws, char opa, char opb, int m, int n, int k, const void* alpha, const void* a, long lda,
                const void* b, long ldb, const void* beta, bool use_beta, const void* cin, long ldcin,
                bool accumulate, void* c, long ldc, hipStream_t s, void* fixed_work = nullptr) {
  if (ozaki_active(h) && k > 0) {  // k = 0: C = beta*C, nothing to emulate
    void* work = fixed_work ? fixed_work : h.ozwork_acquire(ws, ozaki_ws_bytes(h, m, n, k), s);
    int nprod = 0;
    const double al = *(const double*)alpha, be = use_beta ? *(const double*)beta : 0.0;
    if (h.ozaki_moduli > 0) {
      hip_check(ozaki2_dgemm(opa, opb, m, n, k, al, (const double*)a, lda, (const double*)b, ldb, be,
                             (const double*)cin, ldcin, accumulate, (double*)c, ldc, h.ozaki_moduli, work, s,
                             &nprod),
                "Ozaki (CRT) DGEMM launch");
      // products per full inner dimension: k above the CRT pass length runs in several passes
      // of nmod products over parts of k
      if (nprod > 0) nprod = h.ozaki_moduli;
    } else
      hip_check(ozaki_dgemm(opa, opb, m, n, k, al, (const double*)a, lda, (const double*)b, ldb, be,
                            (const double*)cin, ldcin, accumulate, (double*)c, ldc, h.ozaki_slices, work, s, &nprod),
                "Ozaki DGEMM launch");
    if (!fixed_work) h.ozwork_release(ws, s);
    return nprod;
  }
  launch(h.dtype, opa, opb, m, n, k, alpha, a, lda, b, ldb, beta, use_beta, cin, ldcin, accumulate, c, ldc, s,
         h.dgemm_variant);
  return 0;
}

bool scalar_is_zero(int dtype, const void* v) {
  switch (dtype) {
    case 0: return ((const float*)v)[0] == 0.0f;
    case 1: return ((const double*)v)[0] == 0.0;
    case 2: return ((const float*)v)[0] == 0.0f && ((const float*)v)[1] == 0.0f;
    default: return ((const double*)v)[0] == 0.0 && ((const double*)v)[1] == 0.0;
  }
}

// Copy `height` columns of `width` bytes with arbitrary pitches, split over the pool.
void strided_copy_parallel(ThreadPool& tp, char* dst, size_t dpitch, const char* src, size_t spitch, size_t width,
                           size_t height) {
  const size_t total = width * height;
  if (total == 0) return;
  const size_t min_task = 1 << 20;
  int ntask = (int)std::min<size_t>(std::max<size_t>(1, total / min_task), (size_t)(tp.size() + 1) * 2);
  tp.parallel_for(ntask, [&](int t) {
    size_t b0 = total * (size_t)t / ntask, b1 = total * (size_t)(t + 1) / ntask;
    while (b0 < b1) {
      const size_t col = b0 / width, off = b0 % width;
      const size_t len = std::min(width - off, b1 - b0);
      std::memcpy(dst + col * dpitch + off, src + col * spitch + off, len);
      b0 += len;
    }
  });
}

// Worker that scatters device-to-host staging slots into the caller's pageable C.
struct Scatterer {
  struct Item {
    hipEvent_t ev;
    int slot;
    char* dst;
    size_t dpitch;
    const char* src;
    size_t width, height;
  };
  ThreadPool& tp;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Item> q;
  std::vector<bool> busy;
  bool stop = false;
  std::string error;
  std::thread th;

  Scatterer(ThreadPool& t, int nslots) : tp(t), busy(nslots, false) {
    th = std::thread([this]() { loop(); });
  }
  void loop() {
    for (;;) {
      Item it;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return stop || !q.empty(); });
        if (q.empty()) return;
        it = q.front();
        q.pop_front();
      }
      try {
        hip_check(hipEventSynchronize(it.ev), "hipEventSynchronize (D2H staging)");
        strided_copy_parallel(tp, it.dst, it.dpitch, it.src, it.width, it.width, it.height);
      } catch (std::exception& e) {
        std::lock_guard<std::mutex> lk(mu);
        if (error.empty()) error = e.what();
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        busy[it.slot] = false;
      }
      cv.notify_all();
    }
  }
  void acquire(int slot) {
    std::unique_lock<std::mutex> lk(mu);
    cv.wait(lk, [&] { return !busy[slot]; });
    busy[slot] = true;
  }
  void push(const Item& it) {
    {
      std::lock_guard<std::mutex> lk(mu);
      q.push_back(it);
    }
    cv.notify_all();
  }
  void finish() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv.notify_all();
    if (th.joinable()) th.join();
  }
  ~Scatterer() { finish(); }
};

// Staging ring for pageable memory: 8 pinned slots of 128 MiB per direction. Pageable Ozaki
// scheme II DGEMM TN 40000^3 (host-link bound): 0.90-0.99 s per call with 6 x 64 MiB, 0.834 s with
// 6 x 128 MiB, 0.838 with 4 x 256, 0.841 with 12 x 64, 0.825 with 12 x 128; the native call
// (kernel bound) ran 1.856-1.862 s with every ring.
constexpr size_t kStageBytes = 128u << 20;
constexpr int kStageSlots = 8;

// Cross-thread state of one plan execution (one thread per copy queue, the caller's thread for
// the GEMMs). Every op is "marked" once it has been enqueued and, when another queue depends on
// it, its completion event recorded; a thread that enqueues an op with such a dependency first
// waits for the mark, so no wait ever sees an event of an earlier call. Ops are enqueued per
// thread in plan order and dependencies point to earlier ops, so the lowest unmarked op can
// always be enqueued; a copy thread's host wait for a GEMM's completion ends because the GEMM
// was enqueued after waiting only for marked (enqueued) work: the waits cannot deadlock.
struct ExecSync {
  std::mutex mu;
  std::condition_variable cv;
  std::vector<char> marked;
  bool failed = false;
  int code = 0;
  std::string error;
  explicit ExecSync(size_t n) : marked(n, 0) {}
  bool wait(int d) {
    std::unique_lock<std::mutex> lk(mu);
    cv.wait(lk, [&] { return failed || marked[d]; });
    return !failed;
  }
  void mark(size_t i) {
    {
      std::lock_guard<std::mutex> lk(mu);
      marked[i] = 1;
    }
    cv.notify_all();
  }
  void fail(int c, const char* what) {
    {
      std::lock_guard<std::mutex> lk(mu);
      if (!failed) {
        failed = true;
        code = c;
        error = what;
      }
    }
    cv.notify_all();
  }
};

void execute_plan(Handle& h, const Plan& P, const GemmCall& call, const bool pinned[3], bool use_beta) {
  const int es = h.elem_bytes;
  const size_t nops = P.ops.size();
  std::vector<char> need(nops, 0);
  for (const Op& op : P.ops)
    for (int d : op.deps)
      if (d >= 0 && P.ops[d].queue != op.queue) need[d] = 1;

  auto stream_of = [&](int q) -> hipStream_t {
    if (q == 0) return h.h2d;
    if (q == 1) return h.d2h;
    return h.compute_stream((q - 2) % (int)h.compute.size());
  };
  auto devptr = [&](const DevRef& r) -> char* {
    char* base = (char*)(r.region == Region::POOL ? h.pool.ptr : h.fullc.ptr);
    return base + r.off * es;
  };
  const char* host_in[3] = {(const char*)call.a, (const char*)call.b, (const char*)call.c};
  char* host_c = (char*)call.c;

  const bool any_pageable_in = !pinned[0] || !pinned[1] || (use_beta && !pinned[2]);
  const bool pageable_out = call.copy_back && !pinned[2];
  if (any_pageable_in) h.stage_in.ensure(kStageSlots, kStageBytes);
  if (pageable_out) h.stage_out.ensure(kStageSlots, kStageBytes);
  ThreadPool* tp = (any_pageable_in || pageable_out) ? &h.pool_threads() : nullptr;
  std::unique_ptr<Scatterer> scat;
  if (pageable_out) scat.reset(new Scatterer(*tp, kStageSlots));

  for (size_t i = 0; i < nops; ++i) h.event(i);
  const bool crt_reuse = h.dtype == 1 && h.ozaki_moduli > 0 && h.ozaki_reuse;
  OperandReuse reuse;
  size_t reuse_slot_bytes = 0;
  if (ozaki_active(h)) {
    // reserve the slice workspaces up front: growing one later would free memory that an
    // earlier, still running tile GEMM of the same queue uses
    std::vector<size_t> need_ws(P.n_queues, 0);
    for (const Op& op : P.ops)
      if (op.kind == OpKind::GEMM)
        need_ws[op.queue - 2] = std::max(need_ws[op.queue - 2], ozaki_ws_bytes(h, op.m, op.n, op.k));
    for (size_t q = 0; q < need_ws.size(); ++q)
      if (need_ws[q]) h.ozwork_acquire(q, need_ws[q], nullptr);
    if (crt_reuse) {
      // operand cache (scheme II): one slot holds the exponents and residues of one A or B
      // block; slots take at most half of the free device memory (and 64 GB), shared by the
      // compute queues that run scheme II products (big tiles use one)
      std::vector<bool> crt_queue(P.n_queues, false);
      for (const Op& op : P.ops)
        if (op.kind == OpKind::GEMM && op.k > 0 && op.k <= kCrtMaxK) {
          reuse_slot_bytes = std::max({reuse_slot_bytes, ozaki2_operand_bytes(op.m, op.k, h.ozaki_moduli),
                                       ozaki2_operand_bytes(op.n, op.k, h.ozaki_moduli)});
          crt_queue[op.queue - 2] = true;
        }
      const int n_crt_queues = std::max(1, (int)std::count(crt_queue.begin(), crt_queue.end(), true));
      size_t free_b = 0, total_b = 0, held = 0;
      hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
      for (const DeviceBuffer& b : h.ozcache) held += b.bytes;
      const size_t budget = std::min((free_b + held) / 2, (size_t)64 << 30);
      const int cap = reuse_slot_bytes ? (int)std::min<size_t>(budget / reuse_slot_bytes / n_crt_queues, 256) : 0;
      reuse = plan_operand_reuse(P, cap, kCrtMaxK);
      if (h.ozcache.size() < (size_t)P.n_queues) h.ozcache.resize(P.n_queues);
      for (int q = 0; q < P.n_queues; ++q)
        if (reuse.slots[q] > 0) h.ozcache[q].reserve((size_t)reuse.slots[q] * reuse_slot_bytes);
    }
  }
  // Native tile GEMMs: grouped launches of the tile GEMMs of each steady k-chunk
  // (plan_tile_groups), one partially filled last wave per group instead of per tile.
  const bool cplx = h.dtype >= 2, single = h.dtype == 0 || h.dtype == 2;
  auto tile_params = [&](const Op& op) {
    GemmParams<double> p = make_params<double>(call.opa, call.opb, op.m, op.n, op.k, (const double*)call.alpha,
                                               cplx, devptr(op.a), op.a.ld, devptr(op.b), op.b.ld,
                                               (const double*)call.beta, op.use_cin,
                                               op.use_cin ? devptr(op.cin) : nullptr, op.cin.ld, op.accumulate,
                                               devptr(op.c), op.c.ld);
    p.variant = h.dgemm_variant;
    return p;
  };
  auto tile_params_f = [&](const Op& op) {
    GemmParams<float> p = make_params<float>(call.opa, call.opb, op.m, op.n, op.k, (const float*)call.alpha, cplx,
                                             devptr(op.a), op.a.ld, devptr(op.b), op.b.ld, (const float*)call.beta,
                                             op.use_cin, op.use_cin ? devptr(op.cin) : nullptr, op.cin.ld,
                                             op.accumulate, devptr(op.c), op.c.ld);
    p.variant = h.dgemm_variant;
    return p;
  };
  auto groupable = [&](const Op& op) {
    switch (h.dtype) {
      case 0: return sgemm_groupable(tile_params_f(op));
      case 1: return dgemm_groupable(tile_params(op));
      case 2: return cgemm_groupable(tile_params_f(op));
      default: return zgemm_groupable(tile_params(op));
    }
  };
  std::vector<int> group_of(nops, -1);
  std::vector<std::vector<size_t>> groups;
  if (!ozaki_active(h) && h.tile_groups) {
    groups = plan_tile_groups(P, kDgemmGroupMax, groupable, h.tile_groups >= 2 ? h.tile_groups : 0);
    for (size_t g = 0; g < groups.size(); ++g)
      for (size_t i : groups[g]) group_of[i] = (int)g;
  }

  size_t n_gemm = 0;
  for (const Op& op : P.ops) n_gemm += op.kind == OpKind::GEMM;
  if (h.profile) {
    while (h.prof_events.size() < 2 * n_gemm) {
      hipEvent_t e;
      hip_check(hipEventCreate(&e), "hipEventCreate");
      h.prof_events.push_back(e);
    }
  }
  std::vector<hipStream_t> prof_stream;
  std::vector<double> prof_flops;
  std::vector<char> prof_grouped;  // 1: a grouped launch (several tiles in one kernel)
  std::vector<double> prof_bytes;  // grouped launches: algorithmic HBM bytes (distinct A and B
                                   // blocks read once, each C tile read and written once)
  double int8_ops = 0;

  // Enqueue one op on its stream. Staged copies block the calling thread on host work (packing
  // into a pinned slot, waiting for a free slot), so with pageable host memory the H2D and D2H
  // queues run on threads of their own and the GEMM launches never wait behind them.
  ExecSync sync(nops);
  // Copy ops wait for their dependencies on the host (an event synchronisation) right before
  // their first DMA is enqueued, instead of with hipStreamWaitEvent: no barrier packet then
  // sits in a copy stream, where, should that stream share a hardware queue with a compute
  // stream, it would hold back the GEMMs submitted after it.
  auto wait_deps_host = [&](const Op& op) {
    for (int d : op.deps) {
      if (d < 0 || P.ops[d].queue == op.queue) continue;
      if (!sync.wait(d)) throw Error(2, "aborted after a failure in another queue");
      hip_check(hipEventSynchronize(h.events[d]), "hipEventSynchronize (copy dependency)");
    }
  };
  auto enqueue = [&](size_t oi) {
    const Op& op = P.ops[oi];
    hipStream_t s = stream_of(op.queue);
    if (op.kind == OpKind::H2D && op.rows > 0 && op.cols > 0) {
      const int mi = (int)op.mat;
      const char* src = host_in[mi] + op.host_off * es;
      char* dst = devptr(op.dev);
      const size_t width = (size_t)op.rows * es;
      if (pinned[mi]) {
        wait_deps_host(op);
        hip_check(hipMemcpy2DAsync(dst, op.dev.ld * es, src, op.host_ld * es, width, op.cols, hipMemcpyHostToDevice, s),
                  "hipMemcpy2DAsync H2D");
      } else {
        StagingRing& R = h.stage_in;
        if (width > R.slot_bytes) R.ensure(kStageSlots, width);
        const int cpp = (int)std::max<size_t>(1, R.slot_bytes / width);
        for (int c0 = 0; c0 < op.cols; c0 += cpp) {
          const int nc = std::min(cpp, op.cols - c0);
          const int sl = R.next;
          R.next = (R.next + 1) % (int)R.slot.size();
          if (R.ev_pending[sl]) hip_check(hipEventSynchronize(R.ev[sl]), "hipEventSynchronize (H2D staging)");
          strided_copy_parallel(*tp, (char*)R.slot[sl], width, src + (size_t)c0 * op.host_ld * es, op.host_ld * es,
                                width, nc);
          if (c0 == 0) wait_deps_host(op);  // the first piece is packed while the dependency runs
          hip_check(hipMemcpy2DAsync(dst + (size_t)c0 * op.dev.ld * es, op.dev.ld * es, R.slot[sl], width, width, nc,
                                     hipMemcpyHostToDevice, s),
                    "hipMemcpy2DAsync H2D (staged)");
          hip_check(hipEventRecord(R.ev[sl], s), "hipEventRecord");
          R.ev_pending[sl] = true;
        }
      }
    } else if (op.kind == OpKind::GEMM) {
      const size_t pi = prof_stream.size();
      if (h.profile) hip_check(hipEventRecord(h.prof_events[2 * pi], s), "hipEventRecord");
      int np;
      if (crt_reuse && op.k > 0 && op.k <= kCrtMaxK) {
        // scheme II with shared operands: the exponents and residues of a cached A or B block
        // are computed by the first GEMM that reads it; the workspace of the queue holds the
        // uncached operands and the product's residue planes
        const int q = op.queue - 2, nmod = h.ozaki_moduli;
        char* ws = (char*)h.ozwork[q].ptr;
        char* cache = (char*)(q < (int)h.ozcache.size() ? h.ozcache[q].ptr : nullptr);
        const bool a_kc = call.opa != 'N', b_kc = call.opb == 'N';
        auto operand = [&](int slot, bool prep, char* own, const DevRef& ref, bool kc, int rows) -> const void* {
          char* dst = slot >= 0 ? cache + (size_t)slot * reuse_slot_bytes : own;
          if (slot < 0 || prep)
            hip_check(ozaki2_prepare((const double*)devptr(ref), ref.ld, kc, rows, op.k, nmod, dst, s),
                      "Ozaki (CRT) operand preparation");
          return dst;
        };
        const size_t a_bytes = ozaki2_operand_bytes(op.m, op.k, nmod), b_bytes = ozaki2_operand_bytes(op.n, op.k, nmod);
        const void* pa = operand(reuse.slot_a[oi], reuse.prep_a[oi], ws, op.a, a_kc, op.m);
        const void* pb = operand(reuse.slot_b[oi], reuse.prep_b[oi], ws + a_bytes, op.b, b_kc, op.n);
        const double be = op.use_cin ? *(const double*)call.beta : 0.0;
        void* prod = ws + a_bytes + b_bytes;
        hip_check(ozaki2_products(op.m, op.n, op.k, pa, pb, nmod, prod, s), "Ozaki (CRT) products");
        if (op.late_dep >= 0 && P.ops[op.late_dep].queue != op.queue) {
          // C_in copied after this GEMM's operands (PlanRequest::defer_cin): only the
          // reconstruction reads it
          if (!sync.wait(op.late_dep)) throw Error(2, "aborted after a failure in another queue");
          hip_check(hipStreamWaitEvent(s, h.events[op.late_dep], 0), "hipStreamWaitEvent");
        }
        hip_check(ozaki2_reconstruct(op.m, op.n, op.k, *(const double*)call.alpha, pa, pb, be,
                                     op.use_cin ? (const double*)devptr(op.cin) : nullptr, op.cin.ld, op.accumulate,
                                     (double*)devptr(op.c), op.c.ld, nmod, prod, s),
                  "Ozaki (CRT) reconstruction");
        // blocks with an Inf, a NaN or a huge entry: the native kernel on the FP64 blocks, still
        // in place (a slot is refilled only after every GEMM that reads it)
        hip_check(ozaki2_fallback(call.opa, call.opb, op.m, op.n, op.k, *(const double*)call.alpha,
                                  (const double*)devptr(op.a), op.a.ld, (const double*)devptr(op.b), op.b.ld, be,
                                  op.use_cin ? (const double*)devptr(op.cin) : nullptr, op.cin.ld, op.accumulate,
                                  (double*)devptr(op.c), op.c.ld, pa, pb, s),
                  "Ozaki (CRT) native fallback");
        np = op.m > 0 && op.n > 0 ? nmod : 0;
      } else {
        np = launch_mode(h, op.queue - 2, call.opa, call.opb, op.m, op.n, op.k, call.alpha, devptr(op.a), op.a.ld,
                         devptr(op.b), op.b.ld, call.beta, op.use_cin, op.use_cin ? devptr(op.cin) : nullptr,
                         op.cin.ld, op.accumulate, devptr(op.c), op.c.ld, s);
      }
      int8_ops += 2.0 * np * op.m * op.n * op.k;
      if (h.profile) {
        hip_check(hipEventRecord(h.prof_events[2 * pi + 1], s), "hipEventRecord");
        prof_stream.push_back(s);
        prof_flops.push_back((h.dtype >= 2 ? 8.0 : 2.0) * op.m * op.n * op.k);
        prof_grouped.push_back(0);
        prof_bytes.push_back(0);
      }
    } else if (op.kind == OpKind::D2H && op.rows > 0 && op.cols > 0) {
      char* dst = host_c + op.host_off * es;
      const char* src = devptr(op.dev);
      const size_t width = (size_t)op.rows * es;
      wait_deps_host(op);
      h.stats[19] = 1;  // the caller's C is about to change
      if (pinned[2]) {
        hip_check(hipMemcpy2DAsync(dst, op.host_ld * es, src, op.dev.ld * es, width, op.cols, hipMemcpyDeviceToHost, s),
                  "hipMemcpy2DAsync D2H");
      } else {
        StagingRing& R = h.stage_out;
        if (width > R.slot_bytes) throw Error(5, "C tile column larger than a staging slot");
        const int cpp = (int)std::max<size_t>(1, R.slot_bytes / width);
        for (int c0 = 0; c0 < op.cols; c0 += cpp) {
          const int nc = std::min(cpp, op.cols - c0);
          const int sl = R.next;
          R.next = (R.next + 1) % (int)R.slot.size();
          scat->acquire(sl);
          hip_check(hipMemcpy2DAsync(R.slot[sl], width, src + (size_t)c0 * op.dev.ld * es, op.dev.ld * es, width, nc,
                                     hipMemcpyDeviceToHost, s),
                    "hipMemcpy2DAsync D2H (staged)");
          hip_check(hipEventRecord(R.ev[sl], s), "hipEventRecord");
          R.ev_pending[sl] = true;
          scat->push(Scatterer::Item{R.ev[sl], sl, dst + (size_t)c0 * op.host_ld * es, (size_t)op.host_ld * es,
                                     (const char*)R.slot[sl], width, (size_t)nc});
        }
      }
    }
    if (need[oi]) hip_check(hipEventRecord(h.events[oi], s), "hipEventRecord");
  };

  // Run the ops of one thread in plan order: sel = -2 the compute queues, 0 / 1 the H2D / D2H
  // queue. GEMMs wait for other queues on the device (hipStreamWaitEvent), copies on the host
  // (inside enqueue). Returns early once any thread has failed.
  auto run_ops = [&](bool worker, int sel) {
    try {
      if (worker) hip_check(hipSetDevice(h.device), "hipSetDevice");
      for (size_t oi = 0; oi < nops; ++oi) {
        const Op& op = P.ops[oi];
        if (!(sel == -2 ? op.queue >= 2 : op.queue == sel)) continue;
        if (op.queue >= 2 && group_of[oi] >= 0) {
          const std::vector<size_t>& g = groups[group_of[oi]];
          if (oi != g[0]) continue;  // enqueued and marked with its group
          hipStream_t s = stream_of(op.queue);
          std::vector<GemmParams<double>> tiles;
          std::vector<GemmParams<float>> tiles_f;
          for (size_t gi : g) {
            for (int d : P.ops[gi].deps) {
              if (d < 0 || P.ops[d].queue == op.queue) continue;
              if (!sync.wait(d)) return;
              hip_check(hipStreamWaitEvent(s, h.events[d], 0), "hipStreamWaitEvent");
            }
            if (single) tiles_f.push_back(tile_params_f(P.ops[gi]));
            else tiles.push_back(tile_params(P.ops[gi]));
          }
          const size_t pi = prof_stream.size();
          if (h.profile) hip_check(hipEventRecord(h.prof_events[2 * pi], s), "hipEventRecord");
          hipError_t ge;
          switch (h.dtype) {
            case 0: ge = launch_gemm_s_group(tiles_f.data(), (int)tiles_f.size(), s); break;
            case 1: ge = launch_gemm_d_group(tiles.data(), (int)tiles.size(), s); break;
            case 2: ge = launch_gemm_c_group(tiles_f.data(), (int)tiles_f.size(), s); break;
            default: ge = launch_gemm_z_group(tiles.data(), (int)tiles.size(), s); break;
          }
          hip_check(ge, "grouped tile GEMM launch");
          if (h.profile) {
            hip_check(hipEventRecord(h.prof_events[2 * pi + 1], s), "hipEventRecord");
            prof_stream.push_back(s);
            double fl = 0, by = 0;
            std::vector<const char*> seen;
            for (size_t gi : g) {
              const Op& t = P.ops[gi];
              fl += (cplx ? 8.0 : 2.0) * t.m * t.n * t.k;
              const char* pa = devptr(t.a);
              const char* pb = devptr(t.b);
              if (std::find(seen.begin(), seen.end(), pa) == seen.end()) {
                seen.push_back(pa);
                by += (double)t.m * t.k * es;
              }
              if (std::find(seen.begin(), seen.end(), pb) == seen.end()) {
                seen.push_back(pb);
                by += (double)t.n * t.k * es;
              }
              by += (double)t.m * t.n * es * (1 + (t.accumulate ? 1 : 0) + (t.use_cin ? 1 : 0));
            }
            prof_flops.push_back(fl);
            prof_grouped.push_back(1);
            prof_bytes.push_back(by);
          }
          for (size_t gi : g) {
            if (need[gi]) hip_check(hipEventRecord(h.events[gi], s), "hipEventRecord");
            sync.mark(gi);
          }
          continue;
        }
        if (op.queue >= 2) {
          // a scheme II GEMM waits for a late C_in copy between its products and reconstruction
          const int late = crt_reuse && op.k > 0 && op.k <= kCrtMaxK ? op.late_dep : -1;
          for (int d : op.deps) {
            if (d < 0 || d == late || P.ops[d].queue == op.queue) continue;
            if (!sync.wait(d)) return;
            hip_check(hipStreamWaitEvent(stream_of(op.queue), h.events[d], 0), "hipStreamWaitEvent");
          }
        }
        enqueue(oi);
        sync.mark(oi);
      }
    } catch (Error& e) {
      sync.fail(e.code, e.what());
    } catch (std::exception& e) {
      sync.fail(2, e.what());
    }
  };
  {
    std::thread th_in([&] { run_ops(true, 0); });
    std::thread th_out([&] { run_ops(true, 1); });
    run_ops(false, -2);
    th_in.join();
    th_out.join();
  }

  // Drain every stream, also after a failure: staging slots and device buffers outlive the call.
  hipError_t serr = hipSuccess;
  for (auto st : h.compute) {
    if (!st) continue;
    const hipError_t e = hipStreamSynchronize(st);
    if (serr == hipSuccess) serr = e;
  }
  for (hipStream_t st : {h.h2d, h.d2h}) {
    const hipError_t e = hipStreamSynchronize(st);
    if (serr == hipSuccess) serr = e;
  }
  if (scat) scat->finish();
  if (sync.failed) throw Error(sync.code, sync.error);
  hip_check(serr, "hipStreamSynchronize");
  if (scat && !scat->error.empty()) throw Error(2, scat->error);
  // kernel timing: [9] summed GEMM kernel seconds, [10] launches timed, [11] their flops,
  // [12] longest launch in seconds; of these the grouped launches: [16] count, [17] seconds,
  // [18] flops, [23] algorithmic HBM bytes
  h.stats[9] = h.stats[10] = h.stats[11] = h.stats[12] = 0;
  h.stats[16] = h.stats[17] = h.stats[18] = h.stats[23] = 0;
  h.stats[13] = int8_ops;
  for (size_t i = 0; i < prof_stream.size(); ++i) {
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, h.prof_events[2 * i], h.prof_events[2 * i + 1]), "hipEventElapsedTime");
    h.stats[9] += ms * 1e-3;
    h.stats[10] += 1;
    h.stats[11] += prof_flops[i];
    h.stats[12] = std::max(h.stats[12], ms * 1e-3);
    if (prof_grouped[i]) {
      h.stats[16] += 1;
      h.stats[17] += ms * 1e-3;
      h.stats[18] += prof_flops[i];
      h.stats[23] += prof_bytes[i];
    }
  }
}

void validate(const GemmCall& c) {
  if (c.m < 0 || c.n < 0 || c.k < 0) throw Error(1, "matrix dimensions must be non-negative");
  const int a_rows = c.opa == 'N' ? c.m : c.k;
  const int b_rows = c.opb == 'N' ? c.k : c.n;
  // The reference builds these errors but does not throw them (tiled_mm.cpp:516-527); an
  // out-of-range leading dimension would make the strided copies read outside the matrix.
  if (c.lda < std::max(1, a_rows)) throw Error(1, "Leading dimensions for matrix A must be >= the number of rows.");
  if (c.ldb < std::max(1, b_rows)) throw Error(1, "Leading dimensions for matrix B must be >= the number of rows.");
  if (c.ldc < std::max(1, c.m)) throw Error(1, "Leading dimensions for matrix C must be >= the number of rows.");
}

}  // namespace

char normalize_op(char t) {
  const char u = (char)std::toupper((unsigned char)t);
  return (u == 'T' || u == 'C') ? u : 'N';
}

void gemm_host(Handle& h, const GemmCall& in) {
  auto t0 = std::chrono::steady_clock::now();
  for (double& s : h.stats) s = 0;
  GemmCall call = in;
  call.opa = normalize_op(in.opa);
  call.opb = normalize_op(in.opb);
  if (h.dtype < 2) {  // for real types conjugate transpose is transpose
    if (call.opa == 'C') call.opa = 'T';
    if (call.opb == 'C') call.opb = 'T';
  }
  validate(call);
  if (call.m == 0 || call.n == 0) return;
  // an asynchronous Ozaki device call may still use the workspaces this call's queues take
  h.ozwork_quiesce();
  const int es = h.elem_bytes;
  const bool use_beta = !scalar_is_zero(h.dtype, call.beta);
  // alpha = 0: C = beta*C without reading A or B (BLAS semantics: Inf and NaN in A or B do not
  // reach C), planned as a product with an empty inner dimension
  const int plan_k = scalar_is_zero(h.dtype, call.alpha) ? 0 : call.k;

  // Reference tile policy (tiled_mm.cpp:556-560): tile sizes become the handle's current ones.
  int tm = optimal_tile_size(call.m, h.cur_tm, h.max_tm);
  int tn = optimal_tile_size(call.n, h.cur_tn, h.max_tn);
  int tk = optimal_tile_size(std::max(call.k, 1), h.cur_tk, h.max_tk);
  h.cur_tm = tm;
  h.cur_tn = tn;
  h.cur_tk = tk;
  if (h.coarsen) {
    // Merge small tiles so one tile GEMM launches ~1024 workgroups.
    const long target = h.dtype >= 2 ? 2048L * 2048 : 4096L * 4096;
    while ((long)tm * tn < target && (tm < call.m || tn < call.n)) {
      if ((tm <= tn && tm < call.m) || tn >= call.n) tm = std::min(call.m, tm * 2);
      else tn = std::min(call.n, tn * 2);
    }
    while (tk < 1024 && tk < call.k) tk = std::min(call.k, tk * 2);
    // Block edges on the tile kernel's workgroup grid: as many blocks per dimension, each a
    // multiple of the workgroup tile but the last. 40000 = 4 x 8064 + 7744 instead of 5 x 8000
    // (whose last workgroup row of 128 is half empty in every block) for the 128 x 128 FP64
    // workgroup tiles, 4 x 8192 + 7232 for the 256-wide INT8 tiles of the Ozaki schemes, the
    // 256 x 256 FP32 workgroup tiles (complex FP32: 128 x 128) and the columns of the 128 x 256
    // FP64 tiles, which blocks of at least 2^27 elements run on. Results do not change: the k order
    // of every element is the kernel's, whatever block it lies in.
    const bool direct = h.dgemm_variant == 1;
    const bool fp64_big = direct && (long)tm * tn >= (1L << 27);
    const int align_m = h.dtype == 1 ? (ozaki_active(h) ? 256 : 128) : h.dtype == 3 ? 64 : !direct ? 0 : h.dtype == 0 ? 256 : 128;
    const int align_n = h.dtype == 1 ? (ozaki_active(h) || fp64_big ? 256 : 128) : h.dtype == 3 ? 64 : !direct ? 0 : h.dtype == 0 ? 256 : 128;
    auto align_edge = [](int d, int t, int align) {
      if (align <= 0 || t >= d) return t;
      const long nb = (d + (long)t - 1) / t;
      const long ta = ((d + nb - 1) / nb + align - 1) / align * align;
      return (nb - 1) * ta < d ? (int)ta : t;
    };
    tm = align_edge(call.m, tm, align_m);
    tn = align_edge(call.n, tn, align_n);
  }

  PlanRequest rq;
  rq.opa = call.opa;
  rq.opb = call.opb;
  rq.m = call.m;
  rq.n = call.n;
  rq.k = plan_k;
  rq.lda = call.lda;
  rq.ldb = call.ldb;
  rq.ldc = call.ldc;
  rq.beta_nonzero = use_beta;
  rq.copy_back = call.copy_back;
  rq.tile_m = tm;
  rq.tile_n = tn;
  rq.tile_k = tk;
  {
    // One tile GEMM of >= 1024 workgroups fills the GPU on its own. Two queues, so that the next
    // tile fills the CUs the last partial wave of the previous one leaves idle, measured worse:
    // DGEMM TN 40000^3 host-to-host 66.8 -> 61.5 TFLOP/s (concurrent tiles slow each other to
    // 19.7 ms from 12.7), ZGEMM CN 24000x24000x32000 2.24 -> 2.18-2.28 s.
    const long wg_edge = h.dtype >= 2 ? 64 : 128;
    const long wgs = ((tm + wg_edge - 1) / wg_edge) * ((tn + wg_edge - 1) / wg_edge);
    const int nq = (int)h.compute.size();
    rq.n_compute_queues = wgs >= 1024 ? std::min(nq, h.big_tile_queues) : nq;
  }
  bool pinned[3];
  const long a_rows = call.opa == 'N' ? call.m : call.k, a_cols = call.opa == 'N' ? call.k : call.m;
  const long b_rows = call.opb == 'N' ? call.k : call.n, b_cols = call.opb == 'N' ? call.n : call.k;
  auto span = [&](long rows, long cols, long ld) -> size_t {
    return rows == 0 || cols == 0 ? 0 : (size_t)((cols - 1) * ld + rows) * es;
  };
  pinned[0] = is_pinned_range(call.a, span(a_rows, a_cols, call.lda));
  pinned[1] = is_pinned_range(call.b, span(b_rows, b_cols, call.ldb));
  pinned[2] = is_pinned_range(call.c, span(call.m, call.n, call.ldc));

  // a first chunk on the tile kernel's k-step (16 for FP32, 8 otherwise), so that neither it nor
  // the last chunk leaves a remainder launch per tile
  const int kstep = h.dtype == 0 ? 16 : 8;
  rq.first_chunk_k =
      h.first_chunk >= 0 ? (int)h.first_chunk : (tk >= 4096 ? std::max(512, tk / 8) / kstep * kstep : 0);
  // Ozaki scheme II with prepared operands reads the caller's C only in the reconstruction, so
  // the last chunk's C copies can follow the next tiles' A / B copies (which the products wait
  // for). Pinned memory only: a staged (pageable) copy occupies the copy thread, which the
  // GEMM would then wait for in the middle of its work.
  if (h.dtype == 1 && h.ozaki_moduli > 0 && h.ozaki_reuse && pinned[0] && pinned[1] && pinned[2])
    rq.defer_cin = h.defer_cin;
  {
    size_t free_b = 0, total_b = 0;
    hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    size_t budget = h.pool_limit_mb > 0 ? (size_t)h.pool_limit_mb << 20 : (size_t)((free_b + h.pool.bytes) * 0.85);
    if (!call.copy_back) {
      const size_t need_c = (size_t)call.m * call.n * es;
      budget = budget > need_c ? budget - need_c : 0;
    }
    rq.pool_budget_elems = budget / es;
  }
  Plan P;
  try {
    P = make_plan(rq);
    if (h.dtype == 1 && h.ozaki_moduli > 0 && h.coarsen && h.first_chunk < 0 && plan_k > tk) {
      // Ozaki scheme II: every k-chunk costs one reconstruction pass over each C tile, so the
      // k-chunks are merged up to the CRT pass length when the work pool holds the larger
      // blocks without extra host-link traffic; a shorter first chunk (1/4) still lets the
      // first products start early and covers the copy of the first long-chunk blocks
      // (40000^3: 2 chunks, 10000 + 30000, instead of 6; pinned, per call at 1/8, 1/6, 1/5, 1/4,
      // 1/3: 846, 831, 818, 800, 804 ms)
      PlanRequest r2 = rq;
      const int kk = std::min(plan_k, kCrtMaxK);
      r2.first_chunk_k = kk >= 4096 ? std::max(512, kk / 4) : 0;
      r2.tile_k = std::max(1, kk - r2.first_chunk_k);
      Plan P2 = make_plan(r2);
      if (P2.h2d_elems <= P.h2d_elems && P2.pool_elems <= rq.pool_budget_elems) P = std::move(P2);
    }
  } catch (std::invalid_argument& e) {
    throw Error(3, std::string("planning failed: ") + e.what());
  }

  h.pool.reserve(P.pool_elems * es);
  if (!call.copy_back) {
    // mm_handle::set_full_sizes (mm_handle.cpp:72-80): resident C of m*n elements.
    const size_t need = (size_t)call.m * call.n;
    h.fullc.reserve(need * es);
    h.fullc_elems = need;
  }

  execute_plan(h, P, call, pinned, use_beta);

  int ngemm = 0;
  for (auto& op : P.ops) ngemm += op.kind == OpKind::GEMM;
  h.stats[0] = (double)P.h2d_elems * es;
  h.stats[1] = (double)P.d2h_elems * es;
  h.stats[2] = ngemm;
  h.stats[3] = P.tile_m;
  h.stats[4] = P.tile_n;
  h.stats[5] = P.tile_k;
  h.stats[6] = P.nTk;
  h.stats[7] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  h.stats[8] = pinned[0] && pinned[1] && pinned[2] ? 1 : 0;
}

void gemm_device(Handle& h, const GemmCall& in, hipStream_t s) {
  GemmCall call = in;
  call.opa = normalize_op(in.opa);
  call.opb = normalize_op(in.opb);
  if (h.dtype < 2) {
    if (call.opa == 'C') call.opa = 'T';
    if (call.opb == 'C') call.opb = 'T';
  }
  validate(call);
  if (call.m == 0 || call.n == 0) return;
  const bool use_beta = !scalar_is_zero(h.dtype, call.beta);
  if (scalar_is_zero(h.dtype, call.alpha)) call.k = 0;  // C = beta*C; A and B are not read
  // In-place C: C_in and C_out are the same matrix, so the beta term reads each element
  // before the same thread overwrites it.
  const int np = launch_mode(h, 0, call.opa, call.opb, call.m, call.n, call.k, call.alpha, call.a, call.lda, call.b,
                             call.ldb, call.beta, use_beta, call.c, call.ldc, false, call.c, call.ldc,
                             s ? s : h.compute[0]);
  h.stats[13] = 2.0 * np * call.m * call.n * call.k;
}

void gemm_device_noalloc(Handle& h, const GemmCall& in, hipStream_t s) {
  GemmCall call = in;
  call.opa = normalize_op(in.opa);
  call.opb = normalize_op(in.opb);
  if (h.dtype < 2) {
    if (call.opa == 'C') call.opa = 'T';
    if (call.opb == 'C') call.opb = 'T';
  }
  validate(call);
  if (call.m == 0 || call.n == 0) return;
  const bool use_beta = !scalar_is_zero(h.dtype, call.beta);
  if (scalar_is_zero(h.dtype, call.alpha)) call.k = 0;
  void* work = nullptr;
  if (ozaki_active(h) && call.k > 0) {
    const size_t need = ozaki_ws_bytes(h, call.m, call.n, call.k);
    if (need > kOzakiScratchBytes) {
      std::ostringstream os;
      os << "tmm_device_gemm_noalloc: the emulated product of " << call.m << " x " << call.n << " x " << call.k
         << " needs " << need << " bytes of workspace; the fixed scratch holds " << kOzakiScratchBytes;
      throw Error(1, os.str());
    }
    work = ozaki_scratch(h.device);
    if (!work) throw Error(2, "tmm_device_gemm_noalloc: the device scratch was not set up on this device");
  }
  const int np = launch_mode(h, 0, call.opa, call.opb, call.m, call.n, call.k, call.alpha, call.a, call.lda, call.b,
                             call.ldb, call.beta, use_beta, call.c, call.ldc, false, call.c, call.ldc,
                             s ? s : h.compute[0], work);
  h.stats[13] = 2.0 * np * call.m * call.n * call.k;
}

}  // namespace tmm

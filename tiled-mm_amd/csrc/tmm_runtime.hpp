// Host runtime of the out-of-core GEMM: handle state, thread pool, pinned staging and the
// GPU executor of a Plan (tmm_plan.hpp).
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "tmm_plan.hpp"

namespace tmm {

// Error carrying a tmm_status code.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what);

// Fixed pool of worker threads; parallel_for may be called from several threads at once.
class ThreadPool {
 public:
  explicit ThreadPool(int n);
  ~ThreadPool();
  int size() const { return (int)workers_.size(); }
  // Run fn(0..n-1) on the pool and the calling thread; returns when all are done.
  void parallel_for(int n, const std::function<void(int)>& fn);

 private:
  struct Job {
    const std::function<void(int)>* fn;
    int n;
    std::atomic<int> next{0};
    std::atomic<int> done{0};
  };
  void run_job(Job& j);
  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::deque<std::shared_ptr<Job>> jobs_;
  bool stop_ = false;
};

// Grow-only device allocation. Move-only: a copy would free the memory twice (vectors of
// buffers move their elements when they grow).
struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : ptr(o.ptr), bytes(o.bytes) {
    o.ptr = nullptr;
    o.bytes = 0;
  }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
      release();
      ptr = o.ptr;
      bytes = o.bytes;
      o.ptr = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  void reserve(size_t b);  // contents are not preserved on growth
  void release();
  ~DeviceBuffer() { release(); }
};

// Ring of pinned host staging slots for pageable host matrices.
struct StagingRing {
  std::vector<void*> slot;
  std::vector<hipEvent_t> ev;     // completion of the DMA that last used the slot
  std::vector<bool> ev_pending;
  size_t slot_bytes = 0;
  int next = 0;
  void ensure(int n, size_t bytes);
  void release();
  ~StagingRing() { release(); }
};

// Size summary of a batched call, and the pipeline configuration chosen for it.
struct BatchShape {
  int count = 0;
  double flops = 0, h2d_bytes = 0, d2h_bytes = 0;
  bool pinned = false;
  std::vector<int> m, n, k;  // per problem
};
struct BatchConfig {
  int groups = 4;          // pipeline groups (H2D / kernel / D2H overlap between groups)
  int workgroups = 512;    // persistent grid size
  bool use_graph = true;   // capture the batch into one hipGraph (pinned host memory only)
  int body = 1;            // persistent tile body: 0 register-staged, 1 direct-to-LDS ring
  int queues = 1;          // compute queues the groups' kernels alternate over
  int emulated = 0;        // problems run as Ozaki scheme II products (handle with moduli set)
  std::vector<char> emulate;  // per problem: 1 = Ozaki scheme II, 0 = the persistent kernel
  double predicted_s = 0;  // model time of the chosen configuration
};
// Rates measured on the device by the autotuner (0 = not measured yet).
struct TuneRates {
  double h2d_Bps = 0, d2h_Bps = 0, kernel_flops = 0, group_overhead_s = 0;
  double kernel_flops_dlds = 0;  // the direct-to-LDS persistent body
  // Ozaki scheme II at emu_moduli moduli: FP64-equivalent flop rate of a 2048^3 call and the
  // fixed cost of one emulated product (a 128^3 call), measured when a handle first asks for it
  int emu_moduli = 0;
  double emu_flops = 0, emu_overhead_s = 0;
};

struct Handle {
  int dtype = 1;
  int elem_bytes = 8;
  int device = 0;
  int n_streams = 2;
  int max_tm = 5000, max_tn = 5000, max_tk = 5000;
  int cur_tm = 0, cur_tn = 0, cur_tk = 0;  // current tile sizes (set_tile_sizes)
  // options
  bool coarsen = true;
  long first_chunk = -1;
  int host_threads = 0;
  long pool_limit_mb = 0;
  bool profile = false;  // time every tile GEMM launch with events on its stream
  int ozaki_slices = 0;  // > 0: FP64 tile GEMMs run as Ozaki split-precision on INT8 MFMA
  int ozaki_moduli = 0;  // > 0: FP64 tile GEMMs run as Ozaki scheme II (CRT), takes precedence
  bool ozaki_reuse = true;  // scheme II host-to-host: prepare each shared A / B block once
  int batch_graph = -1;  // batched GEMM, pinned memory: -1 autotuner's choice, 1 hipGraph, 0 queues
  int batch_body = -1;   // batched GEMM tile body: -1 autotuner's choice, 0 register-staged, 1 direct-to-LDS
  int batch_queues = 0;  // batched GEMM compute queues: 0 autotuner's choice, 1 or 2
  int batch_emulation = -1;  // batched GEMM with Ozaki moduli set: -1 per problem by the model, 0 none, 1 all (k > 0)
  int defer_cin = 1;  // scheme II host-to-host, pinned: C copies of the last chunk go this many tiles later
  int tile_groups = 2;  // native FP64 / complex FP64 host-to-host: grouped launches of steady k-chunks;
                        // >= 2: also runs of at most this many tiles in the other chunks after the first
  int big_tile_queues = 1;  // compute queues for tiles of >= 1024 workgroups
  int dgemm_variant = 1;  // FP64 / complex FP64 / FP32 tile kernel variant (GemmParams::variant): direct-to-LDS

  std::vector<hipStream_t> compute;  // n_streams entries; only [0] is created eagerly
  hipStream_t h2d = nullptr, d2h = nullptr;
  hipStream_t h2d_aux = nullptr;  // second H2D queue of the batched path, created on first use
  std::vector<hipEvent_t> events;  // pool, grows per plan
  DeviceBuffer pool, fullc, tilebuf[3];
  std::vector<DeviceBuffer> ozwork;   // Ozaki slice workspace, one per compute queue
  // last use of ozwork[i] by an asynchronous device call, recorded on the caller's stream: the next
  // user waits for it (another stream) and a reallocation synchronises with it first
  std::vector<hipEvent_t> ozwork_ev;
  std::vector<char> ozwork_pending;
  std::vector<DeviceBuffer> ozcache;  // scheme II prepared-operand slots, one buffer per compute queue
  DeviceBuffer batch_meta;           // batched path: problem descriptors, work items, counters
  TuneRates rates;
  BatchConfig last_batch;
  size_t fullc_elems = 0;
  size_t tilebuf_elems[3] = {0, 0, 0};
  StagingRing stage_in, stage_out;
  ThreadPool* threads = nullptr;
  std::string last_error;
  double stats[24] = {0};
  std::vector<hipEvent_t> prof_events;  // timing events, pairs per GEMM launch

  Handle(int dtype, int streams, int tm, int tn, int tk);
  ~Handle();
  void set_num_streams(int s);
  hipStream_t compute_stream(int i);  // compute stream i, created on first use
  hipStream_t h2d_aux_stream();       // h2d_aux, created on first use
  void sync_compute();
  ThreadPool& pool_threads();
  hipEvent_t event(size_t i);
  // Ozaki workspace i of at least `bytes`, ordered after its previous asynchronous use: on
  // stream s (device calls), or, with s = nullptr, after that use has finished (host calls).
  void* ozwork_acquire(size_t i, size_t bytes, hipStream_t s);
  // Record that stream s has enqueued work reading or writing Ozaki workspace i.
  void ozwork_release(size_t i, hipStream_t s);
  // Wait on the host until no asynchronous device call uses an Ozaki workspace any more.
  void ozwork_quiesce();
};

struct GemmCall {
  char opa, opb;
  int m, n, k;
  const void* alpha;  // one scalar of the handle's type
  const void* a;
  long lda;
  const void* b;
  long ldb;
  const void* beta;
  void* c;
  long ldc;
  bool copy_back;
};

// Host-to-host GEMM on a handle (the gpu::gemm path).
void gemm_host(Handle& h, const GemmCall& call);
// Device-resident tile GEMM on a stream.
void gemm_device(Handle& h, const GemmCall& call, hipStream_t s);
// The same without any device allocation: an Ozaki emulation works in the device's fixed
// scratch (kOzakiScratchBytes), not in the handle's workspace.
void gemm_device_noalloc(Handle& h, const GemmCall& call, hipStream_t s);
// Batched host-to-host FP64 GEMM (tmm_batch.cpp).
void gemm_batched_host(Handle& h, char opa, char opb, int count, const int* m, const int* n, const int* k,
                       const double* alpha, const double* const* a, const int* lda, const double* const* b,
                       const int* ldb, const double* beta, double* const* c, const int* ldc);
// Autotuner (tmm_autotune.cpp).
BatchConfig choose_batch_config(Handle& h, const BatchShape& shape);
void calibrate_rates(Handle& h);

// True when the whole byte range [p, p+bytes) lies in one pinned (page-locked) allocation.
bool is_pinned_range(const void* p, size_t bytes);

}  // namespace tmm

// Out-of-core GEMM planner: turns one gemm call into an ordered list of copy / compute
// operations with explicit dependencies. Pure host C++ (no HIP), so that the same plan can be
// executed by the GPU executor and, in tests, by a CPU executor that replays the queues in a
// random order allowed by the dependencies.
//
// Relation to the reference (Tiled-MM src/Tiled-MM/tiled_mm.cpp):
//   * tile sizes come from the reference policy (mm_handle.cpp:89-133) unless coarsening is on;
//   * the block geometry (tile counts, short edge tiles, transposed tile coordinates) follows
//     tiled_matrix.cpp:7-90 and tiled_mm.cpp:32-42,126-165, with 64-bit offsets;
//   * the schedule is NOT the reference's round_robin (tiled_mm.cpp:270-454), which re-copies
//     every A and B tile once per C tile: here each A and B block crosses PCIe exactly once
//     (k-chunk outer, C tiles inner in "shell" order), the caller's C is added on the last
//     k-chunk (beta*C_in term of the kernel epilogue), and each C tile is copied back as soon
//     as its last k-chunk is done.
#pragma once
#include <cstddef>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace tmm {

enum class Mat : int { A = 0, B = 1, C = 2 };
enum class OpKind : int { H2D = 0, GEMM = 1, D2H = 2 };
// Device regions an op can address. POOL is the handle's work pool; FULLC the resident C.
enum class Region : int { POOL = 0, FULLC = 1 };

struct DevRef {
  Region region = Region::POOL;
  size_t off = 0;  // in elements
  long ld = 0;     // in elements
};

struct Op {
  OpKind kind;
  int queue;  // 0 = h2d, 1 = d2h, 2.. = compute queues
  std::vector<int> deps;  // indices of ops that must complete first (earlier ops, but late_dep)
  // copies
  Mat mat = Mat::A;
  size_t host_off = 0;  // element offset in the host matrix
  long host_ld = 0;
  int rows = 0, cols = 0;  // extent of the copied block (elements)
  DevRef dev;
  // gemm
  int m = 0, n = 0, k = 0;
  DevRef a, b, c, cin;
  bool accumulate = false;  // C += ...
  bool use_cin = false;     // + beta*C_in
  // the copy of C_in when the plan issues it after this GEMM (PlanRequest::defer_cin); it is
  // also in deps, and an executor that starts the GEMM's products before it must wait for it
  // before the epilogue that reads C_in
  int late_dep = -1;
  // bookkeeping for tests / tracing
  int ti = -1, tj = -1, tc = -1;
};

struct PlanRequest {
  char opa = 'N', opb = 'N';
  int m = 0, n = 0, k = 0;
  long lda = 0, ldb = 0, ldc = 0;
  bool beta_nonzero = false;
  bool copy_back = true;
  int tile_m = 0, tile_n = 0, tile_k = 0;  // block sizes (already chosen)
  int n_compute_queues = 1;
  size_t pool_budget_elems = (size_t)-1;  // device work pool limit in elements
  int first_chunk_k = 0;  // 0: first k-chunk has the regular size
  // > 0: in the last k-chunk, the copy of a tile's caller C (read only by that chunk's GEMM
  // epilogue) is issued after the A / B copies of the tile defer_cin places later in the
  // chunk, so that the blocks the next products wait for go over the host link first; the
  // GEMM then lists it as late_dep
  int defer_cin = 0;
};

struct Plan {
  int tile_m = 0, tile_n = 0, tile_k = 0;
  int nTm = 0, nTn = 0, nTk = 0;
  std::vector<int> kbounds;  // nTk+1 k-chunk boundaries
  size_t pool_elems = 0;      // work pool size required
  int n_queues = 0;
  std::vector<Op> ops;
  // host-link bytes moved, in elements
  size_t h2d_elems = 0, d2h_elems = 0;
};

// Reference tile policy for one dimension (mm_handle.cpp:89-110): if dim fits, use it;
// otherwise the largest divisor of dim not above max, provided it is within max/2 of max,
// else max itself.
int optimal_tile_size(int dim, int curr_tile_size, int max_tile_size);

// Tiles of the C grid in "shell" order: shell s holds the tiles with max(i, j) == s; column s
// above the diagonal first, then row s. Each new shell needs one new A block-row and one new
// B block-column and reuses all earlier ones.
std::vector<std::pair<int, int>> shell_order(int nTm, int nTn);

// Build the plan. Throws std::invalid_argument on inconsistent input.
Plan make_plan(const PlanRequest& rq);

// Index of the H2D op that filled the A (mat == Mat::A) or B operand block a GEMM op reads, or
// -1 when none of its dependencies did (the operand then has no identity to share).
int operand_producer(const Plan& P, size_t gemm, Mat mat);

// Static cache plan for GEMM operands that several GEMMs of one compute queue read from the same
// H2D-filled block (emulated FP64: their exponents and residues are then computed once). Per
// queue, the GEMMs with k <= max_k in issue order; an operand block read by at least two of them
// gets a cache slot, filled ("prepared") before its first use, replaced by farthest next use
// (Belady) when all `capacity` slots are taken and freed after its last use. Both operands of a
// GEMM hold their slots together, so capacity < 2 disables the cache.
struct OperandReuse {
  std::vector<int> slot_a, slot_b;   // per op: slot of the A / B operand, -1 = not cached
  std::vector<char> prep_a, prep_b;  // per op: fill the slot before this GEMM
  std::vector<int> slots;            // per compute queue: slots in use (highest index + 1)
};
OperandReuse plan_operand_reuse(const Plan& P, int capacity, int max_k);

// Groups of tile GEMMs to launch together: per compute queue, runs of consecutive GEMMs (in
// issue order) of one k-chunk, with equal k, each `groupable`. Chunks other than the first and
// the last whose preceding chunk is at least as long take runs of at most max_group: a group
// waits for all blocks it reads, which the preceding chunk's compute then covers. The other
// chunks after the first take runs of at most pair_max (0 or 1: one launch per tile there):
// after the short first chunk the copies are only a little ahead, and in the last chunk each C
// tile is copied back once its run is done; the queue's last GEMM always runs alone. Runs of
// one are not groups. The executor starts a group once every dependency of its members is
// enqueued, so a run admits an op only if each of its dependencies with an index above the
// run's first op depends only on ops before that first op: the copy threads that enqueue those
// dependencies then never wait for the group itself.
std::vector<std::vector<size_t>> plan_tile_groups(const Plan& P, int max_group,
                                                  const std::function<bool(const Op&)>& groupable, int pair_max = 0);

// Rows / cols of op(X) storage for a transpose flag ('N' or not).
inline int stored_rows(char op, int r, int c) { return op == 'N' ? r : c; }

}  // namespace tmm

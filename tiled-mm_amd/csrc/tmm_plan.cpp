// Out-of-core GEMM planner (see tmm_plan.hpp).
#include "tmm_plan.hpp"

#include <algorithm>
#include <map>
#include <stdexcept>

namespace tmm {

int optimal_tile_size(int dim, int curr_tile_size, int max_tile_size) {
  // Tiled-MM src/Tiled-MM/mm_handle.cpp:89-110.
  if (curr_tile_size < max_tile_size && dim <= max_tile_size) return dim;
  const int limit = std::min(max_tile_size, dim);
  int tile = 1;
  for (int i = 1; i <= limit; ++i)
    if (dim % i == 0) tile = i;
  if (std::abs(max_tile_size - tile) <= max_tile_size / 2) return tile;
  return limit;
}

std::vector<std::pair<int, int>> shell_order(int nTm, int nTn) {
  std::vector<std::pair<int, int>> out;
  out.reserve((size_t)nTm * nTn);
  const int S = std::max(nTm, nTn);
  for (int s = 0; s < S; ++s) {
    if (s < nTn)
      for (int i = 0; i < std::min(s, nTm); ++i) out.emplace_back(i, s);
    if (s < nTm)
      for (int j = 0; j <= std::min(s, nTn - 1); ++j) out.emplace_back(s, j);
  }
  return out;
}

namespace {

struct SlotState {
  std::vector<int> users;  // ops that touched the current occupant
};

int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

}  // namespace

Plan make_plan(const PlanRequest& rq) {
  if (rq.m < 0 || rq.n < 0 || rq.k < 0) throw std::invalid_argument("negative matrix dimension");
  if (rq.tile_m <= 0 || rq.tile_n <= 0 || rq.tile_k <= 0) throw std::invalid_argument("tile sizes must be positive");
  if (rq.n_compute_queues <= 0) throw std::invalid_argument("need at least one compute queue");

  Plan P;
  P.n_queues = 2 + rq.n_compute_queues;
  if (rq.m == 0 || rq.n == 0) return P;

  const int tm = std::min(rq.tile_m, rq.m), tn = std::min(rq.tile_n, rq.n);
  const int tk = std::max(1, std::min(rq.tile_k, std::max(rq.k, 1)));
  P.tile_m = tm;
  P.tile_n = tn;
  P.tile_k = tk;
  P.nTm = ceil_div(rq.m, tm);
  P.nTn = ceil_div(rq.n, tn);

  // k-chunk boundaries (k == 0 gives one empty chunk: C = beta*C).
  P.kbounds.push_back(0);
  if (rq.k == 0) {
    P.kbounds.push_back(0);
  } else {
    int pos = 0;
    if (rq.first_chunk_k > 0 && rq.first_chunk_k < tk && rq.k > tk) {
      pos = rq.first_chunk_k;
      P.kbounds.push_back(pos);
    }
    while (pos < rq.k) {
      pos = (int)std::min<long>((long)pos + tk, rq.k);
      P.kbounds.push_back(pos);
    }
  }
  P.nTk = (int)P.kbounds.size() - 1;
  int tkmax = 0;
  for (int c = 0; c < P.nTk; ++c) tkmax = std::max(tkmax, P.kbounds[c + 1] - P.kbounds[c]);
  tkmax = std::max(tkmax, 1);

  const size_t SA = (size_t)tm * tkmax, SB = (size_t)tkmax * tn, SC = (size_t)tm * tn;
  const bool cin_needed = rq.beta_nonzero;

  // Super-block (group) of gm x gn C tiles that fits the pool budget.
  auto pool_need = [&](int gm, int gn) {
    size_t need = 2 * (size_t)gm * SA + 2 * (size_t)gn * SB;
    if (rq.copy_back) need += (size_t)gm * gn * SC;
    if (cin_needed) need += (size_t)gm * gn * SC;
    return need;
  };
  int gm = P.nTm, gn = P.nTn;
  while (pool_need(gm, gn) > rq.pool_budget_elems && (gm > 1 || gn > 1)) {
    if (gm >= gn && gm > 1) gm = (gm + 1) / 2;
    else gn = (gn + 1) / 2;
  }
  if (pool_need(gm, gn) > rq.pool_budget_elems)
    throw std::invalid_argument("device memory budget too small for one tile");

  const size_t offA = 0;
  const size_t offB = offA + 2 * (size_t)gm * SA;
  const size_t offC = offB + 2 * (size_t)gn * SB;
  const size_t offCin = offC + (rq.copy_back ? (size_t)gm * gn * SC : 0);
  P.pool_elems = offCin + (cin_needed ? (size_t)gm * gn * SC : 0);

  std::vector<SlotState> slotA(2 * gm), slotB(2 * gn), slotC(gm * gn), slotCin(gm * gn);
  std::vector<Op>& ops = P.ops;

  auto emit = [&](Op op) {
    ops.push_back(std::move(op));
    return (int)ops.size() - 1;
  };
  auto add_deps = [](Op& op, const std::vector<int>& d) { op.deps.insert(op.deps.end(), d.begin(), d.end()); };

  const int lastc = P.nTk - 1;
  const char opa = rq.opa, opb = rq.opb;

  for (int gi0 = 0; gi0 < P.nTm; gi0 += gm) {
    for (int gj0 = 0; gj0 < P.nTn; gj0 += gn) {
      const int GM = std::min(gm, P.nTm - gi0), GN = std::min(gn, P.nTn - gj0);
      auto order = shell_order(GM, GN);
      const int G = (int)order.size();
      // A / B block loads of this group, per chunk (-1 = not yet loaded)
      std::vector<int> a_op(GM), b_op(GN);
      std::vector<int> last_gemm(G, -1);
      std::vector<int> cin_op(G, -1);
      // deferred C_in copies of the last chunk: (tile position, its op, the GEMM that reads it)
      struct Deferred {
        int p;
        Op op;
        int gemm;
      };
      std::vector<Deferred> deferred;
      auto flush_deferred = [&](int upto) {  // issue the deferred copies of tiles <= upto
        size_t w = 0;
        for (size_t i = 0; i < deferred.size(); ++i) {
          Deferred& d = deferred[i];
          if (d.p > upto) {
            if (w != i) deferred[w] = std::move(d);
            ++w;
            continue;
          }
          const int li = order[d.p].first, lj = order[d.p].second;
          const int slot_c = li * GN + lj;
          const int id = emit(std::move(d.op));
          cin_op[d.p] = id;
          ops[d.gemm].deps.push_back(id);
          ops[d.gemm].late_dep = id;
          slotCin[slot_c].users.push_back(id);
        }
        deferred.resize(w);
      };

      for (int c = 0; c < P.nTk; ++c) {
        const int k0 = P.kbounds[c], kc = P.kbounds[c + 1] - k0;
        std::fill(a_op.begin(), a_op.end(), -1);
        std::fill(b_op.begin(), b_op.end(), -1);
        for (int p = 0; p < G; ++p) {
          const int li = order[p].first, lj = order[p].second;
          const int i = gi0 + li, j = gj0 + lj;
          const int r0 = i * tm, rows = std::min(tm, rq.m - r0);
          const int c0 = j * tn, cols = std::min(tn, rq.n - c0);
          const int queue = 2 + (p % rq.n_compute_queues);

          // A(i, c)
          if (a_op[li] < 0) {
            const int s = (c % 2) * gm + li;
            Op op{OpKind::H2D, 0};
            op.mat = Mat::A;
            if (opa == 'N') {
              op.host_off = (size_t)r0 + (size_t)k0 * rq.lda;
              op.rows = rows;
              op.cols = kc;
            } else {
              op.host_off = (size_t)k0 + (size_t)r0 * rq.lda;
              op.rows = kc;
              op.cols = rows;
            }
            op.host_ld = rq.lda;
            op.dev = DevRef{Region::POOL, offA + (size_t)s * SA, (long)op.rows};
            op.ti = i; op.tc = c;
            add_deps(op, slotA[s].users);
            slotA[s].users.clear();
            a_op[li] = emit(op);
            slotA[s].users.push_back(a_op[li]);
            P.h2d_elems += (size_t)op.rows * op.cols;
          }
          // B(c, j)
          if (b_op[lj] < 0) {
            const int s = (c % 2) * gn + lj;
            Op op{OpKind::H2D, 0};
            op.mat = Mat::B;
            if (opb == 'N') {
              op.host_off = (size_t)k0 + (size_t)c0 * rq.ldb;
              op.rows = kc;
              op.cols = cols;
            } else {
              op.host_off = (size_t)c0 + (size_t)k0 * rq.ldb;
              op.rows = cols;
              op.cols = kc;
            }
            op.host_ld = rq.ldb;
            op.dev = DevRef{Region::POOL, offB + (size_t)s * SB, (long)op.rows};
            op.tj = j; op.tc = c;
            add_deps(op, slotB[s].users);
            slotB[s].users.clear();
            b_op[lj] = emit(op);
            slotB[s].users.push_back(b_op[lj]);
            P.h2d_elems += (size_t)op.rows * op.cols;
          }

          // the deferred C_in copies whose tile lies defer_cin places back go after this
          // tile's A / B copies
          if (!deferred.empty()) flush_deferred(p - rq.defer_cin);

          // caller's C for this tile: loaded during chunk c* (1 + spread), never in chunk 0
          // unless there is only one chunk.
          const int slot_c = li * GN + lj;
          const bool defer = rq.defer_cin > 0 && P.nTk > 1 && c == lastc;
          Op deferred_cin{OpKind::H2D, 0};
          if (cin_needed && cin_op[p] < 0) {
            const int cstar = P.nTk == 1 ? 0 : 1 + (int)(((long)p * (P.nTk - 1)) / G);
            if (c == cstar) {
              Op op{OpKind::H2D, 0};
              op.mat = Mat::C;
              op.host_off = (size_t)r0 + (size_t)c0 * rq.ldc;
              op.host_ld = rq.ldc;
              op.rows = rows;
              op.cols = cols;
              op.dev = DevRef{Region::POOL, offCin + (size_t)slot_c * SC, (long)rows};
              op.ti = i; op.tj = j;
              add_deps(op, slotCin[slot_c].users);  // readers of the slot's previous tile
              slotCin[slot_c].users.clear();
              P.h2d_elems += (size_t)rows * cols;
              if (defer) {
                deferred_cin = std::move(op);  // issued by flush_deferred, after this tile's GEMM
              } else {
                cin_op[p] = emit(op);
                slotCin[slot_c].users.push_back(cin_op[p]);
              }
            }
          }

          // GEMM(i, j, c)
          Op g{OpKind::GEMM, queue};
          g.m = rows; g.n = cols; g.k = kc;
          const int sa = (c % 2) * gm + li, sb = (c % 2) * gn + lj;
          g.a = DevRef{Region::POOL, offA + (size_t)sa * SA, opa == 'N' ? (long)rows : (long)std::max(kc, 1)};
          g.b = DevRef{Region::POOL, offB + (size_t)sb * SB, opb == 'N' ? (long)std::max(kc, 1) : (long)cols};
          if (rq.copy_back) g.c = DevRef{Region::POOL, offC + (size_t)slot_c * SC, (long)rows};
          else g.c = DevRef{Region::FULLC, (size_t)r0 + (size_t)c0 * rq.m, (long)rq.m};
          g.accumulate = c > 0;
          g.use_cin = cin_needed && c == lastc;
          if (g.use_cin) g.cin = DevRef{Region::POOL, offCin + (size_t)slot_c * SC, (long)rows};
          g.ti = i; g.tj = j; g.tc = c;
          g.deps.push_back(a_op[li]);
          g.deps.push_back(b_op[lj]);
          if (c > 0) g.deps.push_back(last_gemm[p]);
          const bool cin_later = g.use_cin && cin_op[p] < 0;  // deferred: added on issue
          if (g.use_cin && !cin_later) g.deps.push_back(cin_op[p]);
          if (c == 0 && rq.copy_back) {
            // acc slot reuse across groups: previous occupant must be copied out
            add_deps(g, slotC[slot_c].users);
            slotC[slot_c].users.clear();
          }
          const int gid = emit(g);
          last_gemm[p] = gid;
          if (cin_later) deferred.push_back(Deferred{p, std::move(deferred_cin), gid});
          slotA[sa].users.push_back(gid);
          slotB[sb].users.push_back(gid);
          if (g.use_cin) slotCin[slot_c].users.push_back(gid);

          if (c == lastc && rq.copy_back) {
            Op d{OpKind::D2H, 1};
            d.mat = Mat::C;
            d.host_off = (size_t)r0 + (size_t)c0 * rq.ldc;
            d.host_ld = rq.ldc;
            d.rows = rows;
            d.cols = cols;
            d.dev = g.c;
            d.ti = i; d.tj = j;
            d.deps.push_back(gid);
            const int did = emit(d);
            slotC[slot_c].users.push_back(did);
            P.d2h_elems += (size_t)rows * cols;
          }
        }
        if (!deferred.empty()) flush_deferred(G);
      }
    }
  }
  return P;
}

int operand_producer(const Plan& P, size_t gemm, Mat mat) {
  const Op& g = P.ops[gemm];
  const DevRef& want = mat == Mat::A ? g.a : g.b;
  for (int d : g.deps) {
    const Op& p = P.ops[d];
    if (p.kind == OpKind::H2D && p.mat == mat && p.dev.region == want.region && p.dev.off == want.off &&
        p.dev.ld == want.ld)
      return d;
  }
  return -1;
}

OperandReuse plan_operand_reuse(const Plan& P, int capacity, int max_k) {
  const size_t nops = P.ops.size();
  OperandReuse R;
  R.slot_a.assign(nops, -1);
  R.slot_b.assign(nops, -1);
  R.prep_a.assign(nops, 0);
  R.prep_b.assign(nops, 0);
  R.slots.assign(std::max(P.n_queues, 0), 0);
  if (capacity < 2) return R;
  for (int q = 0; q < P.n_queues; ++q) {
    // the queue's GEMMs in issue order and the operand keys they read (2 * producer + mat)
    std::vector<size_t> seq;
    std::vector<long> key_a, key_b;
    for (size_t i = 0; i < nops; ++i) {
      const Op& op = P.ops[i];
      if (op.kind != OpKind::GEMM || op.queue - 2 != q || op.k > max_k) continue;
      seq.push_back(i);
      const int pa = operand_producer(P, i, Mat::A), pb = operand_producer(P, i, Mat::B);
      key_a.push_back(pa < 0 ? -1 : 2L * pa);
      key_b.push_back(pb < 0 ? -1 : 2L * pb + 1);
    }
    std::map<long, std::vector<size_t>> uses;  // key -> positions in seq
    for (size_t p = 0; p < seq.size(); ++p) {
      if (key_a[p] >= 0) uses[key_a[p]].push_back(p);
      if (key_b[p] >= 0) uses[key_b[p]].push_back(p);
    }
    auto next_use = [&](long key, size_t p) -> size_t {
      const std::vector<size_t>& u = uses[key];
      auto it = std::upper_bound(u.begin(), u.end(), p);
      return it == u.end() ? (size_t)-1 : *it;
    };
    std::vector<long> held(capacity, -1);  // slot -> key
    std::map<long, int> where;             // key -> slot
    for (size_t p = 0; p < seq.size(); ++p) {
      const size_t i = seq[p];
      const long keys[2] = {key_a[p], key_b[p]};
      for (int o = 0; o < 2; ++o) {
        const long key = keys[o];
        if (key < 0 || uses[key].size() < 2) continue;
        int slot;
        char prep = 0;
        auto it = where.find(key);
        if (it != where.end()) {
          slot = it->second;
        } else {
          slot = -1;
          for (int s = 0; s < capacity && slot < 0; ++s)
            if (held[s] < 0) slot = s;
          if (slot < 0) {  // evict the held key used farthest ahead, never the other operand's
            size_t far = 0;
            for (int s = 0; s < capacity; ++s) {
              if (held[s] == keys[1 - o]) continue;
              const size_t nu = next_use(held[s], p);
              if (slot < 0 || nu > far) {
                slot = s;
                far = nu;
              }
            }
            where.erase(held[slot]);
          }
          held[slot] = key;
          where[key] = slot;
          prep = 1;
        }
        (o == 0 ? R.slot_a : R.slot_b)[i] = slot;
        (o == 0 ? R.prep_a : R.prep_b)[i] = prep;
        R.slots[q] = std::max(R.slots[q], slot + 1);
      }
      for (int o = 0; o < 2; ++o) {  // free the slots of keys with no later use
        const long key = keys[o];
        auto it = key < 0 ? where.end() : where.find(key);
        if (it != where.end() && next_use(key, p) == (size_t)-1) {
          held[it->second] = -1;
          where.erase(it);
        }
      }
    }
  }
  return R;
}

std::vector<std::vector<size_t>> plan_tile_groups(const Plan& P, int max_group,
                                                  const std::function<bool(const Op&)>& groupable, int pair_max) {
  std::vector<std::vector<size_t>> groups;
  for (int q = 0; q < P.n_queues; ++q) {
    std::vector<size_t> cur;
    auto flush = [&] {
      if (cur.size() >= 2) groups.push_back(cur);
      cur.clear();
    };
    size_t last_gemm = P.ops.size();
    for (size_t oi = 0; oi < P.ops.size(); ++oi)
      if (P.ops[oi].kind == OpKind::GEMM && P.ops[oi].queue - 2 == q) last_gemm = oi;
    for (size_t oi = 0; oi < P.ops.size(); ++oi) {
      const Op& op = P.ops[oi];
      if (op.kind != OpKind::GEMM || op.queue - 2 != q) continue;
      // a chunk is grouped when the chunk before it is at least as long: that chunk's compute
      // then covers this chunk's block transfers (a group waits for all blocks it reads); the
      // other chunks after the first take runs of at most pair_max, and the queue's last GEMM
      // stays alone so that the last C tile is copied back as soon as it is done
      const int c = op.tc;
      const bool steady = c >= 1 && c + 1 < P.nTk &&
                          P.kbounds[c] - P.kbounds[c - 1] >= P.kbounds[c + 1] - P.kbounds[c];
      const int limit = steady ? max_group : (c >= 1 && oi != last_gemm ? pair_max : 0);
      if (limit < 2 || !groupable(op)) {
        flush();
        continue;
      }
      bool fits = !cur.empty() && P.ops[cur[0]].tc == op.tc && P.ops[cur[0]].k == op.k && (int)cur.size() < limit;
      for (int d : op.deps)
        if (fits && d > (int)cur[0])
          for (int e : P.ops[d].deps) fits = fits && e < (int)cur[0];
      if (!fits) flush();
      cur.push_back(oi);
    }
    flush();
  }
  return groups;
}

}  // namespace tmm

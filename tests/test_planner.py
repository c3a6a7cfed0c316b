"""Host-side scheduling logic, without a GPU.

The planner (tiled-mm_amd/csrc/tmm_plan.cpp) turns one gemm call into copy / GEMM operations
on an H2D queue, a D2H queue and n compute queues with explicit dependencies. These tests run
the real planner through libtmm_host.so, and a CPU executor that replays the queues in random
orders allowed by the dependencies (device memory poisoned with NaN); with integer data the
result must equal the oracle exactly for every seed. They also check the traffic claim: every
A and B block crosses the host link once (the reference re-copies them per C tile,
Tiled-MM src/Tiled-MM/tiled_mm.cpp:292-330).
"""
import ctypes
import os

import numpy as np
import pytest

import oracle
import tiledmm
from helpers import int_matrix, group_members, grouped_algorithmic_bytes

_host = None


def host():
    global _host
    if _host is None:
        L = ctypes.CDLL(tiledmm.HOST_PATH)
        i, c, vp, l, u = ctypes.c_int, ctypes.c_char, ctypes.c_void_p, ctypes.c_long, ctypes.c_uint
        L.tmmh_plan_info.argtypes = [c, c, i, i, i, i, i, i, i, i, i, i, i, i, l, i, ctypes.POINTER(l)]
        L.tmmh_simulate_gemm.argtypes = [i, c, c, i, i, i, vp, vp, i, vp, i, vp, vp, i, i, vp, i, i, i, i, l, i, u]
        L.tmmh_shell_order.argtypes = [i, i, ctypes.POINTER(i)]
        L.tmmh_set_defer_cin.argtypes = [i]
        L.tmmh_operand_reuse.argtypes = [c, c, i, i, i, i, i, i, i, i, i, i, i, i, l, i, i, i, vp]
        L.tmmh_operand_reuse.restype = l
        L.tmmh_tile_groups.argtypes = [c, c, i, i, i, i, i, i, i, i, i, i, i, i, l, i, i, i, ctypes.POINTER(l)]
        L.tmmh_set_pair_max.argtypes = [i]
        L.tmmh_last_error.restype = ctypes.c_char_p
        _host = L
    return _host


def plan_info(ta, tb, m, n, k, beta_nz=1, copy_back=1, tiles=(4, 4, 4), nq=2, budget=0, first=0):
    info = (ctypes.c_long * 11)()
    lda = m if ta == "N" else k
    ldb = k if tb == "N" else n
    st = host().tmmh_plan_info(ta.encode(), tb.encode(), m, n, k, max(lda, 1), max(ldb, 1), m, beta_nz, copy_back,
                               *tiles, nq, budget, first, info)
    if st != 0:
        raise RuntimeError(host().tmmh_last_error().decode())
    keys = ["ops", "gemm", "h2d", "d2h", "pool", "nTm", "nTn", "nTk", "h2d_ops", "d2h_ops", "late_deps"]
    return dict(zip(keys, list(info)))


def simulate(dtype, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, copy_back, tiles, nq, budget, first, seed):
    code = tiledmm.DTYPE_CODE[np.dtype(dtype)]
    real = np.float32 if code in (0, 2) else np.float64
    al = np.array([complex(alpha).real, complex(alpha).imag], dtype=real)
    be = np.array([complex(beta).real, complex(beta).imag], dtype=real)
    full = np.full(m * n, np.nan, dtype=dtype)
    st = host().tmmh_simulate_gemm(code, ta.encode(), tb.encode(), m, n, k, al.ctypes.data, A.ctypes.data, lda,
                                   B.ctypes.data, ldb, be.ctypes.data, C.ctypes.data, ldc, int(copy_back),
                                   full.ctypes.data, *tiles, nq, budget, first, seed)
    if st != 0:
        raise RuntimeError(host().tmmh_last_error().decode())
    return full.reshape(m, n, order="F")


def test_shell_order_covers_grid():
    for nTm in range(1, 7):
        for nTn in range(1, 7):
            out = (ctypes.c_int * (2 * nTm * nTn))()
            cnt = host().tmmh_shell_order(nTm, nTn, out)
            tiles = [(out[2 * i], out[2 * i + 1]) for i in range(cnt)]
            assert sorted(tiles) == [(i, j) for i in range(nTm) for j in range(nTn)]
            # shells in non-decreasing order: each prefix needs only the A/B blocks of its shells
            shells = [max(i, j) for i, j in tiles]
            assert shells == sorted(shells)


@pytest.mark.parametrize("flags", ["NN", "TN", "NT", "TT"])
@pytest.mark.parametrize("beta_nz,copy_back", [(1, 1), (0, 1), (1, 0), (0, 0)])
def test_each_block_crosses_pcie_once(flags, beta_nz, copy_back):
    m, n, k = 50, 200, 21
    p = plan_info(flags[0], flags[1], m, n, k, beta_nz, copy_back, (4, 4, 4), 2)
    assert p["h2d"] == m * k + k * n + beta_nz * m * n
    assert p["d2h"] == copy_back * m * n
    assert (p["nTm"], p["nTn"], p["nTk"]) == (13, 50, 6)
    assert p["gemm"] == p["nTm"] * p["nTn"] * p["nTk"]
    # A blocks: nTm*nTk, B blocks: nTk*nTn, C_in blocks: nTm*nTn when beta != 0
    assert p["h2d_ops"] == 13 * 6 + 6 * 50 + beta_nz * 13 * 50
    assert p["d2h_ops"] == copy_back * 13 * 50


def test_reference_schedule_traffic_comparison():
    """At 40000^3 with 8000^3 tiles the reference's round_robin copies each A tile once per
    C-tile column and each B tile once per C-tile row (tiled_mm.cpp:292-330): 140.8 GB H2D.
    The plan moves each block once: 38.4 GB."""
    p = plan_info("T", "N", 40000, 40000, 40000, 1, 1, (8000, 8000, 8000), 4)
    assert p["h2d"] * 8 == 3 * 40000 * 40000 * 8  # 38.4 GB
    ref_h2d = (5 * 5 * 5 * 2 * 8000 * 8000 + 25 * 8000 * 8000) * 8
    assert ref_h2d == 140_800_000_000
    assert p["d2h"] * 8 == 40000 * 40000 * 8


def test_budget_groups_and_errors():
    big = plan_info("N", "N", 40, 40, 40, 1, 1, (10, 10, 10), 2)
    small = plan_info("N", "N", 40, 40, 40, 1, 1, (10, 10, 10), 2, budget=1400)
    assert small["pool"] <= 1400 < big["pool"]
    # with fewer resident tiles, A/B blocks are re-sent per group
    assert small["h2d"] > big["h2d"]
    with pytest.raises(RuntimeError):
        plan_info("N", "N", 40, 40, 40, 1, 1, (10, 10, 10), 2, budget=100)


def test_first_chunk_split():
    p = plan_info("T", "N", 100, 100, 100, 1, 1, (50, 50, 40), 2, first=8)
    # k chunks: 8, then 40-sized: 8, 48, 88, 100 -> 4 chunks
    assert p["nTk"] == 4


CASES = [
    # m, n, k, tiles, n_queues, budget, first_chunk
    (50, 200, 21, (4, 4, 4), 2, 0, 0),
    (5, 2, 2, (4, 4, 4), 2, 0, 0),
    (37, 29, 41, (10, 7, 9), 3, 0, 0),
    (37, 29, 41, (10, 7, 9), 1, 0, 3),
    (64, 48, 70, (16, 16, 16), 4, 3000, 0),    # budget forces several groups, slot reuse
    (64, 48, 70, (16, 16, 16), 2, 1800, 5),
    (9, 9, 0, (4, 4, 4), 2, 0, 0),              # k = 0
]


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.complex128, np.complex64])
@pytest.mark.parametrize("case", CASES)
def test_random_interleavings_match_oracle(dtype, case):
    m, n, k, tiles, nq, budget, first = case
    cplx = np.issubdtype(dtype, np.complexfloating)
    ops = ["N", "T", "C"] if cplx else ["N", "T"]
    rng = np.random.default_rng(m + n + k)
    for ta in ops:
        for tb in ops:
            ar, ac = oracle.op_shape(ta, m, k)
            br, bc = oracle.op_shape(tb, k, n)
            A = int_matrix(rng, ar, ac, dtype, ld=ar + 1)
            B = int_matrix(rng, br, bc, dtype, ld=br + 2)
            C0 = int_matrix(rng, m, n, dtype, ld=m + 3)
            for beta in (0, 2):
                ref = C0.copy(order="F")
                oracle.tiled_gemm_c(ta, tb, m, n, k, 3, A, ar + 1, B, br + 2, beta, ref, m + 3, *tiles)
                for copy_back in (True, False):
                    for seed in range(4):
                        C = C0.copy(order="F")
                        if beta == 0:
                            C[:m] = np.nan  # must not be read
                        full = simulate(dtype, ta, tb, m, n, k, 3, A, ar + 1, B, br + 2, beta, C, m + 3, copy_back,
                                        tiles, nq, budget, first, seed)
                        got = C[:m] if copy_back else full
                        np.testing.assert_array_equal(got, ref[:m], err_msg=f"{ta}{tb} beta={beta} cb={copy_back}")
                        if copy_back:
                            # rows between m and ldc are never written
                            np.testing.assert_array_equal(C[m:], C0[m:])


def reuse_plan(ta, tb, m, n, k, tiles, nq, capacity, max_k=1 << 30, budget=0, copy_back=1, beta_nz=1):
    lda = m if ta == "N" else k
    ldb = k if tb == "N" else n
    args = (ta.encode(), tb.encode(), m, n, k, max(lda, 1), max(ldb, 1), m, beta_nz, copy_back, *tiles, nq, budget, 0,
            capacity, max_k)
    nops = host().tmmh_operand_reuse(*args, None)
    assert nops >= 0, host().tmmh_last_error().decode()
    out = np.zeros(8 * nops, dtype=np.int64)
    assert host().tmmh_operand_reuse(*args, out.ctypes.data) == nops
    keys = ["kind", "queue", "prod_a", "slot_a", "prep_a", "prod_b", "slot_b", "prep_b"]
    return [dict(zip(keys, map(int, out[8 * i:8 * i + 8]))) for i in range(nops)]


def check_reuse(ops, capacity):
    """Replays the slot plan per compute queue: a GEMM reads a slot only while it holds the block
    the GEMM's operand came from, the two operands never share a slot, and slots stay below the
    capacity. Returns (preparations, distinct shared blocks) summed over the queues."""
    GEMM = 1
    state, preps, shared = {}, 0, 0
    uses = {}
    for op in ops:
        if op["kind"] == GEMM:
            for mat in "ab":
                if op["prod_" + mat] >= 0:
                    key = (op["queue"], mat, op["prod_" + mat])
                    uses[key] = uses.get(key, 0) + 1
    for op in ops:
        if op["kind"] != GEMM:
            assert op["slot_a"] == op["slot_b"] == -1
            continue
        q = op["queue"]
        if op["slot_a"] >= 0 and op["slot_b"] >= 0:
            assert op["slot_a"] != op["slot_b"]
        for mat in "ab":
            slot, prep, prod = op["slot_" + mat], op["prep_" + mat], op["prod_" + mat]
            key = (q, mat, prod)
            if slot < 0:
                assert prep == 0
                # only blocks read once (or without a producer) stay uncached with room to spare
                continue
            assert 0 <= slot < capacity and prod >= 0
            if prep:
                state[(q, slot)] = key
                preps += 1
            assert state.get((q, slot)) == key, "slot read while it holds another block"
    shared = sum(1 for v in uses.values() if v >= 2)
    return preps, shared


@pytest.mark.parametrize("flags", ["NN", "TN", "NT", "TT"])
@pytest.mark.parametrize("nq", [1, 2])
def test_operand_reuse_plan(flags, nq):
    """Scheme II host-to-host calls prepare each shared A / B block once (plan_operand_reuse):
    with room for every block, one preparation per block read by two or more tile GEMMs of a
    queue and every such read cached; with 2 or 3 slots the plan stays valid (more
    preparations); with 1 slot or k above the limit nothing is cached."""
    ta, tb = flags
    for (m, n, k, tiles) in [(40, 36, 28, (8, 8, 8)), (33, 50, 70, (8, 8, 16)), (64, 8, 64, (8, 8, 8)),
                             (17, 17, 5, (4, 4, 4))]:
        ops = reuse_plan(ta, tb, m, n, k, tiles, nq, capacity=64)
        preps, shared = check_reuse(ops, 64)
        assert preps == shared
        for op in ops:
            if op["kind"] == 1:
                for mat in "ab":
                    prod = op["prod_" + mat]
                    assert prod >= 0
        for cap in (2, 3):
            p_small, _ = check_reuse(reuse_plan(ta, tb, m, n, k, tiles, nq, capacity=cap), cap)
            assert p_small >= preps
        none = reuse_plan(ta, tb, m, n, k, tiles, nq, capacity=1)
        assert all(op["slot_a"] == op["slot_b"] == -1 for op in none)
        limited = reuse_plan(ta, tb, m, n, k, tiles, nq, capacity=64, max_k=0)
        assert all(op["slot_a"] == op["slot_b"] == -1 for op in limited if op["kind"] == 1)


@pytest.mark.parametrize("defer", [1, 2, 5])
@pytest.mark.parametrize("case", [c for c in CASES if c[2] > 0] + [(40, 36, 20, (8, 8, 10), 2, 0, 0),
                                                                     (40, 36, 20, (8, 8, 10), 1, 0, 6)])
def test_deferred_cin_copies_match_oracle(defer, case):
    """PlanRequest::defer_cin: the last chunk's C_in copies go after later tiles' A / B copies
    (a forward late_dep of their GEMM); random interleavings of the queues still finish (no
    deadlock), give the oracle's result exactly and move the same bytes."""
    m, n, k, tiles, nq, budget, first = case
    rng = np.random.default_rng(7 * m + n + k + defer)
    base = plan_info("T", "N", m, n, k, tiles=tiles, nq=nq, budget=budget, first=first)
    host().tmmh_set_defer_cin(defer)
    try:
        info = plan_info("T", "N", m, n, k, tiles=tiles, nq=nq, budget=budget, first=first)
        assert info["h2d"] == base["h2d"] and info["ops"] == base["ops"] and base["late_deps"] == 0
        # with two chunks every tile's C_in copy falls in the last chunk and is read late
        if info["nTk"] == 2 and budget == 0:
            assert info["late_deps"] == info["nTm"] * info["nTn"]
        for ta, tb in (("N", "N"), ("T", "N"), ("N", "T")):
            ar, ac = oracle.op_shape(ta, m, k)
            br, bc = oracle.op_shape(tb, k, n)
            A = int_matrix(rng, ar, ac, np.float64, ld=ar + 1)
            B = int_matrix(rng, br, bc, np.float64, ld=br + 2)
            C0 = int_matrix(rng, m, n, np.float64, ld=m + 3)
            ref = C0.copy(order="F")
            oracle.tiled_gemm_c(ta, tb, m, n, k, 3, A, ar + 1, B, br + 2, 2, ref, m + 3, *tiles)
            for copy_back in (True, False):
                for seed in range(4):
                    C = C0.copy(order="F")
                    full = simulate(np.float64, ta, tb, m, n, k, 3, A, ar + 1, B, br + 2, 2, C, m + 3, copy_back,
                                    tiles, nq, budget, first, seed)
                    got = C[:m] if copy_back else full
                    np.testing.assert_array_equal(got, ref[:m], err_msg=f"{ta}{tb} cb={copy_back} seed={seed}")
    finally:
        host().tmmh_set_defer_cin(0)


def test_operand_reuse_counts_on_the_40000_grid():
    """The 40000^3 TN schedule with 8000^3 tiles on one compute queue: 5 A and 5 B blocks per
    k-chunk, each read by 5 tile GEMMs, so 50 preparations instead of 250."""
    ops = reuse_plan("T", "N", 40000, 40000, 40000, (8000, 8000, 8000), 1, capacity=10)
    preps, shared = check_reuse(ops, 10)
    assert shared == 50 and preps == 50
    assert sum(op["kind"] == 1 for op in ops) == 125


def tile_groups(ta, tb, m, n, k, tiles, nq, budget=0, first=0, max_group=16, unsafe=0, beta_nz=1, copy_back=1):
    info = (ctypes.c_long * 3)()
    lda = m if ta == "N" else k
    ldb = k if tb == "N" else n
    st = host().tmmh_tile_groups(ta.encode(), tb.encode(), m, n, k, max(lda, 1), max(ldb, 1), m, beta_nz, copy_back,
                                 *tiles, nq, budget, first, max_group, unsafe, info)
    assert st == 0, host().tmmh_last_error().decode()
    return {"groups": info[0], "grouped": info[1], "finishes": bool(info[2])}


@pytest.mark.parametrize("flags", ["NN", "TN", "NT", "TT"])
@pytest.mark.parametrize("nq", [1, 2])
def test_tile_groups_progress(flags, nq):
    """Grouped tile launches (plan_tile_groups): the executor's threads always finish, with and
    without a pool budget, and the 40000^3 schedule groups the tile GEMMs of the k-chunks that
    follow a chunk at least as long, except the last."""
    ta, tb = flags
    for (m, n, k, tiles, first) in [(40, 36, 28, (8, 8, 8), 0), (33, 50, 70, (8, 8, 16), 4), (64, 8, 64, (8, 8, 8), 0),
                                    (17, 17, 5, (4, 4, 4), 0), (100, 90, 80, (16, 16, 16), 2)]:
        for budget in (0, 8 * 8 * 6, 8 * 8 * 10):
            try:
                r = tile_groups(ta, tb, m, n, k, tiles, nq, budget=budget, first=first, max_group=4)
            except AssertionError as e:
                assert "budget" in str(e) or "pool" in str(e)
                continue
            assert r["finishes"], (m, n, k, tiles, budget)
    # chunks 1000, 7800 x 5: the tile GEMMs of chunks 2..4 (3 x 25) are grouped, 2 groups each
    big = tile_groups("T", "N", 40000, 40000, 40000, (8000, 8000, 8000), 1, first=1000)
    assert big["finishes"] and big["grouped"] == 75 and big["groups"] == 6


def test_tile_group_members_40000():
    """The grouped launches of the 40000^3 TN call on 8064-wide blocks (TMM_OPT_TILE_GROUPS 2):
    chunks 2-4 as one 16-block and one 9-block launch each, chunks 1 and 5 in pairs of tiles that
    follow each other in the plan, the last tile of each of those chunks alone; the algorithmic
    bytes of a 16-block launch are 4 A and 4 B blocks plus 16 C blocks read and written."""
    groups = group_members("T", "N", 40000, 40000, 40000, 40064, 40064, 40064, 1, 1, (8064, 8064, 8000), 1,
                           first=1000, max_group=16, pair_max=2)
    sizes = {}
    for g in groups:
        assert len({t["tc"] for t in g}) == 1
        sizes.setdefault(g[0]["tc"], []).append(len(g))
    assert sizes == {1: [2] * 12, 2: [16, 9], 3: [16, 9], 4: [16, 9], 5: [2] * 12}
    grouped = {(t["ti"], t["tj"], t["tc"]) for g in groups for t in g}
    assert (4, 4, 1) not in grouped and (4, 4, 5) not in grouped  # each of those chunks' last tile
    big = next(g for g in groups if len(g) == 16)
    b = 8064 * 8000
    assert grouped_algorithmic_bytes([big], 8) == 8 * (8 * b + 16 * 2 * 8064 * 8064)


@pytest.mark.parametrize("flags", ["NN", "TN", "NT", "TT"])
@pytest.mark.parametrize("nq", [1, 2, 3])
@pytest.mark.parametrize("pair_max", [2, 3])
def test_tile_groups_short_runs_progress(flags, nq, pair_max):
    """Short runs in the chunks that are not steady (TMM_OPT_TILE_GROUPS >= 2): the executor's
    threads always finish; at 40000^3 chunk 1 and the last chunk run as pairs, the last GEMM alone."""
    ta, tb = flags
    host().tmmh_set_pair_max(pair_max)
    try:
        for (m, n, k, tiles, first) in [(40, 36, 28, (8, 8, 8), 0), (33, 50, 70, (8, 8, 16), 4),
                                        (64, 8, 64, (8, 8, 8), 0), (17, 17, 5, (4, 4, 4), 0),
                                        (100, 90, 80, (16, 16, 16), 2), (50, 200, 21, (4, 4, 4), 0)]:
            for budget in (0, 8 * 8 * 6, 8 * 8 * 10):
                try:
                    r = tile_groups(ta, tb, m, n, k, tiles, nq, budget=budget, first=first, max_group=4)
                except AssertionError as e:
                    assert "budget" in str(e) or "pool" in str(e)
                    continue
                assert r["finishes"], (m, n, k, tiles, budget)
        # chunk 1: 25 tiles in runs of pair_max; the last chunk: 24 tiles in runs, then its last GEMM alone
        big = tile_groups("T", "N", 40000, 40000, 40000, (8000, 8000, 8000), 1, first=1000)
        assert big["finishes"]
        assert big["groups"] == 6 + 25 // pair_max + 24 // pair_max
        assert big["grouped"] == 75 + pair_max * (25 // pair_max + 24 // pair_max)
    finally:
        host().tmmh_set_pair_max(0)


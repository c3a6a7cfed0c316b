"""Batched host-to-host DGEMM (tmm_dgemm_batched) against the oracle.

Problems of uneven size in one call, through the persistent kernel and (with pinned host
memory) one hipGraph: integer data must come back exact for every problem; pageable host
memory takes the staged path with the same results. The autotuner's choice is recorded and
its measured rates must be positive.
"""
import numpy as np
import pytest

import oracle
import tiledmm

pytestmark = pytest.mark.gpu


def _batch(rng, count, lo, hi, flags, pinned_torch=None):
    ta, tb = flags
    probs = []
    keep = []
    for _ in range(count):
        m, n, k = (int(round(np.exp(rng.uniform(np.log(lo), np.log(hi))))) for _ in range(3))
        ar, ac = oracle.op_shape(ta, m, k)
        br, bc = oracle.op_shape(tb, k, n)
        lda, ldb, ldc = ar + int(rng.integers(0, 3)), br + int(rng.integers(0, 3)), m + int(rng.integers(0, 3))
        mats = []
        for rows, ld, cols in ((ar, lda, ac), (br, ldb, bc), (m, ldc, n)):
            X = np.zeros((ld, cols), order="F")
            X[:rows] = rng.integers(0, 10, (rows, cols))
            if pinned_torch is not None:
                t = pinned_torch.empty(X.size, dtype=pinned_torch.float64, pin_memory=True)
                arr = t.numpy().reshape(X.shape, order="F")
                arr[...] = X
                keep.append(t)
                X = arr
            mats.append(X)
        alpha, beta = float(rng.integers(1, 3)), float(rng.integers(0, 3))
        probs.append(dict(m=m, n=n, k=k, A=mats[0], B=mats[1], C=mats[2], lda=lda, ldb=ldb, ldc=ldc, ar=ar, br=br,
                          alpha=alpha, beta=beta))
    return probs, keep


def _run(h, flags, probs):
    h.gemm_batched(flags[0], flags[1], [p["m"] for p in probs], [p["n"] for p in probs], [p["k"] for p in probs],
                   [p["alpha"] for p in probs], [p["A"] for p in probs], [p["lda"] for p in probs],
                   [p["B"] for p in probs], [p["ldb"] for p in probs], [p["beta"] for p in probs],
                   [p["C"] for p in probs], [p["ldc"] for p in probs])


@pytest.mark.parametrize("flags", ["NN", "TN", "NT"])
@pytest.mark.parametrize("pinned,graph", [(True, 1), (True, 0), (True, -1), (False, -1)])
def test_batched_integer_exact(torch_gpu, flags, pinned, graph):
    """Exact on integer data through each way of running the pipeline: the captured hipGraph
    (TMM_OPT_BATCH_GRAPH 1), the three queues (0), the autotuner's choice (-1), pageable memory."""
    rng = np.random.default_rng(11)
    probs, keep = _batch(rng, 40, 8, 700, flags, torch_gpu if pinned else None)
    expected = []
    for p in probs:
        ref = p["C"][:p["m"]].copy(order="F")
        oracle.tiled_gemm_np(flags[0], flags[1], p["m"], p["n"], p["k"], p["alpha"], p["A"], p["lda"], p["B"],
                             p["ldb"], p["beta"], ref, p["m"], 1 << 20, 1 << 20, 1 << 20)
        expected.append(ref)
    with tiledmm.Handle(tiledmm.F64, 2) as h:
        h.set_option(tiledmm.OPT_BATCH_GRAPH, graph)
        _run(h, flags, probs)
        st = h.stats()
        cfg = h.batch_config()
    for p, ref in zip(probs, expected):
        np.testing.assert_array_equal(p["C"][:p["m"]], ref)
    want = {1: 1.0, 0: 0.0}.get(graph, float(cfg["graph"])) if pinned else 0.0
    assert st["batch_graph"] == want
    assert 1 <= cfg["groups"] <= 40 and cfg["workgroups"] >= 256
    assert cfg["h2d_Bps"] > 1e9 and cfg["d2h_Bps"] > 1e9 and cfg["kernel_flops"] > 1e12
    del keep


def test_batched_uniform_bound(torch_gpu):
    rng = np.random.default_rng(12)
    probs = []
    for _ in range(12):
        m, n, k = (int(x) for x in rng.integers(64, 600, 3))
        probs.append(dict(m=m, n=n, k=k, A=np.asfortranarray(rng.uniform(-1, 1, (m, k))),
                          B=np.asfortranarray(rng.uniform(-1, 1, (k, n))),
                          C=np.asfortranarray(rng.uniform(-1, 1, (m, n))), lda=m, ldb=k, ldc=m, alpha=1.0, beta=1.0))
    C0 = [p["C"].copy(order="F") for p in probs]
    with tiledmm.Handle(tiledmm.F64, 1) as h:
        _run(h, "NN", probs)
    for p, c0 in zip(probs, C0):
        ref = p["A"] @ p["B"] + c0
        scale = np.abs(p["A"]) @ np.abs(p["B"]) + np.abs(c0)
        assert oracle.normwise_error(p["C"], ref, scale) <= 2 * p["k"] * 2.0 ** -53


def test_batched_errors(torch_gpu):
    with tiledmm.Handle(tiledmm.F64, 1) as h:
        A = np.zeros((4, 4), order="F")
        with pytest.raises(tiledmm.TmmError):
            h.gemm_batched("N", "N", [4], [4], [4], [1.0], [A], [3], [A], [4], [0.0], [A], [4])
    with tiledmm.Handle(tiledmm.F32, 1) as h:
        A = np.zeros((4, 4), order="F")
        with pytest.raises(tiledmm.TmmError) as e:
            h.gemm_batched("N", "N", [4], [4], [4], [1.0], [A], [4], [A], [4], [0.0], [A], [4])
        assert e.value.status == 4


def _expected(flags, probs):
    out = []
    for p in probs:
        ref = p["C"][:p["m"]].copy(order="F")
        oracle.tiled_gemm_np(flags[0], flags[1], p["m"], p["n"], p["k"], p["alpha"], p["A"], p["lda"], p["B"],
                             p["ldb"], p["beta"], ref, p["m"], 1 << 20, 1 << 20, 1 << 20)
        out.append(ref)
    return out


@pytest.mark.parametrize("flags", ["NN", "TN", "NT", "TT"])
@pytest.mark.parametrize("body,queues,graph", [(0, 1, 0), (1, 1, 0), (1, 2, 0), (0, 2, 1), (1, 2, 1)])
def test_batched_tile_bodies_and_queues_exact(torch_gpu, flags, body, queues, graph):
    """Every choice the autotuner can make, forced: the register-staged and the direct-to-LDS
    persistent body (odd m, n and k not a multiple of 8: the arena over-read of a row-contiguous
    operand and the k remainder of the epilogue), one or two compute queues, graph or queues; exact
    on integer data."""
    rng = np.random.default_rng(13)
    probs, keep = _batch(rng, 40, 1, 700, flags, torch_gpu)
    expected = _expected(flags, probs)
    with tiledmm.Handle(tiledmm.F64, 2) as h:
        h.set_option(tiledmm.OPT_BATCH_BODY, body)
        h.set_option(tiledmm.OPT_BATCH_QUEUES, queues)
        h.set_option(tiledmm.OPT_BATCH_GRAPH, graph)
        _run(h, flags, probs)
        st = h.stats()
    for p, ref in zip(probs, expected):
        np.testing.assert_array_equal(p["C"][:p["m"]], ref)
    assert st["batch_body"] == body and st["batch_queues"] == queues and st["batch_graph"] == graph
    del keep


@pytest.mark.parametrize("emulation", [1, -1, 0])
def test_batched_emulation_per_problem(torch_gpu, emulation):
    """A handle with Ozaki moduli set: the autotuner runs a problem as scheme II products when its
    model says so (-1), or always (1), or never (0). Large problems are emulated, small ones are
    not; integer data: exact either way."""
    rng = np.random.default_rng(14)
    probs, keep = _batch(rng, 12, 16, 300, "NN", torch_gpu)
    for m, n, k in ((4096, 4096, 4096), (3072, 4096, 4096)):  # large enough to favour emulation
        A = np.asfortranarray(rng.integers(0, 10, (m, k)).astype(np.float64))
        B = np.asfortranarray(rng.integers(0, 10, (k, n)).astype(np.float64))
        C = np.asfortranarray(rng.integers(0, 10, (m, n)).astype(np.float64))
        probs.append(dict(m=m, n=n, k=k, A=A, B=B, C=C, lda=m, ldb=k, ldc=m, ar=m, br=k, alpha=1.0, beta=1.0))
    expected = _expected("NN", probs)
    with tiledmm.Handle(tiledmm.F64, 2) as h:
        h.set_option(tiledmm.OPT_OZAKI_MODULI, 16)
        h.set_option(tiledmm.OPT_BATCH_EMULATION, emulation)
        _run(h, "NN", probs)
        st = h.stats()
        cfg = h.batch_config()
    print(f"[batched] emulation {emulation}: {cfg}, emulated {st['batch_emulated']}", flush=True)
    for p, ref in zip(probs, expected):
        np.testing.assert_array_equal(p["C"][:p["m"]], ref)
    if emulation == 1:
        assert st["batch_emulated"] == len(probs)
    elif emulation == 0:
        assert st["batch_emulated"] == 0
    else:
        assert 2 <= st["batch_emulated"] < len(probs), st
    del keep

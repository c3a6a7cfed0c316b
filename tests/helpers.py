"""Shared helpers of the test-suite: data in the reference test's style and error checks."""
import numpy as np

import oracle

NP = {"s": np.float32, "d": np.float64, "c": np.complex64, "z": np.complex128}


def int_matrix(rng, rows, cols, dtype, ld=None):
    """Integer entries in [0, 9], as tests/test-multiply.cpp:58-66 fills its matrices."""
    ld = rows if ld is None else ld
    X = np.zeros((ld, cols), dtype=dtype, order="F")
    X[:rows, :] = rng.integers(0, 10, size=(rows, cols))
    if np.iscomplexobj(X):
        X[:rows, :] += 1j * rng.integers(0, 10, size=(rows, cols))
    return X


def uniform_matrix(rng, rows, cols, dtype, ld=None, lo=-1.0, hi=1.0):
    ld = rows if ld is None else ld
    X = np.zeros((ld, cols), dtype=dtype, order="F")
    X[:rows, :] = rng.uniform(lo, hi, size=(rows, cols))
    if np.iscomplexobj(X):
        X[:rows, :] += 1j * rng.uniform(lo, hi, size=(rows, cols))
    if ld > rows:  # padding rows must never be read: poison them
        X[rows:, :] = np.nan
    return X


def bound(dtype, k):
    """Stated normwise bound: ||C - C_ref||_F <= 2 k u || |alpha||A||B| + |beta||C0| ||_F."""
    return 2 * max(k, 1) * oracle.unit_roundoff(np.dtype(dtype).type(0).real.dtype)


def check_normwise(C, Cref, ta, tb, alpha, A, B, beta, C0, k, dtype):
    scale = oracle.error_scale(ta, tb, alpha, A, B, beta, C0)
    err = oracle.normwise_error(C, Cref, scale)
    assert err <= bound(dtype, k), f"normwise error {err:.3e} above bound {bound(dtype, k):.3e}"
    return err


def group_members(ta, tb, m, n, k, lda, ldb, ldc, beta_nz, copy_back, tiles, nq, budget=0, first=0, max_group=16,
                  pair_max=0):
    """Members of the grouped tile launches the host planner makes for a plan (tmmh_group_members):
    a list of groups, each a list of dicts (ti, tj, tc, m, n, k, accumulate, use_cin)."""
    import ctypes

    import tiledmm
    L = ctypes.CDLL(tiledmm.HOST_PATH)
    c, i, l = ctypes.c_char, ctypes.c_int, ctypes.c_long
    L.tmmh_group_members.argtypes = [c, c, i, i, i, i, i, i, i, i, i, i, i, i, l, i, i, i, ctypes.c_void_p]
    L.tmmh_group_members.restype = l
    args = (ta.encode(), tb.encode(), m, n, k, lda, ldb, ldc, beta_nz, copy_back, *tiles, nq, budget, first,
            max_group, pair_max)
    cnt = L.tmmh_group_members(*args, None)
    assert cnt >= 0
    buf = (l * (9 * max(cnt, 1)))()
    L.tmmh_group_members(*args, buf)
    keys = ["ti", "tj", "tc", "m", "n", "k", "accumulate", "use_cin"]
    groups = {}
    for r in range(cnt):
        row = list(buf[9 * r:9 * r + 9])
        groups.setdefault(row[0], []).append(dict(zip(keys, row[1:])))
    return [groups[g] for g in sorted(groups)]


def grouped_algorithmic_bytes(groups, elem_bytes):
    """Per grouped launch: its distinct A blocks (ti, tc) and B blocks (tj, tc) read once, each C
    tile written once and read once per operand use (accumulation, the caller's C)."""
    total = 0
    for g in groups:
        seen_a, seen_b = set(), set()
        for t in g:
            if (t["ti"], t["tc"]) not in seen_a:
                seen_a.add((t["ti"], t["tc"]))
                total += t["m"] * t["k"]
            if (t["tj"], t["tc"]) not in seen_b:
                seen_b.add((t["tj"], t["tc"]))
                total += t["n"] * t["k"]
            total += t["m"] * t["n"] * (1 + t["accumulate"] + t["use_cin"])
    return total * elem_bytes

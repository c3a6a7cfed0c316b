"""Pins the oracle (oracle/tmm_oracle.c, oracle/oracle.py) before it is trusted as the checker.

  * tile policy and tile geometry: golden tables produced by the reference's own sources
    (tests/golden/make_golden.py via oracle/_ref) and, when oracle/_ref is built, a live sweep;
  * the 4x4 packing example of Tiled-MM src/Tiled-MM/tiled_mm.hpp:20-35;
  * tiled GEMM: exact integer known answers computed with Python integers, exact dyadic
    brute force, numpy's BLAS on uniform data (stated bound), algebraic identities, and planted
    mistakes in a copy of the restatement that these checks must catch.
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle
from helpers import bound, int_matrix, uniform_matrix

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def test_tile_policy_golden():
    for dim, cur, mx, expected in _load("tile_policy.json"):
        assert oracle.optimal_tile_size(dim, cur, mx) == expected, (dim, cur, mx)


def test_tile_policy_configurations():
    # the tile sizes the reference picks at the benchmark configurations
    assert oracle.optimal_tile_size(2048, 0, 512) == 512
    assert oracle.optimal_tile_size(40000, 0, 8192) == 8000
    assert [oracle.optimal_tile_size(d, 0, 6144) for d in (24000, 24000, 32000)] == [6000, 6000, 4000]
    # closed form: if dim <= max -> dim
    for d in range(1, 200):
        assert oracle.optimal_tile_size(d, 0, 200) == d


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
def test_tile_policy_matches_reference_sweep():
    R = oracle.ref()
    for mx in (1, 2, 3, 4, 7, 64, 100, 257):
        for dim in range(1, 700):
            for cur in (0, mx // 2, mx):
                assert oracle.optimal_tile_size(dim, cur, mx) == R.ref_optimal_tile_size(dim, cur, mx)


def test_geometry_golden():
    L = oracle.lib()
    for g in _load("geometry.json"):
        rows, cols, ld = g["rows"], g["cols"], g["ld"]
        tr, tc = g["tile"]
        ti, tj = g["coord"]
        assert [L.oracle_num_tiles(rows, tr), L.oracle_num_tiles(cols, tc)] == g["num_tiles"]
        assert [L.oracle_tile_extent(rows, tr, ti), L.oracle_tile_extent(cols, tc, tj)] == g["extent"]
        assert L.oracle_tile_offset(rows, cols, ld, tr, tc, ti, tj) == g["offset"]


def test_packing_example():
    import ctypes

    g = _load("packing_4x4.json")
    host = np.array(g["host"], dtype=np.float64)
    out = np.zeros(4)
    r, c = ctypes.c_int(), ctypes.c_int()
    for key, expected in g["tiles"].items():
        ti, tj = int(key[0]), int(key[1])
        oracle.lib().oracle_pack_tile_d(host.ctypes.data, 4, 4, 4, 2, 2, ti, tj, out.ctypes.data, ctypes.byref(r),
                                        ctypes.byref(c))
        assert (r.value, c.value) == (2, 2)
        assert out.tolist() == expected, key


@pytest.mark.parametrize("tiles", [(2, 3, 1), (4, 4, 4), (100, 100, 100)])
def test_gemm_kat_golden(tiles):
    for case in _load("gemm_kat.json"):
        m, n, k = case["m"], case["n"], case["k"]
        A = np.asfortranarray(np.array(case["A"], dtype=np.float64))
        B = np.asfortranarray(np.array(case["B"], dtype=np.float64))
        C = np.asfortranarray(np.array(case["C"], dtype=np.float64))
        for fn in (oracle.tiled_gemm_c, oracle.tiled_gemm_np):
            Cw = C.copy(order="F")
            fn(case["ta"], case["tb"], m, n, k, case["alpha"], A, A.shape[0], B, B.shape[0], case["beta"], Cw, m,
               *tiles)
            assert Cw.tolist() == case["expected"], (fn.__name__, case["ta"], case["tb"], m, n, k)


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.complex128, np.complex64])
@pytest.mark.parametrize("shape,tiles", [((50, 200, 21), (4, 4, 4)), ((5, 2, 2), (4, 4, 4)),
                                         ((37, 29, 41), (10, 7, 9)), ((64, 64, 64), (64, 64, 64))])
def test_integer_exact_all_types(dtype, shape, tiles):
    """Reference-test style data (integers in [0, 9]): every tiled result is exact."""
    m, n, k = shape
    rng = np.random.default_rng(0)
    cplx = np.issubdtype(dtype, np.complexfloating)
    ops = ["N", "T", "C"] if cplx else ["N", "T"]
    for ta in ops:
        for tb in ops:
            ar, ac = oracle.op_shape(ta, m, k)
            br, bc = oracle.op_shape(tb, k, n)
            A = int_matrix(rng, ar, ac, dtype)
            B = int_matrix(rng, br, bc, dtype)
            C = int_matrix(rng, m, n, dtype)
            Ai = A.astype(np.complex128 if cplx else np.int64)
            Bi = B.astype(np.complex128 if cplx else np.int64)
            exact = 2 * (oracle.apply_op(ta, Ai) @ oracle.apply_op(tb, Bi)) + 3 * C
            for fn in (oracle.tiled_gemm_c, oracle.tiled_gemm_np):
                Cw = C.copy(order="F")
                fn(ta, tb, m, n, k, 2, A, ar, B, br, 3, Cw, m, *tiles)
                np.testing.assert_array_equal(Cw, exact.astype(dtype), err_msg=f"{fn.__name__} {ta}{tb}")


def test_dyadic_brute_force():
    """Inputs with few significant bits: every product and sum is exact in double, so the
    oracle must equal exact rational arithmetic."""
    rng = np.random.default_rng(1)
    m, n, k = 6, 5, 7
    A = np.asfortranarray(rng.integers(-64, 64, (m, k)) / 8.0)
    B = np.asfortranarray(rng.integers(-64, 64, (k, n)) / 16.0)
    C = np.asfortranarray(rng.integers(-64, 64, (m, n)) / 4.0)
    alpha, beta = 0.75, -1.5
    exact = [[Fraction(alpha) * sum(Fraction(A[i, p]) * Fraction(B[p, j]) for p in range(k))
              + Fraction(beta) * Fraction(C[i, j]) for j in range(n)] for i in range(m)]
    Cw = C.copy(order="F")
    oracle.tiled_gemm_c("N", "N", m, n, k, alpha, A, m, B, k, beta, Cw, m, 4, 3, 2)
    assert [[Fraction(v) for v in row] for row in Cw.tolist()] == exact


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.complex128])
def test_uniform_vs_numpy_blas(dtype):
    m, n, k = 120, 97, 150
    rng = np.random.default_rng(2)
    cplx = np.issubdtype(dtype, np.complexfloating)
    wide = np.complex128 if cplx else np.float64
    alpha, beta = ((0.5 + 2j), (1 - 1j)) if cplx else (0.5, -2.0)
    for ta, tb in (("N", "N"), ("T", "N"), ("N", "T")) + ((("C", "N"), ("N", "C")) if cplx else ()):
        ar, ac = oracle.op_shape(ta, m, k)
        br, bc = oracle.op_shape(tb, k, n)
        A = uniform_matrix(rng, ar, ac, dtype, ld=ar + 2)
        B = uniform_matrix(rng, br, bc, dtype)
        C = uniform_matrix(rng, m, n, dtype)
        Cw = C.copy(order="F")
        oracle.tiled_gemm_c(ta, tb, m, n, k, alpha, A, ar + 2, B, br, beta, Cw, m, 32, 40, 50)
        ref = oracle.host_blas_gemm(ta, tb, alpha, A[:ar].astype(wide), B.astype(wide), beta, C.astype(wide))
        scale = oracle.error_scale(ta, tb, alpha, A[:ar], B, beta, C)
        assert oracle.normwise_error(Cw, ref, scale) <= bound(dtype, k), (ta, tb)


def test_identities():
    rng = np.random.default_rng(3)
    m = 23
    A = np.asfortranarray(rng.uniform(-1, 1, (m, m)))
    I = np.asfortranarray(np.eye(m))
    Cw = np.full((m, m), np.nan, order="F")
    oracle.tiled_gemm_c("N", "N", m, m, m, 1.0, A, m, I, m, 0.0, Cw, m, 5, 7, 3)
    np.testing.assert_array_equal(Cw, A)  # A*I = A exactly, C (NaN) not read with beta = 0
    Cw = np.full((m, m), np.nan, order="F")
    oracle.tiled_gemm_c("T", "N", m, m, m, 1.0, A, m, I, m, 0.0, Cw, m, 5, 7, 3)
    np.testing.assert_array_equal(Cw, A.T)
    C = np.asfortranarray(rng.uniform(-1, 1, (m, m)))
    Cw = C.copy(order="F")
    oracle.tiled_gemm_c("N", "N", m, m, m, 0.0, A, m, A, m, 2.0, Cw, m, 5, 7, 3)
    np.testing.assert_array_equal(Cw, 2.0 * C)  # alpha = 0: beta*C
    Cw = C.copy(order="F")
    oracle.tiled_gemm_c("N", "N", m, m, 0, 1.0, A, m, A, 1, 2.0, Cw, m, 5, 7, 3)
    np.testing.assert_array_equal(Cw, 2.0 * C)  # k = 0: beta*C
    Z = np.asfortranarray(rng.uniform(-1, 1, (m, m)) + 1j * rng.uniform(-1, 1, (m, m)))
    ZI = np.asfortranarray(np.eye(m, dtype=np.complex128))
    Cw = np.zeros((m, m), dtype=np.complex128, order="F")
    oracle.tiled_gemm_c("C", "N", m, m, m, 1.0, Z, m, ZI, m, 0.0, Cw, m, 5, 7, 3)
    np.testing.assert_array_equal(Cw, Z.conj().T)


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.complex128])
def test_alpha_zero_reads_neither_a_nor_b(dtype):
    """BLAS: alpha = 0 gives C = beta*C without reading A or B, so Inf and NaN there do not reach C
    (the quick path of the reference BLAS dgemm, kept by the vendor GEMM each tile calls); with
    beta = 0 as well C becomes 0 without being read."""
    rng = np.random.default_rng(9)
    m, n, k = 13, 11, 17
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k))).astype(dtype)
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n))).astype(dtype)
    A[3, 5], B[7, 2], A[0, 0] = np.nan, np.inf, -np.inf
    C = np.asfortranarray(rng.uniform(-1, 1, (m, n))).astype(dtype)
    for fn in (oracle.tiled_gemm_c, oracle.tiled_gemm_np):
        Cw = C.copy(order="F")
        fn("N", "N", m, n, k, 0.0, A, m, B, k, 1.5, Cw, m, 5, 4, 6)
        np.testing.assert_array_equal(Cw, (1.5 * C).astype(dtype))
        Cw = np.full((m, n), np.nan, dtype=dtype, order="F")
        fn("N", "N", m, n, k, 0.0, A, m, B, k, 0.0, Cw, m, 5, 4, 6)
        np.testing.assert_array_equal(Cw, np.zeros((m, n), dtype=dtype))
    np.testing.assert_array_equal(oracle.host_blas_gemm("N", "N", 0.0, A, B, 1.5, C), 1.5 * C)
    # alpha != 0: the NaN and Inf propagate (row 3 meets a NaN, column 2 an Inf, row 0 a -Inf)
    Cw = C.copy(order="F")
    oracle.tiled_gemm_c("N", "N", m, n, k, 1.0, A, m, B, k, 1.0, Cw, m, 5, 4, 6)
    assert np.all(~np.isfinite(Cw[3, :])) and np.all(~np.isfinite(Cw[:, 2])) and np.all(~np.isfinite(Cw[0, :]))


def _planted(kind, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tm, tn, tk):
    """A copy of the numpy restatement with one deliberate mistake (kind)."""
    Av = oracle.apply_op(ta, oracle._view(A, lda, *oracle.op_shape(ta, m, k)))
    Bv = oracle.apply_op(tb, oracle._view(B, ldb, *oracle.op_shape(tb, k, n)))
    Cv = oracle._view(C, ldc, m, n)
    for i0 in range(0, m, tm):
        for j0 in range(0, n, tn):
            acc = Cv[i0:i0 + tm, j0:j0 + tn].copy()
            for kt, k0 in enumerate(range(0, k, tk)):
                prod = Av[i0:i0 + tm, k0:k0 + tk] @ Bv[k0:k0 + tk, j0:j0 + tn]
                if kind == "beta_every_ktile":
                    bp = beta
                elif kind == "drop_last_ktile" and k0 + tk >= k and kt > 0:
                    continue
                else:
                    bp = beta if kt == 0 else 1
                acc = alpha * prod + bp * acc
            if kind == "swapped_tile_index":
                Cv[j0:j0 + tn, i0:i0 + tm] = acc.T if acc.shape[::-1] == Cv[j0:j0 + tn, i0:i0 + tm].shape else 0
            else:
                Cv[i0:i0 + tm, j0:j0 + tn] = acc


@pytest.mark.parametrize("kind", ["beta_every_ktile", "drop_last_ktile", "swapped_tile_index"])
def test_planted_mistakes_are_caught(kind):
    rng = np.random.default_rng(4)
    m, n, k = 12, 8, 10
    A = int_matrix(rng, m, k, np.float64)
    B = int_matrix(rng, k, n, np.float64)
    C = int_matrix(rng, m, n, np.float64)
    exact = A.astype(np.int64) @ B.astype(np.int64) + 2 * C.astype(np.int64)
    good = C.copy(order="F")
    oracle.tiled_gemm_np("N", "N", m, n, k, 1.0, A, m, B, k, 2.0, good, m, 4, 4, 3)
    np.testing.assert_array_equal(good, exact)
    bad = C.copy(order="F")
    _planted(kind, "N", "N", m, n, k, 1.0, A, m, B, k, 2.0, bad, m, 4, 4, 3)
    assert not np.array_equal(bad, exact), kind

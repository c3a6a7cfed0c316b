"""Host-to-host GEMM (gpu::gemm through tmm_{s,d,c,z}gemm) against the oracle.

Cases follow the reference's own tests: the CTest size matrix (Tiled-MM tests/CMakeLists.txt:11-14,
the largest scaled down), the CI tiny-tile cases (ci/daint-alps.yml:51,60), integer data in [0, 9]
with exact comparison (tests/test-multiply.cpp:58-98), both copy_c_back modes
(test-multiply.cpp:303-346) and the NN/TT/NT/TN flags; plus what the reference does not test:
'C' flags, complex and single precision, ld > rows, beta = 0 with NaN in C, pinned and pageable
host memory, k = 0, and the reference library itself (built from its sources) on the same inputs.
"""
import numpy as np
import pytest

import oracle
import tiledmm
from helpers import bound, group_members, grouped_algorithmic_bytes, int_matrix, uniform_matrix

pytestmark = pytest.mark.gpu

CODES = {np.float32: tiledmm.F32, np.float64: tiledmm.F64, np.complex64: tiledmm.C32, np.complex128: tiledmm.C64}


def _pinned_like(torch, X):
    """Copy X into page-locked host memory (hipHostMalloc through torch), same layout."""
    flat = np.ascontiguousarray(X.reshape(-1, order="F"))
    t = torch.empty(flat.size, dtype={np.dtype(np.float64): torch.float64, np.dtype(np.float32): torch.float32,
                                      np.dtype(np.complex128): torch.complex128,
                                      np.dtype(np.complex64): torch.complex64}[X.dtype], pin_memory=True)
    arr = t.numpy()
    arr[...] = flat
    return t, arr.reshape(X.shape, order="F")


def _gemm(h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, copy_back=True):
    h.gemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, pin=False, copy_back=copy_back)
    if not copy_back:
        import torch

        p, cnt = h.full_c()
        assert cnt == m * n
        out = torch.empty(m * n, dtype={np.dtype(np.float64): torch.float64, np.dtype(np.float32): torch.float32,
                                        np.dtype(np.complex128): torch.complex128,
                                        np.dtype(np.complex64): torch.complex64}[C.dtype])
        import ctypes

        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        assert hip.hipMemcpy(out.data_ptr(), p, out.numel() * out.element_size(), 2) == 0  # D2H
        return out.numpy().reshape(m, n, order="F")
    return C[:m, :]


# (m, n, k, tile_m, tile_n, tile_k, streams, coarsen)
CASES = [
    (1000, 1000, 1000, 5000, 5000, 5000, 2, 1),     # CTest square-small
    (1234, 4567, 1357, 5000, 5000, 5000, 2, 1),     # CTest non-square-small
    (1234, 1567, 1357, 500, 400, 300, 3, 0),        # reference tile grid with short edge tiles
    (50, 200, 21, 4, 4, 4, 2, 0),                   # CI test1, reference tiles kept
    (50, 200, 21, 4, 4, 4, 2, 1),                   # CI test1, coarsened
    (5, 2, 2, 4, 4, 4, 2, 0),                       # CI test2
    (2048, 2048, 2048, 512, 512, 512, 2, 1),        # 2048^3 NN configuration
]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("flags", ["NN", "TT", "NT", "TN"])
def test_integer_kat_double(torch_gpu, case, flags):
    m, n, k, tm, tn, tk, ns, coarsen = case
    ta, tb = flags
    rng = np.random.default_rng(42)
    ar, ac = oracle.op_shape(ta, m, k)
    br, bc = oracle.op_shape(tb, k, n)
    A = int_matrix(rng, ar, ac, np.float64)
    B = int_matrix(rng, br, bc, np.float64)
    C = int_matrix(rng, m, n, np.float64)
    ref = C.copy(order="F")
    oracle.tiled_gemm_np(ta, tb, m, n, k, 1.0, A, ar, B, br, 1.0, ref, m, tm, tn, tk)
    with tiledmm.Handle(tiledmm.F64, ns, tm, tn, tk) as h:
        h.set_option(tiledmm.OPT_COARSEN, coarsen)
        for copy_back in (True, False):
            Cw = C.copy(order="F")
            out = _gemm(h, ta, tb, m, n, k, 1.0, A, ar, B, br, 1.0, Cw, m, copy_back=copy_back)
            np.testing.assert_array_equal(out, ref, err_msg=f"copy_back={copy_back}")
        # the reference policy's tile sizes become the handle's current ones
        assert h.tile_sizes() == tuple(oracle.optimal_tile_size(d, 0, t) for d, t in ((m, tm), (n, tn), (k, tk)))


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.complex128, np.complex64])
@pytest.mark.parametrize("pinned", [False, True])
def test_uniform_all_types(torch_gpu, dtype, pinned):
    torch = torch_gpu
    m, n, k = 700, 530, 610
    rng = np.random.default_rng(5)
    cplx = np.issubdtype(dtype, np.complexfloating)
    alpha = (1 + 0.5j) if cplx else 1.0
    beta = (0.5 - 1j) if cplx else 1.0
    wide = np.complex128 if cplx else np.float64
    flag_sets = ["CN", "NC", "TT", "NN"] if cplx else ["TN", "NT", "NN", "TT"]
    for flags in flag_sets:
        ta, tb = flags
        ar, ac = oracle.op_shape(ta, m, k)
        br, bc = oracle.op_shape(tb, k, n)
        A = uniform_matrix(rng, ar, ac, dtype, ld=ar + 7)
        B = uniform_matrix(rng, br, bc, dtype, ld=br + 1)
        C = uniform_matrix(rng, m, n, dtype, ld=m + 3)
        C0 = C[:m, :].copy()
        keep = []
        if pinned:
            tA, A = _pinned_like(torch, A)
            tB, B = _pinned_like(torch, B)
            tC, C = _pinned_like(torch, C)
            keep = [tA, tB, tC]
        with tiledmm.Handle(CODES[dtype], 3, 256, 192, 160) as h:
            h.set_option(tiledmm.OPT_COARSEN, 0)
            h.gemm(ta, tb, m, n, k, alpha, A, ar + 7, B, br + 1, beta, C, m + 3)
            assert h.stats()["all_pinned"] == (1.0 if pinned else 0.0)
        ref = oracle.host_blas_gemm(ta, tb, alpha, A[:ar, :].astype(wide), B[:br, :].astype(wide), beta,
                                    C0.astype(wide))
        scale = oracle.error_scale(ta, tb, alpha, A[:ar, :], B[:br, :], beta, C0)
        err = oracle.normwise_error(C[:m, :], ref, scale)
        assert err <= bound(dtype, k), f"{flags}: {err:.3e}"
        assert np.all(np.isnan(C[m:, :])), "padding rows of C were written"
        del keep


def test_beta_zero_nan_c_and_alpha(torch_gpu):
    """beta == 0: C is not read (NaN stays out); every byte of C is written."""
    m, n, k = 333, 211, 123
    rng = np.random.default_rng(9)
    A = uniform_matrix(rng, m, k, np.float64)
    B = uniform_matrix(rng, k, n, np.float64)
    C = np.full((m, n), np.nan, order="F")
    with tiledmm.Handle(tiledmm.F64, 2, 100, 100, 50) as h:
        h.set_option(tiledmm.OPT_COARSEN, 0)
        h.gemm("N", "N", m, n, k, -2.0, A, m, B, k, 0.0, C, m)
    ref = -2.0 * (A @ B)
    assert not np.any(np.isnan(C))
    err = oracle.normwise_error(C, ref, 2.0 * np.abs(A) @ np.abs(B))
    assert err <= bound(np.float64, k)


def test_k_zero_scales_c(torch_gpu):
    m, n = 77, 55
    C = np.asfortranarray(np.arange(m * n, dtype=np.float64).reshape(m, n))
    A = np.zeros((m, 1), order="F")
    B = np.zeros((1, n), order="F")
    with tiledmm.Handle(tiledmm.F64, 2, 64, 64, 64) as h:
        Cw = C.copy(order="F")
        h.gemm("N", "N", m, n, 0, 1.0, A, m, B, 1, 2.0, Cw, m)
        np.testing.assert_array_equal(Cw, 2.0 * C)


def test_invalid_arguments(torch_gpu):
    with tiledmm.Handle(tiledmm.F64, 2, 64, 64, 64) as h:
        A = np.zeros((10, 10), order="F")
        with pytest.raises(tiledmm.TmmError) as e:
            h.gemm("N", "N", 10, 10, 10, 1.0, A, 9, A, 10, 0.0, A, 10)
        assert e.value.status == 1 and "matrix A" in str(e.value)
        with pytest.raises(tiledmm.TmmError):
            h.gemm("N", "N", -1, 10, 10, 1.0, A, 10, A, 10, 0.0, A, 10)
        with pytest.raises(tiledmm.TmmError):
            h.set_tile_sizes(65, 64, 64)
    with tiledmm.Handle(tiledmm.F32, 1) as h:
        A = np.zeros((4, 4), order="F")
        # a float handle called through the double entry point
        st = tiledmm.load().tmm_dgemm(h.ptr, b"N", b"N", 4, 4, 4, tiledmm._scalar(1, 1.0), A.ctypes.data, 4,
                                      A.ctypes.data, 4, tiledmm._scalar(1, 0.0), A.ctypes.data, 4, 0, 1)
        assert st == 4


@pytest.mark.skipif(not oracle.ref_available(), reason="reference library not built (oracle/_ref)")
@pytest.mark.parametrize("flags,shape,tiles", [
    ("NN", (1234, 1567, 1357), (500, 400, 300)),   # short edge tiles in every dimension
    ("TN", (1200, 1600, 1500), (400, 400, 300)),   # transposed A: tile grid without short k-tiles
    ("NT", (1234, 1567, 1357), (500, 400, 300)),
])
def test_against_reference_library(torch_gpu, flags, shape, tiles):
    """The reference's own gemm (built from its sources, rocBLAS per tile) on the same inputs."""
    m, n, k = shape
    ta, tb = flags
    rng = np.random.default_rng(42)
    ar, ac = oracle.op_shape(ta, m, k)
    br, bc = oracle.op_shape(tb, k, n)
    R = oracle.ref()
    for data in ("int", "uniform"):
        mk = int_matrix if data == "int" else uniform_matrix
        A, B, C = mk(rng, ar, ac, np.float64), mk(rng, br, bc, np.float64), mk(rng, m, n, np.float64)
        alpha, beta = (1.0, 1.0) if data == "int" else (0.5, -1.0)
        for copy_back in (1, 0):
            mine, theirs = C.copy(order="F"), C.copy(order="F")
            assert R.ref_dgemm(ta.encode(), tb.encode(), m, n, k, alpha, A.ctypes.data, ar, B.ctypes.data, br, beta,
                               theirs.ctypes.data, m, 2, *tiles, copy_back) == 0
            with tiledmm.Handle(tiledmm.F64, 2, *tiles) as h:
                out = _gemm(h, ta, tb, m, n, k, alpha, A, ar, B, br, beta, mine, m, copy_back=bool(copy_back))
            if data == "int":
                np.testing.assert_array_equal(out, theirs)
            else:
                ref = oracle.host_blas_gemm(ta, tb, alpha, A, B, beta, C)
                scale = oracle.error_scale(ta, tb, alpha, A, B, beta, C)
                assert oracle.normwise_error(out, ref, scale) <= bound(np.float64, k)
                assert oracle.normwise_error(theirs, ref, scale) <= bound(np.float64, k)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference library not built (oracle/_ref)")
def test_reference_defect_transposed_a_short_k_tile(torch_gpu):
    """Documented deviation: for op(A) = 'T' the reference reads the k-extent of a tile from A's
    tile (m_tile, k_tile) instead of (k_tile, m_tile) (Tiled-MM src/Tiled-MM/tiled_mm.cpp:148-165,
    get_tile_sizes), so a short last k-tile is computed with the wrong k. The library follows BLAS
    semantics (and the oracle); the reference does not on this input."""
    m, n, k = 1234, 1567, 1357
    tiles = (500, 400, 300)
    rng = np.random.default_rng(42)
    A = int_matrix(rng, k, m, np.float64)
    B = int_matrix(rng, k, n, np.float64)
    C = int_matrix(rng, m, n, np.float64)
    mine, theirs, ref = C.copy(order="F"), C.copy(order="F"), C.copy(order="F")
    assert oracle.ref().ref_dgemm(b"T", b"N", m, n, k, 1.0, A.ctypes.data, k, B.ctypes.data, k, 1.0,
                                  theirs.ctypes.data, m, 2, *tiles, 1) == 0
    with tiledmm.Handle(tiledmm.F64, 2, *tiles) as h:
        h.gemm("T", "N", m, n, k, 1.0, A, k, B, k, 1.0, mine, m)
    oracle.tiled_gemm_np("T", "N", m, n, k, 1.0, A, k, B, k, 1.0, ref, m, *tiles)
    np.testing.assert_array_equal(mine, ref)
    np.testing.assert_array_equal(ref, A.T @ B + C)
    assert not np.array_equal(theirs, ref)


@pytest.mark.parametrize("flags", ["NN", "TN", "NT", "TT"])
@pytest.mark.parametrize("pinned,streams", [(True, 1), (False, 2)])
def test_tile_groups_equal_per_tile_launches(torch_gpu, flags, pinned, streams):
    """FP64 host-to-host with the tile GEMMs of the steady k-chunks in grouped launches
    (TMM_OPT_TILE_GROUPS 1), and also the last chunk's in runs of two (2), equals, bit for bit, one
    launch per tile (0); with edge
    tiles in both dimensions, several k-chunks, beta != 0 and both copy-back modes; exact on
    integer data."""
    ta, tb = flags
    m, n, k = 1118, 1034, 1280  # 256-tiles with short edge tiles (94, 10), 5 k-chunks of 256
    rng = np.random.default_rng(31)
    ar, ac = oracle.op_shape(ta, m, k)
    br, bc = oracle.op_shape(tb, k, n)
    A, B = uniform_matrix(rng, ar, ac, np.float64), uniform_matrix(rng, br, bc, np.float64)
    C = uniform_matrix(rng, m, n, np.float64)
    keep = []
    if pinned:
        (tA, A), (tB, B), (tC, C) = _pinned_like(torch_gpu, A), _pinned_like(torch_gpu, B), _pinned_like(torch_gpu, C)
        keep = [tA, tB, tC]
    outs, launches = [], []
    for groups in (2, 1, 0):
        for copy_back in (True, False):
            with tiledmm.Handle(tiledmm.F64, streams, 256, 256, 256) as h:
                h.set_option(tiledmm.OPT_COARSEN, 0)
                h.set_option(tiledmm.OPT_TILE_GROUPS, groups)
                h.set_option(tiledmm.OPT_PROFILE, 1)
                Cw = C.copy(order="F")
                out = _gemm(h, ta, tb, m, n, k, 1.5, A, ar, B, br, -0.5, Cw, m, copy_back=copy_back)
                st = h.stats()
                assert st["k_chunks"] > 2
                launches.append(st["kernel_launches"])
                outs.append(np.array(out))
                # the grouped launches' algorithmic bytes (stats[23]) equal the host planner's
                plan = group_members(ta, tb, m, n, k, ar, br, m, 1, int(copy_back), (256, 256, 256), streams,
                                     max_group=16, pair_max=groups if groups >= 2 else 0) if groups else []
                assert st["group_launches"] == len(plan)
                assert st["group_bytes"] == grouped_algorithmic_bytes(plan, 8)
    # grouped launches happened: short runs in the last chunk (2) < steady chunks only (1) < none (0)
    assert launches[0] < launches[2] < launches[4] and launches[1] < launches[3] < launches[5]
    for o in outs[1:]:
        np.testing.assert_array_equal(outs[0], o)
    Ai, Bi, Ci = int_matrix(rng, ar, ac, np.float64), int_matrix(rng, br, bc, np.float64), int_matrix(rng, m, n, np.float64)
    ref = Ci.copy(order="F")
    oracle.tiled_gemm_c(ta, tb, m, n, k, 2, Ai, ar, Bi, br, 3, ref, m, m, n, k)
    with tiledmm.Handle(tiledmm.F64, streams, 256, 256, 256) as h:
        h.set_option(tiledmm.OPT_COARSEN, 0)
        out = _gemm(h, ta, tb, m, n, k, 2.0, Ai, ar, Bi, br, 3.0, Ci, m)
    np.testing.assert_array_equal(out, ref)
    del keep


@pytest.mark.parametrize("tiles", [(256, 256, 256), (300, 300, 256)])
@pytest.mark.parametrize("dtype,flags", [(np.float32, "NN"), (np.float32, "TN"), (np.float32, "NT"),
                                         (np.float32, "TT"), (np.complex64, "CN"), (np.complex64, "NC"),
                                         (np.complex64, "TT"), (np.complex64, "NN")])
def test_tile_groups_equal_per_tile_launches_single(torch_gpu, dtype, flags, tiles):
    """FP32 and complex FP32 host-to-host: grouped launches (TMM_OPT_TILE_GROUPS 2 and 1) equal one
    launch per tile (0) bit for bit, with fewer launches; with 256-tiles edge tiles of 94 rows and
    10 columns (FP32 groups take only tiles of more than 128 rows and columns), with 300-tiles
    every tile grouped and partial 256 x 256 workgroup tiles inside each member; exact on integer
    data against the oracle."""
    ta, tb = flags
    m, n, k = 1118, 1034, 1280
    rng = np.random.default_rng(41)
    ar, ac = oracle.op_shape(ta, m, k)
    br, bc = oracle.op_shape(tb, k, n)
    cplx = np.issubdtype(dtype, np.complexfloating)
    alpha, beta = (2 - 1j, 1 + 2j) if cplx else (2.0, 3.0)
    A, B, C = int_matrix(rng, ar, ac, dtype), int_matrix(rng, br, bc, dtype), int_matrix(rng, m, n, dtype)
    outs, launches = [], []
    for groups in (2, 1, 0):
        with tiledmm.Handle(CODES[dtype], 2, *tiles) as h:
            h.set_option(tiledmm.OPT_COARSEN, 0)
            h.set_option(tiledmm.OPT_TILE_GROUPS, groups)
            h.set_option(tiledmm.OPT_PROFILE, 1)
            Cw = C.copy(order="F")
            out = _gemm(h, ta, tb, m, n, k, alpha, A, ar, B, br, beta, Cw, m)
            st = h.stats()
            assert st["k_chunks"] > 2
            launches.append(st["kernel_launches"])
            outs.append(np.array(out))
    assert launches[0] < launches[1] < launches[2]
    np.testing.assert_array_equal(outs[0], outs[1])
    np.testing.assert_array_equal(outs[0], outs[2])
    ref = C.copy(order="F")
    oracle.tiled_gemm_c(ta, tb, m, n, k, alpha, A, ar, B, br, beta, ref, m, m, n, k)
    np.testing.assert_array_equal(outs[0], ref)


@pytest.mark.parametrize("flags", ["CN", "NT", "CC", "TC"])
def test_cgemm_pipelined_groups_host_to_host(torch_gpu, flags):
    """Complex FP32 host-to-host with blocks large enough that the grouped launches of the steady
    k-chunks take the software-pipelined kernel (over two workgroup tiles per CU), B op 'N' and
    row-contiguous ('T'/'C'); one launch per tile (the 3-stage ring kernel at this block size) gives
    the same result, exact on integer data against numpy's complex128 product."""
    ta, tb = flags
    m, n, k = 4224, 4160, 1024
    rng = np.random.default_rng(43)
    ar, ac = oracle.op_shape(ta, m, k)
    br, bc = oracle.op_shape(tb, k, n)
    alpha, beta = 2 - 1j, 1 + 2j
    A, B, C = (int_matrix(rng, ar, ac, np.complex64), int_matrix(rng, br, bc, np.complex64),
               int_matrix(rng, m, n, np.complex64))
    A64, B64 = A.astype(np.complex128), B.astype(np.complex128)
    opA = {"N": A64, "T": A64.T, "C": A64.conj().T}[ta]
    opB = {"N": B64, "T": B64.T, "C": B64.conj().T}[tb]
    ref = alpha * (opA @ opB) + beta * C.astype(np.complex128)
    for groups in (2, 0):
        with tiledmm.Handle(tiledmm.C32, 2, 2112, 2112, 256) as h:
            h.set_option(tiledmm.OPT_COARSEN, 0)
            h.set_option(tiledmm.OPT_TILE_GROUPS, groups)
            h.set_option(tiledmm.OPT_PROFILE, 1)
            Cw = C.copy(order="F")
            out = _gemm(h, ta, tb, m, n, k, alpha, A, ar, B, br, beta, Cw, m)
            st = h.stats()
            assert (st["group_launches"] > 0) == (groups > 0)
            np.testing.assert_array_equal(out.astype(np.complex128), ref, err_msg=f"{flags} groups={groups}")


@pytest.mark.parametrize("flags", ["NN", "CN", "NC", "TC", "CT", "TN"])
def test_tile_groups_equal_per_tile_launches_complex(torch_gpu, flags):
    """Complex FP64 host-to-host: grouped launches of the tile GEMMs of the steady k-chunks equal,
    bit for bit, one launch per tile; complex alpha and beta, conjugated operands, edge tiles of
    width 94 and 10 (the kernel's 64-wide workgroup tiles leave partial ones); exact on integer
    data against the oracle."""
    ta, tb = flags
    m, n, k = 1118, 1034, 1280
    rng = np.random.default_rng(37)
    ar, ac = oracle.op_shape(ta, m, k)
    br, bc = oracle.op_shape(tb, k, n)
    A, B = uniform_matrix(rng, ar, ac, np.complex128), uniform_matrix(rng, br, bc, np.complex128)
    C = uniform_matrix(rng, m, n, np.complex128)
    outs, launches = [], []
    for groups in (2, 1, 0):
        with tiledmm.Handle(tiledmm.C64, 2, 256, 256, 256) as h:
            h.set_option(tiledmm.OPT_COARSEN, 0)
            h.set_option(tiledmm.OPT_TILE_GROUPS, groups)
            h.set_option(tiledmm.OPT_PROFILE, 1)
            Cw = C.copy(order="F")
            out = _gemm(h, ta, tb, m, n, k, 1.5 - 0.25j, A, ar, B, br, -0.5 + 2j, Cw, m)
            st = h.stats()
            assert st["k_chunks"] > 2
            launches.append(st["kernel_launches"])
            outs.append(np.array(out))
    assert launches[0] < launches[1] < launches[2]
    np.testing.assert_array_equal(outs[0], outs[1])
    np.testing.assert_array_equal(outs[0], outs[2])
    Ai, Bi = int_matrix(rng, ar, ac, np.complex128), int_matrix(rng, br, bc, np.complex128)
    Ci = int_matrix(rng, m, n, np.complex128)
    ref = Ci.copy(order="F")
    oracle.tiled_gemm_c(ta, tb, m, n, k, 2 - 1j, Ai, ar, Bi, br, 3 + 1j, ref, m, m, n, k)
    with tiledmm.Handle(tiledmm.C64, 2, 256, 256, 256) as h:
        h.set_option(tiledmm.OPT_COARSEN, 0)
        out = _gemm(h, ta, tb, m, n, k, 2 - 1j, Ai, ar, Bi, br, 3 + 1j, Ci, m)
    np.testing.assert_array_equal(out, ref)


@pytest.mark.parametrize("dtype,m,n,k,tile,block", [
    (np.float64, 9000, 8800, 2000, 4500, (4608, 4480)),   # reference tiles 4500 x 4400
    (np.complex128, 6100, 6100, 1000, 3100, (3072, 3072)),  # reference tiles 3050 x 3050
    (np.float32, 9000, 8800, 2000, 4500, (4608, 4608)),   # 256 x 256 FP32 workgroup tiles
])
def test_workgroup_aligned_blocks_equal_reference_tiles(torch_gpu, dtype, m, n, k, tile, block):
    """With coarsening on, the block edges move onto the tile kernel's workgroup grid (same block
    count, each a multiple of the workgroup tile but the last: 128 x 128 FP64, 64 x 64 complex
    FP64, 256 x 256 FP32); where no tiles merge,
    that is the only difference from TMM_OPT_COARSEN 0, and the results are equal bit for bit
    (every element keeps its k order); integer data exact."""
    rng = np.random.default_rng(m + k)
    A = int_matrix(rng, k, m, dtype)  # op(A) = 'T'
    B = int_matrix(rng, k, n, dtype)
    C = int_matrix(rng, m, n, dtype)
    outs, blocks = [], []
    for coarsen in (1, 0):
        with tiledmm.Handle(CODES[dtype], 2, tile, tile, tile) as h:
            h.set_option(tiledmm.OPT_COARSEN, coarsen)
            Cw = C.copy(order="F")
            h.gemm("T", "N", m, n, k, 2.0, A, k, B, k, -1.0, Cw, m)
            st = h.stats()
            blocks.append((int(st["tile_m"]), int(st["tile_n"])))
            outs.append(Cw)
    assert blocks[0] == block and blocks[1] != block, blocks
    np.testing.assert_array_equal(outs[0], outs[1])
    exact = 2.0 * (A.T @ B) - C  # small integers: exact in FP64 (and complex FP64)
    np.testing.assert_array_equal(outs[0], exact)

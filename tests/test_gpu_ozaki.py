"""Ozaki split-precision DGEMM (TMM_OPT_OZAKI_SLICES) against the oracle.

Properties checked:
  * exactness: integers in [-9, 9] have at most 4 significant bits after the power-of-two row /
    column scaling, so one 7-bit slice holds them exactly and every slice product is exact in
    int32: the emulated result must equal exact integer arithmetic bit for bit, for every op
    combination and edge shape (m, n not multiples of the 256 tile, k not a multiple of 64);
  * accuracy: on uniform[-1,1] data the normwise error against the host BLAS falls with the slice
    count, and reaches the north-star bound 1e-15 * || |A||B| || at 7 slices;
  * the out-of-core host-to-host path with Ozaki tile GEMMs (several k-chunks, beta != 0) keeps
    the same exactness and bound.
"""
import numpy as np
import pytest

import oracle
import tiledmm

pytestmark = pytest.mark.gpu


def _dev(torch, X):
    return torch.from_numpy(np.ascontiguousarray(X.reshape(-1, order="F"))).cuda()


def _device_gemm(torch, slices, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    dA, dB, dC = _dev(torch, A), _dev(torch, B), _dev(torch, C)
    with tiledmm.Handle(tiledmm.F64, 1) as h:
        h.set_option(tiledmm.OPT_OZAKI_SLICES, slices)
        h.device_gemm(ta, tb, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc)
        h.synchronize()
        st = h.stats()
    return dC.cpu().numpy().reshape(C.shape, order="F"), st


@pytest.mark.parametrize("shape", [(1, 1, 1), (17, 33, 5), (256, 256, 64), (300, 257, 200), (513, 129, 1000)])
@pytest.mark.parametrize("flags", ["NN", "TN", "NT", "TT"])
def test_integer_exact(torch_gpu, shape, flags):
    m, n, k = shape
    ta, tb = flags
    rng = np.random.default_rng(m + 7 * n + 13 * k)
    ar, ac = oracle.op_shape(ta, m, k)
    br, bc = oracle.op_shape(tb, k, n)
    A = np.asfortranarray(rng.integers(-9, 10, (ar, ac)).astype(np.float64))
    B = np.asfortranarray(rng.integers(-9, 10, (br, bc)).astype(np.float64))
    C = np.asfortranarray(rng.integers(-9, 10, (m, n)).astype(np.float64))
    exact = 2 * (oracle.apply_op(ta, A.astype(np.int64)) @ oracle.apply_op(tb, B.astype(np.int64))) - 3 * C.astype(np.int64)
    for slices in (1, 3, 7):
        out, st = _device_gemm(torch_gpu, slices, ta, tb, m, n, k, 2.0, A, ar, B, br, -3.0, C, m)
        np.testing.assert_array_equal(out, exact.astype(np.float64), err_msg=f"slices={slices}")
        assert st["int8_ops"] == 2.0 * m * n * k * slices * (slices + 1) / 2


def test_error_falls_with_slices_and_meets_bound(torch_gpu):
    m, n, k = 640, 520, 1500
    rng = np.random.default_rng(3)
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
    C = np.zeros((m, n), order="F")
    ref = A @ B
    scale = np.abs(A) @ np.abs(B)
    errs = {}
    for slices in (4, 5, 6, 7, 8):
        out, _ = _device_gemm(torch_gpu, slices, "N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
        errs[slices] = oracle.normwise_error(out, ref, scale)
    assert errs[4] > errs[5] > errs[6] > errs[7], errs
    assert errs[7] <= 1e-15, errs
    assert errs[8] <= 1e-15, errs


@pytest.mark.parametrize("phi", [0.5, 2.0])
def test_wide_range_bound_with_more_slices(torch_gpu, phi):
    """entries (u - 0.5) * exp(phi * g): 9 slices keep the normwise error within 1e-15."""
    m, n, k = 512, 384, 1024
    u = tiledmm.util()
    A = np.empty((m, k), order="F")
    B = np.empty((k, n), order="F")
    u.tmmu_fill_range_d(A.ctypes.data, m, k, m, 0, 0, m, 42, phi, 0)
    u.tmmu_fill_range_d(B.ctypes.data, k, n, k, 0, 0, k, 43, phi, 0)
    C = np.zeros((m, n), order="F")
    out, _ = _device_gemm(torch_gpu, 9, "N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
    err = oracle.normwise_error(out, A @ B, np.abs(A) @ np.abs(B))
    assert err <= 1e-15, err


def test_host_to_host_ozaki_pipeline(torch_gpu):
    """gpu::gemm with Ozaki tile GEMMs: several k-chunks (accumulate) and beta != 0 on the last."""
    m, n, k = 700, 650, 900
    rng = np.random.default_rng(8)
    A = np.asfortranarray(rng.integers(-9, 10, (k, m)).astype(np.float64))  # op(A) = 'T'
    B = np.asfortranarray(rng.integers(-9, 10, (k, n)).astype(np.float64))
    C = np.asfortranarray(rng.integers(-9, 10, (m, n)).astype(np.float64))
    exact = (A.T.astype(np.int64) @ B.astype(np.int64)) + 2 * C.astype(np.int64)
    with tiledmm.Handle(tiledmm.F64, 2, 256, 320, 200) as h:
        h.set_option(tiledmm.OPT_COARSEN, 0)
        h.set_option(tiledmm.OPT_OZAKI_SLICES, 7)
        Cw = C.copy(order="F")
        h.gemm("T", "N", m, n, k, 1.0, A, k, B, k, 2.0, Cw, m)
        st = h.stats()
        assert st["k_chunks"] > 1 and st["int8_ops"] > 0
        np.testing.assert_array_equal(Cw, exact.astype(np.float64))
        # uniform data: within the north-star bound
        Au = np.asfortranarray(rng.uniform(-1, 1, (k, m)))
        Bu = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
        Cu = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
        C0 = Cu.copy(order="F")
        h.gemm("T", "N", m, n, k, 1.0, Au, k, Bu, k, 1.0, Cu, m)
    ref = Au.T @ Bu + C0
    err = oracle.normwise_error(Cu, ref, np.abs(Au.T) @ np.abs(Bu) + np.abs(C0))
    assert err <= 1e-15, err


# ---------------------------------------------------------------------------------------------
# Ozaki scheme II (CRT, TMM_OPT_OZAKI_MODULI)
# ---------------------------------------------------------------------------------------------
def _device_gemm_crt(torch, nmod, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    dA, dB, dC = _dev(torch, A), _dev(torch, B), _dev(torch, C)
    with tiledmm.Handle(tiledmm.F64, 1) as h:
        h.set_option(tiledmm.OPT_OZAKI_MODULI, nmod)
        h.device_gemm(ta, tb, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc)
        h.synchronize()
        st = h.stats()
    return dC.cpu().numpy().reshape(C.shape, order="F"), st


@pytest.mark.parametrize("shape", [(1, 1, 1), (17, 33, 5), (256, 256, 64), (300, 257, 200), (513, 129, 1000)])
@pytest.mark.parametrize("flags", ["NN", "TN", "NT", "TT"])
def test_crt_integer_exact(torch_gpu, shape, flags):
    """integers in [-9, 9] scale to integers exactly, so the CRT result is exact."""
    m, n, k = shape
    ta, tb = flags
    rng = np.random.default_rng(m + 5 * n + 11 * k)
    ar, ac = oracle.op_shape(ta, m, k)
    br, bc = oracle.op_shape(tb, k, n)
    A = np.asfortranarray(rng.integers(-9, 10, (ar, ac)).astype(np.float64))
    B = np.asfortranarray(rng.integers(-9, 10, (br, bc)).astype(np.float64))
    C = np.asfortranarray(rng.integers(-9, 10, (m, n)).astype(np.float64))
    exact = 2 * (oracle.apply_op(ta, A.astype(np.int64)) @ oracle.apply_op(tb, B.astype(np.int64))) - 3 * C.astype(np.int64)
    for nmod in (8, 13, 20):
        out, st = _device_gemm_crt(torch_gpu, nmod, ta, tb, m, n, k, 2.0, A, ar, B, br, -3.0, C, m)
        np.testing.assert_array_equal(out, exact.astype(np.float64), err_msg=f"nmod={nmod}")
        assert st["int8_ops"] == 2.0 * m * n * k * nmod


@pytest.mark.parametrize("nmod", [6, 8, 12, 16, 20])
def test_crt_matches_exact_model(torch_gpu, nmod):
    """Wide-range data: the device result equals, bit for bit, the exact-integer model of the
    same scaling, residues and CRT sum in FP64 limbs (tests/crt_model.py); 6 and 12 moduli have
    an even limb count (2, 4), 8, 16 and 20 an odd one (3, 5, 7)."""
    import crt_model

    m, n, k = 19, 13, 37
    rng = np.random.default_rng(40 + nmod)
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)) * np.exp(2.0 * rng.standard_normal((m, k))))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)) * np.exp(2.0 * rng.standard_normal((k, n))))
    A[3, :] = 0.0  # a row without nonzero entries
    C = np.zeros((m, n), order="F")
    out, _ = _device_gemm_crt(torch_gpu, nmod, "N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
    model, *_ = crt_model.emulate(A, B, nmod)
    np.testing.assert_array_equal(out, model)


def test_crt_error_falls_with_moduli(torch_gpu):
    m, n, k = 640, 520, 1500
    rng = np.random.default_rng(3)
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
    C = np.zeros((m, n), order="F")
    ref = A @ B
    scale = np.abs(A) @ np.abs(B)
    errs = {}
    for nmod in (10, 12, 14, 16, 18):
        out, _ = _device_gemm_crt(torch_gpu, nmod, "N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
        errs[nmod] = oracle.normwise_error(out, ref, scale)
    assert errs[10] > errs[12] > errs[14] > errs[16], errs
    assert errs[16] <= 1e-15, errs
    assert errs[18] <= 1e-16, errs


@pytest.mark.parametrize("phi", [0.5, 2.0])
def test_crt_wide_range_bound(torch_gpu, phi):
    """entries (u - 0.5) * exp(phi * g): 18 moduli keep the normwise error within 1e-15."""
    m, n, k = 512, 384, 1024
    u = tiledmm.util()
    A = np.empty((m, k), order="F")
    B = np.empty((k, n), order="F")
    u.tmmu_fill_range_d(A.ctypes.data, m, k, m, 0, 0, m, 42, phi, 0)
    u.tmmu_fill_range_d(B.ctypes.data, k, n, k, 0, 0, k, 43, phi, 0)
    C = np.zeros((m, n), order="F")
    out, _ = _device_gemm_crt(torch_gpu, 18, "N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
    err = oracle.normwise_error(out, A @ B, np.abs(A) @ np.abs(B))
    assert err <= 1e-15, err


def test_host_to_host_crt_pipeline(torch_gpu):
    """gpu::gemm with CRT tile GEMMs: several k-chunks (accumulate) and beta != 0 on the last."""
    m, n, k = 700, 650, 900
    rng = np.random.default_rng(9)
    A = np.asfortranarray(rng.integers(-9, 10, (k, m)).astype(np.float64))  # op(A) = 'T'
    B = np.asfortranarray(rng.integers(-9, 10, (k, n)).astype(np.float64))
    C = np.asfortranarray(rng.integers(-9, 10, (m, n)).astype(np.float64))
    exact = (A.T.astype(np.int64) @ B.astype(np.int64)) + 2 * C.astype(np.int64)
    with tiledmm.Handle(tiledmm.F64, 2, 256, 320, 200) as h:
        h.set_option(tiledmm.OPT_COARSEN, 0)
        h.set_option(tiledmm.OPT_OZAKI_MODULI, 16)
        Cw = C.copy(order="F")
        h.gemm("T", "N", m, n, k, 1.0, A, k, B, k, 2.0, Cw, m)
        st = h.stats()
        assert st["k_chunks"] > 1 and st["int8_ops"] > 0
        np.testing.assert_array_equal(Cw, exact.astype(np.float64))
        Au = np.asfortranarray(rng.uniform(-1, 1, (k, m)))
        Bu = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
        Cu = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
        C0 = Cu.copy(order="F")
        h.gemm("T", "N", m, n, k, 1.0, Au, k, Bu, k, 1.0, Cu, m)
    ref = Au.T @ Bu + C0
    err = oracle.normwise_error(Cu, ref, np.abs(Au.T) @ np.abs(Bu) + np.abs(C0))
    assert err <= 1e-15, err


@pytest.mark.parametrize("flags", ["NN", "TN", "NT", "TT"])
@pytest.mark.parametrize("nq", [1, 2])
def test_host_to_host_crt_operand_reuse(torch_gpu, flags, nq):
    """Scheme II host-to-host with shared A / B blocks prepared once per call (the default) equals,
    bit for bit, the same call preparing them per tile product, on wide-range data, for all
    layouts, several k-chunks, edge tiles and beta != 0; and is exact on integer data."""
    ta, tb = flags
    m, n, k = 610, 530, 700
    rng = np.random.default_rng(17 + nq)
    ar, ac = oracle.op_shape(ta, m, k)
    br, bc = oracle.op_shape(tb, k, n)
    A = np.asfortranarray((rng.uniform(-0.5, 0.5, (ar, ac)) * np.exp(2 * rng.standard_normal((ar, ac)))))
    B = np.asfortranarray((rng.uniform(-0.5, 0.5, (br, bc)) * np.exp(2 * rng.standard_normal((br, bc)))))
    C = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    outs = []
    for reuse in (1, 0):
        with tiledmm.Handle(tiledmm.F64, nq, 256, 200, 250) as h:
            h.set_option(tiledmm.OPT_COARSEN, 0)
            h.set_option(tiledmm.OPT_OZAKI_MODULI, 16)
            h.set_option(tiledmm.OPT_OZAKI_REUSE, reuse)
            Cw = C.copy(order="F")
            h.gemm(ta, tb, m, n, k, 1.5, A, ar, B, br, -0.5, Cw, m)
            assert h.stats()["k_chunks"] > 1
            outs.append(Cw)
    np.testing.assert_array_equal(outs[0], outs[1])
    ref = 1.5 * (oracle.apply_op(ta, A) @ oracle.apply_op(tb, B)) - 0.5 * C
    scale = 1.5 * (np.abs(oracle.apply_op(ta, A)) @ np.abs(oracle.apply_op(tb, B))) + 0.5 * np.abs(C)
    assert oracle.normwise_error(outs[0], ref, scale) <= 1e-13
    Ai = np.asfortranarray(rng.integers(-9, 10, (ar, ac)).astype(np.float64))
    Bi = np.asfortranarray(rng.integers(-9, 10, (br, bc)).astype(np.float64))
    Ci = np.asfortranarray(rng.integers(-9, 10, (m, n)).astype(np.float64))
    exact = 2 * (oracle.apply_op(ta, Ai.astype(np.int64)) @ oracle.apply_op(tb, Bi.astype(np.int64))) - 3 * Ci.astype(np.int64)
    with tiledmm.Handle(tiledmm.F64, nq, 256, 200, 250) as h:
        h.set_option(tiledmm.OPT_COARSEN, 0)
        h.set_option(tiledmm.OPT_OZAKI_MODULI, 16)
        h.gemm(ta, tb, m, n, k, 2.0, Ai, ar, Bi, br, -3.0, Ci, m)
    np.testing.assert_array_equal(Ci, exact.astype(np.float64))


def test_host_to_host_crt_merged_k_chunks(torch_gpu):
    """With coarsening on (the default), a scheme II host-to-host call merges the k-chunks of the
    reference tile policy (each costs a reconstruction pass per C tile) into a short first chunk
    and one long one: exact on integer data, reuse on / off equal bit for bit."""
    m, n, k = 600, 520, 5000
    rng = np.random.default_rng(23)
    A = np.asfortranarray(rng.integers(-9, 10, (k, m)).astype(np.float64))  # op(A) = 'T'
    B = np.asfortranarray(rng.integers(-9, 10, (k, n)).astype(np.float64))
    C = np.asfortranarray(rng.integers(-9, 10, (m, n)).astype(np.float64))
    exact = 3 * (A.T.astype(np.int64) @ B.astype(np.int64)) - C.astype(np.int64)
    outs = []
    for reuse in (1, 0):
        with tiledmm.Handle(tiledmm.F64, 2, 256, 256, 512) as h:
            h.set_option(tiledmm.OPT_OZAKI_MODULI, 16)
            h.set_option(tiledmm.OPT_OZAKI_REUSE, reuse)
            Cw = C.copy(order="F")
            h.gemm("T", "N", m, n, k, 3.0, A, k, B, k, -1.0, Cw, m)
            assert h.stats()["k_chunks"] == 2
            outs.append(Cw)
        with tiledmm.Handle(tiledmm.F64, 2, 256, 256, 512) as h:  # native: the reference chunks
            Cn = C.copy(order="F")
            h.gemm("T", "N", m, n, k, 3.0, A, k, B, k, -1.0, Cn, m)
            assert h.stats()["k_chunks"] > 2
    np.testing.assert_array_equal(outs[0], exact.astype(np.float64))
    np.testing.assert_array_equal(outs[0], outs[1])


@pytest.mark.parametrize("nq,copy_back,pool_mb", [(1, True, 0), (2, True, 0), (1, False, 0), (2, True, 900)])
def test_host_to_host_crt_deferred_c_copies(torch_gpu, nq, copy_back, pool_mb):
    """Pinned host memory: the last chunk's C tile copies go after later tiles' A / B copies and
    each tile's products start before its C arrives (TMM_OPT_DEFER_CIN 1 and 2, 0 = off); the
    results are equal bit for bit, exact on integer data, also on a call whose uniform data have
    rounding in every product; with C copied back or left in device memory, and with a work pool
    too small for all blocks at once (tile groups, A / B re-sent per group)."""
    import ctypes
    torch = torch_gpu
    m, n, k = 8400, 8400, 4200  # 2 x 3 blocks after coarsening (4352 / 4048 x 2816 / 2768), 2 merged k-chunks
    rng = np.random.default_rng(29 + nq)

    def pinned(X):
        t = torch.empty(X.size, dtype=torch.float64, pin_memory=True)
        a = t.numpy().reshape(X.shape, order="F")
        a[...] = X
        return t, a

    for integer in (True, False):
        if integer:
            A0 = np.asfortranarray(rng.integers(-9, 10, (k, m)).astype(np.float64))  # op(A) = 'T'
            B0 = np.asfortranarray(rng.integers(-9, 10, (k, n)).astype(np.float64))
            C0 = np.asfortranarray(rng.integers(-9, 10, (m, n)).astype(np.float64))
        else:
            A0, B0, C0 = (np.asfortranarray(rng.uniform(-1, 1, s)) for s in ((k, m), (k, n), (m, n)))
        (tA, A), (tB, B) = pinned(A0), pinned(B0)
        outs = []
        for defer in (0, 1, 2):
            tC, Cw = pinned(C0)
            with tiledmm.Handle(tiledmm.F64, nq, 2048, 2048, 2048) as h:
                h.set_option(tiledmm.OPT_OZAKI_MODULI, 16)
                h.set_option(tiledmm.OPT_DEFER_CIN, defer)
                if pool_mb:
                    h.set_option(tiledmm.OPT_POOL_LIMIT_MB, pool_mb)
                h.gemm("T", "N", m, n, k, 3.0, A, k, B, k, -1.0, Cw, m, pin=False, copy_back=copy_back)
                st = h.stats()
                assert (st["k_chunks"] >= 2 if pool_mb else st["k_chunks"] == 2) and st["all_pinned"] == 1.0
                assert st["tile_m"] < m and st["tile_n"] < n, st
                if pool_mb:
                    assert st["h2d_bytes"] > 8.0 * (2 * m * k + m * n), st  # A / B sent once per group
                if copy_back:
                    outs.append(Cw.copy())
                else:
                    p, cnt = h.full_c()
                    assert cnt == m * n
                    got = torch.empty(m * n, dtype=torch.float64)
                    hip = ctypes.CDLL("libamdhip64.so")
                    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
                    assert hip.hipMemcpy(got.data_ptr(), p, m * n * 8, 2) == 0  # D2H
                    outs.append(got.numpy().reshape(m, n, order="F").copy())
            del tC
        if integer:
            exact = 3.0 * (A0.T @ B0) - C0  # sums of products of integers below 2^53: exact in FP64
            np.testing.assert_array_equal(outs[0], exact)
        np.testing.assert_array_equal(outs[1], outs[0])
        np.testing.assert_array_equal(outs[2], outs[0])
        del tA, tB


@pytest.mark.parametrize("shape", [(8400, 260, 8500), (300, 8300, 200)])
def test_crt_large_integer_exact(torch_gpu, shape):
    """Scheme II on larger shapes (many 256-row tiles, long k): exact on integer data."""
    m, n, k = shape
    rng = np.random.default_rng(m + n + k)
    A = np.asfortranarray(rng.integers(-9, 10, (k, m)).astype(np.float64))  # op(A) = 'T'
    B = np.asfortranarray(rng.integers(-9, 10, (k, n)).astype(np.float64))
    C = np.asfortranarray(rng.integers(-9, 10, (m, n)).astype(np.float64))
    exact = A.T @ B + 2.0 * C  # products and sums of small integers: exact in FP64
    out, st = _device_gemm_crt(torch_gpu, 16, "T", "N", m, n, k, 1.0, A, k, B, k, 2.0, C, m)
    np.testing.assert_array_equal(out, exact)
    assert st["int8_ops"] == 2.0 * m * n * k * 16

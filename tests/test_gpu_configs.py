"""Full-size GPU tests of the workloads the path is measured on, and of the reference's own large
test cases.

* DGEMM TN 40000^3, leading dimensions 40064 with NaN in the padding rows,
  make_context(4, 8192, 8192, 8192): pinned and pageable host memory, copy_c_back true and false,
  and the Ozaki scheme II call (16 moduli) on the same pinned data. 64 sampled columns of C
  against the host BLAS: normwise error within the stated bound 2 k u, the emulation's error not
  above native's.
* ZGEMM CN 24000 x 24000 x 32000, alpha = (1, 0.5), beta = (0.5, -1), make_context(3, 6144^3),
  both copy_c_back modes, sampled columns.
* Ozaki scheme II DGEMM NN 40000^3 on HBM-resident data with entries (u - 0.5) exp(phi g),
  phi = 0.5, 2, 4 at 16, 17, 18 moduli: normwise error <= 1e-15 and not above native FP64's; the
  componentwise error is printed beside it.
* The 512-problem batch (m, n, k log-uniform in [64, 4096], seed 42, alpha = beta = 1): every
  problem against the host BLAS.
* Tiled-MM tests/CMakeLists.txt:12,14: 10000^3 and 12345 x 23456 x 67891 with the reference test's
  integer data (tests/test-multiply.cpp:58-66), exact. Native through the C++ driver
  (tests/cpp/test_multiply.cpp: the reference's flags, default tiles 5000 and 2 streams, both
  copy_c_back modes); the Ozaki scheme II path host-to-host and on device data, where
  k = 67891 > 65536 runs the emulation in two k-passes.

Large host matrices live in hipHostMalloc'd or numpy memory that is released at the end of each
test, so the suite's host memory stays near one workload's (<= 40 GB).
"""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

import tiledmm

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(tiledmm.LIB_DIR, "test-multiply")
U = 2.0 ** -53


def say(*a):
    print("[configs]", *a, flush=True)


def _hip():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


def device_columns(ptr, ld, rows, cols, dtype=np.float64):
    """Columns `cols` (rows entries each) of a column-major device matrix with leading dimension ld."""
    hip = _hip()
    es = np.dtype(dtype).itemsize
    out = np.empty((rows, len(cols)), dtype=dtype, order="F")
    for q, j in enumerate(cols):
        assert hip.hipMemcpy(out[:, q].ctypes.data, ptr + int(j) * ld * es, rows * es, 2) == 0  # D2H
    return out


def sampled_ref(ta, A, Bcols, m, k, C0cols=None, alpha=1.0, beta=0.0, chunk=4096):
    """alpha op(A) Bcols + beta C0cols and the error scale |alpha||op(A)||Bcols| + |beta||C0cols|, with
    op(A) in row chunks of the m x k operand (A stored m x k for 'N', k x m for 'T'/'C')."""
    dt = np.result_type(A.dtype, Bcols.dtype)
    ref = np.empty((m, Bcols.shape[1]), dtype=dt)
    scale = np.empty((m, Bcols.shape[1]))
    absB = np.abs(Bcols)
    for i0 in range(0, m, chunk):
        i1 = min(m, i0 + chunk)
        blk = A[i0:i1, :k] if ta == "N" else (A[:k, i0:i1].T if ta == "T" else A[:k, i0:i1].conj().T)
        ref[i0:i1] = blk @ Bcols
        scale[i0:i1] = np.abs(blk) @ absB
    ref *= alpha
    scale *= abs(alpha)
    if C0cols is not None and beta != 0:
        ref += beta * C0cols
        scale += abs(beta) * np.abs(C0cols)
    return ref, scale


def errors(got, ref, scale):
    nw = float(np.linalg.norm(got - ref) / np.linalg.norm(scale))
    cw = float(np.max(np.abs(got - ref) / scale))
    return nw, cw


# ------------------------------------------------------------------------------------------------
# DGEMM TN 40000^3, ld 40064, make_context(4, 8192^3)
# ------------------------------------------------------------------------------------------------
def _tn_problem(size, ld, pinned):
    mats = []
    for seed in (42, 43, 44):
        M = tiledmm.pinned_empty((ld, size)) if pinned else np.empty((ld, size), order="F")
        tiledmm.fill_uniform(M, seed, -1.0, 1.0, rows=size, cols=size, ld=ld)
        M[size:, :] = np.nan  # padding rows must never be read
        mats.append(M)
    return mats


@pytest.mark.parametrize("pinned", [True, False], ids=["pinned", "pageable"])
def test_dgemm_tn_40000(torch_gpu, pinned):
    size, ld = 40000, 40064
    cols = np.sort(np.random.default_rng(3).choice(size, 64, replace=False))
    say(f"DGEMM TN {size}^3 ld {ld}, {'pinned' if pinned else 'pageable'} host memory: filling")
    A, B, C = _tn_problem(size, ld, pinned)
    Bc = np.ascontiguousarray(B[:size, cols])
    bound = 2 * size * U
    errs = {}
    with tiledmm.Handle(tiledmm.F64, 4, 8192, 8192, 8192) as h:
        # copy_c_back = true: C <- A^T B + C0
        C0 = np.ascontiguousarray(C[:size, cols])
        say("copy_c_back=true")
        h.gemm("T", "N", size, size, size, 1.0, A, ld, B, ld, 1.0, C, ld, copy_back=True)
        assert h.tile_sizes() == (8000, 8000, 8000)  # the reference policy for (40000, 8192)
        assert h.stats()["all_pinned"] == (1.0 if pinned else 0.0)
        ref, scale = sampled_ref("T", A, Bc, size, size, C0, 1.0, 1.0)
        errs["copy_back"] = errors(C[:size, cols], ref, scale)
        assert np.all(np.isnan(C[size:, :8])), "padding rows of C were written"
        # copy_c_back = false on the updated C: the resident device C (ld = m) holds A^T B + C1
        C1 = np.ascontiguousarray(C[:size, cols])
        say("copy_c_back=false")
        h.gemm("T", "N", size, size, size, 1.0, A, ld, B, ld, 1.0, C, ld, copy_back=False)
        np.testing.assert_array_equal(C[:size, cols], C1)  # host C untouched
        p, cnt = h.full_c()
        assert cnt == size * size
        got = device_columns(p, size, size, cols)
        ref, scale = sampled_ref("T", A, Bc, size, size, C1, 1.0, 1.0)
        errs["resident"] = errors(got, ref, scale)
    if pinned:
        # Ozaki scheme II on the same data: error not above native's (the metric's equal error)
        C1 = np.ascontiguousarray(C[:size, cols])
        with tiledmm.Handle(tiledmm.F64, 4, 8192, 8192, 8192) as h:
            h.set_option(tiledmm.OPT_OZAKI_MODULI, 16)
            say("Ozaki scheme II, 16 moduli")
            h.gemm("T", "N", size, size, size, 1.0, A, ld, B, ld, 1.0, C, ld, copy_back=True)
        ref, scale = sampled_ref("T", A, Bc, size, size, C1, 1.0, 1.0)
        errs["ozaki16"] = errors(C[:size, cols], ref, scale)
    say(f"normwise / componentwise errors: {errs} (bound {bound:.2e})")
    for key, (nw, _) in errs.items():
        assert nw <= bound, (key, nw)
    if pinned:
        assert errs["ozaki16"][0] <= errs["copy_back"][0], errs
    del A, B, C


# ------------------------------------------------------------------------------------------------
# ZGEMM CN 24000 x 24000 x 32000, make_context(3, 6144^3)
# ------------------------------------------------------------------------------------------------
def test_zgemm_cn_24000x24000x32000(torch_gpu):
    m, n, k = 24000, 24000, 32000
    alpha, beta = complex(1, 0.5), complex(0.5, -1)
    say(f"ZGEMM CN {m}x{n}x{k}: filling")
    mats = []
    for seed, rows, ncols in ((42, k, m), (43, k, n), (44, m, n)):
        M = tiledmm.pinned_empty((rows, ncols), np.complex128)
        tiledmm.util().tmmu_fill_uniform_z(M.ctypes.data, rows, ncols, rows, 0, 0, rows, seed, -1.0, 1.0, 0)
        mats.append(M)
    A, B, C = mats
    cols = np.sort(np.random.default_rng(4).choice(n, 48, replace=False))
    Bc = np.ascontiguousarray(B[:, cols])
    bound = 2 * k * U
    with tiledmm.Handle(tiledmm.C64, 3, 6144, 6144, 6144) as h:
        C0 = np.ascontiguousarray(C[:, cols])
        say("copy_c_back=true")
        h.gemm("C", "N", m, n, k, alpha, A, k, B, k, beta, C, m, copy_back=True)
        assert h.tile_sizes() == (6000, 6000, 4000)
        ref, scale = sampled_ref("C", A, Bc, m, k, C0, alpha, beta)
        e1 = errors(C[:, cols], ref, scale)
        C1 = np.ascontiguousarray(C[:, cols])
        say("copy_c_back=false")
        h.gemm("C", "N", m, n, k, alpha, A, k, B, k, beta, C, m, copy_back=False)
        p, cnt = h.full_c()
        assert cnt == m * n
        got = device_columns(p, m, m, cols, np.complex128)
        ref, scale = sampled_ref("C", A, Bc, m, k, C1, alpha, beta)
        e2 = errors(got, ref, scale)
    say(f"normwise / componentwise errors: copy back {e1}, resident {e2} (bound {bound:.2e})")
    assert e1[0] <= bound and e2[0] <= bound
    del A, B, C, mats


# ------------------------------------------------------------------------------------------------
# Ozaki scheme II NN 40000^3, entries (u - 0.5) exp(phi g)
# ------------------------------------------------------------------------------------------------
def test_ozaki_nn_40000_phi(torch_gpu):
    torch = torch_gpu
    size = 40000
    cols = np.arange(0, size, size // 32)
    u = tiledmm.util()
    dA = torch.empty(size * size, dtype=torch.float64, device="cuda")
    dB = torch.empty(size * size, dtype=torch.float64, device="cuda")
    dC = torch.empty(size * size, dtype=torch.float64, device="cuda")
    host = np.empty((size, size), order="F")
    stream = torch.cuda.Stream()
    report = []
    for phi, moduli in ((0.5, 16), (2.0, 17), (4.0, 18)):
        say(f"phi {phi}: filling")
        u.tmmu_fill_range_d(host.ctypes.data, size, size, size, 0, 0, size, 43, phi, 0)
        Bc = np.ascontiguousarray(host[:, cols])
        dB.copy_(torch.from_numpy(host.reshape(-1, order="F")))
        u.tmmu_fill_range_d(host.ctypes.data, size, size, size, 0, 0, size, 42, phi, 0)
        dA.copy_(torch.from_numpy(host.reshape(-1, order="F")))
        ref, scale = sampled_ref("N", host, Bc, size, size)
        res = {}
        for mode, count in (("native", 0), ("crt", moduli)):
            dC.fill_(float("nan"))  # beta = 0: C must not be read
            torch.cuda.synchronize()  # the fill runs on torch's stream, the product on another
            with tiledmm.Handle(tiledmm.F64, 1) as h:
                h.set_option(tiledmm.OPT_OZAKI_MODULI, count)
                say(f"phi {phi}: {mode} {count}")
                h.device_gemm("N", "N", size, size, size, 1.0, dA, size, dB, size, 0.0, dC, size,
                              stream=stream.cuda_stream)
                stream.synchronize()
            got = dC.view(size, size)[torch.as_tensor(cols, device="cuda")].t().cpu().numpy()
            res[mode] = errors(got, ref, scale)
        report.append((phi, moduli, res))
        say(f"phi {phi}: normwise / componentwise: native {res['native']}, {moduli} moduli {res['crt']}")
    del dA, dB, dC
    torch.cuda.empty_cache()
    for phi, moduli, res in report:
        assert res["crt"][0] <= 1e-15, (phi, moduli, res)
        assert res["crt"][0] <= res["native"][0], (phi, moduli, res)


# ------------------------------------------------------------------------------------------------
# 512 uneven problems, m, n, k log-uniform in [64, 4096] (seed 42)
# ------------------------------------------------------------------------------------------------
def test_batched_512(torch_gpu):
    rng = np.random.default_rng(42)
    dims = np.round(np.exp(rng.uniform(np.log(64), np.log(4096), size=(512, 3)))).astype(int)
    say("batched 512: filling")
    As, Bs, Cs, C0s = [], [], [], []
    for i, (m, n, k) in enumerate(dims):
        for rows, ncols, lst, seed in ((m, k, As, 3 * i), (k, n, Bs, 3 * i + 1), (m, n, Cs, 3 * i + 2)):
            M = tiledmm.pinned_empty((int(rows), int(ncols)))
            tiledmm.fill_uniform(M, 1000 + seed)
            lst.append(M)
        C0s.append(Cs[-1].copy(order="F"))
    ms_, ns_, ks_ = ([int(x) for x in dims[:, j]] for j in range(3))
    with tiledmm.Handle(tiledmm.F64, 2) as h:
        say("batched 512: call")
        h.gemm_batched("N", "N", ms_, ns_, ks_, [1.0] * 512, As, ms_, Bs, ks_, [1.0] * 512, Cs, ms_)
        cfg = h.batch_config()
    worst = 0.0
    for i, (m, n, k) in enumerate(dims):
        ref = As[i] @ Bs[i] + C0s[i]
        scale = np.abs(As[i]) @ np.abs(Bs[i]) + np.abs(C0s[i])
        err = float(np.linalg.norm(Cs[i] - ref) / np.linalg.norm(scale))
        assert err <= 2 * k * U, (i, m, n, k, err)
        worst = max(worst, err / (2 * k * U))
    say(f"batched 512: all problems within the bound (worst {worst:.2e} of it), autotuner {cfg}")


# ------------------------------------------------------------------------------------------------
# The reference's large CTest cases (Tiled-MM tests/CMakeLists.txt:12,14)
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("args", ["-m 10000 -n 10000 -k 10000",
                                  "-m 12345 -n 23456 -k 67891 --check_cols 64"])
def test_reference_large_ctest_driver(torch_gpu, args):
    assert os.path.exists(DRIVER), "build the library first (make -C tiled-mm_amd)"
    say(f"test-multiply {args}")
    r = subprocess.run([DRIVER] + args.split(), capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:], r.stderr[-2000:], file=sys.stderr)
    assert r.returncode == 0 and "The result is CORRECT" in r.stdout, r.stdout + r.stderr


@pytest.mark.parametrize("shape", [(10000, 10000, 10000), (12345, 23456, 67891)], ids=["10000^3", "12345x23456x67891"])
def test_reference_large_ozaki_exact(torch_gpu, shape):
    """Integer data in [0, 9] (the reference test's), NN, alpha = 1, beta = 0: the Ozaki scheme II
    result is exact, host-to-host (make_context(2, 5000^3) defaults) and on device data (one
    tmm_device_gemm; at k = 67891 two k-passes of at most 65536)."""
    torch = torch_gpu
    m, n, k = shape
    say(f"Ozaki exact {m}x{n}x{k}: filling")
    A = tiledmm.pinned_empty((m, k))
    B = tiledmm.pinned_empty((k, n))
    C = tiledmm.pinned_empty((m, n))
    tiledmm.fill_int(A, 1)
    tiledmm.fill_int(B, 2)
    C[...] = np.nan  # beta = 0: never read
    cols = np.unique(np.linspace(0, n - 1, 64).astype(int))
    Bc = np.ascontiguousarray(B[:, cols])
    ref = np.empty((m, len(cols)))
    for i0 in range(0, m, 4096):  # integer dot products below 2^53: exact in any order
        ref[i0:i0 + 4096] = A[i0:i0 + 4096] @ Bc
    with tiledmm.Handle(tiledmm.F64, 2, 5000, 5000, 5000) as h:
        h.set_option(tiledmm.OPT_OZAKI_MODULI, 16)
        say("host-to-host")
        h.gemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
        np.testing.assert_array_equal(C[:, cols], ref)
    dA = torch.from_numpy(A.reshape(-1, order="F")).cuda()
    dB = torch.from_numpy(B.reshape(-1, order="F")).cuda()
    dC = torch.full((m * n,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()  # the product runs on the handle's stream
    with tiledmm.Handle(tiledmm.F64, 1) as h:
        h.set_option(tiledmm.OPT_OZAKI_MODULI, 16)
        say("device data")
        h.device_gemm("N", "N", m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m)
        h.synchronize()
        assert h.stats()["int8_ops"] == 2.0 * 16 * m * n * k  # 16 products over every k (in one or two passes)
    got = dC.view(n, m)[torch.as_tensor(cols, device="cuda")].t().cpu().numpy()
    np.testing.assert_array_equal(got, ref)
    del dA, dB, dC, A, B, C
    torch.cuda.empty_cache()

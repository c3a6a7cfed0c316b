"""tmm_device_gemm_noalloc: the device GEMM entry point that allocates nothing (an Ozaki emulation
works in the library's fixed per-device scratch instead of the handle's workspace).

Checked through the C ABI against tmm_device_gemm on the same inputs, bit for bit (the same kernels
run on the same data): two calls in a row, the second one smaller (the scratch keeps the first
call's contents, which must not leak into the second), for Ozaki scheme II, the slice scheme and
native FP64; a product too large for the scratch is refused with TMM_ERR_INVALID_ARG before any
launch, and the handle still works after it.
"""
import ctypes

import numpy as np
import pytest

import tiledmm

pytestmark = pytest.mark.gpu

SHAPES = [(1536, 1280, 2048, "NN"), (700, 513, 900, "TN")]  # second call smaller
MODES = [("crt16", tiledmm.OPT_OZAKI_MODULI, 16), ("slices6", tiledmm.OPT_OZAKI_SLICES, 6), ("native", None, 0)]


def _call(lib, h, fn, ta, tb, m, n, k, A, lda, B, ldb, C, ldc, stream):
    al, be = ctypes.c_double(1.5), ctypes.c_double(-0.5)
    return getattr(lib, fn)(h.ptr, ta.encode(), tb.encode(), m, n, k, ctypes.byref(al), A.data_ptr(), lda,
                            B.data_ptr(), ldb, ctypes.byref(be), C.data_ptr(), ldc, stream)


@pytest.mark.parametrize("name,opt,val", MODES, ids=[m[0] for m in MODES])
def test_noalloc_matches_device_gemm(torch_gpu, name, opt, val):
    torch = torch_gpu
    lib = tiledmm.load()
    rng = np.random.default_rng(31)
    stream = torch.cuda.Stream()
    with tiledmm.Handle(tiledmm.F64, 1) as h:
        if opt is not None:
            h.set_option(opt, val)
        for m, n, k, flags in SHAPES:
            ta, tb = flags
            ar, ac = (m, k) if ta == "N" else (k, m)
            dA = torch.from_numpy(rng.uniform(-1, 1, ar * ac)).cuda()
            dB = torch.from_numpy(rng.uniform(-1, 1, k * n)).cuda()
            C0 = torch.from_numpy(rng.uniform(-1, 1, m * n)).cuda()
            ref, got = C0.clone(), C0.clone()
            torch.cuda.synchronize()
            print(f"[noalloc] {name} {m}x{n}x{k} {flags}", flush=True)
            assert _call(lib, h, "tmm_device_gemm", ta, tb, m, n, k, dA, ar, dB, k, ref, m, stream.cuda_stream) == 0
            stream.synchronize()
            st = _call(lib, h, "tmm_device_gemm_noalloc", ta, tb, m, n, k, dA, ar, dB, k, got, m, stream.cuda_stream)
            assert st == 0, lib.tmm_last_error(h.ptr).decode()
            stream.synchronize()
            assert torch.equal(got, ref), f"{name} {m}x{n}x{k}: noalloc differs from tmm_device_gemm"
            assert torch.isfinite(got).all()


def test_noalloc_refuses_products_beyond_the_scratch(torch_gpu):
    torch = torch_gpu
    lib = tiledmm.load()
    m = n = k = 4096  # 16 moduli: 2 x 256 MiB of residues alone
    dA = torch.zeros(m * k, dtype=torch.float64, device="cuda")
    dB = torch.zeros(k * n, dtype=torch.float64, device="cuda")
    dC = torch.full((m * n,), 7.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    with tiledmm.Handle(tiledmm.F64, 1) as h:
        h.set_option(tiledmm.OPT_OZAKI_MODULI, 16)
        st = _call(lib, h, "tmm_device_gemm_noalloc", "N", "N", m, n, k, dA, m, dB, k, dC, m, None)
        assert st == 1  # TMM_ERR_INVALID_ARG
        assert "fixed scratch" in lib.tmm_last_error(h.ptr).decode()
        h.synchronize()
        assert bool((dC == 7.0).all())  # nothing ran
        # a product that fits still runs on the same handle
        small = 256
        st = _call(lib, h, "tmm_device_gemm_noalloc", "N", "N", small, small, small, dA, small, dB, small, dC, small,
                   None)
        assert st == 0
        h.synchronize()
        # 1.5 * 0 + (-0.5) * 7 on the small block
        assert bool((dC[:small * small] == -3.5).all())

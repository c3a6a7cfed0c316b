"""Per-tile GEMM kernels on device-resident data (tmm_device_gemm) against the oracle.

The kernel replaces the vendor {s,d,c,z}gemm the reference calls for every tile
(Tiled-MM src/Tiled-MM/tiled_mm.cpp:181-268). Checks: exact results on integer data (all partial
sums are exactly representable), and the stated normwise bound on uniform[-1,1] data, for every
op combination, edge sizes below/above the 128x128 (real) / 64x64 (complex) workgroup tile, and
ld > rows. C holds NaN where beta == 0 to check that C is then not read.
"""
import numpy as np
import pytest

import oracle
import tiledmm
from helpers import bound, int_matrix, uniform_matrix

pytestmark = pytest.mark.gpu

CODES = {np.float32: tiledmm.F32, np.float64: tiledmm.F64, np.complex64: tiledmm.C32, np.complex128: tiledmm.C64}

SHAPES = [(1, 1, 1), (5, 2, 2), (17, 33, 5), (64, 64, 16), (128, 128, 16), (129, 127, 31), (300, 260, 200),
          (257, 1, 130), (1, 300, 77), (96, 80, 0)]


def _dev(torch, X):
    return torch.from_numpy(np.ascontiguousarray(X.reshape(-1, order="F"))).cuda()


def _run(torch, dtype, ta, tb, m, n, k, alpha, beta, A, lda, B, ldb, C, ldc):
    dA, dB, dC = _dev(torch, A), _dev(torch, B), _dev(torch, C)
    with tiledmm.Handle(CODES[dtype], 1, 256, 256, 256) as h:
        h.device_gemm(ta, tb, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc)
        h.synchronize()
    return dC.cpu().numpy().reshape(C.shape, order="F")


def _ops(dtype):
    ops = ["N", "T"] + (["C"] if np.issubdtype(dtype, np.complexfloating) else [])
    return [(a, b) for a in ops for b in ops]


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.complex128, np.complex64])
@pytest.mark.parametrize("shape", SHAPES)
def test_integer_exact(torch_gpu, dtype, shape):
    m, n, k = shape
    rng = np.random.default_rng(m * 131 + n * 7 + k)
    for ta, tb in _ops(dtype):
        ar, ac = oracle.op_shape(ta, m, k)
        br, bc = oracle.op_shape(tb, k, n)
        A = int_matrix(rng, ar, ac, dtype)
        B = int_matrix(rng, br, bc, dtype)
        C = int_matrix(rng, m, n, dtype)
        ref = C.copy(order="F")
        oracle.tiled_gemm_c(ta, tb, m, n, k, 2, A, max(ar, 1), B, max(br, 1), 3, ref, m, m, n, max(k, 1))
        out = _run(torch_gpu, dtype, ta, tb, m, n, k, 2, 3, A, max(ar, 1), B, max(br, 1), C, m)
        np.testing.assert_array_equal(out, ref, err_msg=f"{ta}{tb} {shape}")


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.complex128, np.complex64])
@pytest.mark.parametrize("shape", [(129, 127, 31), (300, 260, 200), (513, 129, 260)])
def test_uniform_bound_and_padding(torch_gpu, dtype, shape):
    m, n, k = shape
    rng = np.random.default_rng(11)
    cplx = np.issubdtype(dtype, np.complexfloating)
    alpha = (1.25 - 0.5j) if cplx else 1.25
    for beta in ([0.0, (0.5 + 1j) if cplx else -0.75]):
        for ta, tb in _ops(dtype):
            ar, ac = oracle.op_shape(ta, m, k)
            br, bc = oracle.op_shape(tb, k, n)
            A = uniform_matrix(rng, ar, ac, dtype, ld=ar + 3)
            B = uniform_matrix(rng, br, bc, dtype, ld=br + 5)
            C = uniform_matrix(rng, m, n, dtype, ld=m + 2)
            if beta == 0:
                C[:m, :] = np.nan  # must not be read
            C0 = C[:m, :].copy()
            out = _run(torch_gpu, dtype, ta, tb, m, n, k, alpha, beta, A, ar + 3, B, br + 5, C, m + 2)
            ref = oracle.host_blas_gemm(ta, tb, alpha, A[:ar, :].astype(np.complex128 if cplx else np.float64),
                                        B[:br, :].astype(np.complex128 if cplx else np.float64), beta,
                                        C0.astype(np.complex128 if cplx else np.float64))
            scale = oracle.error_scale(ta, tb, alpha, A[:ar, :], B[:br, :], beta, np.nan_to_num(C0))
            err = oracle.normwise_error(out[:m, :], ref, scale)
            assert err <= bound(dtype, k), f"{ta}{tb} beta={beta} err {err:.3e}"
            # padding rows of C untouched
            if beta != 0:
                assert np.all(np.isnan(out[m:, :]))


def test_large_square_dgemm(torch_gpu):
    """One 2048^3 tile (the 2048^3 NN configuration as a single device GEMM)."""
    torch = torch_gpu
    m = n = k = 2048
    g = torch.Generator(device="cuda").manual_seed(3)
    A = torch.rand(m * k, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    B = torch.rand(k * n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    C = torch.rand(m * n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    C0 = C.clone()
    with tiledmm.Handle(tiledmm.F64, 1) as h:
        h.device_gemm("N", "N", m, n, k, 1.0, A, m, B, k, 1.0, C, m)
        h.synchronize()
    An = A.cpu().numpy().reshape(m, k, order="F")
    Bn = B.cpu().numpy().reshape(k, n, order="F")
    C0n = C0.cpu().numpy().reshape(m, n, order="F")
    ref = An @ Bn + C0n
    scale = np.abs(An) @ np.abs(Bn) + np.abs(C0n)
    err = oracle.normwise_error(C.cpu().numpy().reshape(m, n, order="F"), ref, scale)
    assert err <= bound(np.float64, k), err


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("shape", [(1, 1, 16), (4, 4, 16), (17, 33, 5), (128, 128, 32), (129, 127, 31), (300, 260, 200),
                                   (256, 132, 130), (96, 80, 0), (1000, 600, 1001), (516, 388, 48)])
def test_sgemm_kernel_variants_integer_exact(torch_gpu, shape, variant):
    """Both FP32 tile kernels (TMM_OPT_DGEMM_KERNEL 1 = direct global->LDS ring with 16-byte chunks
    of 4 floats, the default; 0 = register-staged double buffer): exact on integer data for all
    layouts, K not a multiple of 16 (the remainder goes to the register-staged kernel), row counts
    not a multiple of 4 (those layouts fall back) and leading dimensions that leave rows only
    4-byte aligned."""
    torch = torch_gpu
    m, n, k = shape
    rng = np.random.default_rng(m * 29 + n * 13 + k)
    for ta, tb in _ops(np.float32):
        ar, ac = oracle.op_shape(ta, m, k)
        br, bc = oracle.op_shape(tb, k, n)
        A = int_matrix(rng, ar, ac, np.float32, ld=ar + 1)
        B = int_matrix(rng, br, bc, np.float32, ld=br + 3)
        C = int_matrix(rng, m, n, np.float32)
        ref = C.copy(order="F")
        oracle.tiled_gemm_c(ta, tb, m, n, k, 2, A, ar + 1, B, br + 3, 3, ref, m, m, n, max(k, 1))
        dA, dB, dC = _dev(torch, A), _dev(torch, B), _dev(torch, C)
        with tiledmm.Handle(tiledmm.F32, 1) as h:
            h.set_option(tiledmm.OPT_DGEMM_KERNEL, variant)
            h.device_gemm(ta, tb, m, n, k, 2.0, dA, ar + 1, dB, br + 3, 3.0, dC, m)
            h.synchronize()
        out = dC.cpu().numpy().reshape(C.shape, order="F")
        np.testing.assert_array_equal(out, ref, err_msg=f"{ta}{tb} {shape}")


@pytest.mark.parametrize("shape", [(2900, 2952, 272), (3000, 2944, 1001), (2944, 2944, 16)])
def test_sgemm_pipelined_integer_exact(torch_gpu, shape):
    """The software-pipelined FP32 kernel (128 x 128 tiles, used from two tiles per CU on): exact on
    integer data for all layouts, against the float64 product (exact on these small integers and
    in float32 range). Tiles cut by m and n, an odd number of 16-deep stages, a K remainder added by
    the register-staged kernel, a single stage; alpha and beta applied, C read in place."""
    torch = torch_gpu
    m, n, k = shape
    rng = np.random.default_rng(7 * m + n + k)
    for ta, tb in _ops(np.float32):
        ar, ac = oracle.op_shape(ta, m, k)
        br, bc = oracle.op_shape(tb, k, n)
        A = int_matrix(rng, ar, ac, np.float32, ld=ar + 4)
        B = int_matrix(rng, br, bc, np.float32, ld=br + 8)
        C = int_matrix(rng, m, n, np.float32)
        opA = (A[:ar, :ac] if ta == "N" else A[:ar, :ac].T).astype(np.float64)
        opB = (B[:br, :bc] if tb == "N" else B[:br, :bc].T).astype(np.float64)
        ref = 2.0 * (opA @ opB) + 3.0 * C.astype(np.float64)
        dA, dB, dC = _dev(torch, A), _dev(torch, B), _dev(torch, C)
        with tiledmm.Handle(tiledmm.F32, 1) as h:
            h.device_gemm(ta, tb, m, n, k, 2.0, dA, ar + 4, dB, br + 8, 3.0, dC, m)
            h.synchronize()
        out = dC.cpu().numpy().reshape(C.shape, order="F").astype(np.float64)
        np.testing.assert_array_equal(out, ref, err_msg=f"{ta}{tb} {shape}")


@pytest.mark.parametrize("shape", [(2900, 2952, 272), (3000, 2944, 1001), (2944, 2944, 8)])
def test_cgemm_pipelined_integer_exact(torch_gpu, shape):
    """The software-pipelined complex FP32 kernel (128 x 128 complex tiles, used from two tiles per
    CU on; with B op 'T'/'C' one B fragment set refilled in place): exact on integer data for all
    nine layouts including conjugation, against numpy's complex128 product (exact on these small
    integers, in float32 range). Tiles cut by m and n, an odd number of
    8-deep stages, a K remainder added by the register-staged kernel, a single stage; complex alpha
    and beta, C read in place."""
    torch = torch_gpu
    m, n, k = shape
    rng = np.random.default_rng(3 * m + 7 * n + k)
    alpha, beta = 2 - 1j, 1 + 3j
    for ta, tb in _ops(np.complex64):
        ar, ac = oracle.op_shape(ta, m, k)
        br, bc = oracle.op_shape(tb, k, n)
        A = int_matrix(rng, ar, ac, np.complex64, ld=ar + 2)
        B = int_matrix(rng, br, bc, np.complex64, ld=br + 4)
        C = int_matrix(rng, m, n, np.complex64)
        A64, B64 = A[:ar, :ac].astype(np.complex128), B[:br, :bc].astype(np.complex128)
        opA = {"N": A64, "T": A64.T, "C": A64.conj().T}[ta]
        opB = {"N": B64, "T": B64.T, "C": B64.conj().T}[tb]
        ref = alpha * (opA @ opB) + beta * C.astype(np.complex128)
        dA, dB, dC = _dev(torch, A), _dev(torch, B), _dev(torch, C)
        with tiledmm.Handle(tiledmm.C32, 1) as h:
            h.device_gemm(ta, tb, m, n, k, alpha, dA, ar + 2, dB, br + 4, beta, dC, m)
            h.synchronize()
        out = dC.cpu().numpy().reshape(C.shape, order="F").astype(np.complex128)
        np.testing.assert_array_equal(out, ref, err_msg=f"{ta}{tb} {shape}")


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("shape", [(1, 1, 8), (2, 2, 8), (17, 33, 5), (128, 128, 16), (129, 127, 31), (300, 260, 200),
                                   (258, 130, 130), (96, 80, 0), (640, 388, 1001)])
def test_cgemm_kernel_variants_integer_exact(torch_gpu, shape, variant):
    """Both complex FP32 tile kernels (TMM_OPT_DGEMM_KERNEL 1 = direct global->LDS ring with
    16-byte chunks of two complex floats, the default; 0 = register-staged): exact on integer data
    for all layouts including conjugation, K not a multiple of 8, odd row counts of row-contiguous
    operands (those fall back) and leading dimensions that leave rows 8-byte aligned; complex alpha
    and beta."""
    torch = torch_gpu
    m, n, k = shape
    rng = np.random.default_rng(m * 23 + n * 11 + k)
    for ta, tb in _ops(np.complex64):
        ar, ac = oracle.op_shape(ta, m, k)
        br, bc = oracle.op_shape(tb, k, n)
        A = int_matrix(rng, ar, ac, np.complex64, ld=ar + 1)
        B = int_matrix(rng, br, bc, np.complex64, ld=br + 3)
        C = int_matrix(rng, m, n, np.complex64)
        ref = C.copy(order="F")
        oracle.tiled_gemm_c(ta, tb, m, n, k, 2 - 1j, A, ar + 1, B, br + 3, 1 + 2j, ref, m, m, n, max(k, 1))
        dA, dB, dC = _dev(torch, A), _dev(torch, B), _dev(torch, C)
        with tiledmm.Handle(tiledmm.C32, 1) as h:
            h.set_option(tiledmm.OPT_DGEMM_KERNEL, variant)
            h.device_gemm(ta, tb, m, n, k, 2 - 1j, dA, ar + 1, dB, br + 3, 1 + 2j, dC, m)
            h.synchronize()
        out = dC.cpu().numpy().reshape(C.shape, order="F")
        np.testing.assert_array_equal(out, ref, err_msg=f"{ta}{tb} {shape}")


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("shape", [(1, 1, 8), (2, 2, 8), (17, 33, 5), (128, 128, 16), (129, 127, 31), (300, 260, 200),
                                   (257, 130, 130), (96, 80, 0), (1000, 600, 1001)])
def test_dgemm_kernel_variants_integer_exact(torch_gpu, shape, variant):
    """Both FP64 tile kernels (TMM_OPT_DGEMM_KERNEL 1 = direct global->LDS ring, the default;
    0 = register-staged double buffer): exact on integer data for all layouts, K not a multiple
    of 8 (variant 1 hands the remainder to the register-staged kernel) and odd row counts."""
    torch = torch_gpu
    m, n, k = shape
    rng = np.random.default_rng(m * 31 + n * 17 + k)
    for ta, tb in _ops(np.float64):
        ar, ac = oracle.op_shape(ta, m, k)
        br, bc = oracle.op_shape(tb, k, n)
        A = int_matrix(rng, ar, ac, np.float64, ld=ar + 1)
        B = int_matrix(rng, br, bc, np.float64, ld=br + 2)
        C = int_matrix(rng, m, n, np.float64)
        ref = C.copy(order="F")
        oracle.tiled_gemm_c(ta, tb, m, n, k, 2, A, ar + 1, B, br + 2, 3, ref, m, m, n, max(k, 1))
        dA, dB, dC = _dev(torch, A), _dev(torch, B), _dev(torch, C)
        with tiledmm.Handle(tiledmm.F64, 1) as h:
            h.set_option(tiledmm.OPT_DGEMM_KERNEL, variant)
            h.device_gemm(ta, tb, m, n, k, 2.0, dA, ar + 1, dB, br + 2, 3.0, dC, m)
            h.synchronize()
        out = dC.cpu().numpy().reshape(C.shape, order="F")
        np.testing.assert_array_equal(out, ref, err_msg=f"{ta}{tb} {shape}")


@pytest.mark.parametrize("shape", [(4100, 4150, 264), (4096, 4096, 1001), (4098, 4352, 8), (11800, 11650, 24)])
def test_dgemm_big_tiles_integer_exact(torch_gpu, shape):
    """The software-pipelined FP64 kernels: 128 x 128 tiles at two per CU (at least two tiles per
    CU; the first three shapes) and 128 x 256 tiles (at least 2^27 elements of C; the last):
    exact on integer data for all layouts, against numpy's product, which is exact on these small
    integers. Covers tiles cut by m and n, an odd number of 8-deep stages (the loop's unpaired
    last stage), a K remainder added by the register-staged kernel and a single stage; alpha and
    beta both applied, C read in place."""
    torch = torch_gpu
    m, n, k = shape
    rng = np.random.default_rng(m + 3 * n + 7 * k)
    for ta, tb in _ops(np.float64):
        ar, ac = oracle.op_shape(ta, m, k)
        br, bc = oracle.op_shape(tb, k, n)
        A = int_matrix(rng, ar, ac, np.float64, ld=ar + 2)
        B = int_matrix(rng, br, bc, np.float64, ld=br + 4)
        C = int_matrix(rng, m, n, np.float64)
        opA = A[:ar, :ac] if ta == "N" else A[:ar, :ac].T
        opB = B[:br, :bc] if tb == "N" else B[:br, :bc].T
        ref = 2.0 * (opA @ opB) + 3.0 * C
        dA, dB, dC = _dev(torch, A), _dev(torch, B), _dev(torch, C)
        with tiledmm.Handle(tiledmm.F64, 1) as h:
            h.device_gemm(ta, tb, m, n, k, 2.0, dA, ar + 2, dB, br + 4, 3.0, dC, m)
            h.synchronize()
        out = dC.cpu().numpy().reshape(C.shape, order="F")
        np.testing.assert_array_equal(out, ref, err_msg=f"{ta}{tb} {shape}")


@pytest.mark.parametrize("shape", [(1500, 1530, 264), (1600, 1472, 1001), (1536, 1536, 8)])
def test_zgemm_pipelined_integer_exact(torch_gpu, shape):
    """The software-pipelined complex FP64 kernel (64 x 64 complex tiles, used from two tiles per
    CU on): exact on integer data for all nine layouts including conjugation, against numpy's
    product (exact on these small Gaussian integers). Tiles cut by m and n, an odd number of 8-deep
    stages, a K remainder added by the register-staged kernel, a single stage; complex alpha and
    beta, C read in place."""
    torch = torch_gpu
    m, n, k = shape
    rng = np.random.default_rng(m + 5 * n + 3 * k)
    alpha, beta = 2 - 1j, 1 + 3j
    for ta, tb in _ops(np.complex128):
        ar, ac = oracle.op_shape(ta, m, k)
        br, bc = oracle.op_shape(tb, k, n)
        A = int_matrix(rng, ar, ac, np.complex128, ld=ar + 1)
        B = int_matrix(rng, br, bc, np.complex128, ld=br + 3)
        C = int_matrix(rng, m, n, np.complex128)
        opA = {"N": A[:ar, :ac], "T": A[:ar, :ac].T, "C": A[:ar, :ac].conj().T}[ta]
        opB = {"N": B[:br, :bc], "T": B[:br, :bc].T, "C": B[:br, :bc].conj().T}[tb]
        ref = alpha * (opA @ opB) + beta * C
        dA, dB, dC = _dev(torch, A), _dev(torch, B), _dev(torch, C)
        with tiledmm.Handle(tiledmm.C64, 1) as h:
            h.device_gemm(ta, tb, m, n, k, alpha, dA, ar + 1, dB, br + 3, beta, dC, m)
            h.synchronize()
        out = dC.cpu().numpy().reshape(C.shape, order="F")
        np.testing.assert_array_equal(out, ref, err_msg=f"{ta}{tb} {shape}")


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("shape", [(1, 1, 8), (2, 3, 8), (17, 33, 5), (64, 64, 16), (65, 63, 31), (300, 260, 200),
                                   (129, 70, 130), (96, 80, 0)])
def test_zgemm_kernel_variants_integer_exact(torch_gpu, shape, variant):
    """Both complex FP64 tile kernels (TMM_OPT_DGEMM_KERNEL 1 = direct global->LDS ring with
    conjugation in registers, 0 = register-staged): exact on complex integer data for all nine
    op combinations, K not a multiple of 8 and edge tiles, with complex alpha and beta."""
    torch = torch_gpu
    m, n, k = shape
    rng = np.random.default_rng(m * 37 + n * 11 + k)
    for ta, tb in _ops(np.complex128):
        ar, ac = oracle.op_shape(ta, m, k)
        br, bc = oracle.op_shape(tb, k, n)
        A = int_matrix(rng, ar, ac, np.complex128, ld=ar + 1)
        B = int_matrix(rng, br, bc, np.complex128, ld=br + 2)
        C = int_matrix(rng, m, n, np.complex128)
        alpha, beta = 2 - 1j, 1 + 3j
        ref = C.copy(order="F")
        oracle.tiled_gemm_c(ta, tb, m, n, k, alpha, A, ar + 1, B, br + 2, beta, ref, m, m, n, max(k, 1))
        dA, dB, dC = _dev(torch, A), _dev(torch, B), _dev(torch, C)
        with tiledmm.Handle(tiledmm.C64, 1) as h:
            h.set_option(tiledmm.OPT_DGEMM_KERNEL, variant)
            h.device_gemm(ta, tb, m, n, k, alpha, dA, ar + 1, dB, br + 2, beta, dC, m)
            h.synchronize()
        out = dC.cpu().numpy().reshape(C.shape, order="F")
        np.testing.assert_array_equal(out, ref, err_msg=f"{ta}{tb} {shape}")

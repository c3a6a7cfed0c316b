"""BLAS semantics at the edges of the FP64 range, and the Ozaki workspaces under concurrent calls.

* Inf, NaN and huge entries: the reference's per-tile GEMM is the vendor dgemm
  (Tiled-MM src/Tiled-MM/tiled_mm.cpp:200-217), which propagates Inf and NaN and overflows to Inf.
  The Ozaki paths cannot emulate such entries (their exponents leave the integer range), so a tile
  whose operands hold an Inf, a NaN or an entry of at least 2^480 runs on the native FP64 kernel,
  decided on the device. Checked: the Inf / NaN pattern of C equals native DGEMM's, tiles with such
  operands equal native bit for bit, tiles without them stay emulated (they differ from native and
  are more accurate).
* alpha = 0: C = beta*C without reading A or B (BLAS), on the host-to-host and the device path,
  native and emulated; with beta = 0 as well C becomes 0 without being read.
* Two asynchronous emulated device calls on two streams share the handle's workspace: the second
  waits for the first (both results equal the same calls made one at a time).
"""
import numpy as np
import pytest

import oracle
import tiledmm

pytestmark = pytest.mark.gpu
MODES = [("crt16", tiledmm.OPT_OZAKI_MODULI, 16), ("slices7", tiledmm.OPT_OZAKI_SLICES, 7)]


def _dev(torch, X):
    return torch.from_numpy(np.ascontiguousarray(X.reshape(-1, order="F"))).cuda()


def _device(torch, opt, val, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    dA, dB, dC = _dev(torch, A), _dev(torch, B), _dev(torch, C)
    torch.cuda.synchronize()
    with tiledmm.Handle(tiledmm.F64, 1) as h:
        if opt is not None:
            h.set_option(opt, val)
        h.device_gemm(ta, tb, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc)
        h.synchronize()
    return dC.cpu().numpy().reshape(C.shape, order="F")


def _host(opt, val, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tiles, coarsen=0):
    Cw = C.copy(order="F")
    with tiledmm.Handle(tiledmm.F64, 2, *tiles) as h:
        h.set_option(tiledmm.OPT_COARSEN, coarsen)
        if opt is not None:
            h.set_option(opt, val)
        h.gemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, Cw, ldc)
    return Cw


def _poisoned(rng, ta, tb, m, n, k):
    """Uniform data with a NaN, +Inf, -Inf and huge (1e300) entries in op(A) row 5 / 9 / 40 and
    op(B) column 7 / 30, placed in the storage layout of each op."""
    ar, ac = oracle.op_shape(ta, m, k)
    br, bc = oracle.op_shape(tb, k, n)
    A = np.asfortranarray(rng.uniform(-1, 1, (ar, ac)))
    B = np.asfortranarray(rng.uniform(-1, 1, (br, bc)))
    opA = A if ta == "N" else A.T  # views: writes land in A
    opB = B if tb == "N" else B.T
    opA[5, 3] = np.nan
    opA[9, 11] = np.inf
    opA[40, 0] = 1e300
    opB[17, 7] = -np.inf
    opB[2, 30] = 1e300
    return A, B, ar, br


def _masks(X):
    return np.isnan(X), np.isposinf(X), np.isneginf(X)


@pytest.mark.parametrize("name,opt,val", MODES, ids=[m[0] for m in MODES])
@pytest.mark.parametrize("flags", ["NN", "TN", "NT"])
def test_device_nonfinite_matches_native(torch_gpu, name, opt, val, flags):
    ta, tb = flags
    m, n, k = 300, 257, 200
    rng = np.random.default_rng(21)
    A, B, ar, br = _poisoned(rng, ta, tb, m, n, k)
    C = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    nat = _device(torch_gpu, None, 0, ta, tb, m, n, k, 1.5, A, ar, B, br, 0.5, C, m)
    emu = _device(torch_gpu, opt, val, ta, tb, m, n, k, 1.5, A, ar, B, br, 0.5, C, m)
    for x, y in zip(_masks(nat), _masks(emu)):
        np.testing.assert_array_equal(x, y)
    assert np.isnan(nat[5]).all() and np.isinf(nat[9, 0]) and not np.isfinite(nat[:, 7]).any()
    # one tile on the device path: the whole product ran on the native kernel
    np.testing.assert_array_equal(emu, nat)


@pytest.mark.parametrize("name,opt,val", MODES, ids=[m[0] for m in MODES])
def test_device_overflow_matches_native(torch_gpu, name, opt, val):
    """Finite but huge entries whose products overflow: Inf with the sign native DGEMM gives."""
    m, n, k = 64, 48, 96
    rng = np.random.default_rng(22)
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
    A[3, 10], B[10, 4] = 1e300, 1e300
    A[7, 20], B[20, 9] = -1e200, 1e200
    C = np.zeros((m, n), order="F")
    nat = _device(torch_gpu, None, 0, "N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
    emu = _device(torch_gpu, opt, val, "N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
    assert np.isposinf(nat[3, 4]) and np.isneginf(nat[7, 9])
    np.testing.assert_array_equal(emu, nat)


@pytest.mark.parametrize("name,opt,val", MODES, ids=[m[0] for m in MODES])
def test_host_to_host_nonfinite_per_tile(torch_gpu, name, opt, val):
    """Tiles of 128 x 128 x 96 (reference grid kept): only the tiles whose A or B block holds a
    poisoned entry fall back; the C pattern equals native's, the other tiles are emulated."""
    ta, tb = "N", "N"
    m, n, k = 384, 320, 288
    rng = np.random.default_rng(23)
    A, B, ar, br = _poisoned(rng, ta, tb, m, n, k)
    C = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    tiles = (128, 128, 96)
    nat = _host(None, 0, ta, tb, m, n, k, 1.0, A, ar, B, br, 1.0, C, m, tiles)
    emu = _host(opt, val, ta, tb, m, n, k, 1.0, A, ar, B, br, 1.0, C, m, tiles)
    for x, y in zip(_masks(nat), _masks(emu)):
        np.testing.assert_array_equal(x, y)
    # C tile rows 128..255 x columns 128..255 meet no poisoned block: emulated, and within the bound
    blk = (slice(128, 256), slice(128, 256))
    A2, B2 = A[128:256, :], B[:, 128:256]
    ref = A2 @ B2 + C[blk]
    scale = np.abs(A2) @ np.abs(B2) + np.abs(C[blk])
    assert oracle.normwise_error(emu[blk], ref, scale) <= 1e-15
    assert not np.array_equal(emu[blk], nat[blk]), "a clean tile ran on the native kernel"
    # every finite entry (its row of op(A) and column of op(B) free of Inf and NaN; the huge
    # entries' k-chunk ran natively, the others emulated) within the bound of native's
    fin = np.isfinite(nat)
    scale = np.abs(np.nan_to_num(A, nan=0, posinf=0, neginf=0)) @ np.abs(np.nan_to_num(B, nan=0, posinf=0, neginf=0))
    scale += np.abs(C)
    assert np.all(np.abs(emu[fin] - nat[fin]) <= 2 * k * 2.0 ** -53 * scale[fin])
    assert np.abs(emu[40, :]).max() > 1e298  # the huge entry's row


def test_finite_inputs_stay_emulated(torch_gpu):
    """No flag on ordinary data: the emulated result is not the native kernel's, and closer to the
    host BLAS (so the gated native launch did no work)."""
    m, n, k = 512, 384, 640
    rng = np.random.default_rng(24)
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
    C = np.zeros((m, n), order="F")
    nat = _device(torch_gpu, None, 0, "N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
    emu = _device(torch_gpu, tiledmm.OPT_OZAKI_MODULI, 18, "N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
    ref = A @ B
    scale = np.abs(A) @ np.abs(B)
    assert not np.array_equal(emu, nat)
    assert oracle.normwise_error(emu, ref, scale) < oracle.normwise_error(nat, ref, scale)


@pytest.mark.parametrize("mode", [None, MODES[0], MODES[1]], ids=["native", "crt16", "slices7"])
def test_alpha_zero_reads_neither_a_nor_b(torch_gpu, mode):
    opt, val = (None, 0) if mode is None else mode[1:]
    m, n, k = 200, 150, 300
    rng = np.random.default_rng(25)
    A = np.full((m, k), np.nan, order="F")
    B = np.full((k, n), np.inf, order="F")
    C = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    for beta in (0.5, 0.0):
        Cin = C if beta != 0 else np.full((m, n), np.nan, order="F")
        want = 0.5 * C if beta != 0 else np.zeros((m, n))
        got = _host(opt, val, "N", "N", m, n, k, 0.0, A, m, B, k, beta, Cin, m, (96, 96, 64))
        np.testing.assert_array_equal(got, want, err_msg=f"host-to-host beta={beta}")
        got = _device(torch_gpu, opt, val, "N", "N", m, n, k, 0.0, A, m, B, k, beta, Cin, m)
        np.testing.assert_array_equal(got, want, err_msg=f"device beta={beta}")


def test_two_streams_share_the_emulation_workspace(torch_gpu):
    """A large and then a smaller emulated product on two streams, enqueued back to back: the
    second call waits for the first one's use of the workspace; a third, larger call on the
    first stream waits for the second before the workspace grows."""
    torch = torch_gpu
    rng = np.random.default_rng(26)
    shapes = [(1536, 1280, 2048), (700, 900, 1100), (2048, 1536, 2304)]
    probs = []
    for m, n, k in shapes:
        A = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
        B = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
        probs.append((m, n, k, _dev(torch, A), _dev(torch, B)))
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    # one at a time
    alone = []
    with tiledmm.Handle(tiledmm.F64, 1) as h:
        h.set_option(tiledmm.OPT_OZAKI_MODULI, 16)
        for m, n, k, dA, dB in probs:
            dC = torch.zeros(m * n, dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            h.device_gemm("N", "N", m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, stream=s1.cuda_stream)
            s1.synchronize()
            alone.append(dC.cpu().numpy())
    # back to back on two streams, one handle
    outs = [torch.zeros(m * n, dtype=torch.float64, device="cuda") for m, n, *_ in probs]
    torch.cuda.synchronize()
    with tiledmm.Handle(tiledmm.F64, 1) as h:
        h.set_option(tiledmm.OPT_OZAKI_MODULI, 16)
        for (m, n, k, dA, dB), dC, s in zip(probs, outs, (s1, s2, s1)):
            h.device_gemm("N", "N", m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, stream=s.cuda_stream)
        torch.cuda.synchronize()
    for want, got in zip(alone, outs):
        np.testing.assert_array_equal(got.cpu().numpy(), want)

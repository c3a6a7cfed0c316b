"""The C++ drop-in API on the GPU: tests/cpp/test_multiply.cpp follows the reference's test
program (Tiled-MM tests/test-multiply.cpp:116-351: integer data, both copy_c_back modes,
get_full_device_buffer_c, |difference| <= 1e-6) and is run with the reference's CI tiny-tile
arguments (ci/daint-alps.yml:51,60) and further shapes, tiles and transposes."""
import os
import subprocess

import pytest

import tiledmm

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(tiledmm.LIB_DIR, "test-multiply")

CASES = [
    "-m 50 -n 200 -k 21 --tile_m 4 --tile_n 4 --tile_k 4 --beta 0 --n_streams 2",
    "-m 5 -n 2 -k 2 --tile_m 4 --tile_n 4 --tile_k 4 --beta 0 --n_streams 2",
    "-m 500 -n 600 -k 700",
    "-m 500 -n 600 -k 700 --transpose TN --beta 1 --alpha 2",
    "-m 321 -n 123 -k 300 --transpose NT --tile_m 100 --tile_n 50 --tile_k 64 --beta 1 --n_streams 3",
    "-m 321 -n 123 -k 300 --transpose TT --tile_m 128 --tile_n 128 --tile_k 128 --beta 2 --ld_a 400 --ld_b 200 --ld_c 350",
]


def test_context_streams_events_buffers(torch_gpu):
    """tests/cpp/test_context.cpp: get_stream / get_device_stream / get_result_stream and their
    range checks (Tiled-MM src/Tiled-MM/gpu_context.cpp:21-37), enqueue_event / wait_on_event across
    two streams (device_stream.hpp:58-72), device_event::time_since (device_event.hpp:66-72),
    device_buffer::stream_buffer (device_buffer.hpp:44-54), set_num_streams growth
    (gpu_context.cpp:49-53) and a GEMM on the grown context."""
    drv = os.path.join(tiledmm.LIB_DIR, "test-context")
    assert os.path.exists(drv), "build the library first (make -C tiled-mm_amd)"
    r = subprocess.run([drv], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "CONTEXT OK" in r.stdout, r.stdout + r.stderr


@pytest.mark.parametrize("args", CASES)
def test_reference_driver_flow(torch_gpu, args):
    assert os.path.exists(DRIVER), "build the library first (make -C tiled-mm_amd)"
    r = subprocess.run([DRIVER] + args.split(), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "CORRECT" in r.stdout, r.stdout + r.stderr


def test_blas_interposer_dgemm_zgemm(torch_gpu):
    """libtmm_blas.so: Fortran-ABI dgemm_/zgemm_ routed through the out-of-core path."""
    import ctypes

    import numpy as np

    os.environ["TMM_BLAS_MIN_FLOPS"] = "0"
    lib = ctypes.CDLL(os.path.join(tiledmm.LIB_DIR, "libtmm_blas.so"))
    i = ctypes.c_int
    rng = np.random.default_rng(5)
    m, n, k = 300, 200, 250
    A = np.asfortranarray(rng.integers(0, 10, (k, m)).astype(np.float64))  # 'T'
    B = np.asfortranarray(rng.integers(0, 10, (k, n)).astype(np.float64))
    C = np.asfortranarray(rng.integers(0, 10, (m, n)).astype(np.float64))
    ref = 2.0 * (A.T @ B) + 3.0 * C
    al, be = ctypes.c_double(2.0), ctypes.c_double(3.0)
    lib.dgemm_(b"T", b"N", ctypes.byref(i(m)), ctypes.byref(i(n)), ctypes.byref(i(k)), ctypes.byref(al),
               A.ctypes.data_as(ctypes.c_void_p), ctypes.byref(i(k)), B.ctypes.data_as(ctypes.c_void_p),
               ctypes.byref(i(k)), ctypes.byref(be), C.ctypes.data_as(ctypes.c_void_p), ctypes.byref(i(m)))
    np.testing.assert_array_equal(C, ref)
    Z = np.asfortranarray(rng.integers(0, 10, (k, m)) + 1j * rng.integers(0, 10, (k, m)))
    W = np.asfortranarray(rng.integers(0, 10, (k, n)) + 1j * rng.integers(0, 10, (k, n)))
    Y = np.zeros((m, n), dtype=np.complex128, order="F")
    za, zb = (ctypes.c_double * 2)(1.0, 1.0), (ctypes.c_double * 2)(0.0, 0.0)
    lib.zgemm_(b"C", b"N", ctypes.byref(i(m)), ctypes.byref(i(n)), ctypes.byref(i(k)), za,
               Z.ctypes.data_as(ctypes.c_void_p), ctypes.byref(i(k)), W.ctypes.data_as(ctypes.c_void_p),
               ctypes.byref(i(k)), zb, Y.ctypes.data_as(ctypes.c_void_p), ctypes.byref(i(m)))
    np.testing.assert_array_equal(Y, (1 + 1j) * (Z.conj().T @ W))


def test_blas_interposer_sgemm_cgemm(torch_gpu):
    """libtmm_blas.so: Fortran-ABI sgemm_/cgemm_ routed through the out-of-core path (the FP32 and
    complex FP32 direct-to-LDS tile kernels); exact on small integers, beta = 0 never reads C."""
    import ctypes

    import numpy as np

    os.environ["TMM_BLAS_MIN_FLOPS"] = "0"
    lib = ctypes.CDLL(os.path.join(tiledmm.LIB_DIR, "libtmm_blas.so"))
    i = ctypes.c_int
    rng = np.random.default_rng(6)
    m, n, k = 520, 300, 272
    A = np.asfortranarray(rng.integers(0, 10, (m, k)).astype(np.float32))  # 'N'
    B = np.asfortranarray(rng.integers(0, 10, (n, k)).astype(np.float32))  # 'T'
    C = np.asfortranarray(rng.integers(0, 10, (m, n)).astype(np.float32))
    ref = 2.0 * (A.astype(np.float64) @ B.T.astype(np.float64)) - 1.0 * C
    al, be = ctypes.c_float(2.0), ctypes.c_float(-1.0)
    lib.sgemm_(b"N", b"T", ctypes.byref(i(m)), ctypes.byref(i(n)), ctypes.byref(i(k)), ctypes.byref(al),
               A.ctypes.data_as(ctypes.c_void_p), ctypes.byref(i(m)), B.ctypes.data_as(ctypes.c_void_p),
               ctypes.byref(i(n)), ctypes.byref(be), C.ctypes.data_as(ctypes.c_void_p), ctypes.byref(i(m)))
    np.testing.assert_array_equal(C.astype(np.float64), ref)
    Z = np.asfortranarray((rng.integers(0, 10, (k, m)) + 1j * rng.integers(0, 10, (k, m))).astype(np.complex64))
    W = np.asfortranarray((rng.integers(0, 10, (n, k)) + 1j * rng.integers(0, 10, (n, k))).astype(np.complex64))
    Y = np.full((m, n), np.nan, dtype=np.complex64, order="F")  # beta = 0: must not be read
    za, zb = (ctypes.c_float * 2)(1.0, -1.0), (ctypes.c_float * 2)(0.0, 0.0)
    lib.cgemm_(b"C", b"C", ctypes.byref(i(m)), ctypes.byref(i(n)), ctypes.byref(i(k)), za,
               Z.ctypes.data_as(ctypes.c_void_p), ctypes.byref(i(k)), W.ctypes.data_as(ctypes.c_void_p),
               ctypes.byref(i(n)), zb, Y.ctypes.data_as(ctypes.c_void_p), ctypes.byref(i(m)))
    ref_z = (1 - 1j) * (Z.astype(np.complex128).conj().T @ W.astype(np.complex128).conj().T)
    np.testing.assert_array_equal(Y.astype(np.complex128), ref_z)

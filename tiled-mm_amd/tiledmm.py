"""ctypes binding of libtiledmm.so (C ABI in include/tiledmm.h) for tests and the benchmark.

The Python layer is plumbing: it loads the native library and marshals pointers. It has no
fallback: if lib/libtiledmm.so is missing, `load()` raises, so nothing can silently run
without the HIP code.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libtiledmm.so")
UTIL_PATH = os.path.join(LIB_DIR, "libtmm_hostutil.so")
HOST_PATH = os.path.join(LIB_DIR, "libtmm_host.so")

F32, F64, C32, C64 = 0, 1, 2, 3
DTYPE_CODE = {np.dtype(np.float32): F32, np.dtype(np.float64): F64,
              np.dtype(np.complex64): C32, np.dtype(np.complex128): C64}
REAL_OF = {F32: np.float32, F64: np.float64, C32: np.float32, C64: np.float64}
NP_OF = {F32: np.float32, F64: np.float64, C32: np.complex64, C64: np.complex128}

OPT_COARSEN, OPT_FIRST_CHUNK, OPT_HOST_THREADS, OPT_POOL_LIMIT_MB, OPT_PROFILE, OPT_OZAKI_SLICES = 1, 2, 3, 4, 5, 6
OPT_DGEMM_KERNEL = 7
OPT_OZAKI_MODULI = 8
OPT_OZAKI_REUSE = 9
OPT_BATCH_GRAPH = 10
OPT_TILE_GROUPS = 11
OPT_DEFER_CIN = 12
OPT_BATCH_BODY = 13
OPT_BATCH_QUEUES = 14
OPT_BATCH_EMULATION = 15
STATUS = {0: "TMM_SUCCESS", 1: "TMM_ERR_INVALID_ARG", 2: "TMM_ERR_HIP", 3: "TMM_ERR_ALLOC",
          4: "TMM_ERR_DTYPE", 5: "TMM_ERR_INTERNAL"}

_lib = None
_util = None


class TmmError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


def load():
    """Load libtiledmm.so (raises OSError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} not found: build it with `make -C tiled-mm_amd`")
        lib = ctypes.CDLL(LIB_PATH)
        vp, i, c, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_char, ctypes.c_double
        lib.tmm_create.argtypes = [ctypes.POINTER(vp), i, i, i, i, i]
        lib.tmm_destroy.argtypes = [vp]
        lib.tmm_last_error.argtypes = [vp]
        lib.tmm_last_error.restype = ctypes.c_char_p
        for name in ("tmm_sgemm", "tmm_dgemm", "tmm_cgemm", "tmm_zgemm"):
            getattr(lib, name).argtypes = [vp, c, c, i, i, i, vp, vp, i, vp, i, vp, vp, i, i, i]
        lib.tmm_device_gemm.argtypes = [vp, c, c, i, i, i, vp, vp, i, vp, i, vp, vp, i, vp]
        lib.tmm_device_gemm_noalloc.argtypes = [vp, c, c, i, i, i, vp, vp, i, vp, i, vp, vp, i, vp]
        lib.tmm_synchronize.argtypes = [vp]
        lib.tmm_set_option.argtypes = [vp, i, ctypes.c_long]
        lib.tmm_get_option.argtypes = [vp, i, ctypes.POINTER(ctypes.c_long)]
        lib.tmm_last_call_stats.argtypes = [vp, ctypes.POINTER(d), i]
        lib.tmm_full_c_device_ptr.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_size_t)]
        lib.tmm_tile_buffer.argtypes = [vp, i, i, ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_size_t)]
        lib.tmm_get_stream.argtypes = [vp, i, ctypes.POINTER(vp)]
        lib.tmm_set_num_streams.argtypes = [vp, i]
        lib.tmm_get_num_streams.argtypes = [vp, ctypes.POINTER(i)]
        lib.tmm_set_tile_sizes.argtypes = [vp, i, i, i]
        lib.tmm_get_tile_sizes.argtypes = [vp] + [ctypes.POINTER(i)] * 3
        lib.tmm_get_max_tile_sizes.argtypes = [vp] + [ctypes.POINTER(i)] * 3
        lib.tmm_set_streams_and_tiles.argtypes = [vp, i, i, i, i]
        lib.tmm_set_full_sizes.argtypes = [vp, i, i, i]
        lib.tmm_optimal_tile_sizes.argtypes = [vp, i, i, i] + [ctypes.POINTER(i)] * 3
        lib.tmm_optimal_tile_size.argtypes = [i, i, i]
        lib.tmm_get_dtype.argtypes = [vp, ctypes.POINTER(i)]
        lib.tmm_dgemm_batched.argtypes = [vp, c, c, i] + [vp] * 11
        lib.tmm_batch_config.argtypes = [vp, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i),
                                         ctypes.POINTER(d), ctypes.POINTER(d)]
        lib.tmm_batch_rates.argtypes = [vp, ctypes.POINTER(d), i]
        _lib = lib
    return _lib


def util():
    """libtmm_hostutil.so: deterministic multi-threaded fills."""
    global _util
    if _util is None:
        u = ctypes.CDLL(UTIL_PATH)
        vp, l, d, i = ctypes.c_void_p, ctypes.c_long, ctypes.c_double, ctypes.c_int
        for name in ("tmmu_fill_uniform_d", "tmmu_fill_uniform_s", "tmmu_fill_uniform_z"):
            getattr(u, name).argtypes = [vp, l, l, l, l, l, l, ctypes.c_uint64, d, d, i]
        u.tmmu_fill_range_d.argtypes = [vp, l, l, l, l, l, l, ctypes.c_uint64, d, i]
        u.tmmu_fill_int_d.argtypes = [vp, l, l, l, l, l, l, ctypes.c_uint64, l, l, i]
        u.tmmu_copy_block.argtypes = [vp, l, vp, l, l, l, i, i]
        _util = u
    return _util


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if isinstance(x, np.ndarray):
        return x.ctypes.data
    if hasattr(x, "data_ptr"):  # torch tensor
        return x.data_ptr()
    raise TypeError(f"cannot take a pointer of {type(x)}")


def _scalar(dtype, v):
    if dtype in (F32, C32):
        arr = (ctypes.c_float * 2)()
    else:
        arr = (ctypes.c_double * 2)()
    v = complex(v)
    arr[0], arr[1] = v.real, v.imag
    return arr


class Handle:
    """Owns one tmm_handle (the C counterpart of gpu::mm_handle<T>)."""

    def __init__(self, dtype=F64, n_streams=2, tile_m=5000, tile_n=5000, tile_k=5000):
        lib = load()
        self.dtype = dtype
        self._h = ctypes.c_void_p()
        st = lib.tmm_create(ctypes.byref(self._h), dtype, n_streams, tile_m, tile_n, tile_k)
        if st != 0:
            raise TmmError(st, lib.tmm_last_error(None).decode())

    def close(self):
        if self._h:
            load().tmm_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def ptr(self):
        return self._h

    def _check(self, st):
        if st != 0:
            raise TmmError(st, load().tmm_last_error(self._h).decode())

    def gemm(self, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, pin=False, copy_back=True):
        """Host-to-host GEMM (gpu::gemm). a, b, c: numpy arrays / pointers of host memory."""
        fn = {F32: "tmm_sgemm", F64: "tmm_dgemm", C32: "tmm_cgemm", C64: "tmm_zgemm"}[self.dtype]
        st = getattr(load(), fn)(self._h, ta.encode() if isinstance(ta, str) else ta,
                                 tb.encode() if isinstance(tb, str) else tb, m, n, k,
                                 _scalar(self.dtype, alpha), _ptr(a), lda, _ptr(b), ldb,
                                 _scalar(self.dtype, beta), _ptr(c), ldc, int(pin), int(copy_back))
        self._check(st)

    def device_gemm(self, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream=None):
        """Device-resident tile GEMM (asynchronous); a, b, c are device pointers or tensors."""
        st = load().tmm_device_gemm(self._h, ta.encode(), tb.encode(), m, n, k, _scalar(self.dtype, alpha),
                                    _ptr(a), lda, _ptr(b), ldb, _scalar(self.dtype, beta), _ptr(c), ldc,
                                    stream)
        self._check(st)

    def synchronize(self):
        self._check(load().tmm_synchronize(self._h))

    def set_option(self, opt, value):
        self._check(load().tmm_set_option(self._h, opt, value))

    def get_option(self, opt):
        v = ctypes.c_long()
        self._check(load().tmm_get_option(self._h, opt, ctypes.byref(v)))
        return v.value

    def stats(self):
        arr = (ctypes.c_double * 24)()
        self._check(load().tmm_last_call_stats(self._h, arr, 24))
        keys = ["h2d_bytes", "d2h_bytes", "gemm_launches", "tile_m", "tile_n", "tile_k", "k_chunks",
                "wall_s", "all_pinned", "kernel_s", "kernel_launches", "kernel_flops", "kernel_max_s", "int8_ops",
                "batch_graph", "batch_groups", "group_launches", "group_kernel_s", "group_flops", "host_c_written", "batch_body", "batch_queues",
                "batch_emulated", "group_bytes"]
        return dict(zip(keys, list(arr)))

    def gemm_batched(self, ta, tb, ms, ns, ks, alphas, As, ldas, Bs, ldbs, betas, Cs, ldcs):
        """Batched host-to-host DGEMM (tmm_dgemm_batched); As/Bs/Cs: lists of arrays or pointers."""
        cnt = len(ms)
        I = ctypes.c_int * cnt
        D = ctypes.c_double * cnt
        P = ctypes.c_void_p * cnt
        st = load().tmm_dgemm_batched(self._h, ta.encode(), tb.encode(), cnt, I(*ms), I(*ns), I(*ks), D(*alphas),
                                      P(*[_ptr(x) for x in As]), I(*ldas), P(*[_ptr(x) for x in Bs]), I(*ldbs),
                                      D(*betas), P(*[_ptr(x) for x in Cs]), I(*ldcs))
        self._check(st)

    def batch_config(self):
        g, w, gr = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        pred = ctypes.c_double()
        rates = (ctypes.c_double * 4)()
        self._check(load().tmm_batch_config(self._h, ctypes.byref(g), ctypes.byref(w), ctypes.byref(gr),
                                            ctypes.byref(pred), rates))
        allr = (ctypes.c_double * 7)()
        self._check(load().tmm_batch_rates(self._h, allr, 7))
        return {"groups": g.value, "workgroups": w.value, "graph": gr.value, "predicted_s": pred.value,
                "h2d_Bps": rates[0], "d2h_Bps": rates[1], "kernel_flops": rates[2], "group_overhead_s": rates[3],
                "kernel_flops_direct_lds": allr[4], "emulated_flops": allr[5], "emulated_overhead_s": allr[6]}

    def full_c(self):
        p, n = ctypes.c_void_p(), ctypes.c_size_t()
        self._check(load().tmm_full_c_device_ptr(self._h, ctypes.byref(p), ctypes.byref(n)))
        return p.value, n.value

    def stream(self, i):
        s = ctypes.c_void_p()
        self._check(load().tmm_get_stream(self._h, i, ctypes.byref(s)))
        return s.value

    def set_num_streams(self, n):
        self._check(load().tmm_set_num_streams(self._h, n))

    def num_streams(self):
        v = ctypes.c_int()
        self._check(load().tmm_get_num_streams(self._h, ctypes.byref(v)))
        return v.value

    def set_tile_sizes(self, tm, tn, tk):
        self._check(load().tmm_set_tile_sizes(self._h, tm, tn, tk))

    def tile_sizes(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(load().tmm_get_tile_sizes(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def max_tile_sizes(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(load().tmm_get_max_tile_sizes(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def optimal_tile_sizes(self, m, n, k):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(load().tmm_optimal_tile_sizes(self._h, m, n, k, ctypes.byref(a), ctypes.byref(b),
                                                  ctypes.byref(c)))
        return a.value, b.value, c.value

    def set_full_sizes(self, m, n, k):
        self._check(load().tmm_set_full_sizes(self._h, m, n, k))

    def tile_buffer(self, which, stream_id):
        p, n = ctypes.c_void_p(), ctypes.c_size_t()
        self._check(load().tmm_tile_buffer(self._h, which, stream_id, ctypes.byref(p), ctypes.byref(n)))
        return p.value, n.value


_hip = None


class _PinnedBlock:
    """Owner of one hipHostMalloc allocation; hipHostFree runs when the last view goes away."""

    def __init__(self, nbytes):
        global _hip
        if _hip is None:
            _hip = ctypes.CDLL("libamdhip64.so")
            _hip.hipHostMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_uint]
            _hip.hipHostFree.argtypes = [ctypes.c_void_p]
        self._hip = _hip  # module globals may already be gone when the last view dies at exit
        self.ptr = ctypes.c_void_p()
        st = _hip.hipHostMalloc(ctypes.byref(self.ptr), max(int(nbytes), 1), 0)
        if st != 0 or not self.ptr.value:
            raise MemoryError(f"hipHostMalloc of {nbytes} bytes failed (hipError {st})")

    def __del__(self):
        ptr = self.ptr
        if ptr is not None and ptr.value:
            self.ptr = None  # module globals (ctypes among them) may be gone at interpreter exit
            self._hip.hipHostFree(ptr)


def pinned_empty(shape, dtype=np.float64, order="F"):
    """Uninitialised numpy array in page-locked host memory (hipHostMalloc). The memory goes back
    to the system (hipHostFree) when the array and every view of it are gone."""
    dt = np.dtype(dtype)
    shape = (shape,) if isinstance(shape, int) else tuple(shape)
    nbytes = int(np.prod(shape)) * dt.itemsize
    blk = _PinnedBlock(nbytes)
    raw = (ctypes.c_char * max(nbytes, 1)).from_address(blk.ptr.value)
    raw._owner = blk
    flat = np.frombuffer(raw, dtype=dt, count=int(np.prod(shape)))
    return flat.reshape(shape, order=order)


def fill_uniform(arr, seed, lo=-1.0, hi=1.0, rows=None, cols=None, ld=None, i0=0, j0=0, total_rows=None,
                 threads=0):
    """Deterministic fill of a column-major block (numpy array, Fortran order, or raw pointer).

    Element (i0+i, j0+j) of a matrix with `total_rows` rows depends only on (seed, i, j)."""
    u = util()
    if isinstance(arr, np.ndarray):
        dt = arr.dtype
        if rows is None:
            rows, cols = arr.shape[0], (arr.shape[1] if arr.ndim > 1 else 1)
        if ld is None:
            ld = arr.shape[0]
    else:
        dt = np.dtype(arr[1])
        arr = arr[0]
    total_rows = rows if total_rows is None else total_rows
    fn = {np.dtype(np.float64): u.tmmu_fill_uniform_d, np.dtype(np.float32): u.tmmu_fill_uniform_s,
          np.dtype(np.complex128): u.tmmu_fill_uniform_z}[np.dtype(dt)]
    fn(_ptr(arr), rows, cols, ld, i0, j0, total_rows, seed, lo, hi, threads)


def fill_int(arr, seed, lo=0, hi=9, rows=None, cols=None, ld=None, i0=0, j0=0, total_rows=None, threads=0):
    """Integers in [lo, hi] as float64, deterministic per (seed, i, j) like fill_uniform."""
    if rows is None:
        rows, cols = arr.shape[0], (arr.shape[1] if arr.ndim > 1 else 1)
    if ld is None:
        ld = arr.shape[0]
    total_rows = rows if total_rows is None else total_rows
    util().tmmu_fill_int_d(_ptr(arr), rows, cols, ld, i0, j0, total_rows, seed, lo, hi, threads)

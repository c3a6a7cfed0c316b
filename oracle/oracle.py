"""Oracle for Tiled-MM's host-to-host GEMM path. TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the benchmark's CPU-baseline leg use this module; it is the
checker, never the thing measured or shipped.

Contents
  * `lib()`              : the C restatement (tmm_oracle.c): tile policy, tile geometry, and the
                           round-robin tiled GEMM semantics with per-tile BLAS arithmetic in
                           working precision (small sizes).
  * `tiled_gemm_np`      : the same round-robin semantics with numpy's BLAS for each tile product
                           (Tiled-MM src/Tiled-MM/tiled_mm.cpp:270-365: beta on k-tile 0, then 1)
                           for medium sizes.
  * `host_blas_gemm`     : C = alpha*op(A)*op(B) + beta*C on the whole matrices with numpy's BLAS
                           (the host BLAS the north-star metric is checked against and timed as the
                           CPU baseline).
  * `normwise_error` etc : error measures ||C - C_ref||_F / || |alpha||op(A)||op(B)| + |beta||C0| ||_F.
  * `ref()`              : the reference library compiled from its own sources (oracle/_ref), when
                           it has been built.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "lib", "libtmm_oracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libtiledmm_ref.so")

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(ORACLE_LIB)
        i, ll, vp, c = ctypes.c_int, ctypes.c_longlong, ctypes.c_void_p, ctypes.c_char
        L.oracle_optimal_tile_size.argtypes = [i, i, i]
        L.oracle_num_tiles.argtypes = [i, i]
        L.oracle_tile_extent.argtypes = [i, i, i]
        L.oracle_tile_offset.argtypes = [i, i, ll, i, i, i, i]
        L.oracle_tile_offset.restype = ll
        L.oracle_pack_tile_d.argtypes = [vp, i, i, ll, i, i, i, i, vp, ctypes.POINTER(i), ctypes.POINTER(i)]
        L.oracle_sgemm.argtypes = [c, c, i, i, i, ctypes.c_float, vp, ll, vp, ll, ctypes.c_float, vp, ll, i, i, i]
        L.oracle_dgemm.argtypes = [c, c, i, i, i, ctypes.c_double, vp, ll, vp, ll, ctypes.c_double, vp, ll, i, i, i]
        L.oracle_cgemm.argtypes = [c, c, i, i, i, vp, vp, ll, vp, ll, vp, vp, ll, i, i, i]
        L.oracle_zgemm.argtypes = [c, c, i, i, i, vp, vp, ll, vp, ll, vp, vp, ll, i, i, i]
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_LIB)


def ref():
    """Reference library built from Tiled-MM's sources (oracle/_ref). Raises if absent."""
    global _ref
    if _ref is None:
        R = ctypes.CDLL(REF_LIB)
        i, d, vp, c = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_char
        R.ref_optimal_tile_size.argtypes = [i, i, i]
        R.ref_tile_geometry.argtypes = [i, i, i, i, i, i, i, ctypes.POINTER(ctypes.c_longlong)]
        R.ref_dgemm.argtypes = [c, c, i, i, i, d, vp, i, vp, i, d, vp, i, i, i, i, i, i]
        R.ref_zgemm.argtypes = [c, c, i, i, i, vp, vp, i, vp, i, vp, vp, i, i, i, i, i, i]
        _ref = R
    return _ref


def optimal_tile_size(dim, curr, max_tile):
    return lib().oracle_optimal_tile_size(dim, curr, max_tile)


def op_shape(op, rows, cols):
    """Storage shape of a matrix whose op() is rows x cols."""
    return (rows, cols) if op.upper() not in ("T", "C") else (cols, rows)


def apply_op(op, X):
    op = op.upper()
    if op == "T":
        return X.T
    if op == "C":
        return X.conj().T
    return X


def _pairs(dtype, v):
    real = np.float32 if dtype in (np.float32, np.complex64) else np.float64
    v = complex(v)
    return np.array([v.real, v.imag], dtype=real)


def tiled_gemm_c(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tm, tn, tk):
    """C restatement of the tiled GEMM (in place on C). A, B, C: 1-D or Fortran numpy buffers."""
    L = lib()
    dt = C.dtype
    ta, tb = ta.encode(), tb.encode()
    if dt == np.float64:
        L.oracle_dgemm(ta, tb, m, n, k, float(np.real(alpha)), A.ctypes.data, lda, B.ctypes.data, ldb,
                       float(np.real(beta)), C.ctypes.data, ldc, tm, tn, tk)
    elif dt == np.float32:
        L.oracle_sgemm(ta, tb, m, n, k, float(np.real(alpha)), A.ctypes.data, lda, B.ctypes.data, ldb,
                       float(np.real(beta)), C.ctypes.data, ldc, tm, tn, tk)
    elif dt == np.complex64:
        al, be = _pairs(np.complex64, alpha), _pairs(np.complex64, beta)
        L.oracle_cgemm(ta, tb, m, n, k, al.ctypes.data, A.ctypes.data, lda, B.ctypes.data, ldb, be.ctypes.data,
                       C.ctypes.data, ldc, tm, tn, tk)
    elif dt == np.complex128:
        al, be = _pairs(np.complex128, alpha), _pairs(np.complex128, beta)
        L.oracle_zgemm(ta, tb, m, n, k, al.ctypes.data, A.ctypes.data, lda, B.ctypes.data, ldb, be.ctypes.data,
                       C.ctypes.data, ldc, tm, tn, tk)
    else:
        raise TypeError(dt)


def _view(X, ld, rows, cols):
    """rows x cols column-major view of a buffer with leading dimension ld."""
    flat = X.reshape(-1, order="F")
    return np.lib.stride_tricks.as_strided(flat, shape=(rows, cols), strides=(flat.itemsize, flat.itemsize * ld))


def tiled_gemm_np(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tm, tn, tk):
    """Round-robin tiled semantics with numpy per tile (in place on C).

    For each C tile and k-tile kt: C_t = alpha*op(A_t)op(B_t) + beta'*C_t, beta' = beta if kt == 0
    else 1, C_t read only when beta' != 0 (Tiled-MM src/Tiled-MM/tiled_mm.cpp:270-365)."""
    ta, tb = ta.upper(), tb.upper()
    ta = ta if ta in "TC" else "N"
    tb = tb if tb in "TC" else "N"
    if m <= 0 or n <= 0:
        return
    ar, ac = op_shape(ta, m, k)
    br, bc = op_shape(tb, k, n)
    Av = apply_op(ta, _view(A, lda, ar, ac))
    Bv = apply_op(tb, _view(B, ldb, br, bc))
    Cv = _view(C, ldc, m, n)
    tm, tn = min(tm, m), min(tn, n)
    tk = min(tk, k) if k > 0 else 1
    for i0 in range(0, m, tm):
        for j0 in range(0, n, tn):
            ct = Cv[i0:i0 + tm, j0:j0 + tn]
            acc = ct.copy() if beta != 0 else np.zeros_like(ct)
            if k == 0:
                acc = beta * acc if beta != 0 else acc
            for kt, k0 in enumerate(range(0, k, tk)):
                bp = beta if kt == 0 else 1
                if alpha == 0:  # BLAS quick path: A and B are not read
                    acc = np.zeros_like(acc) if bp == 0 else bp * acc
                    continue
                prod = Av[i0:i0 + tm, k0:k0 + tk] @ Bv[k0:k0 + tk, j0:j0 + tn]
                acc = alpha * prod if bp == 0 else alpha * prod + bp * acc
            ct[...] = acc


def host_blas_gemm(ta, tb, alpha, A, B, beta, C):
    """alpha*op(A)*op(B) + beta*C on whole (Fortran-ordered) matrices; C not read when beta == 0,
    A and B not read when alpha == 0 (BLAS)."""
    if alpha == 0:
        return np.zeros_like(C) if beta == 0 else beta * C
    prod = apply_op(ta, A) @ apply_op(tb, B)
    return alpha * prod if beta == 0 else alpha * prod + beta * C


def error_scale(ta, tb, alpha, A, B, beta, C0):
    """| |alpha| |op(A)| |op(B)| + |beta| |C0| | elementwise (the BLAS error scale)."""
    s = abs(alpha) * (np.abs(apply_op(ta, A)) @ np.abs(apply_op(tb, B)))
    if beta != 0:
        s = s + abs(beta) * np.abs(C0)
    return s


def normwise_error(C, Cref, scale):
    den = np.linalg.norm(scale)
    return float(np.linalg.norm(C - Cref) / den) if den > 0 else float(np.linalg.norm(C - Cref))


def componentwise_error(C, Cref, scale):
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.abs(C - Cref) / scale
    r = np.where(scale > 0, r, np.abs(C - Cref))
    return float(np.max(r)) if r.size else 0.0


def unit_roundoff(dtype):
    return float(np.finfo(np.dtype(dtype)).eps) / 2

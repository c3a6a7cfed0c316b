"""Shared helpers of the test-suite: data in the reference test's style and error checks."""
import numpy as np

import oracle

NP = {"s": np.float32, "d": np.float64, "c": np.complex64, "z": np.complex128}


def int_matrix(rng, rows, cols, dtype, ld=None):
    """Integer entries in [0, 9], as tests/test-multiply.cpp:58-66 fills its matrices."""
    ld = rows if ld is None else ld
    X = np.zeros((ld, cols), dtype=dtype, order="F")
    X[:rows, :] = rng.integers(0, 10, size=(rows, cols))
    if np.iscomplexobj(X):
        X[:rows, :] += 1j * rng.integers(0, 10, size=(rows, cols))
    return X


def uniform_matrix(rng, rows, cols, dtype, ld=None, lo=-1.0, hi=1.0):
    ld = rows if ld is None else ld
    X = np.zeros((ld, cols), dtype=dtype, order="F")
    X[:rows, :] = rng.uniform(lo, hi, size=(rows, cols))
    if np.iscomplexobj(X):
        X[:rows, :] += 1j * rng.uniform(lo, hi, size=(rows, cols))
    if ld > rows:  # padding rows must never be read: poison them
        X[rows:, :] = np.nan
    return X


def bound(dtype, k):
    """Stated normwise bound: ||C - C_ref||_F <= 2 k u || |alpha||A||B| + |beta||C0| ||_F."""
    return 2 * max(k, 1) * oracle.unit_roundoff(np.dtype(dtype).type(0).real.dtype)


def check_normwise(C, Cref, ta, tb, alpha, A, B, beta, C0, k, dtype):
    scale = oracle.error_scale(ta, tb, alpha, A, B, beta, C0)
    err = oracle.normwise_error(C, Cref, scale)
    assert err <= bound(dtype, k), f"normwise error {err:.3e} above bound {bound(dtype, k):.3e}"
    return err

"""Failure paths of the BLAS-ABI interposer (libtmm_blas.so): an unmodified program that calls
dgemm_ must never continue with a wrong C. When the GPU path fails before C was written, the call
is re-run on the next BLAS in symbol order; with no next BLAS, or once C may be partly written, the
program stops (abort), as the reference BLAS stops in xerbla.

The caller (tests/cpp/blas_caller.c) is linked against libtmm_blas.so alone
(lib/blas-caller-alone) or with a plain host dgemm_ after it (tests/cpp/naive_dgemm.c,
lib/blas-caller-next); it checks C itself and prints "BLAS CALLER OK".
"""
import os
import signal
import subprocess

import pytest

import tiledmm

NEXT = os.path.join(tiledmm.LIB_DIR, "blas-caller-next")
ALONE = os.path.join(tiledmm.LIB_DIR, "blas-caller-alone")
no_gpu_only = pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the path taken where no GPU is present")


def _run(exe, **env):
    assert os.path.exists(exe), "build the library first (make -C tiled-mm_amd)"
    e = dict(os.environ, TMM_BLAS_MIN_FLOPS="0", **env)
    return subprocess.run([exe], capture_output=True, text=True, timeout=120, env=e)


@no_gpu_only
def test_no_gpu_reruns_on_next_blas():
    r = _run(NEXT)
    assert r.returncode == 0 and "BLAS CALLER OK" in r.stdout, r.stdout + r.stderr
    assert "naive dgemm_ called" in r.stdout
    assert "running on the host BLAS" in r.stderr


@no_gpu_only
def test_no_gpu_and_no_next_blas_stops():
    r = _run(ALONE)
    assert r.returncode == -signal.SIGABRT, (r.returncode, r.stdout, r.stderr)
    assert "BLAS CALLER" not in r.stdout  # the caller never saw a return with a wrong C
    assert "no host BLAS" in r.stderr


@pytest.mark.gpu
def test_gpu_runs_the_product(torch_gpu):
    r = _run(NEXT)
    assert r.returncode == 0 and "BLAS CALLER OK" in r.stdout, r.stdout + r.stderr
    assert "naive dgemm_ called" not in r.stdout


@pytest.mark.gpu
def test_gpu_planning_failure_reruns_on_next_blas(torch_gpu):
    """A 1 MiB device work pool cannot hold one tile: the call fails in planning, before any copy,
    so C is unchanged and the next BLAS computes it."""
    r = _run(NEXT, TMM_BLAS_POOL_LIMIT_MB="1")
    assert r.returncode == 0 and "BLAS CALLER OK" in r.stdout, r.stdout + r.stderr
    assert "naive dgemm_ called" in r.stdout
    assert "budget too small" in r.stderr and "C is unchanged" in r.stderr

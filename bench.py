#!/usr/bin/env python3
"""Host-to-host FP64 GEMM throughput of the out-of-core path at 40000^3.

Workload (the reference's headline out-of-core case, sized for MI355X): C = alpha*A^T*B + beta*C
with host-resident column-major FP64 matrices, m = n = k = 40000, leading dimensions 40064 (the
64 padding rows hold NaN and must never be read), alpha = beta = 1, entries uniform[-1, 1] from
fixed seeds, gpu::make_context(4 streams, 8192^3 tiles) -> the reference tile policy picks
8000^3 tiles; copy_c_back = true. One step = one gemm call, timed from entry with the data on
the host to return with C on the host.

  value        : TFLOP/s (2mnk / wall time per call), host matrices in pinned memory
  variants     : the same with pageable host memory (streamed through pinned staging)
  roofline     : the dominant kernel (tile DGEMM): flops per launch / mean launch duration, both
                 measured over the timed steps with HIP events on the kernel's own stream,
                 against the FP64 matrix peak (78.6 TFLOP/s)
  host_roofline: T_lb = max(F/P, H2D bytes/BW, D2H bytes/BW) of the host-to-host call (PCIe 5.0
                 x16 at 64 GB/s per direction), and T_lb / T
  cpu_baseline : the host BLAS (numpy) on the box's cores for a bounded column sample of the
                 same product
  check        : sampled columns of the first call's C against the host BLAS: normwise error
                 relative to |A^T||B| + |C0|
"""
import argparse
import json
import os
import sys
import time

NCORES = len(os.sched_getaffinity(0))
for _v in ("OPENBLAS_NUM_THREADS", "OMP_NUM_THREADS", "MKL_NUM_THREADS"):
    os.environ[_v] = str(NCORES)

import numpy as np  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tiled-mm_amd"))
import tiledmm  # noqa: E402

P_FP64 = 78.6e12  # MI355X dense FP64 matrix peak
PCIE_BW = 64e9  # PCIe 5.0 x16, per direction
METRIC = "host-to-host FP64 TFLOP/s @40000^3; Ozaki-emulated vs native at equal error"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def alloc_host(torch, n_elems, pinned):
    if pinned:
        t = torch.empty(n_elems, dtype=torch.float64, pin_memory=True)
        return t, t.numpy()
    a = np.empty(n_elems, dtype=np.float64)
    return a, a


def make_problem(torch, size, ld, pinned):
    """A (k x m, TN), B (k x n), C (m x n), all with leading dimension ld, padding rows NaN."""
    mats = []
    for seed in (42, 43, 44):
        keep, flat = alloc_host(torch, ld * size, pinned)
        tiledmm.fill_uniform((flat.ctypes.data, np.float64), seed, -1.0, 1.0, rows=size, cols=size, ld=ld)
        M = flat.reshape(size, ld).T  # Fortran view (ld, size)
        M[size:, :] = np.nan
        mats.append((keep, flat, M))
    return mats


def sample_check(A, B, C_after, C0_cols, cols, size):
    """Normwise and componentwise error of C[:, cols] against the host BLAS."""
    At = A[:size, :].T  # m x k view of the stored k x m
    Bc = np.ascontiguousarray(B[:size, cols])
    ref = At @ Bc + C0_cols
    scale = np.abs(At) @ np.abs(Bc) + np.abs(C0_cols)
    got = C_after[:size, cols]
    err = float(np.linalg.norm(got - ref) / np.linalg.norm(scale))
    comp = float(np.max(np.abs(got - ref) / scale))
    return err, comp, bool(np.all(np.isfinite(got)))


def run_variant(torch, size, ld, pinned, steps, warmup, streams, tile, check_cols):
    log(f"[bench] {'pinned' if pinned else 'pageable'}: allocating and filling 3 x {ld * size * 8 / 1e9:.1f} GB")
    t0 = time.time()
    (ka, _, A), (kb, _, B), (kc, _, C) = make_problem(torch, size, ld, pinned)
    log(f"[bench] data ready in {time.time() - t0:.1f}s")
    h = tiledmm.Handle(tiledmm.F64, streams, tile, tile, tile)
    h.set_option(tiledmm.OPT_PROFILE, 1)
    C0_cols = np.ascontiguousarray(C[:size, check_cols])
    res = {}
    for w in range(max(warmup, 1)):
        torch.cuda.synchronize()
        t = time.time()
        h.gemm("T", "N", size, size, size, 1.0, A, ld, B, ld, 1.0, C, ld, pin=False, copy_back=True)
        torch.cuda.synchronize()
        log(f"[bench] warmup call {w}: {time.time() - t:.3f}s  stats {h.stats()}")
        if w == 0:
            err, comp, finite = sample_check(A, B, C, C0_cols, check_cols, size)
            res["check"] = {"columns": len(check_cols), "normwise": err, "componentwise": comp, "finite": finite,
                            "bound": 2 * size * 2.0 ** -53}
            log(f"[bench] sampled-column check: {res['check']}")
    kern_s = kern_n = kern_f = 0.0
    torch.cuda.synchronize()
    t = time.time()
    walls = []
    for _ in range(steps):
        t1 = time.time()
        h.gemm("T", "N", size, size, size, 1.0, A, ld, B, ld, 1.0, C, ld, pin=False, copy_back=True)
        walls.append(time.time() - t1)
        st = h.stats()
        kern_s += st["kernel_s"]
        kern_n += st["kernel_launches"]
        kern_f += st["kernel_flops"]
    torch.cuda.synchronize()
    total = time.time() - t
    st = h.stats()
    res.update({"ms_per_step": 1e3 * total / steps, "walls_s": walls, "stats": st,
                "kernel_mean_s": kern_s / max(kern_n, 1), "kernel_flops_per_launch": kern_f / max(kern_n, 1),
                "kernel_busy_frac": kern_s / total})
    h.close()
    del ka, kb, kc, A, B, C
    return res


def cpu_baseline(size, ld, target_s=12.0):
    """Host BLAS (numpy) on NCORES threads: A^T (size x size) times a column sample of B."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    rng_a = np.empty((size, ld), dtype=np.float64)  # C-order (size, ld) == Fortran (ld, size)
    tiledmm.fill_uniform((rng_a.ctypes.data, np.float64), 42, -1.0, 1.0, rows=size, cols=size, ld=ld)
    A = rng_a.T[:size, :]
    cal = 256
    B = np.empty((size, cal), dtype=np.float64, order="F")
    tiledmm.fill_uniform(B, 43)
    C = np.zeros((size, cal), order="F")
    oracle.host_blas_gemm("T", "N", 1.0, A, B[:, :32], 1.0, C[:, :32])  # warm-up
    t = time.time()
    oracle.host_blas_gemm("T", "N", 1.0, A, B, 1.0, C)
    dt = time.time() - t
    ncols = int(min(size, max(cal, cal * target_s / max(dt, 1e-3))))
    B = np.empty((size, ncols), dtype=np.float64, order="F")
    tiledmm.fill_uniform(B, 43)
    C = np.zeros((size, ncols), order="F")
    t = time.time()
    oracle.host_blas_gemm("T", "N", 1.0, A, B, 1.0, C)
    dt = time.time() - t
    flops = 2.0 * size * size * ncols
    cfg = np.__config__.CONFIG if hasattr(np.__config__, "CONFIG") else {}
    blas = cfg.get("Build Dependencies", {}).get("blas", {}) if isinstance(cfg, dict) else {}
    return {"value": flops / dt / 1e12, "unit": "TFLOP/s", "cores": NCORES, "kind": "port",
            "sample": f"host BLAS dgemm C[:, :{ncols}] += A^T B[:, :{ncols}] with m=k={size} (TN, ld {ld}), "
                      f"{flops / 1e12:.2f} TFLOP in {dt:.1f}s",
            "provider": f"numpy {np.__version__} / {blas.get('name', 'BLAS')} {blas.get('version', '')}".strip()}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=40000)
    ap.add_argument("--ld", type=int, default=None)
    ap.add_argument("--streams", type=int, default=4)
    ap.add_argument("--tile", type=int, default=8192)
    ap.add_argument("--variant", choices=["pinned", "pageable", "both"], default="both")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the out-of-core path runs on one "
              "device and no multi-GPU figure is produced.", file=sys.stderr)
        sys.exit(2)
    size = args.size
    ld = args.ld if args.ld is not None else size + 64
    import torch

    torch.cuda.set_device(0)
    rng = np.random.default_rng(1)
    check_cols = np.sort(rng.choice(size, size=min(16, size), replace=False))

    results = {}
    if args.variant in ("pinned", "both"):
        results["pinned"] = run_variant(torch, size, ld, True, args.steps, args.warmup, args.streams, args.tile,
                                        check_cols)
    if args.variant in ("pageable", "both"):
        results["pageable"] = run_variant(torch, size, ld, False, max(1, args.steps - 1), args.warmup, args.streams,
                                          args.tile, check_cols)
    main_key = "pinned" if "pinned" in results else "pageable"
    R = results[main_key]
    flops = 2.0 * size ** 3
    ms = R["ms_per_step"]
    value = flops / (ms / 1e3) / 1e12
    st = R["stats"]
    achieved = R["kernel_flops_per_launch"] / R["kernel_mean_s"] / 1e12 if R["kernel_mean_s"] > 0 else None
    t_lb = max(flops / P_FP64, st["h2d_bytes"] / PCIE_BW, st["d2h_bytes"] / PCIE_BW)
    alg_h2d = 8.0 * 3 * size * size  # A, B and C (beta != 0) once
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: uniform[-1,1] from fixed seeds (42, 43, 44), padding rows NaN",
        "config": {"workload": f"DGEMM TN m=n=k={size} lda=ldb=ldc={ld} alpha=beta=1, host-to-host, "
                               f"{main_key} host memory, make_context({args.streams}, {args.tile}, {args.tile}, "
                               f"{args.tile}) -> tiles {int(st['tile_m'])}x{int(st['tile_n'])}x{int(st['tile_k'])}, "
                               "copy_c_back=true",
                   "m": size, "n": size, "k": size, "ld": ld, "trans": "TN", "streams": args.streams,
                   "tile_request": args.tile},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3) if achieved else None, "peak": 78.6,
                     "unit": "TFLOP/s", "frac": round(achieved / 78.6, 4) if achieved else None, "traffic": None,
                     "kernel": "tile DGEMM (gemm_real_kernel<double>, TN)",
                     "flops_per_launch": R["kernel_flops_per_launch"], "mean_launch_ms": R["kernel_mean_s"] * 1e3,
                     "launches_timed": int(st["kernel_launches"]) * args.steps},
        "host_roofline": {"t_lb_s": t_lb, "t_s": ms / 1e3, "frac": round(t_lb / (ms / 1e3), 4),
                          "h2d_bytes": st["h2d_bytes"], "d2h_bytes": st["d2h_bytes"],
                          "h2d_algorithmic_bytes": alg_h2d,
                          "reference_schedule_h2d_bytes": 140.8e9 if size == 40000 else None,
                          "h2d_GBps_avg": st["h2d_bytes"] / (ms / 1e3) / 1e9},
        "check": R.get("check"),
        "variants": {k: {"value": round(flops / (v["ms_per_step"] / 1e3) / 1e12, 3), "unit": "TFLOP/s",
                         "ms_per_step": round(v["ms_per_step"], 2), "check": v.get("check"),
                         "kernel_busy_frac": round(v["kernel_busy_frac"], 4)} for k, v in results.items()},
    }
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (host BLAS) ...")
        out["cpu_baseline"] = cpu_baseline(size, ld)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

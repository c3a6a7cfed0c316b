#!/usr/bin/env python3
"""FP64 GEMM throughput of the out-of-core GEMM library at 40000^3.

Workload (the reference's headline out-of-core case, sized for MI355X): C = alpha*A^T*B + beta*C,
column-major FP64, m = n = k = 40000, leading dimensions 40064 (the 64 padding rows hold NaN and
must never be read), alpha = beta = 1, entries uniform[-1, 1] from fixed seeds.

  value         : TFLOP/s of one step = the whole 40000^3 product with A, B and C already resident
                  in HBM (tmm_device_gemm), timed with HIP events on the launching stream
  roofline      : the dominant kernel (the tile DGEMM, here one launch per step): flops per launch /
                  mean launch duration against the FP64 matrix peak (78.6 TFLOP/s)
  host_to_host  : the same product through gpu::gemm with host-resident matrices
                  (make_context(4 streams, 8192^3 tiles) -> 8000^3 tiles, copy_c_back = true), wall
                  time per call from entry with the data on the host to return with C on the host,
                  for pinned and pageable host memory; with the host-link roofline
                  T_lb = max(F/P, H2D bytes/BW, D2H bytes/BW), PCIe 5.0 x16 at 64 GB/s per direction
  cpu_baseline  : the host BLAS (numpy) on the box's cores for a bounded column sample of the product
  check         : sampled columns of C after the first call against the host BLAS: normwise error
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
P_INT8 = 5.0e15  # MI355X dense INT8 matrix peak (ops/s)
PCIE_BW = 64e9  # PCIe 5.0 x16, per direction
METRIC = "host-to-host FP64 TFLOP/s @40000^3; Ozaki-emulated vs native at equal error"


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "dgemm_dlds_tn_40000_device.md")


def measured_traffic(size):
    """HBM bytes per launch of the 40000^3 device-resident kernel, as measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, each scaled by a copy kernel of known size) and recorded in the
    profile summary; None for other sizes or when the summary is absent."""
    if size != 40000 or not os.path.exists(TRAFFIC_PROFILE):
        return None
    for line in open(TRAFFIC_PROFILE):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


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


def sample_check(A, B, got, C0_cols, cols, size, chunk=2048):
    """Normwise and componentwise error of C[:, cols] against the host BLAS (A^T B + C0)."""
    Bc = np.ascontiguousarray(B[:size, cols])
    ref = np.empty((size, len(cols)))
    scale = np.empty((size, len(cols)))
    absB = np.abs(Bc)
    for i0 in range(0, size, chunk):  # rows i0.. of A^T = stored columns of A
        blk = A[:size, i0:i0 + chunk].T
        ref[i0:i0 + chunk] = blk @ Bc
        scale[i0:i0 + chunk] = np.abs(blk) @ absB
    ref += C0_cols
    scale += np.abs(C0_cols)
    err = float(np.linalg.norm(got - ref) / np.linalg.norm(scale))
    comp = float(np.max(np.abs(got - ref) / scale))
    return err, comp, bool(np.all(np.isfinite(got)))


def run_variant(torch, mats, size, ld, pinned, steps, warmup, streams, tile, check_cols, moduli=0):
    (ka, _, A), (kb, _, B), (kc, _, C) = mats
    h = tiledmm.Handle(tiledmm.F64, streams, tile, tile, tile)
    h.set_option(tiledmm.OPT_PROFILE, 1)
    h.set_option(tiledmm.OPT_OZAKI_MODULI, moduli)
    C0_cols = np.ascontiguousarray(C[:size, check_cols])
    res = {}
    for w in range(max(warmup, 1)):
        torch.cuda.synchronize()
        t = time.time()
        h.gemm("T", "N", size, size, size, 1.0, A, ld, B, ld, 1.0, C, ld, pin=False, copy_back=True)
        torch.cuda.synchronize()
        log(f"[bench] warmup call {w}: {time.time() - t:.3f}s  stats {h.stats()}")
        if w == 0:
            err, comp, finite = sample_check(A, B, C[:size, check_cols], C0_cols, check_cols, size)
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
    res.update({"ms_per_step": 1e3 * total / steps, "walls_s": walls, "stats": st, "ref_tiles": h.tile_sizes(),
                "kernel_mean_s": kern_s / max(kern_n, 1), "kernel_flops_per_launch": kern_f / max(kern_n, 1),
                "kernel_busy_frac": kern_s / total})
    h.close()
    del ka, kb, kc, A, B, C
    return res


def run_device(torch, mats, size, ld, steps, warmup, check_cols, moduli=0):
    """The whole product on HBM-resident A, B, C: one tile-kernel launch per step (moduli > 0: the
    Ozaki scheme II emulation with that many INT8 products instead)."""
    (_, fa, A), (_, fb, B), (_, fc, C) = mats
    dev = [torch.empty(ld * size, dtype=torch.float64, device="cuda") for _ in range(3)]
    for d, f in zip(dev, (fa, fb, fc)):
        d.copy_(torch.from_numpy(f))
    dA, dB, dC = dev
    C0_cols = np.ascontiguousarray(C[:size, check_cols])
    stream = torch.cuda.Stream()
    h = tiledmm.Handle(tiledmm.F64, 1)
    h.set_option(tiledmm.OPT_OZAKI_MODULI, moduli)
    res = {}
    for w in range(max(warmup, 1)):
        h.device_gemm("T", "N", size, size, size, 1.0, dA, ld, dB, ld, 1.0, dC, ld, stream=stream.cuda_stream)
        stream.synchronize()
        if w == 0:
            got = dC.view(size, ld)[torch.as_tensor(check_cols, device="cuda"), :size].t().cpu().numpy()
            err, comp, finite = sample_check(A, B, got, C0_cols, check_cols, size)
            res["check"] = {"columns": len(check_cols), "normwise": err, "componentwise": comp, "finite": finite,
                            "bound": 2 * size * 2.0 ** -53}
            log(f"[bench] device-resident check: {res['check']}")
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        h.device_gemm("T", "N", size, size, size, 1.0, dA, ld, dB, ld, 1.0, dC, ld, stream=stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    res["ms_per_step"] = e0.elapsed_time(e1) / steps
    res["int8_ops_per_step"] = h.stats()["int8_ops"]
    h.close()
    del dA, dB, dC, dev
    torch.cuda.empty_cache()
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


# ------------------------------------------------------------------------------------------------
# Other configurations (not the default bench line; `--workload` selects one)
# ------------------------------------------------------------------------------------------------
def workload_zgemm(torch, args):
    """ZGEMM CN m=n=24000 k=32000, alpha=(1,0.5), beta=(0.5,-1), tiles 6144^3, 3 streams."""
    m, n, k = (args.size_m or 24000), (args.size_n or 24000), (args.size_k or 32000)
    alpha, beta = complex(1, 0.5), complex(0.5, -1)
    flops = 8.0 * m * n * k
    mats = []
    for seed, rows, cols in ((42, k, m), (43, k, n), (44, m, n)):
        t = torch.empty(2 * rows * cols, dtype=torch.float64, pin_memory=True)
        flat = t.numpy()
        tiledmm.util().tmmu_fill_uniform_z(flat.ctypes.data, rows, cols, rows, 0, 0, rows, seed, -1.0, 1.0, 0)
        mats.append((t, flat.view(np.complex128).reshape(cols, rows).T))
    (tA, A), (tB, B), (tC, C) = mats
    cols = np.arange(0, n, max(1, n // 8))
    C0 = np.ascontiguousarray(C[:, cols])
    # device-resident
    dA, dB, dC = (torch.from_numpy(x.T.reshape(-1)).cuda() for x in (A, B, C))
    stream = torch.cuda.Stream()
    h = tiledmm.Handle(tiledmm.C64, 1)
    h.device_gemm("C", "N", m, n, k, alpha, dA, k, dB, k, beta, dC, m, stream=stream.cuda_stream)
    stream.synchronize()
    got = dC.view(n, m)[torch.as_tensor(cols, device="cuda")].t().cpu().numpy()
    dC.copy_(torch.from_numpy(C.T.reshape(-1)))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        h.device_gemm("C", "N", m, n, k, alpha, dA, k, dB, k, beta, dC, m, stream=stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    dev_ms = e0.elapsed_time(e1) / args.steps
    h.close()
    del dA, dB, dC
    torch.cuda.empty_cache()
    # host-to-host (the host BLAS reference is computed after all timed calls)
    h = tiledmm.Handle(tiledmm.C64, 3, 6144, 6144, 6144)
    h.set_option(tiledmm.OPT_PROFILE, 1)
    h.set_option(tiledmm.OPT_TILE_GROUPS, args.tile_groups)
    h.gemm("C", "N", m, n, k, alpha, A, k, B, k, beta, C, m)
    got_h2h = np.ascontiguousarray(C[:, cols])
    t0 = time.time()
    for _ in range(args.steps):
        h.gemm("C", "N", m, n, k, alpha, A, k, B, k, beta, C, m)
    h2h_s = (time.time() - t0) / args.steps
    st = h.stats()
    h.close()
    Bc = np.ascontiguousarray(B[:, cols])
    ref = np.empty((m, len(cols)), dtype=np.complex128)
    scale = np.empty((m, len(cols)))
    for i0 in range(0, m, 2048):  # A^H b = conj(A^T conj(b)), in column blocks of A
        blk = A[:, i0:i0 + 2048]
        ref[i0:i0 + 2048] = np.conj(blk.T @ np.conj(Bc))
        scale[i0:i0 + 2048] = np.abs(blk).T @ np.abs(Bc)
    ref = alpha * ref + beta * C0
    scale = abs(alpha) * scale + abs(beta) * np.abs(C0)
    err_dev = float(np.linalg.norm(got - ref) / np.linalg.norm(scale))
    err_h2h = float(np.linalg.norm(got_h2h - ref) / np.linalg.norm(scale))
    return {"workload": f"ZGEMM CN m={m} n={n} k={k} alpha=(1,0.5) beta=(0.5,-1)", "unit": "TFLOP/s",
            "device_resident": {"value": flops / (dev_ms / 1e3) / 1e12, "ms": dev_ms, "frac_of_78.6": flops / (dev_ms / 1e3) / 78.6e12,
                                "normwise_error_sampled": err_dev},
            "host_to_host_pinned": {"value": flops / h2h_s / 1e12, "ms": h2h_s * 1e3, "tiles": [st["tile_m"], st["tile_n"], st["tile_k"]],
                                    "h2d_bytes": st["h2d_bytes"], "d2h_bytes": st["d2h_bytes"],
                                    "tile_kernel_TFLOPs": st["kernel_flops"] / st["kernel_s"] / 1e12 if st["kernel_s"] else None,
                                    "normwise_error_sampled": err_h2h},
            "bound": 2 * k * 2.0 ** -53}


def workload_ozaki(torch, args):
    """NN 40000^3, entries (u-0.5)*exp(phi*g), beta = 0: Ozaki vs native on HBM-resident data."""
    size = args.size
    out = {"workload": f"DGEMM NN {size}^3 alpha=1 beta=0, entries (uniform[-0.5,0.5])*exp(phi*N(0,1))",
           "unit": "TFLOP/s (FP64-equivalent 2mnk)", "results": []}
    dA = torch.empty(size * size, dtype=torch.float64, device="cuda")
    dB = torch.empty(size * size, dtype=torch.float64, device="cuda")
    dC = torch.empty(size * size, dtype=torch.float64, device="cuda")
    host = np.empty((size, size), order="F")
    cols = np.arange(0, size, max(1, size // 8))
    u = tiledmm.util()
    stream = torch.cuda.Stream()
    for phi in args.phis:
        u.tmmu_fill_range_d(host.ctypes.data, size, size, size, 0, 0, size, 42, phi, 0)
        A = host.copy(order="F")
        dA.copy_(torch.from_numpy(A.T.reshape(-1)))
        u.tmmu_fill_range_d(host.ctypes.data, size, size, size, 0, 0, size, 43, phi, 0)
        Bc = np.ascontiguousarray(host[:, cols])
        dB.copy_(torch.from_numpy(host.T.reshape(-1)))
        ref = np.empty((size, len(cols)))
        scale = np.empty((size, len(cols)))
        for i0 in range(0, size, 4096):
            ref[i0:i0 + 4096] = A[i0:i0 + 4096] @ Bc
            scale[i0:i0 + 4096] = np.abs(A[i0:i0 + 4096]) @ np.abs(Bc)
        del A
        runs = [("native", 0)] + [("crt", mm) for mm in args.moduli] + [("slices", sl) for sl in args.slices]
        for mode, count in runs:
            slices = count if mode == "slices" else 0
            h = tiledmm.Handle(tiledmm.F64, 1)
            h.set_option(tiledmm.OPT_OZAKI_SLICES, slices)
            h.set_option(tiledmm.OPT_OZAKI_MODULI, count if mode == "crt" else 0)
            h.device_gemm("N", "N", size, size, size, 1.0, dA, size, dB, size, 0.0, dC, size,
                          stream=stream.cuda_stream)
            stream.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(args.steps):
                h.device_gemm("N", "N", size, size, size, 1.0, dA, size, dB, size, 0.0, dC, size,
                              stream=stream.cuda_stream)
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / args.steps
            int8 = h.stats()["int8_ops"]
            h.close()
            got = dC.view(size, size)[torch.as_tensor(cols, device="cuda")].t().cpu().numpy()
            err = float(np.linalg.norm(got - ref) / np.linalg.norm(scale))
            comp = float(np.max(np.abs(got - ref) / scale))
            r = {"phi": phi, "mode": mode, "slices" if mode == "slices" else "moduli": count, "ms": ms,
                 "TFLOPs_fp64_equiv": 2.0 * size ** 3 / (ms / 1e3) / 1e12, "normwise_error": err,
                 "componentwise_error": comp, "int8_TOPs": int8 / (ms / 1e3) / 1e12 if int8 else None}
            log(f"[bench] ozaki {r}")
            out["results"].append(r)
    # equal error: per phi, the fastest emulation whose normwise error is not above native's
    out["equal_error"] = []
    for phi in args.phis:
        rs = [r for r in out["results"] if r["phi"] == phi]
        nat = next(r for r in rs if r["mode"] == "native")
        ok = [r for r in rs if r["mode"] != "native" and r["normwise_error"] <= nat["normwise_error"]]
        best = max(ok, key=lambda r: r["TFLOPs_fp64_equiv"]) if ok else None
        out["equal_error"].append({"phi": phi, "native_TFLOPs": nat["TFLOPs_fp64_equiv"],
                                   "native_normwise_error": nat["normwise_error"],
                                   "emulation": None if best is None else {k: best[k] for k in best if k != "phi"},
                                   "speedup": None if best is None else best["TFLOPs_fp64_equiv"] / nat["TFLOPs_fp64_equiv"]})
    return out


def workload_batched(torch, args):
    """512 NN problems, m,n,k log-uniform in [64, 4096] (seed 42), alpha=beta=1, pinned host."""
    rng = np.random.default_rng(42)
    dims = np.round(np.exp(rng.uniform(np.log(64), np.log(4096), size=(512, 3)))).astype(int)
    keep, As, Bs, Cs = [], [], [], []
    flops = 0.0
    for i, (m, n, k) in enumerate(dims):
        for (rows, cols, lst, seed) in ((m, k, As, 3 * i), (k, n, Bs, 3 * i + 1), (m, n, Cs, 3 * i + 2)):
            t = torch.empty(int(rows * cols), dtype=torch.float64, pin_memory=True)
            arr = t.numpy()
            tiledmm.fill_uniform((arr.ctypes.data, np.float64), 1000 + seed, -1.0, 1.0, rows=int(rows), cols=int(cols),
                                 ld=int(rows))
            keep.append(t)
            lst.append(arr.reshape(int(cols), int(rows)).T)
        flops += 2.0 * m * n * k
    check = [0, 1, 2, 511]
    C0 = {i: Cs[i].copy() for i in check}
    h = tiledmm.Handle(tiledmm.F64, 2)
    h.set_option(tiledmm.OPT_BATCH_GRAPH, args.batch_graph)
    ms_, ns_, ks_ = [int(x) for x in dims[:, 0]], [int(x) for x in dims[:, 1]], [int(x) for x in dims[:, 2]]
    args_ = (ms_, ns_, ks_, [1.0] * 512, As, ms_, Bs, ks_, [1.0] * 512, Cs, ms_)
    h.gemm_batched("N", "N", *args_)
    errs = []
    for i in check:
        ref = As[i] @ Bs[i] + C0[i]
        scale = np.abs(As[i]) @ np.abs(Bs[i]) + np.abs(C0[i])
        errs.append(float(np.linalg.norm(Cs[i] - ref) / np.linalg.norm(scale)))
    t0 = time.time()
    for _ in range(args.steps):
        h.gemm_batched("N", "N", *args_)
    t = (time.time() - t0) / args.steps
    cfg = h.batch_config()
    st = h.stats()
    h.close()
    h2d = sum(8.0 * (m * k + k * n + m * n) for m, n, k in dims)
    d2h = sum(8.0 * m * n for m, n, k in dims)
    t_lb = max(flops / P_FP64, h2d / PCIE_BW, d2h / PCIE_BW)
    return {"workload": "batched DGEMM: 512 NN problems, m,n,k log-uniform in [64,4096] (seed 42), alpha=beta=1, "
                        "pinned host memory", "unit": "TFLOP/s", "value": flops / t / 1e12, "ms": t * 1e3,
            "flops": flops, "h2d_bytes": h2d, "d2h_bytes": d2h, "t_lb_s": t_lb, "host_roofline_frac": t_lb / t,
            "autotuner": cfg, "graph": st["batch_graph"], "groups": st["batch_groups"],
            "normwise_error_checked_problems": errs}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=40000)
    ap.add_argument("--ld", type=int, default=None)
    ap.add_argument("--streams", type=int, default=4)
    ap.add_argument("--tile", type=int, default=8192)
    ap.add_argument("--variant", choices=["pinned", "pageable", "both", "none"], default="both",
                    help="host memory kinds for the host-to-host line (none: device-resident value only)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=["dgemm", "zgemm", "ozaki", "batched"], default="dgemm",
                    help="dgemm: the default bench line; the others print their own JSON line")
    ap.add_argument("--size-m", type=int, default=None)
    ap.add_argument("--size-n", type=int, default=None)
    ap.add_argument("--size-k", type=int, default=None)
    ap.add_argument("--phis", type=float, nargs="*", default=[0.5, 2.0, 4.0])
    ap.add_argument("--slices", type=int, nargs="*", default=[8, 9])
    ap.add_argument("--moduli", type=int, nargs="*", default=[14, 15, 16, 17, 18, 19, 20])
    ap.add_argument("--batch-graph", type=int, default=-1, choices=[-1, 0, 1],
                    help="batched workload: hipGraph capture chosen by the autotuner (-1), always (1), never (0)")
    ap.add_argument("--tile-groups", type=int, default=1, choices=[0, 1],
                    help="zgemm workload, host-to-host: grouped tile launches of the steady k-chunks (1) or one "
                         "launch per tile (0)")
    ap.add_argument("--crt-moduli", type=int, default=16,
                    help="moduli of the Ozaki scheme II line (0: skip it)")
    ap.add_argument("--copy-probe", type=float, default=None, metavar="GIB",
                    help="only run 3 device-to-device copies of GIB GiB and exit: the known-size kernel that "
                         "scales the HBM byte counters (FETCH_SIZE, WRITE_SIZE) read under rocprofv3 --pmc")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the out-of-core path runs on one "
              "device and no multi-GPU figure is produced.", file=sys.stderr)
        sys.exit(2)
    size = args.size
    ld = args.ld if args.ld is not None else size + 64
    import torch

    torch.cuda.set_device(0)
    if args.copy_probe is not None:
        n = int(args.copy_probe * 2 ** 30) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 8 * n, "copy_probe_bytes_written": 8 * n, "launches": 3}))
        return
    if args.workload != "dgemm":
        fn = {"zgemm": workload_zgemm, "ozaki": workload_ozaki, "batched": workload_batched}[args.workload]
        print(json.dumps(fn(torch, args)), flush=True)
        return
    rng = np.random.default_rng(1)
    check_cols = np.sort(rng.choice(size, size=min(16, size), replace=False))

    flops = 2.0 * size ** 3
    log(f"[bench] allocating and filling 3 x {ld * size * 8 / 1e9:.1f} GB of pinned host memory")
    t0 = time.time()
    pinned_mats = make_problem(torch, size, ld, True)
    log(f"[bench] data ready in {time.time() - t0:.1f}s")

    D = run_device(torch, pinned_mats, size, ld, args.steps, args.warmup, check_cols)
    ms = D["ms_per_step"]
    value = flops / (ms / 1e3) / 1e12
    Z = None
    if args.crt_moduli:
        log(f"[bench] Ozaki scheme II, {args.crt_moduli} moduli, HBM-resident")
        Z = run_device(torch, pinned_mats, size, ld, args.steps, args.warmup, check_cols, args.crt_moduli)

    h2h = {}
    if args.variant in ("pinned", "both"):  # host C is still the initial one: the device run used copies
        h2h["pinned"] = run_variant(torch, pinned_mats, size, ld, True, args.steps, args.warmup, args.streams,
                                    args.tile, check_cols)
    if args.crt_moduli and args.variant != "none":
        h2h["pinned_ozaki_crt"] = run_variant(torch, pinned_mats, size, ld, True, args.steps, args.warmup,
                                              args.streams, args.tile, check_cols, args.crt_moduli)
    del pinned_mats
    if args.variant in ("pageable", "both"):
        log("[bench] pageable host memory")
        pg = make_problem(torch, size, ld, False)
        h2h["pageable"] = run_variant(torch, pg, size, ld, False, max(1, args.steps - 1), args.warmup,
                                      args.streams, args.tile, check_cols)
        if args.crt_moduli:
            h2h["pageable_ozaki_crt"] = run_variant(torch, pg, size, ld, False, max(1, args.steps - 1),
                                                    args.warmup, args.streams, args.tile, check_cols, args.crt_moduli)
        del pg
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
        "config": {"workload": f"DGEMM TN m=n=k={size}, lda=ldb=ldc={ld}, alpha=beta=1, A/B/C resident in HBM "
                               "(tmm_device_gemm, one tile-kernel launch per step)",
                   "m": size, "n": size, "k": size, "ld": ld, "trans": "TN"},
        "roofline": {"bound": "mfma", "achieved": round(value, 3), "peak": 78.6, "unit": "TFLOP/s",
                     "frac": round(value / 78.6, 4), "traffic": measured_traffic(size),
                     "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
                     "algorithmic_bytes_per_launch": 4 * 8 * size * size,
                     "kernel": "gemm_d_dlds_kernel<true, true, 3, 3> (FP64 TN, direct-to-LDS ring), one launch per step",
                     "flops_per_launch": flops, "mean_launch_ms": round(ms, 3),
                     "launches_timed": args.steps},
        "check": D.get("check"),
        "host_to_host": {},
    }
    for key, R in h2h.items():
        s_ = R["stats"]
        t = R["ms_per_step"] / 1e3
        # Compute bound of the call: FP64 MFMA for the native kernel, INT8 MFMA over the products the
        # Ozaki path actually issues (one per modulus) for the emulated one.
        t_compute = s_.get("int8_ops", 0.0) / P_INT8 if "ozaki" in key else flops / P_FP64
        t_lb = max(t_compute, s_["h2d_bytes"] / PCIE_BW, s_["d2h_bytes"] / PCIE_BW)
        out["host_to_host"][key] = {
            "value": round(flops / t / 1e12, 3), "unit": "TFLOP/s", "ms_per_step": round(1e3 * t, 2),
            "config": f"gpu::gemm TN {size}^3 ld {ld}, {key} host memory, make_context({args.streams}, {args.tile}, "
                      f"{args.tile}, {args.tile}) -> tiles {'x'.join(map(str, R['ref_tiles']))} (the reference tile policy), "
                      f"device blocks {int(s_['tile_m'])}x{int(s_['tile_n'])}x{int(s_['tile_k'])} "
                      f"({int(s_['k_chunks'])} k-chunks), copy_c_back=true",
            "t_lb_s": t_lb, "host_roofline_frac": round(t_lb / t, 4),
            "compute_bound": "INT8 MFMA 5.0 POP/s" if "ozaki" in key else "FP64 MFMA 78.6 TFLOP/s",
            "h2d_bytes": s_["h2d_bytes"], "d2h_bytes": s_["d2h_bytes"], "h2d_GBps_avg": s_["h2d_bytes"] / t / 1e9,
            "reference_schedule_h2d_bytes": 140.8e9 if size == 40000 else None,
            "tile_kernel_mean_ms": round(R["kernel_mean_s"] * 1e3, 3),
            "tile_kernel_TFLOPs": round(R["kernel_flops_per_launch"] / R["kernel_mean_s"] / 1e12, 3)
            if R["kernel_mean_s"] > 0 else None,
            "tile_kernel_busy_frac": round(R["kernel_busy_frac"], 4), "check": R.get("check")}
    if Z is not None:
        zt = Z["ms_per_step"] / 1e3
        zo = {"moduli": args.crt_moduli,
              "scheme": "Ozaki scheme II: power-of-two scaled integer operands, one INT8 MFMA product per "
                        "modulus (pairwise coprime moduli <= 256), CRT-sum reconstruction in FP64 limbs",
              "device_resident": {"TFLOPs_fp64_equiv": round(flops / zt / 1e12, 3), "ms_per_step": round(1e3 * zt, 2),
                                  "int8_TOPs": round(Z["int8_ops_per_step"] / zt / 1e12, 1),
                                  "normwise_error": Z["check"]["normwise"],
                                  "componentwise_error": Z["check"]["componentwise"],
                                  "native_normwise_error": D["check"]["normwise"],
                                  "native_TFLOPs": round(value, 3),
                                  "speedup_vs_native": round(ms / (1e3 * zt), 3),
                                  "error_not_above_native": Z["check"]["normwise"] <= D["check"]["normwise"]}}
        if "pinned_ozaki_crt" in out["host_to_host"] and "pinned" in out["host_to_host"]:
            zh, nh = out["host_to_host"]["pinned_ozaki_crt"], out["host_to_host"]["pinned"]
            zo["host_to_host_pinned"] = {
                "TFLOPs_fp64_equiv": zh["value"], "native_TFLOPs": nh["value"],
                "speedup_vs_native": round(zh["value"] / nh["value"], 3),
                "normwise_error": zh["check"]["normwise"], "native_normwise_error": nh["check"]["normwise"],
                "error_not_above_native": zh["check"]["normwise"] <= nh["check"]["normwise"]}
        if "pageable_ozaki_crt" in out["host_to_host"] and "pageable" in out["host_to_host"]:
            zh, nh = out["host_to_host"]["pageable_ozaki_crt"], out["host_to_host"]["pageable"]
            zo["host_to_host_pageable"] = {
                "TFLOPs_fp64_equiv": zh["value"], "native_TFLOPs": nh["value"],
                "speedup_vs_native": round(zh["value"] / nh["value"], 3),
                "normwise_error": zh["check"]["normwise"], "native_normwise_error": nh["check"]["normwise"],
                "error_not_above_native": zh["check"]["normwise"] <= nh["check"]["normwise"]}
        out["ozaki_crt"] = zo
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (host BLAS) ...")
        out["cpu_baseline"] = cpu_baseline(size, ld)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

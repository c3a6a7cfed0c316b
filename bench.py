#!/usr/bin/env python3
"""Host-to-host FP64 GEMM throughput of the out-of-core GEMM library at 40000^3.

Workload (the reference's headline out-of-core case, sized for MI355X): gpu::gemm with
C = alpha*A^T*B + beta*C, column-major FP64, m = n = k = 40000, leading dimensions 40064 (the 64
padding rows hold NaN and must never be read), alpha = beta = 1, entries uniform[-1, 1] from fixed
seeds, A, B and C in pinned host memory, make_context(4 streams, 8192^3 tile request),
copy_c_back = true.

  value        : TFLOP/s of one step = one whole gpu::gemm call, wall time from entry with the data
                 on the host to return with C on the host (2*40000^3 flop per call)
  roofline     : the dominant kernel of that call, the grouped tile DGEMM (gemm_d_pipe_group_kernel):
                 flops per launch / mean launch duration, both measured live with HIP events recorded
                 on the stream each launch runs on, against the FP64 matrix peak (78.6 TFLOP/s);
                 traffic = HBM bytes per launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary
  host_link_roofline : T_lb = max(F/P, H2D bytes/BW, D2H bytes/BW) with PCIe 5.0 x16 at 64 GB/s
                 per direction, and T_lb / T of the call
  ozaki_crt    : the same call through Ozaki scheme II (16 INT8 products), its sampled error beside
                 native's (the metric's "Ozaki-emulated vs native at equal error")
  host_to_host : every host-to-host line (pinned / pageable, native / Ozaki)
  device_resident : the same product on A, B, C already in HBM (one tmm_device_gemm per step)
  workloads    : the other measured configurations, a few steps each: ZGEMM CN 24000x24000x32000,
                 the Ozaki NN 40000^3 dynamic-range sweep (phi = 0.5, 2, 4), the 512-problem batch
  cpu_baseline : the host BLAS (MKL through torch, or numpy's BLAS, whichever is faster) on the box's
                 cores for a bounded column sample of the same product
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

GROUP_TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "tile_dgemm_group_tn_40000_traffic.md")


def profile_value(path, key):
    """A `key: value` line of a profile summary (None when absent)."""
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1])
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_matrix(rows, cols, pinned, dtype=np.float64):
    """Column-major rows x cols host matrix, page-locked (hipHostMalloc) or pageable."""
    if pinned:
        return tiledmm.pinned_empty((rows, cols), dtype)
    return np.empty((rows, cols), dtype=dtype, order="F")


def make_problem(size, ld, pinned):
    """A (k x m, TN), B (k x n), C (m x n), all with leading dimension ld, padding rows NaN."""
    mats = []
    for seed in (42, 43, 44):
        M = host_matrix(ld, size, pinned)
        tiledmm.fill_uniform(M, seed, -1.0, 1.0, rows=size, cols=size, ld=ld)
        M[size:, :] = np.nan
        mats.append(M)
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


def run_host(torch, mats, size, ld, steps, warmup, streams, tile, check_cols, moduli=0):
    """gpu::gemm TN size^3 on host matrices: warm-up calls (the first checked on sampled columns),
    then `steps` timed calls between two device synchronisations."""
    A, B, C = mats
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
    acc = {"kernel_s": 0.0, "kernel_launches": 0.0, "kernel_flops": 0.0, "group_launches": 0.0,
           "group_kernel_s": 0.0, "group_flops": 0.0, "group_bytes": 0.0}
    walls = []
    torch.cuda.synchronize()
    t = time.time()
    for _ in range(steps):
        t1 = time.time()
        h.gemm("T", "N", size, size, size, 1.0, A, ld, B, ld, 1.0, C, ld, pin=False, copy_back=True)
        walls.append(time.time() - t1)
        st = h.stats()
        for key in acc:
            acc[key] += st[key]
    torch.cuda.synchronize()
    total = time.time() - t
    st = h.stats()
    res.update({"ms_per_step": 1e3 * total / steps, "walls_s": walls, "stats": st, "ref_tiles": h.tile_sizes(),
                "kernel": acc, "kernel_busy_frac": acc["kernel_s"] / total})
    h.close()
    return res


def run_device(torch, mats, size, ld, steps, warmup, check_cols, moduli=0):
    """The whole product on HBM-resident A, B, C: one tmm_device_gemm per step (moduli > 0: the
    Ozaki scheme II emulation with that many INT8 products instead)."""
    A, B, C = mats
    dev = [torch.empty(ld * size, dtype=torch.float64, device="cuda") for _ in range(3)]
    for d, M in zip(dev, mats):
        d.copy_(torch.from_numpy(M.reshape(-1, order="F")))
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


def host_line(key, R, size, ld, streams, tile, flops):
    s_ = R["stats"]
    t = R["ms_per_step"] / 1e3
    # Compute bound of the call: FP64 MFMA for the native kernel, INT8 MFMA over the products the
    # Ozaki path actually issues (one per modulus) for the emulated one.
    t_compute = s_.get("int8_ops", 0.0) / P_INT8 if "ozaki" in key else flops / P_FP64
    t_lb = max(t_compute, s_["h2d_bytes"] / PCIE_BW, s_["d2h_bytes"] / PCIE_BW)
    K = R["kernel"]
    return {
        "value": round(flops / t / 1e12, 3), "unit": "TFLOP/s", "ms_per_step": round(1e3 * t, 2),
        "walls_s": [round(w, 4) for w in R["walls_s"]],
        "config": f"gpu::gemm TN {size}^3 ld {ld}, {key} host memory, make_context({streams}, {tile}, {tile}, {tile}) "
                  f"-> tiles {'x'.join(map(str, R['ref_tiles']))} (the reference tile policy), device blocks "
                  f"{int(s_['tile_m'])}x{int(s_['tile_n'])}x{int(s_['tile_k'])} ({int(s_['k_chunks'])} k-chunks), "
                  "copy_c_back=true",
        "t_lb_s": t_lb, "host_roofline_frac": round(t_lb / t, 4),
        "compute_bound": "INT8 MFMA 5.0 POP/s" if "ozaki" in key else "FP64 MFMA 78.6 TFLOP/s",
        "h2d_bytes": s_["h2d_bytes"], "d2h_bytes": s_["d2h_bytes"], "h2d_GBps_avg": s_["h2d_bytes"] / t / 1e9,
        "reference_schedule_h2d_bytes": 140.8e9 if size == 40000 else None,
        "tile_kernel_launches_per_call": K["kernel_launches"] / max(len(R["walls_s"]), 1),
        "tile_kernel_TFLOPs": round(K["kernel_flops"] / K["kernel_s"] / 1e12, 3) if K["kernel_s"] > 0 else None,
        "tile_kernel_busy_frac": round(R["kernel_busy_frac"], 4), "check": R.get("check")}


def cpu_baseline(size, ld, target_s=15.0):
    """Host BLAS on NCORES threads: A^T (size x size, TN, ld) times a column sample of B. The
    faster of MKL (through torch) and numpy's BLAS on a calibration sample is timed."""
    import torch

    torch.set_num_threads(NCORES)
    At = np.empty((size, ld), dtype=np.float64)  # C-order (size, ld) == Fortran (ld, size): At[:, :size] = A^T
    tiledmm.fill_uniform((At.ctypes.data, np.float64), 42, -1.0, 1.0, rows=size, cols=size, ld=ld)
    At = At[:, :size]
    cfg = np.__config__.CONFIG if hasattr(np.__config__, "CONFIG") else {}
    blas = cfg.get("Build Dependencies", {}).get("blas", {}) if isinstance(cfg, dict) else {}
    providers = {
        "torch": (f"torch {torch.__version__} CPU (MKL: {torch.backends.mkl.is_available()})",
                  lambda Bt, Ct: torch.addmm(torch.from_numpy(Ct), torch.from_numpy(At), torch.from_numpy(Bt),
                                             out=torch.from_numpy(Ct))),
        "numpy": (f"numpy {np.__version__} / {blas.get('name', 'BLAS')} {blas.get('version', '')}".strip(),
                  lambda Bt, Ct: np.add(Ct, At @ Bt, out=Ct)),
    }

    def timed(fn, ncols):
        Bt = np.empty((size, ncols), dtype=np.float64)  # B columns as rows-major blocks: A^T @ B
        tiledmm.fill_uniform((Bt.ctypes.data, np.float64), 43, -1.0, 1.0, rows=ncols, cols=size, ld=ncols)
        Ct = np.zeros((size, ncols))
        fn(Bt[:, :min(32, ncols)].copy(), Ct[:, :min(32, ncols)].copy())  # warm-up
        t = time.time()
        fn(Bt, Ct)
        return time.time() - t

    cal = 256
    speed = {name: cal / timed(fn, cal) for name, (_, fn) in providers.items()}
    best = max(speed, key=speed.get)
    ncols = int(min(size, max(cal, speed[best] * target_s)))
    dt = timed(providers[best][1], ncols)
    flops = 2.0 * size * size * ncols
    return {"value": flops / dt / 1e12, "unit": "TFLOP/s", "cores": NCORES, "kind": "port",
            "baseline_kind": "host_blas",
            "sample": f"host BLAS dgemm C[:, :{ncols}] += A^T B[:, :{ncols}] with m=k={size} (TN, ld {ld}), "
                      f"{flops / 1e12:.2f} TFLOP in {dt:.1f}s; the whole {size}^3 product would take "
                      f"{2.0 * size ** 3 / (flops / dt):.0f}s at this rate",
            "provider": providers[best][0],
            "calibration_cols_per_s": {k: round(v, 1) for k, v in speed.items()}}


# ------------------------------------------------------------------------------------------------
# The other configurations (extra keys of the default line; `--workload` runs one on its own)
# ------------------------------------------------------------------------------------------------
def workload_zgemm(torch, args, steps=2):
    """ZGEMM CN m=n=24000 k=32000, alpha=(1,0.5), beta=(0.5,-1), make_context(3, 6144^3), pinned
    host memory: host-to-host (gpu::gemm, copy_c_back) and on HBM-resident data."""
    m, n, k = (args.size_m or 24000), (args.size_n or 24000), (args.size_k or 32000)
    alpha, beta = complex(1, 0.5), complex(0.5, -1)
    flops = 8.0 * m * n * k
    mats = []
    for seed, rows, cols in ((42, k, m), (43, k, n), (44, m, n)):
        M = tiledmm.pinned_empty((rows, cols), np.complex128)
        tiledmm.util().tmmu_fill_uniform_z(M.ctypes.data, rows, cols, rows, 0, 0, rows, seed, -1.0, 1.0, 0)
        mats.append(M)
    A, B, C = mats
    cols = np.arange(0, n, max(1, n // 16))
    C0 = np.ascontiguousarray(C[:, cols])
    # host-to-host (the host BLAS reference is computed after all timed calls)
    h = tiledmm.Handle(tiledmm.C64, 3, 6144, 6144, 6144)
    h.set_option(tiledmm.OPT_PROFILE, 1)
    if args.tile_groups is not None:
        h.set_option(tiledmm.OPT_TILE_GROUPS, args.tile_groups)
    h.gemm("C", "N", m, n, k, alpha, A, k, B, k, beta, C, m)
    got_h2h = np.ascontiguousarray(C[:, cols])
    torch.cuda.synchronize()
    t0 = time.time()
    for _ in range(steps):
        h.gemm("C", "N", m, n, k, alpha, A, k, B, k, beta, C, m)
    torch.cuda.synchronize()
    h2h_s = (time.time() - t0) / steps
    st = h.stats()
    h.close()
    # device-resident, from the initial C (regenerated: the host-to-host calls overwrote it)
    tiledmm.util().tmmu_fill_uniform_z(C.ctypes.data, m, n, m, 0, 0, m, 44, -1.0, 1.0, 0)
    dA, dB, dC = (torch.from_numpy(x.reshape(-1, order="F")).cuda() for x in (A, B, C))
    stream = torch.cuda.Stream()
    h = tiledmm.Handle(tiledmm.C64, 1)
    h.device_gemm("C", "N", m, n, k, alpha, dA, k, dB, k, beta, dC, m, stream=stream.cuda_stream)
    stream.synchronize()
    got = dC.view(n, m)[torch.as_tensor(cols, device="cuda")].t().cpu().numpy()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    dev_steps = 1
    for _ in range(dev_steps):
        h.device_gemm("C", "N", m, n, k, alpha, dA, k, dB, k, beta, dC, m, stream=stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    dev_ms = e0.elapsed_time(e1) / dev_steps
    h.close()
    del dA, dB, dC
    torch.cuda.empty_cache()
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
    t_lb = max(flops / P_FP64, st["h2d_bytes"] / PCIE_BW, st["d2h_bytes"] / PCIE_BW)
    del A, B, C, mats
    return {"workload": f"ZGEMM CN m={m} n={n} k={k} alpha=(1,0.5) beta=(0.5,-1), make_context(3, 6144, 6144, 6144), "
                        "pinned host memory, copy_c_back=true", "unit": "TFLOP/s (8mnk)",
            "host_to_host_pinned": {"value": flops / h2h_s / 1e12, "ms": h2h_s * 1e3, "steps": steps,
                                    "tiles": [st["tile_m"], st["tile_n"], st["tile_k"]],
                                    "h2d_bytes": st["h2d_bytes"], "d2h_bytes": st["d2h_bytes"],
                                    "t_lb_s": t_lb, "host_roofline_frac": t_lb / h2h_s,
                                    "tile_kernel_TFLOPs": st["kernel_flops"] / st["kernel_s"] / 1e12
                                    if st["kernel_s"] else None,
                                    "normwise_error_sampled": err_h2h},
            "device_resident": {"value": flops / (dev_ms / 1e3) / 1e12, "ms": dev_ms, "steps": dev_steps,
                                "frac_of_78.6": flops / (dev_ms / 1e3) / P_FP64, "normwise_error_sampled": err_dev},
            "columns_checked": len(cols), "bound": 2 * k * 2.0 ** -53}


P_FP32 = 157.3e12  # MI355X dense FP32 matrix peak


def workload_fp32(torch, args, steps=2):
    """Host-to-host single precision: SGEMM TN 40000^3 (ld 40064, alpha=beta=1, make_context(4, 8192^3))
    and CGEMM CN m=n=24000 k=32000 (alpha=(1,0.5), beta=(0.5,-1), make_context(3, 6144^3)), pinned
    host memory, copy_c_back. Sampled columns against a float64 host computation of the same product."""
    out = {"unit": "TFLOP/s (2mnk real, 8mnk complex)"}
    cases = [("sgemm_tn_40000", tiledmm.F32, np.float32, "T", "N", 40000, 40000, 40000, 40064, 4, 8192,
              1.0, 1.0, 2.0),
             ("cgemm_cn_24000x24000x32000", tiledmm.C32, np.complex64, "C", "N", 24000, 24000, 32000, None, 3, 6144,
              complex(1, 0.5), complex(0.5, -1), 8.0)]
    u = tiledmm.util()
    for key, code, dt, ta, tb, m, n, k, ld, streams, tile, alpha, beta, fl_per in cases:
        lda = ld or k
        ldb = ld or k
        ldc = ld or m
        mats = []
        for seed, rows, cols, lead in ((42, k, m, lda), (43, k, n, ldb), (44, m, n, ldc)):
            M = tiledmm.pinned_empty((lead, cols), dt)
            w = 2 if dt == np.complex64 else 1  # complex: interleaved parts from one stream of floats
            u.tmmu_fill_uniform_s(M.ctypes.data, w * rows, cols, w * lead, 0, 0, w * rows, seed, -1.0, 1.0, 0)
            mats.append(M)
        A, B, C = mats
        cols = np.arange(0, n, max(1, n // 16))
        C0 = np.ascontiguousarray(C[:m, cols]).astype(np.complex128 if dt == np.complex64 else np.float64)
        h = tiledmm.Handle(code, streams, tile, tile, tile)
        h.set_option(tiledmm.OPT_PROFILE, 1)
        h.gemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
        got = np.ascontiguousarray(C[:m, cols]).astype(C0.dtype)
        torch.cuda.synchronize()
        t0 = time.time()
        for _ in range(steps):
            h.gemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
        torch.cuda.synchronize()
        sec = (time.time() - t0) / steps
        st = h.stats()
        h.close()
        # reference columns in double precision: op(A)^T-style products of k x m stored A
        Bc = np.ascontiguousarray(B[:k, cols]).astype(C0.dtype)
        ref = np.empty((m, len(cols)), dtype=C0.dtype)
        scale = np.empty((m, len(cols)))
        for i0 in range(0, m, 2048):
            blk = A[:k, i0:i0 + 2048].astype(C0.dtype)
            opa = np.conj(blk.T) if ta == "C" else blk.T
            ref[i0:i0 + 2048] = opa @ Bc
            scale[i0:i0 + 2048] = np.abs(blk).T @ np.abs(Bc)
        ref = alpha * ref + beta * C0
        scale = abs(alpha) * scale + abs(beta) * np.abs(C0)
        err = float(np.linalg.norm(got - ref) / np.linalg.norm(scale))
        flops = fl_per * m * n * k
        t_lb = max(flops / P_FP32, st["h2d_bytes"] / PCIE_BW, st["d2h_bytes"] / PCIE_BW)
        out[key] = {"value": flops / sec / 1e12, "ms": sec * 1e3, "steps": steps, "t_lb_s": t_lb,
                    "host_roofline_frac": t_lb / sec, "tiles": [st["tile_m"], st["tile_n"], st["tile_k"]],
                    "tile_kernel_TFLOPs": st["kernel_flops"] / st["kernel_s"] / 1e12 if st["kernel_s"] else None,
                    "normwise_error_sampled": err, "bound": 2 * k * 2.0 ** -24, "columns_checked": len(cols),
                    "config": f"{'SGEMM' if dt == np.float32 else 'CGEMM'} {ta}{tb} m={m} n={n} k={k}, "
                              f"make_context({streams}, {tile}^3), pinned host memory, copy_c_back=true"}
        log(f"[bench] {key}: {out[key]}")
        del A, B, C, mats, Bc, ref, scale
    return out


# moduli tried per phi in the default run: the count that reaches native's error and one fewer
DEFAULT_PHI_MODULI = {0.5: [15, 16], 2.0: [16, 17], 4.0: [17, 18]}


def workload_ozaki(torch, args, moduli_of=None, slices=(), steps=1):
    """NN 40000^3, entries (u-0.5)*exp(phi*g), beta = 0: Ozaki vs native on HBM-resident data."""
    size = args.size
    out = {"workload": f"DGEMM NN {size}^3 alpha=1 beta=0, entries (uniform[-0.5,0.5])*exp(phi*N(0,1)), "
                       "HBM-resident (tmm_device_gemm)",
           "unit": "TFLOP/s (FP64-equivalent 2mnk)", "results": []}
    dA = torch.empty(size * size, dtype=torch.float64, device="cuda")
    dB = torch.empty(size * size, dtype=torch.float64, device="cuda")
    dC = torch.empty(size * size, dtype=torch.float64, device="cuda")
    host = np.empty((size, size), order="F")
    cols = np.arange(0, size, max(1, size // 16))
    u = tiledmm.util()
    stream = torch.cuda.Stream()
    for phi in args.phis:
        u.tmmu_fill_range_d(host.ctypes.data, size, size, size, 0, 0, size, 42, phi, 0)
        dA.copy_(torch.from_numpy(host.reshape(-1, order="F")))
        A = host.copy(order="F")
        u.tmmu_fill_range_d(host.ctypes.data, size, size, size, 0, 0, size, 43, phi, 0)
        Bc = np.ascontiguousarray(host[:, cols])
        dB.copy_(torch.from_numpy(host.reshape(-1, order="F")))
        ref = np.empty((size, len(cols)))
        scale = np.empty((size, len(cols)))
        for i0 in range(0, size, 4096):
            ref[i0:i0 + 4096] = A[i0:i0 + 4096] @ Bc
            scale[i0:i0 + 4096] = np.abs(A[i0:i0 + 4096]) @ np.abs(Bc)
        del A
        moduli = moduli_of(phi) if moduli_of else args.moduli
        runs = [("native", 0)] + [("crt", mm) for mm in moduli] + [("slices", sl) for sl in slices]
        for mode, count in runs:
            h = tiledmm.Handle(tiledmm.F64, 1)
            h.set_option(tiledmm.OPT_OZAKI_SLICES, count if mode == "slices" else 0)
            h.set_option(tiledmm.OPT_OZAKI_MODULI, count if mode == "crt" else 0)
            h.device_gemm("N", "N", size, size, size, 1.0, dA, size, dB, size, 0.0, dC, size,
                          stream=stream.cuda_stream)
            stream.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(steps):
                h.device_gemm("N", "N", size, size, size, 1.0, dA, size, dB, size, 0.0, dC, size,
                              stream=stream.cuda_stream)
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / steps
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
    del dA, dB, dC
    torch.cuda.empty_cache()
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


def workload_vendor(torch, args, sizes=(8192, 16384)):
    """Reference rates of the vendor library on the same box: rocBLAS dgemm TN against this library's
    FP64 tile kernel (tmm_device_gemm) on the same random HBM-resident operands, and rocBLAS gemm_ex
    int8 -> int32 TN (the operation crt_product_kernel performs per modulus; the emulated call with 16
    moduli runs beside it, its per-product kernel time is in a kernel trace of this run). HIP events
    on one stream, best of 3 after a warm-up call."""
    import ctypes
    rb = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "librocblas.so"))
    handle = ctypes.c_void_p()
    assert rb.rocblas_create_handle(ctypes.byref(handle)) == 0
    stream = torch.cuda.Stream()
    assert rb.rocblas_set_stream(handle, ctypes.c_void_p(stream.cuda_stream)) == 0
    NONE, TRANS, F64, I8, I32 = 111, 112, 152, 160, 162
    out = {"workload": "vendor reference: rocBLAS vs this library, HBM-resident random operands, TN",
           "results": []}

    def best_ms(fn):
        fn()
        stream.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        times = []
        for _ in range(3):
            e0.record(stream)
            fn()
            e1.record(stream)
            stream.synchronize()
            times.append(e0.elapsed_time(e1))
        return min(times)

    h = tiledmm.Handle(tiledmm.F64, 1)
    for n in sizes:
        a = torch.empty(n * n, dtype=torch.float64, device="cuda").uniform_(-1, 1)
        b = torch.empty(n * n, dtype=torch.float64, device="cuda").uniform_(-1, 1)
        c = torch.zeros(n * n, dtype=torch.float64, device="cuda")
        one, zero = ctypes.c_double(1.0), ctypes.c_double(0.0)
        torch.cuda.synchronize()
        ms_rb = best_ms(lambda: rb.rocblas_dgemm(handle, TRANS, NONE, n, n, n, ctypes.byref(one), ctypes.c_void_p(a.data_ptr()), n,
                                                 ctypes.c_void_p(b.data_ptr()), n, ctypes.byref(zero),
                                                 ctypes.c_void_p(c.data_ptr()), n))
        ms_tmm = best_ms(lambda: h.device_gemm("T", "N", n, n, n, 1.0, a, n, b, n, 0.0, c, n, stream=stream.cuda_stream))
        h.set_option(tiledmm.OPT_OZAKI_MODULI, 16)  # one INT8 product per modulus (crt_product_kernel)
        ms_oz = best_ms(lambda: h.device_gemm("T", "N", n, n, n, 1.0, a, n, b, n, 0.0, c, n, stream=stream.cuda_stream))
        h.set_option(tiledmm.OPT_OZAKI_MODULI, 0)
        fl = 2.0 * n ** 3
        out["results"].append({"op": "dgemm TN", "n": n, "rocblas_TFLOPs": fl / ms_rb / 1e9,
                               "tiledmm_TFLOPs": fl / ms_tmm / 1e9,
                               "tiledmm_ozaki16_TFLOPs_fp64_equiv": fl / ms_oz / 1e9})
        log(f"[bench] vendor {out['results'][-1]}")
        del a, b, c
        a8 = torch.randint(-128, 128, (n * n,), dtype=torch.int8, device="cuda")
        b8 = torch.randint(-128, 128, (n * n,), dtype=torch.int8, device="cuda")
        c32 = torch.zeros(n * n, dtype=torch.int32, device="cuda")
        i1, i0 = ctypes.c_int32(1), ctypes.c_int32(0)
        torch.cuda.synchronize()
        ms8 = best_ms(lambda: rb.rocblas_gemm_ex(handle, TRANS, NONE, n, n, n, ctypes.byref(i1), ctypes.c_void_p(a8.data_ptr()), I8, n,
                                                 ctypes.c_void_p(b8.data_ptr()), I8, n, ctypes.byref(i0),
                                                 ctypes.c_void_p(c32.data_ptr()), I32, n, ctypes.c_void_p(c32.data_ptr()),
                                                 I32, n, I32, 0, 0, 0))
        out["results"].append({"op": "int8 gemm_ex TN", "n": n, "rocblas_POPs": fl / ms8 / 1e12})
        log(f"[bench] vendor {out['results'][-1]}")
        del a8, b8, c32
        torch.cuda.empty_cache()
    h.close()
    rb.rocblas_destroy_handle(handle)
    return out


BATCH_SETTLE_S = 3.0


def batched_problem(count=512, lo=64, hi=4096, seed=42):
    """m, n, k of the 512-problem batch: each log-uniform in [64, 4096] from the seed."""
    rng = np.random.default_rng(seed)
    return np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), size=(count, 3)))).astype(int)


def workload_batched(torch, args, steps=3):
    """512 NN problems, m,n,k log-uniform in [64, 4096] (seed 42), alpha=beta=1, pinned host."""
    dims = batched_problem()
    As, Bs, Cs = [], [], []
    flops = 0.0
    for i, (m, n, k) in enumerate(dims):
        for (rows, cols, lst, seed) in ((m, k, As, 3 * i), (k, n, Bs, 3 * i + 1), (m, n, Cs, 3 * i + 2)):
            M = tiledmm.pinned_empty((int(rows), int(cols)))
            tiledmm.fill_uniform(M, 1000 + seed)
            lst.append(M)
        flops += 2.0 * m * n * k
    check = [0, 1, 2, 511]
    C0 = {i: Cs[i].copy() for i in check}
    h = tiledmm.Handle(tiledmm.F64, 2)
    h.set_option(tiledmm.OPT_BATCH_GRAPH, args.batch_graph)
    ms_, ns_, ks_ = [int(x) for x in dims[:, 0]], [int(x) for x in dims[:, 1]], [int(x) for x in dims[:, 2]]
    args_ = (ms_, ns_, ks_, [1.0] * 512, As, ms_, Bs, ks_, [1.0] * 512, Cs, ms_)
    # The workloads before this one free tens of GB of device memory; for a few seconds after such
    # frees the driver's own work on the DMA engines slows host-link copies (a handle opened right
    # after another's 15.5 GB arena was freed ran its first calls at 320 / 308 / 246 ms, with 3 s
    # between them at 238 / 213). The pause and the two untimed calls keep that out of the timing.
    torch.cuda.synchronize()
    time.sleep(BATCH_SETTLE_S)
    h.gemm_batched("N", "N", *args_)
    errs = []
    for i in check:
        ref = As[i] @ Bs[i] + C0[i]
        scale = np.abs(As[i]) @ np.abs(Bs[i]) + np.abs(C0[i])
        errs.append(float(np.linalg.norm(Cs[i] - ref) / np.linalg.norm(scale)))
    h.gemm_batched("N", "N", *args_)  # second untimed call
    torch.cuda.synchronize()
    t0 = time.time()
    for _ in range(steps):
        h.gemm_batched("N", "N", *args_)
    torch.cuda.synchronize()
    t = (time.time() - t0) / steps
    cfg = h.batch_config()
    st = h.stats()
    h.close()
    h2d = sum(8.0 * (m * k + k * n + m * n) for m, n, k in dims)
    d2h = sum(8.0 * m * n for m, n, k in dims)
    t_lb = max(flops / P_FP64, h2d / PCIE_BW, d2h / PCIE_BW)
    return {"workload": "batched DGEMM: 512 NN problems, m,n,k log-uniform in [64,4096] (seed 42), alpha=beta=1, "
                        "pinned host memory, one tmm_dgemm_batched call per step", "unit": "TFLOP/s",
            "value": flops / t / 1e12, "ms": t * 1e3, "steps": steps,
            "flops": flops, "h2d_bytes": h2d, "d2h_bytes": d2h, "t_lb_s": t_lb, "host_roofline_frac": t_lb / t,
            "autotuner": cfg, "graph": st["batch_graph"], "groups": st["batch_groups"],
            "untimed_calls": 2, "settle_s_before_first_call": BATCH_SETTLE_S,
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
    ap.add_argument("--variant", choices=["pinned", "both"], default="both",
                    help="host memory kinds of the host-to-host lines (pinned always runs: it is the value)")
    ap.add_argument("--no-device", action="store_true", help="skip the HBM-resident lines")
    ap.add_argument("--no-extra", action="store_true", help="skip the ZGEMM, Ozaki phi and batched workloads")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=["dgemm", "zgemm", "ozaki", "batched", "vendor", "fp32"], default="dgemm",
                    help="dgemm: the default bench line; the others print their own JSON line")
    ap.add_argument("--size-m", type=int, default=None)
    ap.add_argument("--size-n", type=int, default=None)
    ap.add_argument("--size-k", type=int, default=None)
    ap.add_argument("--phis", type=float, nargs="*", default=[0.5, 2.0, 4.0])
    ap.add_argument("--slices", type=int, nargs="*", default=[8, 9])
    ap.add_argument("--moduli", type=int, nargs="*", default=[14, 15, 16, 17, 18, 19, 20])
    ap.add_argument("--batch-graph", type=int, default=-1, choices=[-1, 0, 1],
                    help="batched workload: hipGraph capture chosen by the autotuner (-1), always (1), never (0)")
    ap.add_argument("--tile-groups", type=int, default=None,
                    help="zgemm workload, host-to-host: TMM_OPT_TILE_GROUPS (default: the library's default)")
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
        if args.workload == "zgemm":
            res = workload_zgemm(torch, args, steps=args.steps)
        elif args.workload == "ozaki":
            res = workload_ozaki(torch, args, slices=args.slices, steps=args.steps)
        elif args.workload == "vendor":
            res = workload_vendor(torch, args)
        elif args.workload == "fp32":
            res = workload_fp32(torch, args, steps=args.steps)
        else:
            res = workload_batched(torch, args, steps=args.steps)
        print(json.dumps(res), flush=True)
        return
    rng = np.random.default_rng(1)
    check_cols = np.sort(rng.choice(size, size=min(16, size), replace=False))
    flops = 2.0 * size ** 3

    log(f"[bench] allocating and filling 3 x {ld * size * 8 / 1e9:.1f} GB of pinned host memory")
    t0 = time.time()
    pinned_mats = make_problem(size, ld, True)
    log(f"[bench] data ready in {time.time() - t0:.1f}s")
    h2h = {"pinned": run_host(torch, pinned_mats, size, ld, args.steps, args.warmup, args.streams, args.tile,
                              check_cols)}
    if args.crt_moduli:
        log(f"[bench] Ozaki scheme II, {args.crt_moduli} moduli, host-to-host")
        h2h["pinned_ozaki_crt"] = run_host(torch, pinned_mats, size, ld, args.steps, args.warmup, args.streams,
                                           args.tile, check_cols, args.crt_moduli)
    D = Z = None
    if not args.no_device:  # starts from the host C the calls above left: the checks use its columns
        D = run_device(torch, pinned_mats, size, ld, args.steps, args.warmup, check_cols)
        if args.crt_moduli:
            Z = run_device(torch, pinned_mats, size, ld, args.steps, args.warmup, check_cols, args.crt_moduli)
    del pinned_mats
    if args.variant == "both":
        log("[bench] pageable host memory")
        pg = make_problem(size, ld, False)
        h2h["pageable"] = run_host(torch, pg, size, ld, max(1, args.steps - 1), args.warmup, args.streams, args.tile,
                                   check_cols)
        if args.crt_moduli:
            h2h["pageable_ozaki_crt"] = run_host(torch, pg, size, ld, max(1, args.steps - 1), args.warmup,
                                                 args.streams, args.tile, check_cols, args.crt_moduli)
        del pg

    P = h2h["pinned"]
    ms = P["ms_per_step"]
    value = flops / (ms / 1e3) / 1e12
    K = P["kernel"]
    g_mean_s = K["group_kernel_s"] / K["group_launches"] if K["group_launches"] else None
    g_flops = K["group_flops"] / K["group_launches"] if K["group_launches"] else None
    achieved = g_flops / g_mean_s / 1e12 if g_mean_s else None
    lines = {key: host_line(key, R, size, ld, args.streams, args.tile, flops) for key, R in h2h.items()}
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
        "config": {"workload": f"gpu::gemm DGEMM TN m=n=k={size}, lda=ldb=ldc={ld}, alpha=beta=1, host-to-host: "
                               f"A, B, C in pinned host memory, make_context({args.streams} streams, {args.tile}^3 "
                               "tile request), copy_c_back=true; one step = one whole call",
                   "m": size, "n": size, "k": size, "ld": ld, "trans": "TN", "streams": args.streams,
                   "tile_request": args.tile, "copy_c_back": True, "host_memory": "pinned"},
        "roofline": {"bound": "mfma", "achieved": None if achieved is None else round(achieved, 3), "peak": 78.6,
                     "unit": "TFLOP/s", "frac": None if achieved is None else round(achieved / 78.6, 4),
                     "traffic": profile_value(GROUP_TRAFFIC_PROFILE, "traffic_bytes_per_launch")
                     if size == 40000 else None,
                     "traffic_unit": "HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
                     "algorithmic_bytes_per_launch": K["group_bytes"] / K["group_launches"]
                     if K["group_launches"] else None,
                     "kernel": "gemm_d_pipe_group_kernel<true, true> (FP64 TN, software-pipelined 128 x 128 "
                               "workgroup tiles at two per CU, direct-to-LDS ring, several tile GEMMs of one k-chunk "
                               "per launch), the largest share of the call's kernel time; "
                               "achieved, traffic and algorithmic bytes are means over all its launches in a call",
                     "launches_timed": K["group_launches"],
                     "flops_per_launch": g_flops, "mean_launch_ms": None if g_mean_s is None else round(1e3 * g_mean_s, 3),
                     "all_tile_launches": {"launches": K["kernel_launches"], "TFLOPs":
                                           round(K["kernel_flops"] / K["kernel_s"] / 1e12, 3) if K["kernel_s"] else None,
                                           "busy_frac_of_call": round(P["kernel_busy_frac"], 4)}},
        "host_link_roofline": {"t_lb_s": lines["pinned"]["t_lb_s"], "frac": lines["pinned"]["host_roofline_frac"],
                               "formula": "T_lb = max(2mnk / 78.6e12, H2D bytes / 64e9, D2H bytes / 64e9)",
                               "h2d_bytes": lines["pinned"]["h2d_bytes"], "d2h_bytes": lines["pinned"]["d2h_bytes"]},
        "check": P.get("check"),
        "host_to_host": lines,
    }
    if "pinned_ozaki_crt" in lines:
        zh, nh = lines["pinned_ozaki_crt"], lines["pinned"]
        zo = {"moduli": args.crt_moduli,
              "scheme": "Ozaki scheme II: power-of-two scaled integer operands, one INT8 MFMA product per "
                        "modulus (pairwise coprime moduli <= 256), CRT-sum reconstruction in FP64 limbs",
              "host_to_host_pinned": {
                  "TFLOPs_fp64_equiv": zh["value"], "native_TFLOPs": nh["value"],
                  "speedup_vs_native": round(zh["value"] / nh["value"], 3),
                  "normwise_error": zh["check"]["normwise"], "native_normwise_error": nh["check"]["normwise"],
                  "componentwise_error": zh["check"]["componentwise"],
                  "native_componentwise_error": nh["check"]["componentwise"],
                  "error_not_above_native": zh["check"]["normwise"] <= nh["check"]["normwise"]}}
        if "pageable_ozaki_crt" in lines:
            zp, npg = lines["pageable_ozaki_crt"], lines["pageable"]
            zo["host_to_host_pageable"] = {
                "TFLOPs_fp64_equiv": zp["value"], "native_TFLOPs": npg["value"],
                "speedup_vs_native": round(zp["value"] / npg["value"], 3),
                "normwise_error": zp["check"]["normwise"], "native_normwise_error": npg["check"]["normwise"],
                "error_not_above_native": zp["check"]["normwise"] <= npg["check"]["normwise"]}
        out["ozaki_crt"] = zo
    if D is not None:
        dv = flops / (D["ms_per_step"] / 1e3) / 1e12
        out["device_resident"] = {
            "workload": f"DGEMM TN {size}^3 ld {ld} on A, B, C resident in HBM, one tmm_device_gemm per step "
                        "(gemm_d_big_kernel<true, true>), HIP events on the launching stream",
            "value": round(dv, 3), "ms_per_step": round(D["ms_per_step"], 2), "frac_of_78.6": round(dv / 78.6, 4),
            "check": D["check"]}
        if Z is not None:
            zt = Z["ms_per_step"] / 1e3
            out["device_resident"]["ozaki_crt"] = {
                "moduli": args.crt_moduli, "TFLOPs_fp64_equiv": round(flops / zt / 1e12, 3),
                "ms_per_step": round(1e3 * zt, 2), "int8_TOPs": round(Z["int8_ops_per_step"] / zt / 1e12, 1),
                "speedup_vs_native": round(D["ms_per_step"] / (1e3 * zt), 3),
                "normwise_error": Z["check"]["normwise"], "componentwise_error": Z["check"]["componentwise"],
                "native_normwise_error": D["check"]["normwise"],
                "error_not_above_native": Z["check"]["normwise"] <= D["check"]["normwise"]}
    if not args.no_extra:
        out["workloads"] = {}
        log("[bench] ZGEMM CN 24000x24000x32000")
        out["workloads"]["zgemm_cn_24000x24000x32000"] = workload_zgemm(torch, args, steps=2)
        log("[bench] Ozaki NN 40000^3 dynamic-range sweep")
        out["workloads"]["ozaki_nn_40000_phi"] = workload_ozaki(
            torch, args, moduli_of=lambda phi: DEFAULT_PHI_MODULI.get(phi, [16]), steps=1)
        log("[bench] batched 512 problems")
        out["workloads"]["batched_512"] = workload_batched(torch, args, steps=3)
        log("[bench] SGEMM TN 40000^3 and CGEMM CN 24000x24000x32000, host-to-host")
        out["workloads"]["fp32"] = workload_fp32(torch, args, steps=2)
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (host BLAS) ...")
        out["cpu_baseline"] = cpu_baseline(size, ld)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

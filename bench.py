"""Benchmark of the adaptive FP64/fp16 snowflake stencil path on one MI355X.

Workload (w3914777_amd/paras/snowflake/workload4096.h): Kobayashi snowflake
4096x4096, dx=0.03, dt=1e-4, zero-flux boundaries, 5-point Laplacian,
eps=0.01(1+0.05cos(6(theta-pi/2))), tau=3e-4, alpha=0.9, gamma=10, K=1.6,
Teq=1, radius-5 nucleus, interface noise a=0.01; 32x32 tiles, gradient
monitor |grad phi| dx > 0.05 every 10 steps.

A step advances every cell of the grid by one explicit-Euler step (all of
its fields).  The timed region runs K steps of the adaptive path with the
fields already resident in HBM; the same W+K steps are run on the plain FP64
path for the speed ratio and the L2/Linf error of the adaptive result.

Prints one JSON line (metric, value = adaptive cell-updates/s, roofline of
the dominant kernel of the adaptive step, cpu_baseline of the oracle on a
bounded sample).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
B_FP64 = 32.0   # bytes per FP64 cell-update: 2 fields x (8 B read + 8 B write)
B_FP16 = 8.0    # bytes per fp16 cell-update: 2 fields x (2 B read + 2 B write)
METRIC = "cell-updates/s, AMP vs pure-FP64 on 1 MI355X at stated FP64-oracle L2 error"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


CHUNK = 2000   # steps per amps_step call between syncs (bounds the work in flight)


def step_chunked(ctx, steps):
    done = 0
    while done < steps:
        n = min(CHUNK, steps - done)
        ctx.step(n)
        done += n
        if done < steps:
            ctx.sync()


def run_path(amps, cfg, warmup, steps, kernel_pass=True, with_fields=True):
    """Timed run (graph replay, wall clock and HIP events around the step
    loop), then -- for the roofline -- the same warmup + steps again from the
    same initial state with HIP events bound to every sweep kernel (graph
    replay suspended), so the per-kernel averages cover the same trajectory
    as the timed region."""
    with amps.Context(cfg) as ctx:
        ctx.init()
        ctx.step(warmup)
        ctx.sync()
        t0 = time.perf_counter()
        step_chunked(ctx, steps)
        ev_ms = ctx.sync()
        wall = time.perf_counter() - t0
        info = ctx.info()
        fields = [ctx.get_field(f) for f in range(min(cfg.nfields(), 3))] if with_fields else None
        grain = None
        if with_fields and cfg.nfields() > 3:
            # grain map of an order-parameter model: sum of eta^2 and the index
            # of the largest eta per cell, accumulated field by field
            s2 = np.zeros((cfg.ny, cfg.nx))
            best = np.full((cfg.ny, cfg.nx), -np.inf)
            gid = np.zeros((cfg.ny, cfg.nx), np.int32)
            for f in range(cfg.nfields()):
                v = ctx.get_field(f)
                s2 += v * v
                m = v > best
                best[m] = v[m]
                gid[m] = f
            grain = (s2, gid)
    res = dict(wall=wall, ev_ms=ev_ms, info=info, fields=fields, grain=grain)
    if kernel_pass:
        with amps.Context(cfg) as ctx:
            ctx.init()
            ctx.step(warmup)
            ctx.sync()
            i0 = ctx.info()
            ctx.set_kernel_timing(True)
            step_chunked(ctx, steps)
            ctx.sync()
            res["k_launches"], res["k_fp64_ms"] = ctx.sweep_time(amps.SWEEP_FP64)
            _, res["k_low_ms"] = ctx.sweep_time(amps.SWEEP_LOW)
            i1 = ctx.info()
            res["k_info"] = i1
            # tiles updated by the snowflake's mixed sweep over the timed steps
            res["k_fp64_tiles"] = i1.fp64_tile_updates - i0.fp64_tile_updates
            res["k_full_tiles"] = i1.low_full_tile_updates - i0.low_full_tile_updates
    return res


def copy_probe():
    """A plain device-to-device copy of known size (1 GiB read + 1 GiB write,
    the HIP runtime's blit kernel), the calibration reference of the
    FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_size_t(1 << 30)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), n) != 0 or hip.hipMalloc(ctypes.byref(b), n) != 0:
        raise RuntimeError("copy probe: hipMalloc failed")
    try:
        hip.hipMemset(a, 1, n)
        hip.hipDeviceSynchronize()
        if hip.hipMemcpy(b, a, n, 3) != 0:   # hipMemcpyDeviceToDevice
            raise RuntimeError("copy probe: hipMemcpy failed")
        hip.hipDeviceSynchronize()
    finally:
        hip.hipFree(a)
        hip.hipFree(b)


def threshold_sweep(amps, amp_cfg, f64_cfg, args, f64, cells, log, horizons=(2000, 5000, 10000)):
    """The workload at monitor thresholds 0.01-0.2: rate, speedup over the
    FP64 run, FP64 tile share and the phi / T RMS error against FP64 at the
    end of the run and after shorter horizons (before the noisy dendrite's
    chaotic divergence dominates the difference)."""
    rms = lambda x, y: float(np.sqrt(np.mean((x - y) ** 2)))
    refs = {h: run_path(amps, f64_cfg, args.warmup, h, kernel_pass=False)["fields"] for h in horizons}
    rows = []
    for thr in (0.01, 0.02, 0.05, 0.1, 0.2):
        c = amp_cfg.copy()
        c.threshold = thr
        r = run_path(amps, c, args.warmup, args.steps, kernel_pass=False)
        row = {"threshold": thr, "value": cells * args.steps / r["wall"],
               "speedup_amp_over_fp64": f64["wall"] / r["wall"], "hi_frac_mean": r["info"].hi_frac_mean,
               "phi_l2_rms": rms(r["fields"][0], f64["fields"][0]), "T_l2_rms": rms(r["fields"][1], f64["fields"][1])}
        for h in horizons:
            fs = run_path(amps, c, args.warmup, h, kernel_pass=False)["fields"]
            row[f"phi_l2_rms_at_{h}"] = rms(fs[0], refs[h][0])
            row[f"T_l2_rms_at_{h}"] = rms(fs[1], refs[h][1])
        rows.append(row)
        log(f"[bench] threshold {thr}: {row['speedup_amp_over_fp64']:.2f}x, hi {row['hi_frac_mean']:.4f}")
    return rows


def copy_bandwidth(reps=10):
    """Achieved device-to-device copy bandwidth on this GPU (GB/s, bytes read
    plus written) of a 1 GiB buffer with the HIP runtime's copy kernel, as
    the STREAM-copy reference beside the nominal 8 TB/s peak."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = 1 << 30
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) != 0 or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) != 0:
        raise RuntimeError("copy bandwidth: hipMalloc failed")
    try:
        hip.hipMemset(a, 1, ctypes.c_size_t(n))
        hip.hipEventCreate(ctypes.byref(e0))
        hip.hipEventCreate(ctypes.byref(e1))
        hip.hipMemcpyAsync(b, a, ctypes.c_size_t(n), 3, None)  # hipMemcpyDeviceToDevice, warm-up
        hip.hipEventRecord(e0, None)
        for _ in range(reps):
            hip.hipMemcpyAsync(b, a, ctypes.c_size_t(n), 3, None)
        hip.hipEventRecord(e1, None)
        if hip.hipEventSynchronize(e1) != 0:
            raise RuntimeError("copy bandwidth: copy failed")
        ms = ctypes.c_float()
        hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
        return 2.0 * n * reps / (ms.value * 1e-3) / 1e9
    finally:
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
        hip.hipFree(a)
        hip.hipFree(b)


def traffic_from_profiles(name):
    """HBM bytes per launch of a kernel from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel), or None."""
    path = os.path.join(ROOT, "profiles", name + "_4096_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def other_workloads(amps, log):
    """The other configurations of the path on one MI355X, each adaptive vs
    FP64 on the same workload (wall time of the step loop, fields resident):
    Q=36 grain growth 2048^2 x 10000 steps with the tiling of each path chosen
    by amps_autotune; sintering 1024^2 x 5000 steps; a 64-instance snowflake
    ensemble (grids 128..1024, seeded delta and K) x 5000 steps."""
    P = os.path.join(ROOT, "w3914777_amd", "paras")
    res = {}
    # grain growth
    gg, ts = amps.Config.from_paras(amps.GRAINGROWTH, os.path.join(P, "graingrowth", "workload2048.h"))
    out = {}
    for name, mode in (("adaptive", amps.ADAPTIVE), ("fp64", amps.PURE_FP64)):
        c = gg.copy()
        c.mode = mode
        t = amps.autotune(c, 50)
        log(f"[bench] grain growth {name}: tile {t.tile}, {t.wg_per_cu} wg/CU")
        r = run_path(amps, amps.tuned(c, t), 10, ts, kernel_pass=False)
        out[name] = dict(r=r, tile=t.tile, wg_per_cu=t.wg_per_cu)
    cells = gg.nx * gg.ny
    a, f = out["adaptive"], out["fp64"]
    res["graingrowth_q36_2048"] = {
        "steps": ts, "cells": cells,
        "adaptive": {"value": cells * ts / a["r"]["wall"], "ms_per_step": a["r"]["wall"] * 1e3 / ts,
                     "tile": a["tile"], "wg_per_cu": a["wg_per_cu"], "hi_frac_mean": a["r"]["info"].hi_frac_mean},
        "fp64": {"value": cells * ts / f["r"]["wall"], "ms_per_step": f["r"]["wall"] * 1e3 / ts,
                 "tile": f["tile"], "wg_per_cu": f["wg_per_cu"]},
        "speedup_amp_over_fp64": f["r"]["wall"] / a["r"]["wall"],
        "rel_l2_eta_fields_0_2": [rel_l2(x, y) for x, y in zip(a["r"]["fields"], f["r"]["fields"])],
        "grain_map_vs_fp64": {
            "sum_eta2_rel_l2": rel_l2(a["r"]["grain"][0], f["r"]["grain"][0]),
            "sum_eta2_linf": float(np.max(np.abs(a["r"]["grain"][0] - f["r"]["grain"][0]))),
            "grain_id_agreement": float(np.mean(a["r"]["grain"][1] == f["r"]["grain"][1]))},
        "unit": "cell-updates/s"}
    # sintering
    sn, ts = amps.Config.from_paras(amps.SINTERING, os.path.join(P, "sintering", "workload1024.h"))
    rr = {}
    for name, mode in (("adaptive", amps.ADAPTIVE), ("fp64", amps.PURE_FP64)):
        c = sn.copy()
        c.mode = mode
        rr[name] = run_path(amps, c, 10, ts, kernel_pass=False)
    cells = sn.nx * sn.ny
    res["sintering_1024"] = {
        "steps": ts, "adaptive": {"value": cells * ts / rr["adaptive"]["wall"],
                                  "hi_frac_mean": rr["adaptive"]["info"].hi_frac_mean},
        "fp64": {"value": cells * ts / rr["fp64"]["wall"]},
        "speedup_amp_over_fp64": rr["fp64"]["wall"] / rr["adaptive"]["wall"],
        "rel_l2_c_eta1_eta2": [rel_l2(x, y) for x, y in zip(rr["adaptive"]["fields"], rr["fp64"]["fields"])],
        "unit": "cell-updates/s"}
    # batched snowflake ensemble
    base, _ = amps.Config.from_paras(amps.SNOWFLAKE, os.path.join(P, "snowflake", "workload4096.h"))
    rng = np.random.default_rng(1)
    cfgs = []
    for i in range(64):
        c = base.copy()
        n = int(rng.choice([128, 256, 512, 1024]))
        c.nx = c.ny = n
        c.params[amps.SF["CX"]] = c.params[amps.SF["CY"]] = n // 2
        c.params[amps.SF["DELTA"]] = rng.uniform(0.02, 0.08)
        c.params[amps.SF["KAPPA"]] = rng.uniform(1.2, 2.0)
        c.seed = 1000 + i
        cfgs.append(c)
    steps = 5000
    cells = sum(c.nx * c.ny for c in cfgs)
    walls, phis = {}, {}
    for name, mode in (("adaptive", amps.ADAPTIVE), ("fp64", amps.PURE_FP64)):
        cs = [c.copy() for c in cfgs]
        for c in cs:
            c.mode = mode
        with amps.Context(cs) as ctx:
            ctx.init()
            ctx.step(10)
            ctx.sync()
            t0 = time.perf_counter()
            step_chunked(ctx, steps)
            ctx.sync()
            walls[name] = time.perf_counter() - t0
            phis[name] = [ctx.get_field(0, inst=i) for i in range(len(cs))]
    rms = [float(np.sqrt(np.mean((x - y) ** 2))) for x, y in zip(phis["adaptive"], phis["fp64"])]
    res["snowflake_ensemble_64"] = {
        "steps": steps, "cells": cells, "instances": 64,
        "adaptive": {"value": cells * steps / walls["adaptive"]}, "fp64": {"value": cells * steps / walls["fp64"]},
        "speedup_amp_over_fp64": walls["fp64"] / walls["adaptive"],
        "phi_l2_rms_vs_fp64": {"median": float(np.median(rms)), "max": max(rms)}, "unit": "cell-updates/s"}
    return res


def cpu_baseline(cfg, seconds_budget):
    """Oracle FP64 step (OpenMP C++) on the box's host cores, same workload,
    a bounded number of steps from the initial state."""
    from oracle import oracle as O
    nx, ny = cfg.nx, cfg.ny
    p = np.array(cfg.params[:], dtype=np.float64)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(os.cpu_count() or 1, 16)
    phi, T = O.sf_init(nx, ny, p)
    # calibrate on one step, then run a sample of about `seconds_budget`
    _, _, s1 = O.sf_run(phi, T, p, 1, stencil=cfg.stencil, bc=cfg.bc, seed=cfg.seed, threads=threads)
    n = max(2, int(seconds_budget / max(s1, 1e-3)))
    _, _, sec = O.sf_run(phi, T, p, n, stencil=cfg.stencil, bc=cfg.bc, seed=cfg.seed, threads=threads)
    return dict(value=nx * ny * n / sec, unit="cell-updates/s", cores=threads, kind="port",
                sample=f"oracle FP64 snowflake step (OpenMP C++, g++ -O3) on the {nx}x{ny} workload, "
                       f"{n} steps from the initial state, {sec:.1f} s on {threads} threads")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--copy-probe", action="store_true",
                    help="run a 1 GiB device copy first (counter calibration under rocprofv3 --pmc)")
    ap.add_argument("--threshold", type=float, default=None)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-others", action="store_true", help="skip the grain-growth, sintering and ensemble lines")
    ap.add_argument("--no-sweep", action="store_true", help="skip the monitor-threshold sweep")
    ap.add_argument("--repeats", type=int, default=5,
                    help="timed runs of each path from the initial state; value is the median")
    ap.add_argument("--traffic-run", action="store_true",
                    help="one adaptive run of warmup + steps only, printing the mixed sweep's algorithmic bytes "
                         "per launch over it (the companion of a rocprofv3 --pmc pass)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X; --gpus {args.gpus} is not supported (the path is single-GPU).",
              file=sys.stderr)
        sys.exit(2)

    import w3914777_amd as amps
    cfg, _ = amps.Config.from_paras(amps.SNOWFLAKE, os.path.join(ROOT, "w3914777_amd", "paras", "snowflake",
                                                                 "workload4096.h"))
    if args.threshold is not None:
        cfg.threshold = args.threshold
    cells = cfg.nx * cfg.ny
    amp_cfg = cfg.copy()
    amp_cfg.mode = amps.ADAPTIVE
    f64_cfg = cfg.copy()
    f64_cfg.mode = amps.PURE_FP64

    if args.copy_probe:
        copy_probe()
    if args.traffic_run:
        log(f"[bench] traffic run: adaptive {cfg.nx}x{cfg.ny}, {args.warmup}+{args.steps} steps")
        r = run_path(amps, amp_cfg, args.warmup, args.steps, kernel_pass=False, with_fields=False)
        n = args.warmup + args.steps   # one mixed-sweep launch per step
        i = r["info"]
        tc = cfg.tile * cfg.tile
        print(json.dumps({"traffic_run": "sf_mixed_kernel<32,false>", "launches": n,
                          "fp64_tile_updates": i.fp64_tile_updates, "low_full_tile_updates": i.low_full_tile_updates,
                          "alg_bytes_per_launch": (i.fp64_tile_updates * tc * B_FP64
                                                   + i.low_full_tile_updates * tc * B_FP16) / n}))
        return
    copy_gbs = copy_bandwidth()
    log(f"[bench] device copy: {copy_gbs:.0f} GB/s (read + write)")
    log(f"[bench] adaptive path: {cfg.nx}x{cfg.ny}, warmup {args.warmup}, steps {args.steps}, {args.repeats} runs")
    amp = run_path(amps, amp_cfg, args.warmup, args.steps)
    amp_walls = [amp["wall"]] + [run_path(amps, amp_cfg, args.warmup, args.steps, kernel_pass=False,
                                          with_fields=False)["wall"] for _ in range(args.repeats - 1)]
    amp["wall"] = float(np.median(amp_walls))
    log(f"[bench] adaptive: {amp['wall']:.3f} s wall (median of {len(amp_walls)}), hi_frac {amp['info'].hi_frac_mean:.4f}")
    f64 = None
    if not args.no_fp64:
        log("[bench] FP64 path")
        f64 = run_path(amps, f64_cfg, args.warmup, args.steps)
        f64_walls = [f64["wall"]] + [run_path(amps, f64_cfg, args.warmup, args.steps, kernel_pass=False,
                                              with_fields=False)["wall"] for _ in range(args.repeats - 1)]
        f64["wall"] = float(np.median(f64_walls))
        log(f"[bench] fp64: {f64['wall']:.3f} s wall (median of {len(f64_walls)})")

    value = cells * args.steps / amp["wall"]
    out = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": amp["wall"] * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64+f16", "data": "synthetic",
        "config": {"workload": "Kobayashi snowflake 4096x4096, adaptive FP64/fp16 tiles 32x32, gradient monitor "
                               f"{cfg.threshold}, 5-point, zero flux, noise 0.01, {args.warmup}+{args.steps} steps",
                   "grid": [cfg.nx, cfg.ny], "tile": cfg.tile, "threshold": cfg.threshold},
    }
    # roofline of the dominant kernel of the adaptive step, the mixed sweep
    # (FP64 tiles plus the low tiles outside the bulk-hold state): its
    # algorithmic bytes per launch are 32 B per FP64 cell and 8 B per such low
    # cell, from the device's tile counts over the kernel-timed steps.  The
    # bulk-hold sweep moves state-dependent bytes (T of the tiles that are not
    # T-flat, nothing for the rest) and is reported by its time.
    nl = max(amp["k_launches"], 1)
    tile_cells = cfg.tile * cfg.tile
    mix_s = amp["k_fp64_ms"] / nl / 1e3
    hold_s = amp["k_low_ms"] / nl / 1e3
    fp64_tiles = amp["k_fp64_tiles"] / nl
    full_tiles = amp["k_full_tiles"] / nl
    mix_bytes = fp64_tiles * tile_cells * B_FP64 + full_tiles * tile_cells * B_FP16
    achieved = mix_bytes / mix_s / 1e9
    tr = traffic_from_profiles("sf_mixed_kernel")
    out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS,
                       "traffic": tr["bytes_per_launch"] if tr else None,
                       "kernel": "sf_mixed_kernel<32,false>", "avg_launch_us": mix_s * 1e6,
                       "alg_bytes_per_launch": mix_bytes, "fp64_tiles_per_launch": fp64_tiles,
                       "low_full_tiles_per_launch": full_tiles, "device_copy_gbs": copy_gbs,
                       "frac_of_device_copy": achieved / copy_gbs}
    out["runs"] = {"adaptive_values": [cells * args.steps / w for w in amp_walls],
                   "note": f"value is the median of {len(amp_walls)} timed runs from the initial state"}
    out["adaptive_step_kernels_us"] = {"sf_mixed_kernel<32,false>": mix_s * 1e6,
                                       "sf_hold_kernel<32,1,false>": hold_s * 1e6,
                                       "note": "per-launch averages over the timed steps, HIP events bound "
                                               "to each kernel's dispatch (its execution time)"}
    out["hi_frac_mean"] = amp["info"].hi_frac_mean
    if f64 is not None:
        f64_value = cells * args.steps / f64["wall"]
        kf_s = f64["k_fp64_ms"] / max(f64["k_launches"], 1) / 1e3
        fa = B_FP64 * cells / kf_s / 1e9
        err = {}
        for name, a, b in (("phi", amp["fields"][0], f64["fields"][0]), ("T", amp["fields"][1], f64["fields"][1])):
            d = a - b
            err[name] = {"l2_rms": float(np.sqrt(np.mean(d * d))),
                         "l2_rel": float(np.linalg.norm(d) / max(np.linalg.norm(b), 1e-300)),
                         "linf": float(np.max(np.abs(d)))}
        out["fp64"] = {"value": f64_value, "ms_per_step": f64["wall"] * 1e3 / args.steps,
                       "speedup_amp_over_fp64": value / f64_value,
                       "roofline": {"achieved": fa, "peak": HBM_PEAK_GBS, "frac": fa / HBM_PEAK_GBS,
                                    "frac_of_device_copy": fa / copy_gbs,
                                    "kernel": "sf_high_kernel<32,false,false>", "avg_launch_us": kf_s * 1e6},
                       "runs_values": [cells * args.steps / w for w in f64_walls]}
        out["error_vs_fp64"] = err
        if not args.no_sweep:
            out["threshold_sweep"] = threshold_sweep(amps, amp_cfg, f64_cfg, args, f64, cells, log)
    if not args.no_others:
        out["other_workloads"] = other_workloads(amps, log)
    if not args.no_cpu:
        log("[bench] cpu baseline")
        out["cpu_baseline"] = cpu_baseline(f64_cfg, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()

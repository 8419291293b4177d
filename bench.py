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
the adaptive sweep kernel, cpu_baseline of the oracle on a bounded sample).
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


def run_path(amps, cfg, warmup, steps, kernel_steps):
    """Returns dict with wall/event timings, final fields, info and sweep-kernel timing."""
    with amps.Context(cfg) as ctx:
        ctx.init()
        ctx.step(warmup)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.step(steps)
        ev_ms = ctx.sync()
        wall = time.perf_counter() - t0
        info = ctx.info()
        fields = [ctx.get_field(0), ctx.get_field(1)]
        # sweep-kernel timing pass: events bracket every step kernel on the
        # context's stream (graph replay suspended), same workload state
        ctx.set_kernel_timing(True)
        ctx.step(kernel_steps)
        ctx.sync()
        nk, kms = ctx.kernel_time()
        info2 = ctx.info()
        ctx.set_kernel_timing(False)
        return dict(wall=wall, ev_ms=ev_ms, info=info, fields=fields, k_launches=nk, k_ms=kms,
                    hi_kernel=info2.hi_frac_last)


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
    ap.add_argument("--kernel-steps", type=int, default=200)
    ap.add_argument("--threshold", type=float, default=None)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
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

    log(f"[bench] adaptive path: {cfg.nx}x{cfg.ny}, warmup {args.warmup}, steps {args.steps}")
    amp = run_path(amps, amp_cfg, args.warmup, args.steps, args.kernel_steps)
    log(f"[bench] adaptive: {amp['wall']:.3f} s wall, {amp['ev_ms']:.1f} ms events, hi_frac {amp['info'].hi_frac_mean:.4f}")
    f64 = None
    if not args.no_fp64:
        log("[bench] FP64 path")
        f64 = run_path(amps, f64_cfg, args.warmup, args.steps, args.kernel_steps)
        log(f"[bench] fp64: {f64['wall']:.3f} s wall, {f64['ev_ms']:.1f} ms events")

    value = cells * args.steps / amp["wall"]
    out = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": amp["wall"] * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64+f16", "data": "synthetic",
        "config": {"workload": "Kobayashi snowflake 4096x4096, adaptive FP64/fp16 tiles 32x32, gradient monitor "
                               f"{cfg.threshold}, 5-point, zero flux, noise 0.01, {args.warmup}+{args.steps} steps",
                   "grid": [cfg.nx, cfg.ny], "tile": cfg.tile, "threshold": cfg.threshold},
    }
    # roofline of the adaptive sweep kernel: algorithmic bytes per launch
    hi = amp["hi_kernel"]
    b_alg = hi * B_FP64 + (1 - hi) * B_FP16
    k_avg_s = amp["k_ms"] / max(amp["k_launches"], 1) / 1e3
    achieved = b_alg * cells / k_avg_s / 1e9
    out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                       "kernel": "sf_step_kernel<32,true>", "avg_launch_us": k_avg_s * 1e6,
                       "bytes_per_cell_update": b_alg, "hi_frac": hi}
    out["hi_frac_mean"] = amp["info"].hi_frac_mean
    if f64 is not None:
        f64_value = cells * args.steps / f64["wall"]
        kf_s = f64["k_ms"] / max(f64["k_launches"], 1) / 1e3
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
                                    "kernel": "sf_step_kernel<32,false>", "avg_launch_us": kf_s * 1e6}}
        out["error_vs_fp64"] = err
    if not args.no_cpu:
        log("[bench] cpu baseline")
        out["cpu_baseline"] = cpu_baseline(f64_cfg, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()

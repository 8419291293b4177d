"""Benchmark of the adaptive FP64/fp16 snowflake stencil path on one MI355X.

Workload (w3914777_amd/paras/snowflake/workload4096.h): Kobayashi snowflake
4096x4096, dx=0.03, dt=1e-4, zero-flux boundaries, 5-point Laplacian,
eps=0.01(1+0.05cos(6(theta-pi/2))), tau=3e-4, alpha=0.9, gamma=10, K=1.6,
Teq=1, radius-5 nucleus, interface noise a=0.01; 32x32 tiles, gradient
monitor |grad phi| dx > 0.05 every 10 steps, 20000 steps.

A step advances every cell of the grid by one explicit-Euler step (both
fields).  The adaptive path runs with AMPS_FLAG_NO_SKIP: every low-precision
tile gets the full update every step, as the reference's type-1 branch does,
so every counted cell-update is computed.  Its FP64 set is the monitor's
tiles dilated by `--dilate` tiles.

Timing: each path is advanced untimed to step 20000 - R*K (the developed
dendrite; --warmup W steps at least), then R windows of K = --steps steps are
timed back to back (graph replay, fields resident in HBM, wall clock around
amps_step + amps_sync); value = cells * K / median window time, so the final
window ends at step 20000, where the adaptive fields are compared with the
FP64 path's (phi and T, RMS and max error).

Prints one JSON line: metric, value, roofline of the dominant kernel of the
adaptive step (HIP events bound to each sweep dispatch over the same
windows), cpu_baseline (the oracle on the host cores), the FP64 path, the
error of this run, the dilation frontier, the engine's skipping variant and
the other workloads of the path.
"""
import argparse
import json
import os
import sys
import time

os.environ.setdefault("OMP_PROC_BIND", "close")  # before any OpenMP runtime starts

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
B_FP64 = 32.0   # bytes per FP64 cell-update: 2 fields x (8 B read + 8 B write)
B_FP16 = 8.0    # bytes per fp16 cell-update: 2 fields x (2 B read + 2 B write)
METRIC = "cell-updates/s, AMP vs pure-FP64 on 1 MI355X at stated FP64-oracle L2 error"
PARAS = os.path.join(ROOT, "w3914777_amd", "paras")
CHUNK = 2000   # steps per amps_step call between syncs (bounds the work in flight)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def advance(ctx, steps):
    done = 0
    while done < steps:
        n = min(CHUNK, steps - done)
        ctx.step(n)
        ctx.sync()
        done += n


def rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)))


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def timed_run(amps, cfg, horizon, K, R, W, fields=True, kernel_pass=False):
    """Advance to max(W, horizon - R K) untimed, time R windows of K steps.
    With kernel_pass, the same trajectory again with HIP events bound to each
    sweep dispatch (graph replay off) over the same windows."""
    start = max(W, horizon - R * K)
    res = {"start_step": start, "end_step": start + R * K}
    with amps.Context(cfg) as ctx:
        ctx.init()
        advance(ctx, start)
        walls, evs = [], []
        for _ in range(R):
            t0 = time.perf_counter()
            ctx.step(K)
            evs.append(ctx.sync())
            walls.append(time.perf_counter() - t0)
        res["walls"], res["event_ms"] = walls, evs
        res["info"] = ctx.info()
        if fields:
            res["fields"] = [ctx.get_field(f) for f in range(min(cfg.nfields(), 3))]
    if kernel_pass:
        with amps.Context(cfg) as ctx:
            ctx.init()
            advance(ctx, start)
            i0 = ctx.info()
            ctx.set_kernel_timing(True)
            ctx.step(R * K)
            ctx.sync()
            n, res["k_fp64_ms"] = ctx.sweep_time(amps.SWEEP_FP64)
            _, res["k_low_ms"] = ctx.sweep_time(amps.SWEEP_LOW)
            res["k_launches"] = n
            i1 = ctx.info()
            res["k_fp64_tiles"] = i1.fp64_tile_updates - i0.fp64_tile_updates
            res["k_low_tiles"] = i1.low_full_tile_updates - i0.low_full_tile_updates
            res["k_listed_tiles"] = i1.low_listed_tile_updates - i0.low_listed_tile_updates
    return res


def rate(cells, K, r):
    return cells * K / float(np.median(r["walls"]))


def field_errors(a, b, names=("phi", "T")):
    out = {}
    for name, x, y in zip(names, a, b):
        d = x - y
        out[name] = {"l2_rms": float(np.sqrt(np.mean(d * d))), "l2_rel": rel_l2(x, y),
                     "linf": float(np.max(np.abs(d)))}
    return out


def copy_bandwidth(reps=10):
    """Achieved device-to-device copy bandwidth (GB/s, bytes read plus
    written) of a 1 GiB buffer with the HIP runtime's copy kernel: the
    STREAM-copy reference beside the nominal 8 TB/s peak, and the
    calibration kernel of the FETCH_SIZE / WRITE_SIZE counters."""
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


def traffic_from_profiles(kernel_key, workload_key, start_step, end_step):
    """HBM bytes per launch of a sweep kernel (FETCH_SIZE + WRITE_SIZE,
    calibrated on the copy kernel) from the committed rocprofv3 --pmc
    summary, used only when it was measured on this workload's trajectory
    and covers exactly the steps of this run's windows (the summary keeps
    per-block means of the counters of each step's dispatch)."""
    path = os.path.join(ROOT, "profiles", kernel_key + "_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("workload_key") != workload_key:
        return None
    first, last, blk = t["first_step"], t["end_step"], t["block_steps"]
    if not (first <= start_step < end_step <= last) or (start_step - first) % blk or (end_step - start_step) % blk:
        return None
    i0, i1 = (start_step - first) // blk, (end_step - first) // blk
    fb, wb = t["fetch_size_kib_blocks"][i0:i1], t["write_size_kib_blocks"][i0:i1]
    b = (t["fetch_factor"] * sum(fb) / len(fb) + t["write_factor"] * sum(wb) / len(wb)) * 1024.0
    return {"bytes_per_launch": b, "source": t["source"], "steps": [start_step, end_step]}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(amps, cfg, seconds_budget):
    """The oracle's FP64 snowflake step (OpenMP C++, g++ -O3 -march=x86-64-v3,
    OMP_PROC_BIND=close) on the host cores given to this process: the
    256x256 x 2000-step workload run in full, and a bounded sample of the
    4096x4096 workload (value)."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    p = np.array(cfg.params[:], dtype=np.float64)
    # the 256^2 x 2000-step workload (the same physics), in full
    p0 = p.copy()
    p0[amps.SF["CX"]] = p0[amps.SF["CY"]] = 128
    phi, T = O.sf_init(256, 256, p0)
    _, _, s0 = O.sf_run(phi, T, p0, 2000, stencil=cfg.stencil, bc=cfg.bc, seed=cfg.seed, threads=threads)
    small = {"grid": [256, 256], "steps": 2000, "seconds": s0, "value": 256 * 256 * 2000 / s0}
    nx, ny = cfg.nx, cfg.ny
    phi, T = O.sf_init(nx, ny, p)
    _, _, s1 = O.sf_run(phi, T, p, 1, stencil=cfg.stencil, bc=cfg.bc, seed=cfg.seed, threads=threads)
    n = max(2, int(seconds_budget / max(s1, 1e-3)))
    _, _, sec = O.sf_run(phi, T, p, n, stencil=cfg.stencil, bc=cfg.bc, seed=cfg.seed, threads=threads)
    return dict(value=nx * ny * n / sec, unit="cell-updates/s", cores=threads, kind="port",
                cpu_model=cpu_model(), omp_proc_bind=os.environ.get("OMP_PROC_BIND"),
                sample=f"oracle FP64 snowflake step (OpenMP C++, g++ -O3 -march=x86-64-v3) on the {nx}x{ny} "
                       f"workload, {n} steps from the initial state, {sec:.1f} s on {threads} threads",
                workload_256x256_2000_steps=small)


def other_workloads(amps, args):
    """The other configurations of the path, each adaptive (NO_SKIP) vs FP64
    on the same workload, wall time of the whole step loop from the initial
    state with fields resident: Q=36 grain growth 2048^2 x 10000 steps
    (FP64 tiling chosen by amps_autotune, adaptive on 16^2 tiles), sintering 1024^2 x 5000 steps (and
    its adjoint), a 64-instance snowflake ensemble x 5000 steps."""
    res = {}

    def whole_run(cfg, steps, fields=True):
        with amps.Context(cfg) as ctx:
            ctx.init()
            ctx.step(10)
            ctx.sync()
            t0 = time.perf_counter()
            advance(ctx, steps)
            wall = time.perf_counter() - t0
            info = ctx.info()
            out = dict(wall=wall, info=info)
            if fields:
                out["fields"] = [ctx.get_field(f) for f in range(min(cfg.nfields(), 3))]
                if cfg.nfields() > 3:
                    s2 = np.zeros((cfg.ny, cfg.nx))
                    best = np.full((cfg.ny, cfg.nx), -np.inf)
                    gid = np.zeros((cfg.ny, cfg.nx), np.int32)
                    for f in range(cfg.nfields()):
                        v = ctx.get_field(f)
                        s2 += v * v
                        m = v > best
                        best[m] = v[m]
                        gid[m] = f
                    out["grain"] = (s2, gid)
        return out

    # grain growth: the adaptive run at --gg-threshold (NO_SKIP), a frontier
    # of monitor thresholds, and the engine's skipping variant
    gg, ts = amps.Config.from_paras(amps.GRAINGROWTH, os.path.join(PARAS, "graingrowth", "workload2048.h"))
    cells = gg.nx * gg.ny
    frontier_thr = [t for t in (1e-4, 1e-3, gg.threshold) if t != args.gg_threshold]
    plan = [("fp64", amps.PURE_FP64, 0, gg.threshold), ("adaptive", amps.ADAPTIVE, amps.FLAG_NO_SKIP, args.gg_threshold)]
    plan += [(f"adaptive_thr_{t:g}", amps.ADAPTIVE, amps.FLAG_NO_SKIP, t) for t in frontier_thr]
    plan += [("adaptive_skipping", amps.ADAPTIVE, 0, gg.threshold)]
    # The FP64 run takes the autotuner's tiling.  The adaptive runs use 16^2
    # tiles: the tuner times 50 steps from the initial state, where every
    # field is live noise and 32^2 tiles win, while over the 10000 steps most
    # fields become exact zeros on most tiles, which the 16^2 one-wave sweep
    # skips (measured over the whole run: 16^2 is the faster adaptive tiling).
    runs = {}
    for name, mode, flags, thr in plan:
        c = gg.copy()
        c.mode, c.flags, c.threshold = mode, flags, thr
        if mode == amps.PURE_FP64:
            t = amps.autotune(c, 50)
            c = amps.tuned(c, t)
        else:
            c.tile, c.wg_per_cu = 16, 0
        r = whole_run(c, ts)
        r["tile"], r["wg_per_cu"], r["threshold"] = c.tile, c.wg_per_cu, thr
        runs[name] = r
        log(f"[bench] grain growth {name}: {r['wall'] * 1e3 / ts:.3f} ms/step (tile {c.tile}, {c.wg_per_cu} wg/CU)")
    f = runs["fp64"]
    out = {"steps": ts, "cells": cells, "unit": "cell-updates/s",
           "fp64": {"value": cells * ts / f["wall"], "ms_per_step": f["wall"] * 1e3 / ts, "tile": f["tile"]}}

    def summary(a):
        return {"value": cells * ts / a["wall"], "ms_per_step": a["wall"] * 1e3 / ts, "tile": a["tile"],
                "threshold": a["threshold"], "hi_frac_mean": a["info"].hi_frac_mean,
                "speedup_amp_over_fp64": f["wall"] / a["wall"],
                "rel_l2_eta_fields_0_2": [rel_l2(x, y) for x, y in zip(a["fields"], f["fields"])],
                "grain_map_vs_fp64": {
                    "sum_eta2_rel_l2": rel_l2(a["grain"][0], f["grain"][0]),
                    "grain_id_agreement": float(np.mean(a["grain"][1] == f["grain"][1]))}}

    out["adaptive"] = summary(runs["adaptive"])
    out["adaptive"]["flags"] = "NO_SKIP"
    out["threshold_frontier"] = [summary(runs[n]) for n, *_ in plan if n.startswith("adaptive_thr_") or n == "adaptive"]
    out["threshold_frontier"].sort(key=lambda r: r["threshold"])
    out["adaptive_skipping"] = summary(runs["adaptive_skipping"])
    out["adaptive_skipping"]["note"] = ("not NO_SKIP: absent grains flushed to exact zeros and grain interiors held; "
                                        "the FP64 path has no such skipping")
    res["graingrowth_q36_2048"] = out

    # sintering
    sn, ts = amps.Config.from_paras(amps.SINTERING, os.path.join(PARAS, "sintering", "workload1024.h"))
    cells = sn.nx * sn.ny
    rr = {}
    for name, mode, flags in (("fp64", amps.PURE_FP64, 0), ("adaptive", amps.ADAPTIVE, amps.FLAG_NO_SKIP)):
        c = sn.copy()
        c.mode, c.flags = mode, flags
        rr[name] = whole_run(c, ts)
    res["sintering_1024"] = {
        "steps": ts, "unit": "cell-updates/s",
        "adaptive": {"value": cells * ts / rr["adaptive"]["wall"], "hi_frac_mean": rr["adaptive"]["info"].hi_frac_mean},
        "fp64": {"value": cells * ts / rr["fp64"]["wall"]},
        "speedup_amp_over_fp64": rr["fp64"]["wall"] / rr["adaptive"]["wall"],
        "rel_l2_c_eta1_eta2": [rel_l2(x, y) for x, y in zip(rr["adaptive"]["fields"], rr["fp64"]["fields"])]}
    res["sintering_1024_adjoint"] = sintering_adjoint(amps, sn, ts)

    # batched snowflake ensemble
    base, _ = amps.Config.from_paras(amps.SNOWFLAKE, os.path.join(PARAS, "snowflake", "workload4096.h"))
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
    for name, mode, flags, dil in (("adaptive", amps.ADAPTIVE, amps.FLAG_NO_SKIP, args.dilate),
                                   ("fp64", amps.PURE_FP64, 0, 0)):
        cs = [c.copy() for c in cfgs]
        for c in cs:
            c.mode, c.flags, c.dilate = mode, flags, dil
        with amps.Context(cs) as ctx:
            ctx.init()
            ctx.step(10)
            ctx.sync()
            t0 = time.perf_counter()
            advance(ctx, steps)
            walls[name] = time.perf_counter() - t0
            phis[name] = [ctx.get_field(0, inst=i) for i in range(len(cs))]
    e = [rms(x, y) for x, y in zip(phis["adaptive"], phis["fp64"])]
    res["snowflake_ensemble_64"] = {
        "steps": steps, "cells": cells, "instances": 64, "unit": "cell-updates/s",
        "sides": [c.nx for c in cfgs], "dilate": args.dilate, "flags": "NO_SKIP",
        "adaptive": {"value": cells * steps / walls["adaptive"]}, "fp64": {"value": cells * steps / walls["fp64"]},
        "speedup_amp_over_fp64": walls["fp64"] / walls["adaptive"],
        "phi_l2_rms_vs_fp64": {"median": float(np.median(e)), "max": max(e)}}
    return res


def sintering_adjoint(amps, sn, steps, ckpt=100):
    """Reverse-time adjoint of the 1024^2 sintering workload over the full
    5000 steps with checkpoints every 100: gradient of J = ||c_final -
    c_target||^2 w.r.t. (Dsurf, Dgb, kappa_c), c_target from the same run at
    parameters perturbed by 1%; the Dsurf component is checked against a
    central difference (h = 1e-6 relative)."""
    ids = [amps.GRAD_SN["DSURF"], amps.GRAD_SN["DGB"], amps.GRAD_SN["KC"]]
    slots = [amps.SN["DSURF"], amps.SN["DGB"], amps.SN["KC"]]
    base = sn.copy()
    base.mode = amps.PURE_FP64

    def final_c(cfg):
        with amps.Context(cfg) as ctx:
            ctx.init()
            advance(ctx, steps)
            return ctx.get_field(0)

    tgt = base.copy()
    for s in slots:
        tgt.params[s] *= 1.01
    c_target = final_c(tgt)
    with amps.Context(base) as ctx:
        ctx.init()
        t0 = time.perf_counter()
        ctx.tape_begin(ckpt)
        advance(ctx, steps)
        t_fwd = time.perf_counter() - t0
        c = ctx.get_field(0)
        lam = [np.stack([2.0 * (c - c_target), np.zeros_like(c), np.zeros_like(c)])]
        t0 = time.perf_counter()
        g, _ = ctx.adjoint(lam, ids)
        t_rev = time.perf_counter() - t0
    J = lambda cc: float(np.sum((cc - c_target) ** 2))
    h = 1e-6 * base.params[slots[0]]
    cp, cm = base.copy(), base.copy()
    cp.params[slots[0]] += h
    cm.params[slots[0]] -= h
    fd = (J(final_c(cp)) - J(final_c(cm))) / (2 * h)
    return {"steps": steps, "checkpoint_every": ckpt, "grid": [sn.nx, sn.ny],
            "dJ_d": {"Dsurf": g[0, 0], "Dgb": g[0, 1], "kappa_c": g[0, 2]},
            "central_difference_Dsurf": fd, "rel_diff_Dsurf": abs(g[0, 0] - fd) / max(abs(fd), 1e-300),
            "forward_taped_s": t_fwd, "reverse_s": t_rev,
            "reverse_cell_steps_per_s": sn.nx * sn.ny * steps / t_rev}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="K: steps per timed window")
    ap.add_argument("--warmup", type=int, default=20, help="W: untimed steps at least")
    ap.add_argument("--repeats", type=int, default=5, help="R: timed windows per path; value is the median")
    ap.add_argument("--horizon", type=int, default=None, help="step the last window ends at (workload: 20000)")
    ap.add_argument("--dilate", type=int, default=2, help="FP64-set dilation of the adaptive path (tiles)")
    ap.add_argument("--gg-threshold", type=float, default=3e-5,
                    help="gradient-monitor threshold of the grain-growth adaptive run (NO_SKIP)")
    ap.add_argument("--frontier", default="0,1,2,3,4", help="dilations of the error/speed frontier ('' to skip)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-others", action="store_true", help="skip the grain-growth, sintering and ensemble lines")
    ap.add_argument("--copy-probe", action="store_true",
                    help="with --profile-run: a 1 GiB device copy first (the calibration kernel of FETCH_SIZE / "
                         "WRITE_SIZE counter passes)")
    ap.add_argument("--profile-run", action="store_true",
                    help="only one path's trajectory and timed windows (for rocprofv3 runs); prints the tile "
                         "updates of the run")
    ap.add_argument("--profile-path", choices=("adaptive", "half2", "fp64", "sintering", "graingrowth"),
                    default="adaptive",
                    help="with --profile-run: the path to run (half2: the adaptive path with "
                         "AMPS_FLAG_HALF2; sintering: 300 steps of the 1024^2 sintering "
                         "workload, FP64 then adaptive NO_SKIP; graingrowth: the Q=36 2048^2 workload's "
                         "10000 steps, FP64 on 32^2 tiles then adaptive NO_SKIP at --gg-threshold on 16^2)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X; --gpus {args.gpus} is not supported (the path is single-GPU).",
              file=sys.stderr)
        sys.exit(2)

    import w3914777_amd as amps
    cfg, horizon = amps.Config.from_paras(amps.SNOWFLAKE, os.path.join(PARAS, "snowflake", "workload4096.h"))
    horizon = args.horizon or horizon or 20000
    K, R, W = args.steps, args.repeats, args.warmup
    cells = cfg.nx * cfg.ny
    tile_cells = cfg.tile * cfg.tile
    amp_cfg = cfg.copy()
    amp_cfg.mode, amp_cfg.flags, amp_cfg.dilate = amps.ADAPTIVE, amps.FLAG_NO_SKIP, args.dilate
    f64_cfg = cfg.copy()
    f64_cfg.mode = amps.PURE_FP64
    workload_key = f"snowflake4096_nsk_d{args.dilate}"

    if args.profile_run and args.profile_path == "graingrowth":
        gg, ts = amps.Config.from_paras(amps.GRAINGROWTH, os.path.join(PARAS, "graingrowth", "workload2048.h"))
        for mode, tile, thr, flags in ((amps.PURE_FP64, 32, gg.threshold, 0),
                                       (amps.ADAPTIVE, 16, args.gg_threshold, amps.FLAG_NO_SKIP)):
            c = gg.copy()
            c.mode, c.tile, c.threshold, c.flags = mode, tile, thr, flags
            with amps.Context(c) as ctx:
                ctx.init()
                t0 = time.perf_counter()
                advance(ctx, ts)
                print(json.dumps({"profile_run": "graingrowth2048_q36", "mode": int(mode), "tile": tile,
                                  "threshold": thr, "steps": ts, "us_per_step": (time.perf_counter() - t0) * 1e6 / ts,
                                  "hi_frac_mean": ctx.info().hi_frac_mean}))
        return
    if args.profile_run and args.profile_path == "sintering":
        sn, _ = amps.Config.from_paras(amps.SINTERING, os.path.join(PARAS, "sintering", "workload1024.h"))
        for mode, flags in ((amps.PURE_FP64, 0), (amps.ADAPTIVE, amps.FLAG_NO_SKIP)):
            c = sn.copy()
            c.mode, c.flags = mode, flags
            with amps.Context(c) as ctx:
                ctx.init()
                advance(ctx, 300)
                print(json.dumps({"profile_run": "sintering1024", "mode": int(mode), "steps": 300,
                                  "hi_frac_last": ctx.info().hi_frac_last}))
        return
    if args.profile_run:
        if args.copy_probe:
            copy_bandwidth(reps=1)
        pc = {"adaptive": amp_cfg, "half2": amp_cfg.copy(), "fp64": f64_cfg}[args.profile_path]
        if args.profile_path == "half2":
            pc.flags = amps.FLAG_NO_SKIP | amps.FLAG_HALF2
        r = timed_run(amps, pc, horizon, K, R, W, fields=False)
        i = r["info"]
        n = i.steps_done
        key = {"adaptive": workload_key, "half2": workload_key + "_half2", "fp64": "fp64"}[args.profile_path]
        print(json.dumps({"profile_run": key, "steps": n,
                          "fp64_tile_updates": i.fp64_tile_updates, "low_tile_updates": i.low_full_tile_updates}))
        return

    copy_gbs = copy_bandwidth()
    log(f"[bench] device copy: {copy_gbs:.0f} GB/s (read + write)")
    log(f"[bench] adaptive (NO_SKIP, dilate {args.dilate}): {cfg.nx}x{cfg.ny}, windows of {K} steps x {R} "
        f"ending at step {horizon}")
    amp = timed_run(amps, amp_cfg, horizon, K, R, W, kernel_pass=True)
    value = rate(cells, K, amp)
    log(f"[bench] adaptive: {value:.4g} cell-updates/s, hi_frac {amp['info'].hi_frac_last:.4f}")
    f64 = timed_run(amps, f64_cfg, horizon, K, R, W, kernel_pass=True)
    f64_value = rate(cells, K, f64)
    log(f"[bench] fp64: {f64_value:.4g} cell-updates/s")
    err = field_errors(amp["fields"], f64["fields"])

    out = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": K,
        "warmup": W, "ms_per_step": float(np.median(amp["walls"])) * 1e3 / K, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32 (fp16 storage)", "data": "synthetic",
        "config": {"workload": f"Kobayashi snowflake {cfg.nx}x{cfg.ny}, adaptive FP64/fp16 tiles "
                               f"{cfg.tile}x{cfg.tile} with every low tile fully updated every step (NO_SKIP), "
                               f"gradient monitor {cfg.threshold} dilated by {args.dilate} tiles, 5-point, "
                               f"zero flux, noise 0.01; {R} timed windows of {K} steps ending at step {horizon}",
                   "grid": [cfg.nx, cfg.ny], "tile": cfg.tile, "threshold": cfg.threshold,
                   "dilate": args.dilate, "flags": "NO_SKIP", "start_step": amp["start_step"],
                   "end_step": amp["end_step"]},
    }
    ai = amp["info"]
    out["error_vs_fp64_at_end_step"] = err
    out["hi_frac_last"] = ai.hi_frac_last
    out["speedup_amp_over_fp64"] = value / f64_value
    out["fp64"] = {"value": f64_value, "ms_per_step": float(np.median(f64["walls"])) * 1e3 / K,
                   "window_values": [cells * K / w for w in f64["walls"]]}
    out["runs"] = {"adaptive_window_values": [cells * K / w for w in amp["walls"]],
                   "adaptive_window_event_ms": amp["event_ms"],
                   "note": "value = median of the windows; wall clock around amps_step + amps_sync"}

    # roofline: the two sweeps of the adaptive step, timed with HIP events
    # bound to each dispatch over the same windows; tiles updated from the
    # device's counters.  The flat sweep finishes the low tiles whose
    # neighbourhood decodes to exact zeros at phi 0 or 1 and lists the rest;
    # the mixed sweep updates the FP64 tiles and the listed low tiles.
    # Algorithmic bytes: 32 B per FP64 cell, 8 B per fp16 cell.
    nl = max(amp["k_launches"], 1)
    mix_s, flat_s = amp["k_fp64_ms"] / nl / 1e3, amp["k_low_ms"] / nl / 1e3
    fp64_tiles, low_tiles = amp["k_fp64_tiles"] / nl, amp["k_low_tiles"] / nl
    listed = amp["k_listed_tiles"] / nl
    sweeps = {
        "sf_flat_kernel<32,false>": (flat_s, (low_tiles - listed) * tile_cells * B_FP16, low_tiles - listed),
        "sf_mixed_kernel<32,false,true,false>": (
            mix_s, (fp64_tiles * B_FP64 + listed * B_FP16) * tile_cells, fp64_tiles + listed),
    }
    dom = max(sweeps, key=lambda k: sweeps[k][0])
    t_dom, b_dom, n_dom = sweeps[dom]
    achieved = b_dom / t_dom / 1e9
    tr = traffic_from_profiles("sf_flat_kernel" if "flat" in dom else "sf_mixed_kernel_nsk", workload_key,
                               amp["start_step"], amp["end_step"])
    out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS,
                       "traffic": tr["bytes_per_launch"] if tr else None,
                       "kernel": dom, "avg_launch_us": t_dom * 1e6, "alg_bytes_per_launch": b_dom,
                       "tiles_per_launch": n_dom, "device_copy_gbs": copy_gbs,
                       "frac_of_device_copy": achieved / copy_gbs,
                       "traffic_source": tr["source"] if tr else None}
    out["adaptive_step_kernels_us"] = {k: v[0] * 1e6 for k, v in sweeps.items()}
    out["adaptive_step_kernels_us"]["tiles_per_step"] = {k: v[2] for k, v in sweeps.items()}
    out["adaptive_step_kernels_us"]["fp64_tiles_per_step"] = fp64_tiles
    out["adaptive_step_kernels_us"]["low_tiles_per_step"] = low_tiles
    out["adaptive_step_kernels_us"]["listed_low_tiles_per_step"] = listed
    kf_s = f64["k_fp64_ms"] / max(f64["k_launches"], 1) / 1e3
    fa = B_FP64 * cells / kf_s / 1e9
    out["fp64"]["roofline"] = {"kernel": "sf_f64_tile_kernel<32,false,512>", "avg_launch_us": kf_s * 1e6,
                               "achieved": fa, "peak": HBM_PEAK_GBS, "frac": fa / HBM_PEAK_GBS,
                               "frac_of_device_copy": fa / copy_gbs}

    # frontier: FP64 share -> error at the end step -> speed, by dilation
    if args.frontier:
        rows = []
        for d in [int(x) for x in args.frontier.split(",")]:
            c = amp_cfg.copy()
            c.dilate = d
            r = amp if d == args.dilate else timed_run(amps, c, horizon, K, R, W)
            e = field_errors(r["fields"], f64["fields"])
            v = rate(cells, K, r)
            rows.append({"dilate": d, "hi_frac_last": r["info"].hi_frac_last, "value": v,
                         "speedup_amp_over_fp64": v / f64_value, "phi_l2_rms": e["phi"]["l2_rms"],
                         "T_l2_rms": e["T"]["l2_rms"], "phi_linf": e["phi"]["linf"]})
            log(f"[bench] frontier d={d}: {v / f64_value:.2f}x, hi {rows[-1]['hi_frac_last']:.4f}, "
                f"phi rms {rows[-1]['phi_l2_rms']:.3g}")
        out["dilation_frontier"] = rows
    # the engine's skipping variant (bulk hold, T-flat hold, noise skip)
    c = amp_cfg.copy()
    c.flags = 0
    r = timed_run(amps, c, horizon, K, R, W, kernel_pass=True)
    v = rate(cells, K, r)
    e = field_errors(r["fields"], f64["fields"])
    out["skipping_variant"] = {
        "value": v, "speedup_amp_over_fp64": v / f64_value, "dilate": args.dilate,
        "phi_l2_rms": e["phi"]["l2_rms"], "T_l2_rms": e["T"]["l2_rms"],
        "tiles_fully_updated_per_step": (r["k_fp64_tiles"] + r["k_low_tiles"]) / max(r["k_launches"], 1),
        "note": "not the headline: low tiles in the bulk phase advance T only (phi within 1e-8 of 0 or 1), "
                "T-flat tiles are held, the low-tile noise term is skipped where |phi(1-phi)| <= 1e-6; "
                "the FP64 path has no such skipping"}
    # the packed-half2 form of the low tiles (AMPS_FLAG_HALF2), at every
    # dilation of the frontier
    h2_rows = []
    for d in sorted({args.dilate} | ({int(x) for x in args.frontier.split(",")} if args.frontier else set())):
        c = amp_cfg.copy()
        c.flags, c.dilate = amps.FLAG_NO_SKIP | amps.FLAG_HALF2, d
        r = timed_run(amps, c, horizon, K, R, W, kernel_pass=d == args.dilate)
        v = rate(cells, K, r)
        e = field_errors(r["fields"], f64["fields"])
        row = {"dilate": d, "value": v, "speedup_amp_over_fp64": v / f64_value,
               "speedup_over_fp32_low_tiles": v / value if d == args.dilate else None,
               "phi_l2_rms": e["phi"]["l2_rms"], "T_l2_rms": e["T"]["l2_rms"], "phi_linf": e["phi"]["linf"]}
        if d == args.dilate:
            nl = max(r["k_launches"], 1)
            row["sf_flat_kernel_us"] = r["k_low_ms"] / nl * 1e3
            row["sf_mixed_kernel_h2_us"] = r["k_fp64_ms"] / nl * 1e3
        h2_rows.append(row)
        log(f"[bench] half2 d={d}: {v / f64_value:.2f}x, phi rms {row['phi_l2_rms']:.3g}")
    out["half2_variant"] = {
        "flags": "NO_SKIP|HALF2", "rows": h2_rows,
        "note": "low tiles with the six-fold anisotropy and the driving force's atan in packed half2 "
                "(v_pk_fma_f16), gradients, Laplacians, flux sums and the update in FP32; not the headline, "
                "which keeps the FP32 arithmetic"}
    if not args.no_others:
        out["other_workloads"] = other_workloads(amps, args)
    if not args.no_cpu:
        log("[bench] cpu baseline")
        out["cpu_baseline"] = cpu_baseline(amps, f64_cfg, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()

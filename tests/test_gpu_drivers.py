"""The command-line drivers (w3914777_amd/bin/amps_run) reproduce the
reference applications' surface: arguments, time_tmp.csv and the result CSV
in writetocsv's format, with the same field values as a C-ABI run of the
same parameter files."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "w3914777_amd", "bin", "amps_run")
PARAS = os.path.join(ROOT, "w3914777_amd", "paras")
STEPS = 30


def run_driver(cwd, *args):
    print("[driver]", " ".join(args), flush=True)
    r = subprocess.run([DRIVER, *args], cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r


def read_csv(path):
    with open(path) as f:
        header = f.readline().rstrip("\n").split(",")
        rows = [line.rstrip("\n").split(",") for line in f]
    assert header[0] == "" and [int(h) for h in header[1:]] == list(range(len(header) - 1))
    assert [int(r[0]) for r in rows] == list(range(len(rows)))
    return np.array([[float(v) for v in r[1:]] for r in rows])


def api_field(amps, model, files, mode, field, types=None):
    cfg = amps.Config.default(model)
    for f in files:
        cfg, _ = amps.Config.from_paras(model, f, base=cfg)
    cfg.mode = mode
    with amps.Context(cfg) as ctx:
        ctx.init()
        if types is not None:
            ctx.set_tile_types(types)
        ctx.step(STEPS)
        ctx.sync()
        return ctx.get_field(field), cfg


def same_6_digits(csv_vals, ref):
    # the CSV holds 6 significant digits (default ostream format)
    return np.all(np.abs(csv_vals - ref) <= 5e-6 * np.maximum(np.abs(ref), 1e-300) + 1e-300)


def test_snowflake_purehigh_and_end2end(amps, tmp_path):
    files = [os.path.join(PARAS, "snowflake", "total.h"), os.path.join(PARAS, "snowflake", "para2.h")]
    popt = sum((["--paras", f] for f in files), [])
    run_driver(tmp_path, "snowflake", "purehigh", "4", "phi.csv", *popt, "--steps", str(STEPS))
    t = float(open(tmp_path / "time_tmp.csv").read())
    assert t > 0
    ref, cfg = api_field(amps, amps.SNOWFLAKE, files, amps.PURE_FP64, 0)
    got = read_csv(tmp_path / "phi.csv")
    assert got.shape == (cfg.ny, cfg.nx)
    assert same_6_digits(got, ref)
    # run_idx != "4": timing only, no field
    run_driver(tmp_path, "snowflake", "end2end_mix", "1", "mix.csv", *popt, "--steps", str(STEPS))
    assert not (tmp_path / "mix.csv").exists()
    run_driver(tmp_path, "snowflake", "end2end_mix", "4", "mix.csv", *popt, "--steps", str(STEPS))
    ref, _ = api_field(amps, amps.SNOWFLAKE, files, amps.ADAPTIVE, 0)
    assert same_6_digits(read_csv(tmp_path / "mix.csv"), ref)


def test_graingrowth_purelow_writes_eta2(amps, tmp_path):
    files = [os.path.join(PARAS, "graingrowth", "total.h"), os.path.join(PARAS, "graingrowth", "para1.h")]
    popt = sum((["--paras", f] for f in files), [])
    r = run_driver(tmp_path, "graingrowth", "purelow", *popt, "--steps", str(STEPS))
    assert r.stdout.strip() == "half2result.csv"
    ref, _ = api_field(amps, amps.GRAINGROWTH, files, amps.PURE_HALF2, 1)
    assert same_6_digits(read_csv(tmp_path / "half2result.csv"), ref)


def test_snowflake_puremix_gram_maps(amps, tmp_path):
    files = [os.path.join(PARAS, "snowflake", "total.h"), os.path.join(PARAS, "snowflake", "para1.h")]
    popt = sum((["--paras", f] for f in files), [])
    for kind in (1, 2):
        run_driver(tmp_path, "snowflake", "puremix", *popt, "--steps", str(STEPS), "--gram", str(kind),
                   "--ratio", "0.6")
        cfg = amps.Config.default(amps.SNOWFLAKE)
        for f in files:
            cfg, _ = amps.Config.from_paras(amps.SNOWFLAKE, f, base=cfg)
        n = (cfg.nx // cfg.tile) * (cfg.ny // cfg.tile)
        i = np.arange(n)
        types = np.where((i % 100 < 60) if kind == 2 else (i < n * 0.6), 1, 2).astype(np.int32)
        types = types.reshape(1, cfg.ny // cfg.tile, cfg.nx // cfg.tile)
        ref, _ = api_field(amps, amps.SNOWFLAKE, files, amps.STATIC_MAP, 0, types)
        assert same_6_digits(read_csv(tmp_path / f"gram{kind}.csv"), ref)


def test_sintering_main_result_tree(amps, tmp_path):
    files = [os.path.join(PARAS, "sintering", "para4.h")]
    popt = sum((["--paras", f] for f in files), [])
    run_driver(tmp_path, "sintering", "main", *popt, "--steps", str(STEPS))         # GET_TIME form
    assert float(open(tmp_path / "time_tmp.csv").read()) > 0
    run_driver(tmp_path, "sintering", "main", "4", "double", *popt, "--steps", str(STEPS))
    cfg = amps.Config.default(amps.SINTERING)
    cfg, _ = amps.Config.from_paras(amps.SINTERING, files[0], base=cfg)
    sub = {amps.PURE_FP64: "pure"}.get(cfg.mode, "amstencil")
    found = [os.path.join(d, f) for d, _, fs in os.walk(tmp_path / "result") for f in fs]
    assert len(found) == 1 and f"/end2end/double/{sub}" in found[0] and "/para4/" in found[0], found
    ref, _ = api_field(amps, amps.SINTERING, files, cfg.mode, 0)
    assert same_6_digits(read_csv(found[0]), ref)

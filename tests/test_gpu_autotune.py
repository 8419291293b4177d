"""amps_autotune and the tiling it controls (tile edge, sweep workgroups per
CU): the tuner returns a valid cached choice, and the FP64 result of a run
does not depend on the tiling."""
import numpy as np
import pytest

from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def sf_cfg(amps, n=256):
    c = amps.Config.default(amps.SNOWFLAKE)
    c.nx = c.ny = n
    c.params[amps.SF["CX"]] = c.params[amps.SF["CY"]] = n // 2
    c.params[amps.SF["R2"]] = 60
    c.params[amps.SF["NOISE"]] = 0.01
    return c


def run(amps, cfg, steps):
    with amps.Context(cfg) as ctx:
        ctx.init()
        ctx.step(steps)
        return [ctx.get_field(f) for f in range(cfg.nfields())]


def test_autotune_returns_cached_valid_tiling(amps):
    for mode in (amps.PURE_FP64, amps.ADAPTIVE):
        c = sf_cfg(amps)
        c.mode = mode
        c.monitor = amps.MON_GRADIENT
        c.threshold = 0.05
        t = amps.autotune(c, 30)
        assert t.tile in (16, 32) and 0 <= t.wg_per_cu <= 4 and t.us_per_step > 0
        t2 = amps.autotune(c, 30)
        assert (t2.tile, t2.wg_per_cu, t2.us_per_step) == (t.tile, t.wg_per_cu, t.us_per_step)
        out = run(amps, amps.tuned(c, t), 20)
        assert all(np.isfinite(f).all() for f in out)


def test_fp64_snowflake_is_independent_of_tiling(amps):
    base = sf_cfg(amps)
    ref = run(amps, base, 40)
    for tile, wg in ((32, 0), (16, 1), (32, 3)):
        c = base.copy()
        c.tile, c.wg_per_cu = tile, wg
        got = run(amps, c, 40)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), (tile, wg)


def test_fp64_graingrowth_is_independent_of_tiling(amps, O):
    n, q = 128, 5
    c = amps.Config.default(amps.GRAINGROWTH)
    c.nx = c.ny = n
    c.q, c.init = q, amps.INIT_CONFIG
    p = O.gg_params_config()
    for i in range(12):
        c.params[i] = p[i]
    c.params[amps.GG["AMP"]] = 0.3
    ref = run(amps, c, 30)
    for tile, wg in ((32, 0), (16, 2), (32, 1)):
        c2 = c.copy()
        c2.tile, c2.wg_per_cu = tile, wg
        got = run(amps, c2, 30)
        for a, b in zip(got, ref):
            assert rel_err(a, b) <= 1e-14, (tile, wg, rel_err(a, b))


def test_tiling_arguments_are_validated(amps):
    c = sf_cfg(amps)
    for wg in (-1, 65):
        c.wg_per_cu = wg
        with pytest.raises(amps.AmpsError):
            amps.Context(c)
    c.wg_per_cu = 0
    c.flags = 1 << 7  # no such flag
    with pytest.raises(amps.AmpsError):
        amps.Context(c)
    c.flags = 0
    c.dilate = -1
    with pytest.raises(amps.AmpsError):
        amps.Context(c)

"""Sweep-kernel timing (amps_set_kernel_timing / amps_get_sweep_time): the
counts and the split by sweep (the adaptive step's low sweep, bulk-hold or,
with NO_SKIP, flat, then the mixed sweep), and that timing (eager launches
with events bound to each dispatch) leaves the results bitwise equal to
graph replay."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def workload(amps, O, mode, n=256):
    c = amps.Config.default(amps.SNOWFLAKE)
    c.nx = c.ny = n
    c.tile, c.mode, c.monitor, c.threshold = 32, mode, amps.MON_GRADIENT, 0.05
    p = O.sf_params_config(n, n)
    c.stencil, c.bc = amps.STENCIL_LAP5, amps.BC_NEUMANN
    for i in range(17):
        c.params[i] = p[i]
    return c


@pytest.mark.parametrize("mode", ["adaptive", "noskip", "fp64"])
def test_sweep_timing_counts_split_and_results(amps, O, mode):
    c = workload(amps, O, amps.PURE_FP64 if mode == "fp64" else amps.ADAPTIVE)
    if mode == "noskip":
        c.flags = amps.FLAG_NO_SKIP
    steps = 40
    with amps.Context(c) as ctx:   # reference: graph replay, no timing
        ctx.init()
        ctx.step(10 + steps)
        ctx.sync()
        ref = [ctx.get_field(0), ctx.get_field(1)]
    with amps.Context(c) as ctx:
        ctx.init()
        ctx.step(10)
        ctx.sync()
        ctx.set_kernel_timing(True)
        ctx.step(steps)
        ctx.sync()
        n_all, ms_all = ctx.kernel_time()
        n_hi, ms_hi = ctx.sweep_time(amps.SWEEP_FP64)
        n_lo, ms_lo = ctx.sweep_time(amps.SWEEP_LOW)
        info = ctx.info()
        got = [ctx.get_field(0), ctx.get_field(1)]
        ctx.set_kernel_timing(False)
        assert ctx.kernel_time() == (0, 0.0)   # counters restart
    assert n_all == n_hi == n_lo == steps
    assert ms_all > 0 and ms_hi > 0
    assert abs(ms_hi + ms_lo - ms_all) <= 1e-6 * ms_all
    if mode == "fp64":
        assert ms_lo == 0.0         # one sweep kernel per step
    else:
        assert ms_lo > 0            # the bulk-hold sweep / the flat sweep
    if mode == "noskip":
        # the flat sweep lists the low tiles near the crystal for the mixed sweep
        assert 0 < info.low_listed_tile_updates < info.low_full_tile_updates
    else:
        assert info.low_listed_tile_updates == 0
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)

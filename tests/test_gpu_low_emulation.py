"""The engine's low-precision snowflake tiles against the CPU emulation of
their arithmetic (oracle/low_emulation.cpp), step by step.

From the engine's own raw state (fp16 words, tile offsets and scales, tile
types; FP64 values on FP64 tiles) the emulator advances every low tile by
one step; the engine's next raw state must decode to within one fp16 ulp
of the word (times the tile's scale) of the emulated one on every low cell,
and the
FP64 tiles must match the FP64 oracle step of the decoded state to 1e-12.
Steps are chosen between monitor passes, so the tile types do not change.
The reference branch this stands in for is snowflake/function.h:417-537
(the type-1 half2 update), whose own arithmetic cannot be run here."""
import numpy as np
import pytest

from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

CASES = [
    # mode, n, tile, stencil, bc, dilate, monitor threshold
    ("adaptive", 256, 32, 1, 1, 1, 0.05),
    ("adaptive", 256, 16, 0, 1, 0, 0.05),
    ("low", 128, 16, 0, 0, 0, 0.0),
    ("low", 128, 32, 1, 0, 0, 0.0),
]


def raw_state(ctx):
    w0, o0, s0 = ctx.get_raw_field(0)
    w1, o1, s1 = ctx.get_raw_field(1)
    return np.stack([w0, w1]), np.stack([o0, o1]), np.stack([s0, s1])


@pytest.mark.parametrize("mode,n,tile,stencil,bc,dil,thr", CASES)
def test_low_tiles_match_cpu_emulation(amps, O, mode, n, tile, stencil, bc, dil, thr):
    c = amps.Config.default(amps.SNOWFLAKE)
    c.nx = c.ny = n
    c.tile, c.stencil, c.bc = tile, stencil, bc
    p = O.sf_params_config(n, n)
    for i in range(17):
        c.params[i] = p[i]
    c.params[O.SF["R2"]] = p[O.SF["R2"]] = 60.0
    c.seed = 17
    c.flags, c.dilate = amps.FLAG_NO_SKIP, dil
    if mode == "adaptive":
        c.mode, c.monitor, c.threshold = amps.ADAPTIVE, amps.MON_GRADIENT, thr
    else:
        c.mode = amps.PURE_HALF2
    step0 = 255
    with amps.Context(c) as ctx:
        ctx.init()
        ctx.step(step0)
        ctx.sync()
        n_low = 0
        for k in range(3):
            types = ctx.get_tile_types()
            phi, T = ctx.get_field(0), ctx.get_field(1)
            w, off, scl = raw_state(ctx)
            ctx.step(1)
            phi1, T1 = ctx.get_field(0), ctx.get_field(1)
            w1, off1, scl1 = raw_state(ctx)
            assert np.array_equal(ctx.get_tile_types(), types)  # no monitor pass in between
            ew, eo, es = O.sf_low_step(types, phi, T, w, off, scl, p, tile, stencil, bc, seed=c.seed, step=step0 + k)
            low = np.repeat(np.repeat(types == 1, tile, 0), tile, 1)
            n_low += int(low.sum())
            for f, (got_f, (gw, go, gs), (xw, xo, xs)) in enumerate(
                    zip((phi1, T1), zip(w1, off1, scl1), zip(ew, eo, es))):
                g = O.decode_tiles(gw, go, gs, tile)
                e = O.decode_tiles(xw, xo, xs, tile)
                assert np.array_equal(g[low], got_f[low])  # get_field decodes the same words
                # one fp16 ulp of either word, plus the offsets' difference (the
                # re-centring mid-range in FP32)
                ulp = np.maximum(O.f16_ulp_value(gw, gs, tile), O.f16_ulp_value(xw, xs, tile))
                doff = np.repeat(np.repeat(np.abs(go - xo), tile, 0), tile, 1)
                d = np.abs(g - e)
                tol = 1.001 * ulp + doff + 1e-300
                assert np.all(d[low] <= tol[low]), (k, f, float(np.max((d / tol)[low])))
                # most cells agree to far better than an ulp
                assert np.mean(d[low] <= 0.01 * ulp[low] + doff[low] + 1e-300) > 0.9, (k, f)
            # FP64 tiles: the FP64 oracle step of the decoded state
            po, to = O.sf_step(phi, T, p, stencil=stencil, bc=bc, seed=c.seed, step=step0 + k)
            hi = ~low
            if hi.any():
                assert rel_err(phi1[hi], po[hi]) <= 1e-12
                assert rel_err(T1[hi], to[hi]) <= 1e-12
        assert n_low > 0

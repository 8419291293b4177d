"""Monitor values and tile maps of the grain-growth and sintering models,
bit-exact against the oracle's restatement of the reference monitors, on
the decoded state the engine holds (low tiles read as off + h * scl):

- grain growth, one gradient monitor per field (16^2 tiles: the one-wave
  kernel, including its exact-zero shortcut for fields held at exactly 0 on
  a tile and its edge neighbours; 32^2 tiles: the workgroup kernel);
- grain growth, Monitor1 / Monitor2 on eta2 with its map broadcast to every
  field (graingrowth/tools.h:202-255, graingrowth/end2end_mix.cu:83);
- sintering, Monitor1 / Monitor2 on c (sintering/tools.h:3-56)."""
import numpy as np
import pytest

from tests.helpers import smooth_fields

pytestmark = pytest.mark.gpu


def gg_fields(n, q, seed):
    """Two grains meeting along smooth interfaces (fields 0 and 1), the rest
    absent (1e-12 noise, below the engine's hold bound, so in the default
    step they are flushed to exact zeros and their zero masks are set)."""
    rng = np.random.default_rng(seed)
    a, _ = smooth_fields(n, n, seed=seed, nbump=3)
    f = [a, 1.0 - a]
    for _ in range(q - 2):
        f.append(1e-12 * rng.random((n, n)))
    return f


@pytest.mark.parametrize("tile", [16, 32])
def test_graingrowth_per_field_gradient_monitor_bit_exact(amps, O, tile):
    n, q = 128, 5
    c = amps.Config.default(amps.GRAINGROWTH)
    c.nx = c.ny = n
    c.q, c.tile = q, tile
    c.mode, c.monitor, c.monitor_field, c.threshold = amps.ADAPTIVE, amps.MON_GRADIENT, -1, 0.02
    p = np.array(c.params[:], np.float64)
    thr = c.threshold ** 2
    with amps.Context(c) as ctx:
        ctx.init()
        for f, v in enumerate(gg_fields(n, q, 3)):
            ctx.set_field(f, v)
        ctx.step(25)  # monitor passes at steps 0, 10, 20; absent grains flushed in between
        ctx.sync()
        types0 = ctx.get_tile_types()
        dec = [ctx.get_field(f) for f in range(q)]
        if tile == 16:  # the one-wave sweep flushed the absent grains: exactly zero
            assert not np.any(dec[q - 1])
        ctx.monitor_now()
        md = ctx.get_monitor_values()
        types = ctx.get_tile_types()
    for f in range(q):
        ref = O.monitor_gradient(dec[f], tile, p[O.GG["DX"]], p[O.GG["DY"]], c.bc)
        assert np.array_equal(md[f], ref), f
        _, exp = O.get_type(ref, thr, types0[f])
        assert np.array_equal(types[f], exp), f
    assert 0 < (types == 1).sum() < types.size


@pytest.mark.parametrize("kind", [1, 2])
def test_graingrowth_sample_monitor_on_eta2_bit_exact(amps, O, kind):
    n, tile = 128, 16
    c = amps.Config.default(amps.GRAINGROWTH)
    c.nx = c.ny = n
    c.tile = tile
    c.params[O.GG["R"]] = 30.0
    c.params[O.GG["CX"]] = c.params[O.GG["CY"]] = 64.0
    c.mode, c.monitor, c.monitor_field, c.threshold = amps.ADAPTIVE, kind, 1, 0.3
    p = np.array(c.params[:], np.float64)
    init = O.gg_init(n, n, 2, p)
    with amps.Context(c) as ctx:
        ctx.init()
        f = gg_fields(n, 2, 7)
        for k, v in enumerate(f):
            ctx.set_field(k, v)
        dec1 = ctx.get_field(1)
        ctx.monitor_now()
        md = ctx.get_monitor_values()
        types = ctx.get_tile_types()
    ref = O.monitor_spatial(dec1, tile) if kind == 1 else O.monitor_temporal(np.where(init[1] == 0, 1.0, 0.0), dec1, tile)
    assert np.array_equal(md[0], ref)  # one evaluation, of eta2
    _, exp = O.get_type(ref, c.threshold, np.full(ref.shape, 2, np.int32))
    for g in range(2):  # every field takes eta2's map
        assert np.array_equal(types[g], exp)
    assert 0 < (types[0] == 1).sum() < types[0].size


@pytest.mark.parametrize("kind", [1, 2])
def test_sintering_sample_monitor_on_c_bit_exact(amps, O, kind):
    n, tile = 128, 16
    c = amps.Config.default(amps.SINTERING)
    c.nx = c.ny = n
    c.tile = tile
    p = O.sn_params_ref(1, n, n)
    p[O.SN["R1"]] = 20; p[O.SN["R2"]] = 10; p[O.SN["X1"]] = p[O.SN["X2"]] = 72
    p[O.SN["Y1"]] = 50; p[O.SN["Y2"]] = 80
    for i in range(20):
        c.params[i] = p[i]
    c.mode, c.monitor, c.monitor_field, c.threshold = amps.ADAPTIVE, kind, 0, 0.2 if kind == 1 else 1e-6
    with amps.Context(c) as ctx:
        ctx.init()
        ctx.step(40)
        ctx.sync()
        cc = ctx.get_field(0)
        ctx.monitor_now()
        md = ctx.get_monitor_values()
        types = ctx.get_tile_types()
        cc2 = ctx.get_field(0)
    assert np.array_equal(cc, cc2)
    if kind == 1:
        ref = O.monitor_spatial(cc, tile)
        assert np.array_equal(md, ref)
    else:
        # Monitor2's history holds the samples of the last pass, the one
        # before step 30, i.e. of c after 30 steps
        with amps.Context(c) as ctx:
            ctx.init()
            ctx.step(30)
            ctx.sync()
            c30 = ctx.get_field(0)
        ref = O.monitor_temporal(c30, cc, tile)
        assert np.array_equal(md, ref)
    assert 0 < (types == 1).sum() < types.size

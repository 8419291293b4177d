"""GPU tests of the adaptive machinery: monitor masks (bit-exact against the
oracle's restatement of the reference monitor), tile conversion (bit-exact
against an IEEE round-to-nearest-even fp16 encoder), precision-mode
equivalences (bitwise) and error bounds of the low-precision paths against
the FP64 path."""
import math

import numpy as np
import pytest

from tests.helpers import l2_rms, rel_err, smooth_fields

pytestmark = pytest.mark.gpu


def sf_config(amps, O, n, tile=16, mode=2, monitor=2, thr=0.99, every=10, workload=False):
    c = amps.Config.default(amps.SNOWFLAKE)
    c.nx = c.ny = n
    c.tile, c.mode, c.monitor, c.threshold, c.monitor_every = tile, mode, monitor, thr, every
    p = O.sf_params_config(n, n) if workload else O.sf_params_ref(2, n, n)
    if workload:
        c.stencil, c.bc = amps.STENCIL_LAP5, amps.BC_NEUMANN
    for i in range(17):
        c.params[i] = p[i]
    return c, p


# ----------------------------------------------------------------- monitor
@pytest.mark.parametrize("kind,tile,thr", [(1, 16, 0.3), (2, 16, 0.2), (3, 16, 0.05), (1, 32, 0.5),
                                           (2, 32, 0.1), (3, 32, 0.3)])
def test_monitor_masks_bit_exact(amps, O, kind, tile, thr):
    n = 256
    c, p = sf_config(amps, O, n, tile=tile, monitor=kind, thr=thr)
    phi_init, _ = O.sf_init(n, n, p)
    phi, T = smooth_fields(n, n, seed=kind + tile)
    thr_eff = thr * thr if kind == 3 else thr
    with amps.Context(c) as ctx:
        ctx.init()
        ctx.set_field(0, phi)
        ctx.set_field(1, T)
        ctx.monitor_now()
        md = ctx.get_monitor_values()
        if kind == 1:
            ref = O.monitor_spatial(phi, tile)
        elif kind == 2:
            ref = O.monitor_temporal(np.where(phi_init == 0, 1.0, 0.0), phi, tile)
        else:
            ref = O.monitor_gradient(phi, tile, p[O.SF["DX"]], p[O.SF["DY"]], 0)
        assert np.array_equal(md, ref)
        old, types = O.get_type(ref, thr_eff, np.full(ref.shape, 2, np.int32))
        got = ctx.get_tile_types()
        assert np.array_equal(got, types)
        if kind != 2:
            # (Monitor2's history starts as the complement of the initial field,
            # snowflake/end2end_mix.cu:44-46, so its first pass is mostly FP64)
            assert 0 < (types == 1).sum() < types.size   # the case mixes precisions
        # second pass on a changed field, now partly stored in fp16
        phi2 = np.roll(phi, 9, axis=1)
        ctx.set_field(0, phi2)
        dec = ctx.get_field(0)
        ctx.monitor_now()
        md2 = ctx.get_monitor_values()
        if kind == 1:
            ref2 = O.monitor_spatial(dec, tile)
        elif kind == 2:
            ref2 = O.monitor_temporal(phi, dec, tile)  # history = samples of the previous pass
        else:
            ref2 = O.monitor_gradient(dec, tile, p[O.SF["DX"]], p[O.SF["DY"]], 0)
        assert np.array_equal(md2, ref2)
        _, types2 = O.get_type(ref2, thr_eff, types)
        assert np.array_equal(ctx.get_tile_types(), types2)
        assert 0 < (types2 == 1).sum() < types2.size


# ------------------------------------------------------------- conversion
def encode_decode_cpu(O, v):
    """Per-tile codec: off = mid-range, scale 2^(e-14) with half-range < 2^e,
    h = fp16_rne((v - off) * 2^(14-e)), value = off + h * 2^(e-14)."""
    mn, mx = v.min(), v.max()
    off, hw = 0.5 * (mn + mx), 0.5 * (mx - mn)
    e = math.frexp(hw)[1] if hw > 0 else 0
    e = max(e, -100)
    scaled = (v - off) * 2.0 ** (14 - e)
    h = np.array([O.f16_bits_to_f64(O.f64_to_f16_bits(s)) for s in scaled.ravel()]).reshape(v.shape)
    return off + h * 2.0 ** (e - 14)


def test_tile_conversion_is_exact_rne(amps, O):
    n, t = 64, 16
    c, p = sf_config(amps, O, n, tile=t, mode=3)
    rng = np.random.default_rng(0)
    phi = rng.uniform(0.2, 0.9, (n, n))
    # tile (0,0): exact ties between fp16 neighbours (off = 0, scale 2^-15)
    k = np.arange(t * t).reshape(t, t) % 1000
    tie = (2 * k + 2049) / 2.0 * 2.0 ** -15 * np.where(k % 2 == 0, 1, -1)
    tie[0, 0], tie[0, 1] = -0.375, 0.375
    phi[:t, :t] = tie
    T = rng.uniform(-1e-6, 1e-6, (n, n))
    with amps.Context(c) as ctx:
        ctx.init()
        ctx.set_field(0, phi)
        ctx.set_field(1, T)
        ctx.set_tile_types(np.ones((n // t, n // t), np.int32))
        gp, gt = ctx.get_field(0), ctx.get_field(1)
        for ty in range(n // t):
            for tx in range(n // t):
                s = np.s_[ty * t:(ty + 1) * t, tx * t:(tx + 1) * t]
                assert np.array_equal(gp[s], encode_decode_cpu(O, phi[s]))
                assert np.array_equal(gt[s], encode_decode_cpu(O, T[s]))
        # back to FP64: the decoded values are kept exactly
        ctx.set_tile_types(np.full((n // t, n // t), 2, np.int32))
        assert np.array_equal(ctx.get_field(0), gp) and np.array_equal(ctx.get_field(1), gt)


# ---------------------------------------------------- mode equivalences
def run_mode(amps, c, steps, types=None):
    with amps.Context(c) as ctx:
        ctx.init()
        if types is not None:
            ctx.set_tile_types(types)
        ctx.step(steps)
        ctx.sync()
        return [ctx.get_field(f) for f in range(c.nfields())]


@pytest.mark.parametrize("tile", [16, 32])
def test_all_high_map_equals_fp64_bitwise(amps, O, tile):
    n = 128
    c64, _ = sf_config(amps, O, n, tile=tile, mode=0, workload=True)
    cst, _ = sf_config(amps, O, n, tile=tile, mode=3, workload=True)
    a = run_mode(amps, c64, 60)
    b = run_mode(amps, cst, 60)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_all_low_map_equals_pure_half2_bitwise(amps, O):
    n, t = 128, 32
    ch, _ = sf_config(amps, O, n, tile=t, mode=1, workload=True)
    cst, _ = sf_config(amps, O, n, tile=t, mode=3, workload=True)
    a = run_mode(amps, ch, 60)
    b = run_mode(amps, cst, 60, types=np.ones((n // t, n // t), np.int32))
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("tile", [16, 32])
@pytest.mark.parametrize("amp", [0.0, 2.0 ** -130, 2.0 ** -110])
def test_low_path_subnormal_gradients_stay_finite(amps, O, tile, amp):
    """Low tiles whose phi deviations are zero or subnormal in single
    precision (phi = amp * noise beside a half at 0.5, so the tiles next to
    the step are not in the bulk-hold state): the anisotropy's direction
    must stay finite for gradients far below 1/FLT_MAX, and the step must
    match the FP64 step."""
    n = 64
    rng = np.random.default_rng(5)
    phi = np.where(np.arange(n)[None, :] < n // 2, amp * rng.random((n, n)), 0.5)
    T = np.full((n, n), 0.25)
    out = {}
    for mode in (0, 1):
        c, _ = sf_config(amps, O, n, tile=tile, mode=mode, workload=True)
        c.params[amps.SF["NOISE"]] = 0.0
        with amps.Context(c) as ctx:
            ctx.init()
            ctx.set_field(0, phi)
            ctx.set_field(1, T)
            ctx.step(3)
            ctx.sync()
            out[mode] = [ctx.get_field(0), ctx.get_field(1)]
    for g, r in zip(out[1], out[0]):
        assert np.all(np.isfinite(g))
        assert np.max(np.abs(g - r)) < 2e-3


def test_mixed_static_map_error_between_extremes(amps, O):
    n, t = 128, 16
    c64, _ = sf_config(amps, O, n, tile=t, mode=0, workload=True)
    ref = run_mode(amps, c64, 300)
    ch, _ = sf_config(amps, O, n, tile=t, mode=1, workload=True)
    low = run_mode(amps, ch, 300)
    cst, _ = sf_config(amps, O, n, tile=t, mode=3, workload=True)
    types = np.ones((n // t, n // t), np.int32)
    types[2:6, 2:6] = 2   # the nucleus region in FP64
    mix = run_mode(amps, cst, 300, types=types)
    e_low = l2_rms(low[0], ref[0])
    e_mix = l2_rms(mix[0], ref[0])
    assert e_mix <= e_low + 1e-12
    assert e_low < 1e-2


# ------------------------------------------------------------ error bounds
def test_adaptive_snowflake_error_vs_fp64(amps, O):
    """Workload physics (5-point, zero flux, noise) at 256^2, tile 32,
    gradient monitor 0.05: final-field L2 (RMS) error <= 1e-3 vs FP64."""
    n = 256
    c64, _ = sf_config(amps, O, n, tile=32, mode=0, workload=True)
    ref = run_mode(amps, c64, 1500)
    ca, _ = sf_config(amps, O, n, tile=32, mode=2, monitor=3, thr=0.05, workload=True)
    with amps.Context(ca) as ctx:
        ctx.init()
        ctx.step(1500)
        ctx.sync()
        got = [ctx.get_field(0), ctx.get_field(1)]
        info = ctx.info()
    assert info.monitor_passes == 150
    assert 0.0 < info.hi_frac_mean < 0.5
    for g, r in zip(got, ref):
        assert l2_rms(g, r) <= 1e-3
        assert np.max(np.abs(g - r)) < 0.05


@pytest.mark.parametrize("tile", [16, 32])
def test_adaptive_graingrowth_per_field_types(amps, O, tile):
    n, q = 128, 4
    base = amps.Config.default(amps.GRAINGROWTH)
    base.nx = base.ny = n; base.q = q; base.init = amps.INIT_CONFIG; base.tile = tile
    p = O.gg_params_config()
    for i in range(12):
        base.params[i] = p[i]
    ref = run_mode(amps, base, 400)
    ca = base.copy()
    ca.mode, ca.monitor, ca.monitor_field, ca.threshold = amps.ADAPTIVE, amps.MON_GRADIENT, -1, 0.02
    with amps.Context(ca) as ctx:
        ctx.init()
        ctx.step(400)
        ctx.sync()
        got = [ctx.get_field(f) for f in range(q)]
        types = ctx.get_tile_types()
    assert types.shape == (q, n // tile, n // tile)
    err = max(l2_rms(g, r) for g, r in zip(got, ref))
    assert err < 2e-3, err


def test_adaptive_sintering_keeps_c_fp64(amps, O):
    n = 128
    base = amps.Config.default(amps.SINTERING)
    base.nx = base.ny = n; base.tile = 16
    p = O.sn_params_ref(1, n, n)
    p[O.SN["R1"]] = 20; p[O.SN["R2"]] = 10; p[O.SN["X1"]] = p[O.SN["X2"]] = 72
    p[O.SN["Y1"]] = 50; p[O.SN["Y2"]] = 80
    for i in range(20):
        base.params[i] = p[i]
    ref = run_mode(amps, base, 200)
    ca = base.copy()
    ca.mode, ca.monitor, ca.monitor_field, ca.threshold = amps.ADAPTIVE, amps.MON_GRADIENT, 0, 0.05
    got = run_mode(amps, ca, 200)
    for g, r in zip(got, ref):
        assert l2_rms(g, r) < 1e-3


# ------------------------------------------------- NO_SKIP and dilation
@pytest.mark.parametrize("tile", [16, 32])
def test_noskip_all_high_equals_fp64_bitwise(amps, O, tile):
    """Adaptive with NO_SKIP and a threshold every tile reaches: the FP64
    tiles go through the adaptive FP64 sweep and must equal the plain FP64
    path bit for bit."""
    n = 128
    c64, _ = sf_config(amps, O, n, tile=tile, mode=0, workload=True)
    ca, _ = sf_config(amps, O, n, tile=tile, mode=2, monitor=3, thr=0.0, workload=True)
    ca.flags = amps.FLAG_NO_SKIP
    a = run_mode(amps, c64, 60)
    b = run_mode(amps, ca, 60)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_noskip_updates_every_tile_every_step(amps, O):
    n, t, steps = 256, 32, 200
    for dil in (0, 2):
        c, _ = sf_config(amps, O, n, tile=t, mode=2, monitor=3, thr=0.05, workload=True)
        c.flags, c.dilate = amps.FLAG_NO_SKIP, dil
        with amps.Context(c) as ctx:
            ctx.init()
            ctx.step(steps)
            ctx.sync()
            i = ctx.info()
            types = ctx.get_tile_types()
        assert i.fp64_tile_updates + i.low_full_tile_updates == steps * (n // t) ** 2
        assert 0 < (types == 1).sum() < types.size


def test_dilation_grows_the_fp64_set(amps, O):
    """The FP64 set after a monitor pass with dilation d is the set of tiles
    within Chebyshev distance d (non-periodic) of a tile the reference rule
    (own value or a von-Neumann neighbour's >= threshold) makes FP64."""
    n, t = 256, 16
    phi, T = smooth_fields(n, n, seed=4, nbump=2)
    for d in (1, 2):
        c, p = sf_config(amps, O, n, tile=t, mode=2, monitor=3, thr=0.1, workload=True)
        c.dilate = d
        with amps.Context(c) as ctx:
            ctx.init()
            ctx.set_field(0, phi)
            ctx.set_field(1, T)
            ctx.monitor_now()
            got = ctx.get_tile_types()
        md = O.monitor_gradient(phi, t, p[O.SF["DX"]], p[O.SF["DY"]], c.bc)
        _, base = O.get_type(md, 0.01, np.full(md.shape, 2, np.int32))
        hi = base == 2
        exp = np.zeros_like(hi)
        ty, tx = hi.shape
        for y, x in zip(*np.nonzero(md >= 0.01)):
            exp[max(y - d, 0):y + d + 1, max(x - d, 0):x + d + 1] = True
        assert np.array_equal(got == 2, exp | hi)
        assert (got == 2).sum() > (base == 2).sum()


@pytest.mark.parametrize("flags,dil", [(0, 0), (1, 0), (1, 2)])
def test_adaptive_batch_instances_equal_single_runs(amps, O, flags, dil):
    """Three uneven grids (64, 96, 128; different seeds and delta) batched in
    one adaptive context (work-lists and, without NO_SKIP, hold groups that
    mix instances) give every instance the fields, tile types and monitor
    values of the same configuration run alone, bit for bit."""
    cfgs = []
    for n, seed, delta in ((64, 3, 0.03), (96, 5, 0.05), (128, 9, 0.07)):
        c, _ = sf_config(amps, O, n, tile=16, mode=2, monitor=3, thr=0.05, workload=True)
        c.params[amps.SF["CX"]] = c.params[amps.SF["CY"]] = n // 2
        c.params[amps.SF["DELTA"]] = delta
        c.seed, c.flags, c.dilate = seed, flags, dil
        cfgs.append(c)
    steps = 120
    with amps.Context(cfgs) as ctx:
        ctx.init()
        ctx.step(steps)
        ctx.sync()
        batch = [([ctx.get_field(f, inst=i) for f in range(2)], ctx.get_tile_types(inst=i),
                  ctx.get_monitor_values(inst=i)) for i in range(3)]
    mixed = 0
    for i, c in enumerate(cfgs):
        with amps.Context(c) as ctx:
            ctx.init()
            ctx.step(steps)
            ctx.sync()
            single = ([ctx.get_field(f) for f in range(2)], ctx.get_tile_types(), ctx.get_monitor_values())
        for a, b in zip(batch[i][0], single[0]):
            assert np.array_equal(a, b), i
        assert np.array_equal(batch[i][1], single[1]), i
        assert np.array_equal(batch[i][2], single[2]), i
        mixed += 0 < (single[1] == 1).sum() < single[1].size
    assert mixed > 0  # some instance holds both precisions

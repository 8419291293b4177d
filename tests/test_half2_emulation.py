"""CPU checks of the packed-half2 emulation (AMPS_FLAG_HALF2 low snowflake
tiles; oracle/low_emulation.cpp), before it checks the engine:

- its fp16 rounding of a double equals numpy's (correctly rounded, nearest
  even) on random magnitudes, ties, subnormals and the overflow edge;
- its half2 atan stays within 2e-3 of atan over |x| <= 40 and is odd; its
  six-fold unit vector rotated by 6 theta0 stays within 6e-3 of
  (cos 6(theta - theta0), sin 6(theta - theta0)) over gradients from 1e-30 to
  1e3, and a zero gradient gives (cos, -sin) of 6 theta0 rounded to fp16;
- one emulated half2 step of an all-low state stays within the FP64 oracle
  step plus the FP32 tolerance plus a bound from the two half2 errors above
  (the driving force's atan scaled by alpha/pi |phi(1-phi)|, the anisotropy
  through eps' in P and Q and through eps^2), and differs from the FP32
  emulated step (so the half2 terms are really taken)."""
import numpy as np
import pytest

from tests.helpers import smooth_fields
from tests.test_low_emulation import box_scale, encode_grid

ATAN_BOUND = 2.0e-3
ANISO_BOUND = 6.0e-3


def test_fp16_round_matches_numpy(O):
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.standard_normal(20000) * np.exp(rng.uniform(-25, 11, 20000)),
                         (np.arange(2049) + 0.5) * 2.0 ** -24,           # subnormal ties
                         (2049 + 2 * np.arange(500)) / 2.0 * 2.0 ** -10,  # ties of normal numbers
                         [0.0, 65504.0, 65519.99, 65520.0, 2.0 ** -14, 2.0 ** -25, 3 * 2.0 ** -26]])
    xs = np.concatenate([xs, -xs])
    got = np.array([O.fp16_round(x) for x in xs])
    ref = xs.astype(np.float16).astype(np.float64)
    assert np.array_equal(got, ref)


def test_atan_h2_bound_and_symmetry(O):
    x = np.concatenate([np.linspace(-40.0, 40.0, 400001), [0.0, 1.0, -1.0, 1e-8, -1e-8, 65504.0, 1e6]])
    x = x.astype(np.float32)
    r = O.atan_h2(x)
    err = np.abs(r.astype(np.float64) - np.arctan(x.astype(np.float64)))
    assert err.max() <= ATAN_BOUND, float(err.max())
    assert err.max() > 1e-4  # fp16 arithmetic, not the FP32 atan
    assert np.array_equal(O.atan_h2(-x), -r)
    assert O.atan_h2(np.array([0.0], np.float32))[0] == 0.0


@pytest.mark.parametrize("theta0", [0.0, np.pi / 2, 0.3])
def test_aniso6_h2_bound(O, theta0):
    rng = np.random.default_rng(1)
    th = rng.uniform(-np.pi, np.pi, 200000)
    mag = 10.0 ** rng.uniform(-30, 3, th.size)
    gx, gy = (mag * np.cos(th)).astype(np.float32), (mag * np.sin(th)).astype(np.float32)
    # the engine's exact power-of-two rescale: max(|u|, |v|) in [0.5, 1)
    e = np.frexp(np.maximum(np.abs(gx), np.abs(gy)))[1]
    u, v = np.ldexp(gx, -e).astype(np.float32), np.ldexp(gy, -e).astype(np.float32)
    c0, s0 = np.float32(np.cos(6 * theta0)), np.float32(np.sin(6 * theta0))
    ca, sa = O.aniso6_h2(u, v, c0, s0)
    t = np.arctan2(gy.astype(np.float64), gx.astype(np.float64)) - theta0
    assert np.abs(ca - np.cos(6 * t)).max() <= ANISO_BOUND
    assert np.abs(sa - np.sin(6 * t)).max() <= ANISO_BOUND
    # zero gradient: theta = atan2(0, 0) = 0
    z_ca, z_sa = O.aniso6_h2(np.array([1.0]), np.array([0.0]), c0, s0)
    assert z_ca[0] == O.fp16_round(c0) and z_sa[0] == -O.fp16_round(s0)


def half2_bound(O, phi, p, stencil):
    """Per cell, how far the half2 terms may move one step of phi: dt/tau times
    alpha/pi ATAN_BOUND |phi(1-phi)| (driving force) plus ANISO_BOUND times the
    anisotropic flux's sensitivity, |d/dy| of eps |mepsb| |gx| plus |d/dx| of
    eps |mepsb| |gy| (eps' = mepsb sin, so P = eps eps' gx moves by at most
    eps |mepsb| |gx| ANISO_BOUND) and 2 eps epsb delta |lap phi| ANISO_BOUND
    (eps^2 through eps = epsb (1 + delta cos))."""
    S = O.SF
    dx, dy, dt, tau = p[S["DX"]], p[S["DY"]], p[S["DT"]], p[S["TAU"]]
    epsb, delta, n = p[S["EPSB"]], p[S["DELTA"]], p[S["ANISO"]]
    eps_max = epsb * (1 + delta)
    m_epsd = epsb * n * delta
    gx = (np.roll(phi, -1, 1) - np.roll(phi, 1, 1)) / (2 * dx)
    gy = (np.roll(phi, -1, 0) - np.roll(phi, 1, 0)) / (2 * dy)
    Pm, Qm = eps_max * m_epsd * np.abs(gx), eps_max * m_epsd * np.abs(gy)
    flux = (np.roll(Pm, -1, 0) + np.roll(Pm, 1, 0)) / (2 * dy) + (np.roll(Qm, -1, 1) + np.roll(Qm, 1, 1)) / (2 * dx)
    if stencil == 1:
        lap = sum(np.roll(phi, s, a) for s in (-2, -1, 1, 2) for a in (0, 1)) - 8 * phi
    else:
        lap = sum(np.roll(phi, s, a) for s in (-1, 1) for a in (0, 1)) - 4 * phi
    lap = np.abs(lap) / (dx * dy)
    aop = p[S["ALPHA"]] / p[S["PI"]]
    return dt / tau * (aop * ATAN_BOUND * np.abs(phi * (1 - phi)) + ANISO_BOUND * (flux + 2 * eps_max * epsb * delta * lap))


@pytest.mark.parametrize("stencil,bc", [(1, 1), (0, 1)])
def test_half2_low_step_within_bound_of_fp64_step(O, stencil, bc):
    n, tile = 128, 32
    p = O.sf_params_config(n, n)
    phi, T = smooth_fields(n, n, seed=2, nbump=4)
    wp, op, sp = encode_grid(O, phi, tile)
    wt, ot, st = encode_grid(O, T, tile)
    dphi, dT = O.decode_tiles(wp, op, sp, tile), O.decode_tiles(wt, ot, st, tile)
    types = np.ones((n // tile, n // tile), np.int32)
    W, Of, Sc = np.stack([wp, wt]), np.stack([op, ot]), np.stack([sp, st])
    hw, ho, hs = O.sf_low_step(types, dphi, dT, W, Of, Sc, p, tile, stencil, bc, seed=3, step=5, half2=True)
    fw, fo, fs = O.sf_low_step(types, dphi, dT, W, Of, Sc, p, tile, stencil, bc, seed=3, step=5)
    rp, rt = O.sf_step(dphi, dT, p, stencil=stencil, bc=bc, seed=3, step=5)
    s9 = box_scale(dphi, op, tile) + box_scale(dT, ot, tile)
    bphi = half2_bound(O, dphi, p, stencil)
    kap = p[O.SF["KAPPA"]]
    for f, (ref, b) in enumerate(((rp, bphi), (rt, kap * bphi))):  # T moves by kappa times phi's change
        e = O.decode_tiles(hw[f], ho[f], hs[f], tile)
        ulp = np.maximum(O.f16_ulp_value(hw[f], hs[f], tile), O.f16_ulp_value(fw[f], fs[f], tile))
        tol = 0.5 * ulp + 16 * 2.0 ** -24 * s9 + b
        d = np.abs(e - ref)
        assert np.all(d <= tol), (f, float(np.max(d / tol)))
    # the half2 terms are taken: phi differs from the FP32 emulation
    assert not np.array_equal(hw[0], fw[0])

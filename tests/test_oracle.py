"""Pins of the CPU oracle (oracle/oracle.cpp) by closed forms, invariants,
special cases and brute force.  CPU only.

The reference ships no golden vectors, known-answer tests or fixtures for
this path (its outputs are git-ignored CSVs generated at run time), so each
oracle function is pinned here by what the mathematics fixes, chosen so that
a dropped term, a wrong sign or neighbour index, or a transposed operand fails
one of the checks.
"""
import math

import numpy as np
import pytest

from tests.helpers import rel_err, smooth_fields


# --------------------------------------------------------------- operators
@pytest.mark.parametrize("stencil", [0, 1])
@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("k", [1, 3, 7])
def test_sf_T_fourier_mode_is_eigenvector(O, stencil, axis, k):
    """phi = 0 reduces the snowflake step to T' = T + dt L(T).  A periodic
    Fourier mode is an exact eigenvector: one step scales it by 1 + dt*lam,
    with lam = (2cos w + 2cos 2w - 4)/dxdy for the reference's radius-2 cross
    (snowflake/function.h:27-28) and (2cos w - 2)/dxdy for the 5-point form."""
    n = 64
    p = O.sf_params_ref(1, n, n)
    w = 2 * math.pi * k / n
    idx = np.arange(n)
    mode = np.sin(w * idx + 0.3)
    T = np.tile(mode, (n, 1)) if axis == 1 else np.tile(mode[:, None], (1, n))
    T = np.ascontiguousarray(0.2 * T)
    phi = np.zeros((n, n))
    po, to = O.sf_step(phi, T, p, stencil=stencil)
    dxdy = p[O.SF["DX"]] * p[O.SF["DY"]]
    lam = ((2 * math.cos(w) + 2 * math.cos(2 * w) - 4) if stencil == 0 else (2 * math.cos(w) - 2)) / dxdy
    assert np.all(po == 0.0)
    assert rel_err(to, T * (1 + p[O.SF["DT"]] * lam)) < 1e-12


def test_sf_quadratic_operator_factor(O):
    """On T = x^2 + y^2 (interior cells) the radius-2 cross gives 20/dxdy =
    5 * nabla^2 T; the 5-point form gives 4/dxdy."""
    n = 32
    p = O.sf_params_ref(1, n, n)
    y, x = np.mgrid[0:n, 0:n].astype(float)
    T = np.ascontiguousarray(1e-3 * (x * x + y * y))
    phi = np.zeros((n, n))
    dxdy = p[O.SF["DX"]] * p[O.SF["DY"]]
    for stencil, factor in ((0, 20.0), (1, 4.0)):
        _, to = O.sf_step(phi, T, p, stencil=stencil)
        lap = (to - T)[4:-4, 4:-4] / p[O.SF["DT"]]
        assert np.allclose(lap, 1e-3 * factor / dxdy, rtol=1e-9)


@pytest.mark.parametrize("stencil", [0, 1])
def test_gg_fourier_mode_moore(O, stencil):
    """alpha = beta = 0 and eta_2 = 0 make the grain-growth update linear:
    eta' = eta + dt L kappa lap(eta).  sin(a x + b y) is an eigenvector with
    lam = (2cos a + 2cos b + 4 cos a cos b - 8)/dxdy for the Moore operator
    (graingrowth/function.h:20) and (2cos a + 2cos b - 4)/dxdy for 5-point."""
    n = 48
    p = O.gg_params_ref(1, n, n)
    p[O.GG["ALPHA"]] = 0.0
    p[O.GG["BETA"]] = 0.0
    a, b = 2 * math.pi * 2 / n, 2 * math.pi * 5 / n
    y, x = np.mgrid[0:n, 0:n].astype(float)
    eta = np.zeros((2, n, n))
    eta[0] = 0.3 * np.sin(a * x + b * y)
    out = O.gg_step(eta, p, stencil=stencil)
    dxdy = 0.25
    if stencil == 0:
        lam = (2 * math.cos(a) + 2 * math.cos(b) + 4 * math.cos(a) * math.cos(b) - 8) / dxdy
    else:
        lam = (2 * math.cos(a) + 2 * math.cos(b) - 4) / dxdy
    g = p[O.GG["DT"]] * p[O.GG["L"]] * p[O.GG["KAPPA"]]
    assert rel_err(out[0], eta[0] * (1 + g * lam)) < 1e-12
    assert np.all(out[1] == 0.0)


def test_sn_laplacian_mode(O):
    """With eta = 0 and c = c0 + eps*mode, mu's Laplacian part and c's update
    follow the 5-point eigenvalue: for a small-amplitude mode, c' - c is
    proportional to the mode with factor dt*M(c0)*lam*(f''(c0) - kc/2*lam)
    to first order (checked at 1e-6 relative with eps = 1e-7)."""
    n = 32
    p = O.sn_params_ref(1, n, n)
    w = 2 * math.pi * 3 / n
    x = np.arange(n)
    mode = np.tile(np.cos(w * x), (n, 1))
    c0, eps = 0.4, 1e-7
    c = np.ascontiguousarray(c0 + eps * mode)
    z = np.zeros((n, n))
    co, e1, e2 = O.sn_step(c, z, z, p)
    A, B, kc = p[O.SN["A"]], p[O.SN["B"]], p[O.SN["KC"]]
    lam = (2 * math.cos(w) - 2) / 0.25
    fpp = B * 2 + 2 * A * (1 - 6 * c0 + 6 * c0 * c0)   # d2f/dc2 at eta = 0
    h = c0 ** 3 * (10 - 15 * c0 + 6 * c0 ** 2)
    M = p[O.SN["DVOL"]] * h + p[O.SN["DVAP"]] * (1 - h) + p[O.SN["DSURF"]] * c0 * (1 - c0)
    pred = p[O.SN["DT"]] * M * lam * (fpp - 0.5 * kc * lam) * eps * mode
    assert rel_err(co - c, pred) < 1e-5
    assert np.all(e1 == 0) and np.all(e2 == 0)


# ----------------------------------------------------- special cases / ODEs
def test_sf_fixed_points(O):
    """Uniform phi in {0, 1}, and phi = 1/2 at T = Teq (m = 0), are exact
    fixed points; T is unchanged when phi does not move."""
    n = 16
    p = O.sf_params_ref(2, n, n)
    for ph, t in ((0.0, 0.0), (1.0, 0.3), (0.5, p[O.SF["TEQ"]])):
        phi = np.full((n, n), ph)
        T = np.full((n, n), t)
        po, to = O.sf_step(phi, T, p)
        assert np.all(po == ph) and np.all(to == t)


def test_sf_uniform_relaxation_sign(O):
    """Uniform phi in (0,1): d phi/dt = phi(1-phi)(phi - 1/2 + m)/tau with
    m = alpha/pi atan(gamma (Teq - T)); at T = 0 (undercooled) m > 0, so phi
    grows when phi > 1/2 - m and shrinks below; T follows K*(phi' - phi)."""
    n = 16
    p = O.sf_params_ref(1, n, n)
    m = p[O.SF["ALPHA"]] / p[O.SF["PI"]] * math.atan(p[O.SF["GAMMA"]] * p[O.SF["TEQ"]])
    for ph in (0.5 - m - 0.05, 0.5 - m + 0.05, 0.9):
        phi = np.full((n, n), ph)
        T = np.zeros((n, n))
        po, to = O.sf_step(phi, T, p)
        d = po[0, 0] - ph
        assert np.sign(d) == np.sign(ph - 0.5 + m)
        assert abs(to[0, 0] - p[O.SF["KAPPA"]] * d) < 1e-15


def test_gg_fixed_points(O):
    """(eta_i = 1, others 0) and all-zero are fixed points of the Fan-Chen
    bulk free energy (eta^3 - eta = 0)."""
    n, q = 16, 4
    p = O.gg_params_ref(1, n, n)
    eta = np.zeros((q, n, n))
    eta[2] = 1.0
    out = O.gg_step(eta, p)
    assert np.array_equal(out, eta)


def test_gg_field_permutation(O):
    """Relabelling order parameters relabels the result (q = 3)."""
    n, q = 24, 3
    p = O.gg_params_ref(1, n, n)
    rng = np.random.default_rng(4)
    eta = rng.uniform(-0.2, 1.0, size=(q, n, n))
    out = O.gg_step(eta, p)
    perm = [2, 0, 1]
    out2 = O.gg_step(np.ascontiguousarray(eta[perm]), p)
    assert rel_err(out2, out[perm]) < 1e-14


def test_sn_uniform_and_clamp(O):
    """Uniform fields: lap mu = 0 leaves c unchanged; (c, eta1, eta2) =
    (1, 1, 0) is a fixed point; every update stays in [0, 1] (clamps,
    sintering/function.h:100-101, 131-133)."""
    n = 16
    p = O.sn_params_ref(1, n, n)
    c = np.full((n, n), 0.7); e1 = np.full((n, n), 0.2); e2 = np.full((n, n), 0.1)
    co, _, _ = O.sn_step(c, e1, e2, p)
    assert np.all(co == 0.7)
    c = np.ones((n, n)); e1 = np.ones((n, n)); e2 = np.zeros((n, n))
    co, a, b = O.sn_step(c, e1, e2, p)
    assert np.all(co == 1) and np.all(a == 1) and np.all(b == 0)
    rng = np.random.default_rng(2)
    c, e1, e2 = (rng.uniform(0, 1, (n, n)) for _ in range(3))
    p[O.SN["DT"]] = 0.01   # large enough to overshoot, so the clamps act
    co, a, b = O.sn_step(c, e1, e2, p)
    for f in (co, a, b):
        assert f.min() >= 0.0 and f.max() <= 1.0
    assert np.any(co == 1.0) and np.any(co == 0.0) and np.any(a == 0.0)


# ------------------------------------------------------------ invariants
def test_sf_rotation_and_transpose_equivariance(O):
    """aniso = 4, theta0 = 0 (para1): eps(theta) is invariant under the square
    lattice's rotations and its diagonal reflection; on a periodic square grid
    the step commutes with rot90 and with transposition.  An exchanged x/y
    operand or a flux term with the wrong sign breaks this."""
    n = 48
    p = O.sf_params_ref(1, n, n)
    phi, T = smooth_fields(n, n, seed=3)
    po, to = O.sf_step(phi, T, p)
    for op in (np.rot90, np.transpose):
        pr, tr = O.sf_step(np.ascontiguousarray(op(phi)), np.ascontiguousarray(op(T)), p)
        assert rel_err(pr, op(po)) < 1e-13
        assert rel_err(tr, op(to)) < 1e-13


def test_sf_translation_equivariance_bitwise(O):
    n = 40
    p = O.sf_params_ref(3, n, n)
    phi, T = smooth_fields(n, n, seed=5)
    po, to = O.sf_step(phi, T, p)
    s = (5, 11)
    pr, tr = O.sf_step(np.roll(phi, s, (0, 1)), np.roll(T, s, (0, 1)), p)
    assert np.array_equal(pr, np.roll(po, s, (0, 1)))
    assert np.array_equal(tr, np.roll(to, s, (0, 1)))


def test_sf_neumann_mirror_equivariance(O):
    """Zero flux: the step commutes with the mirror x -> nx-1-x when the
    anisotropy is mirror-symmetric (aniso 4, theta0 0)."""
    n = 32
    p = O.sf_params_ref(1, n, n)
    phi, T = smooth_fields(n, n, seed=8)
    po, to = O.sf_step(phi, T, p, bc=1, stencil=1)
    pr, tr = O.sf_step(np.ascontiguousarray(phi[:, ::-1]), np.ascontiguousarray(T[:, ::-1]), p, bc=1, stencil=1)
    assert rel_err(pr, po[:, ::-1]) < 1e-13 and rel_err(tr, to[:, ::-1]) < 1e-13


def test_sn_reference_init_mirror_symmetry_bitwise(O):
    """The reference's two discs share the column x = dimX/2+8 (1-based,
    sintering/paras/para1.h:8); the periodic 5-point update preserves the
    mirror about that column exactly."""
    n = 128
    p = O.sn_params_ref(1, n, n)
    p[O.SN["R1"]] = 20; p[O.SN["R2"]] = 10; p[O.SN["Y1"]] = 50; p[O.SN["Y2"]] = 80
    c, e1, e2 = O.sn_init(n, n, p)
    c, e1, e2, _ = O.sn_run(c, e1, e2, p, 20)
    x0 = int(p[O.SN["X1"]]) - 1
    idx = (2 * x0 - np.arange(n)) % n
    for f in (c, e1, e2):
        assert np.array_equal(f, f[:, idx])


# --------------------------------------------------------------- init
def test_init_counts(O):
    """snowflake nucleus: (x-264)^2+(y-264)^2 < 10 has 29 cells
    (snowflake/tools.h:27-29); grain-growth circles R = 50 and 80 on 512^2
    hold 7845 and 20081 cells (Gauss circle counts)."""
    phi, T = O.sf_init(512, 512, O.sf_params_ref(1))
    assert phi.sum() == 29 and T.sum() == 0
    assert O.gg_init(512, 512, 2, O.gg_params_ref(1))[1].sum() == 7845
    assert O.gg_init(512, 512, 2, O.gg_params_ref(4))[1].sum() == 20081
    eta = O.gg_init(64, 64, 3, O.gg_params_ref(1, 64, 64))
    assert np.all(eta[0] + eta[1] == 1) and np.all(eta[2] == 0)


def test_sn_init_brute_force(O):
    n = 512
    p = O.sn_params_ref(1, n, n)
    c, e1, e2 = O.sn_init(n, n, p)
    yy, xx = np.mgrid[1:n + 1, 1:n + 1].astype(float)
    d1 = (xx - 264) ** 2 + (yy - 240) ** 2 <= 50 ** 2
    d2 = (xx - 264) ** 2 + (yy - 315) ** 2 <= 25 ** 2
    assert np.array_equal(c, (d1 | d2).astype(float))
    assert np.array_equal(e2, d2.astype(float))
    assert np.array_equal(e1, (d1 & ~d2).astype(float))


def test_config_inits(O):
    p = O.gg_params_config()
    eta = O.gg_init(64, 32, 5, p, init=1, seed=7)
    assert np.abs(eta).max() <= 1e-3 and abs(eta.mean()) < 5e-5
    assert not np.array_equal(eta[0], eta[1])
    assert np.array_equal(eta, O.gg_init(64, 32, 5, p, init=1, seed=7))
    ps = O.sn_params_config(256, 256, R=40.0)
    c, e1, e2 = O.sn_init(256, 256, ps, init=1, seed=3)
    assert c.min() >= 0 and c.max() <= 1
    assert abs(e1[128, 128] - 0.5) < 1e-12  # tanh profile is 1/2 on the radius (d1 = R)


# --------------------------------------------------------------- RNG
def test_rng_statistics_and_streams(O):
    u = np.array([O.rng_u01(1, 0, i) for i in range(20000)])
    assert u.min() >= 0 and u.max() < 1
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.005
    v = np.array([O.rng_u01(1, 1, i) for i in range(2000)])
    assert not np.allclose(u[:2000], v)
    assert O.rng_u01(5, 3, 9) == O.rng_u01(5, 3, 9)


def test_splitmix64_published_vector(O):
    """The mixer is SplitMix64: with seed 0 the generator's first outputs are
    0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4 (Vigna's reference, splitmix64.c)."""
    def sm(z):
        z = (z + 0x9E3779B97F4A7C15) & (2 ** 64 - 1)
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2 ** 64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2 ** 64 - 1)
        return z ^ (z >> 31)
    assert sm(0) == 0xE220A8397B1DCDAF
    assert sm(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4
    key = sm(7 ^ sm((3 + 0x632BE59BD9B4E019) & (2 ** 64 - 1)))
    assert O.rng_u01(7, 3, 11) == (sm(key ^ 11) >> 11) * 2.0 ** -53


def test_noise_hash_statistics_and_definition(O):
    """Per-step noise: u = (lowbias32(key ^ cell) >> 8) * 2^-24 with key the
    low 32 bits of the SplitMix64 keying of (seed, step)."""
    def sm(z):
        z = (z + 0x9E3779B97F4A7C15) & (2 ** 64 - 1)
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2 ** 64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2 ** 64 - 1)
        return z ^ (z >> 31)

    def lb(x):
        x ^= x >> 16; x = (x * 0x7FEB352D) & 0xFFFFFFFF
        x ^= x >> 15; x = (x * 0x846CA68B) & 0xFFFFFFFF
        return x ^ (x >> 16)
    key = sm(3 ^ sm((17 + 0x632BE59BD9B4E019) & (2 ** 64 - 1))) & 0xFFFFFFFF
    assert O.noise_u(3, 17, 12345) == (lb(key ^ 12345) >> 8) * 2.0 ** -24
    u = np.array([O.noise_u(1, 5, i) for i in range(20000)])
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.005
    assert not np.allclose(u[:1000], [O.noise_u(1, 6, i) for i in range(1000)])


# --------------------------------------------------------------- fp16
def test_f16_decode_exhaustive(O):
    bits = np.arange(65536, dtype=np.uint16)
    ref = bits.view(np.float16).astype(np.float64)
    got = np.array([O.f16_bits_to_f64(int(b)) for b in bits])
    ok = np.isnan(ref)
    assert np.array_equal(got[~ok], ref[~ok]) and np.all(np.isnan(got[ok]))


def test_f16_encode_round_trip_and_ties(O):
    bits = np.arange(65536, dtype=np.uint16)
    vals = bits.view(np.float16).astype(np.float64)
    fin = np.isfinite(vals)
    back = np.array([O.f64_to_f16_bits(v) for v in vals[fin]], dtype=np.uint16)
    assert np.array_equal(back, bits[fin])
    # midpoints and near-midpoints between consecutive finite halves, incl. subnormals
    pos = np.sort(vals[fin & (vals >= 0)])
    pos = np.unique(pos)
    mid = (pos[:-1] + pos[1:]) / 2
    tests = np.concatenate([mid, np.nextafter(mid, 0), np.nextafter(mid, 1), -mid, [65519.99, 65520.0, 1e-9, 2.0 ** -25]])
    got = np.array([O.f64_to_f16_bits(v) for v in tests], dtype=np.uint16)
    assert np.array_equal(got, tests.astype(np.float16).view(np.uint16))


# --------------------------------------------------------------- monitor
def test_monitor_spatial_hand_cases(O):
    t = 16
    f = np.zeros((32, 32))
    f[5, 10] = 0.7          # a sample point of tile (0,0)
    f[6, 10] = 5.0          # not a sample point
    f[16 + 15, 16] = -0.25  # sample of tile (1,1)
    md = O.monitor_spatial(f, t)
    assert md[0, 0] == 0.7 and md[1, 1] == 0.25 and md[0, 1] == 0.0
    # min starts at 1 and max at 0 (snowflake/tools.h:185): a constant tile of 2 gives 1
    assert O.monitor_spatial(np.full((16, 16), 2.0), 16)[0, 0] == 1.0
    assert O.sample_offsets(16, 1) == [0, 5, 10, 15] and O.sample_offsets(16, 2) == [2, 6, 10, 14]


def test_monitor_temporal_and_gradient(O):
    f = np.zeros((32, 32)); last = np.zeros((32, 32))
    last[2, 6] = 0.3; last[3, 6] = 9.0  # (2,6) sampled, (3,6) not
    md = O.monitor_temporal(last, f, 16)
    assert md[0, 0] == 0.3 and md[1, 1] == 0.0
    g = np.zeros((32, 32)); g[:, 20:] = 1.0
    mg = O.monitor_gradient(g, 16, bc=1)
    assert mg[0, 0] == 0.0 and mg[0, 1] == 0.25      # (0.5*(1-0))^2 at x = 19, 20
    mp = O.monitor_gradient(g, 16, bc=0)
    assert mp[0, 0] == 0.25                           # periodic wrap x = 0 sees x = 31


def test_get_type_rule(O):
    md = np.zeros((5, 6)); md[0, 0] = 1.0; md[3, 4] = 0.5
    old, t = O.get_type(md, 0.4, np.ones((5, 6), np.int32))
    exp = np.ones((5, 6), np.int32)
    for (y, x) in ((0, 0), (0, 1), (1, 0), (3, 4), (2, 4), (4, 4), (3, 3), (3, 5)):
        exp[y, x] = 2
    assert np.array_equal(t, exp) and np.all(old == 1)    # non-periodic: (4,0), (0,5) untouched
    md2 = np.full((2, 2), np.nan)
    prev = np.array([[1, 2], [2, 1]], np.int32)
    _, t2 = O.get_type(md2, 0.4, prev)
    assert np.array_equal(t2, prev)                       # unordered keeps the old type


# ------------------------------------------------------- forward tangents
def test_sf_tangent_matches_central_difference(O):
    n = 24
    p = O.sf_params_ref(2, n, n)
    phi, T = smooth_fields(n, n, seed=1, nbump=3)
    _, _, dphi, dT = O.sf_tangent(phi, T, p, 5, O.GRAD_SF["DELTA"])
    h = 1e-6
    pp, pm = p.copy(), p.copy()
    pp[O.SF["DELTA"]] += h; pm[O.SF["DELTA"]] -= h
    a, b, _ = O.sf_run(phi, T, pp, 5)
    c, d, _ = O.sf_run(phi, T, pm, 5)
    assert rel_err(dphi, (a - c) / (2 * h)) < 1e-5
    assert rel_err(dT, (b - d) / (2 * h)) < 1e-5


def test_sn_tangent_matches_central_difference(O):
    n = 32
    p = O.sn_params_ref(1, n, n)
    p[O.SN["R1"]] = 6; p[O.SN["R2"]] = 3; p[O.SN["X1"]] = 16; p[O.SN["X2"]] = 16; p[O.SN["Y1"]] = 12; p[O.SN["Y2"]] = 21
    c, e1, e2 = O.sn_init(n, n, p)
    for pid, key in ((O.GRAD_SN["DSURF"], "DSURF"), (O.GRAD_SN["KC"], "KC")):
        _, _, _, dc, _, _ = O.sn_tangent(c, e1, e2, p, 4, pid)
        h = 1e-6 * p[O.SN[key]]
        pp, pm = p.copy(), p.copy()
        pp[O.SN[key]] += h; pm[O.SN[key]] -= h
        cp = O.sn_run(c, e1, e2, pp, 4)[0]
        cm = O.sn_run(c, e1, e2, pm, 4)[0]
        assert rel_err(dc, (cp - cm) / (2 * h)) < 1e-5

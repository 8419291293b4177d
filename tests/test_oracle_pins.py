"""Pins of the oracle's update terms that equivariance and fixed points leave
open, each against a value the mathematics fixes independently of the
oracle's code:

- snowflake: the scalar ODE of a uniform field (every Laplacian and flux
  vanishes), axis-aligned plane waves (the gradient angle is 0/pi or
  +-pi/2, so for even aniso eps is a known constant and the flux vanishes:
  the update is eps^2 * lambda * mode + reaction, lambda the stencil's
  eigenvalue), and the anisotropic flux d_y(eps eps' phi_x) - d_x(eps eps'
  phi_y) against its continuum closed form on a Gaussian bump, converging
  at second order (snowflake/function.h:30-59);
- grain growth: the scalar ODE of a uniform field with three non-zero
  order parameters and non-unit alpha, beta, gamma (the coupling
  2 gamma eta sum_{j!=i} eta_j^2 included), and the embedding of a q-field
  state into q+1 fields with an identically zero field
  (graingrowth/function.h:20-23);
- sintering: the eta reaction ODE of a uniform state at c != 1, and the
  linearisations of the Cahn-Hilliard and Allen-Cahn updates around a
  uniform state, whose growth factors per Fourier mode are closed forms in
  A, B, kappa_c, kappa_eta, L and the mobility (sintering/function.h:67-133).

Each case is chosen so that a dropped term, a wrong coefficient or a flipped
sign changes the compared value by far more than the tolerance."""
import math

import numpy as np
import pytest

N = 64


def sf_p(O, n=N, aniso=6.0, theta0=3.1415926 / 2, delta=0.05, noise=0.0):
    p = O.sf_params_config(n, n)
    p[O.SF["ANISO"]] = aniso
    p[O.SF["THETA0"]] = theta0
    p[O.SF["DELTA"]] = delta
    p[O.SF["NOISE"]] = noise
    return p


def sf_m(O, p, T):
    return p[O.SF["ALPHA"]] / p[O.SF["PI"]] * np.arctan(p[O.SF["GAMMA"]] * (p[O.SF["TEQ"]] - T))


# ------------------------------------------------------------- snowflake
@pytest.mark.parametrize("noise", [0.0, 0.01])
def test_sf_uniform_field_scalar_ode(O, noise):
    p = sf_p(O, noise=noise)
    dt, tau, kap, a = (p[O.SF[k]] for k in ("DT", "TAU", "KAPPA", "NOISE"))
    seed, step = 7, 3
    u_noise = np.array([O.noise_u(seed, step, c) for c in range(N * N)]).reshape(N, N)
    for u, T0 in ((0.3, 0.2), (0.7, 0.9), (0.55, -0.3), (0.02, 1.4)):
        phi, T = np.full((N, N), u), np.full((N, N), T0)
        for stencil in (O.STENCIL_REF, O.STENCIL_LAP5):
            po, to = O.sf_step(phi, T, p, stencil=stencil, bc=O.BC_PERIODIC, seed=seed, step=step)
            br = u * (1 - u) * (u - 0.5 + sf_m(O, p, T0)) + a * u * (1 - u) * (u_noise - 0.5)
            exp_phi = u + dt / tau * br
            exp_T = T0 + kap * (exp_phi - u)
            assert np.max(np.abs(po - exp_phi)) <= 1e-14
            assert np.max(np.abs(to - exp_T)) <= 1e-14


@pytest.mark.parametrize("aniso,theta0", [(6.0, 3.1415926 / 2), (4.0, 0.25), (2.0, 0.4)])
@pytest.mark.parametrize("stencil", [0, 1])
@pytest.mark.parametrize("axis", [0, 1])
def test_sf_plane_wave_eps_squared_laplacian(O, aniso, theta0, stencil, axis):
    """phi = 0.5 + A sin(w x) (or y), T uniform: gy = 0 (gx = 0), theta in
    {0, pi} ({+-pi/2}); for even aniso eps = epsb(1 + delta cos(n(theta -
    theta0))) is the same constant on the whole grid and eps eps' is too,
    so both flux terms vanish and phi' = phi + dt/tau (eps^2 lambda (phi -
    0.5) + reaction) with lambda the stencil's eigenvalue of the mode."""
    p = sf_p(O, aniso=aniso, theta0=theta0, delta=0.07)
    dx, dt, tau, eb, de = (p[O.SF[k]] for k in ("DX", "DT", "TAU", "EPSB", "DELTA"))
    k = 3
    w = 2 * math.pi * k / N
    idx = np.arange(N)
    mode = 0.3 * np.sin(w * idx)
    phi = 0.5 + (mode[None, :] * np.ones((N, 1)) if axis == 0 else mode[:, None] * np.ones((1, N)))
    T0 = 0.25
    T = np.full((N, N), T0)
    po, to = O.sf_step(phi, T, p, stencil=stencil, bc=O.BC_PERIODIC)
    theta = 0.0 if axis == 0 else math.pi / 2
    E = eb * (1 + de * math.cos(aniso * (theta - theta0)))
    lam = (2 * math.cos(w) - 2) / (dx * dx)
    if stencil == O.STENCIL_REF:
        lam += (2 * math.cos(2 * w) - 2) / (dx * dx)
    br = E * E * lam * (phi - 0.5) + phi * (1 - phi) * (phi - 0.5 + sf_m(O, p, T0))
    exp_phi = phi + dt / tau * br
    scale = np.max(np.abs(exp_phi - phi))
    assert np.max(np.abs(po - exp_phi)) <= 1e-10 * scale
    # a wrong power of eps moves the result by far more than the tolerance
    alt = phi + dt / tau * (E * lam * (phi - 0.5) + phi * (1 - phi) * (phi - 0.5 + sf_m(O, p, T0)))
    assert np.max(np.abs(po - alt)) > 1e3 * 1e-10 * scale


def _bump_derivs(n, L, w, amp, x0, y0):
    """Gaussian phi = 0.5 + amp exp(-r^2/w^2) on an n^2 periodic grid of
    side L and its exact first and second derivatives."""
    h = L / n
    c = np.arange(n) * h
    X, Y = np.meshgrid(c - x0, c - y0)
    g = amp * np.exp(-(X * X + Y * Y) / (w * w))
    gx, gy = -2 * X / w ** 2 * g, -2 * Y / w ** 2 * g
    gxx = (4 * X * X / w ** 4 - 2 / w ** 2) * g
    gyy = (4 * Y * Y / w ** 4 - 2 / w ** 2) * g
    gxy = 4 * X * Y / w ** 4 * g
    return h, 0.5 + g, gx, gy, gxx, gyy, gxy, np.hypot(X, Y)


@pytest.mark.parametrize("stencil,factor", [(1, 1.0), (0, 5.0)])
def test_sf_anisotropic_flux_converges_to_continuum(O, stencil, factor):
    """phi' - phi = dt/tau (flux + eps^2 lap + reaction); with T uniform the
    reaction is pointwise, so (phi' - phi) tau/dt - reaction must converge
    at second order to the continuum operator
      f'(theta) (phi_x^2 phi_yy - 2 phi_x phi_y phi_xy + phi_y^2 phi_xx)/|grad phi|^2
      + eps(theta)^2 * factor * lap(phi),
    f = eps eps', theta = atan2(phi_y, phi_x), factor 1 for the 5-point
    operator and 5 for the radius-2 cross (sum of the 8 axis neighbours
    - 8 phi).  The first term is d_y(f phi_x) - d_x(f phi_y) with the chain
    rule through theta_x, theta_y: an independent derivation."""
    n_an, th0, de, eb = 4.0, 0.3, 0.3, 1.0
    L, w, amp = 1.0, 0.12, 0.4
    errs = []
    for n in (96, 192, 384):
        h, phi, gx, gy, gxx, gyy, gxy, r = _bump_derivs(n, L, w, amp, 0.5, 0.5)
        p = sf_p(O, n=n, aniso=n_an, theta0=th0, delta=de)
        p[O.SF["DX"]] = p[O.SF["DY"]] = h
        p[O.SF["EPSB"]] = eb
        p[O.SF["DT"]] = 1e-9
        p[O.SF["TAU"]] = 1.0
        T0 = 0.4
        T = np.full((n, n), T0)
        po, _ = O.sf_step(phi, T, p, stencil=stencil, bc=O.BC_PERIODIC)
        react = phi * (1 - phi) * (phi - 0.5 + sf_m(O, p, T0))
        num = (po - phi) / p[O.SF["DT"]] - react
        th = np.arctan2(gy, gx)
        a = n_an * (th - th0)
        eps = eb * (1 + de * np.cos(a))
        epsd = -eb * n_an * de * np.sin(a)
        epsdd = -eb * n_an * n_an * de * np.cos(a)
        fprime = epsd * epsd + eps * epsdd
        with np.errstate(invalid="ignore", divide="ignore"):  # the bump's centre (excluded below)
            flux = fprime * (gx * gx * gyy - 2 * gx * gy * gxy + gy * gy * gxx) / (gx * gx + gy * gy)
        cont = flux + eps * eps * factor * (gxx + gyy)
        ring = (r > 0.5 * w) & (r < 2.0 * w)
        errs.append(np.max(np.abs(num - cont)[ring]) / np.max(np.abs(cont)[ring]))
    # second order: halving h divides the error by ~4 (3.3 and 3.7 measured
    # on this bump, approaching 4); a wrong coefficient leaves an O(1) error
    # that does not shrink
    assert errs[2] < 0.02
    for e0, e1 in zip(errs, errs[1:]):
        assert 3.0 < e0 / e1 < 5.0


# ----------------------------------------------------------- grain growth
def test_gg_uniform_field_scalar_ode(O):
    p = O.gg_params_config()
    p[O.GG["ALPHA"]], p[O.GG["BETA"]], p[O.GG["GAMMA"]] = 1.3, 0.7, 1.1
    dt, L = p[O.GG["DT"]], p[O.GG["L"]]
    al, be, ga = 1.3, 0.7, 1.1
    v = np.array([0.3, 0.6, 0.1, -0.2])
    eta = np.ascontiguousarray(np.broadcast_to(v[:, None, None], (4, 32, 32)))
    for stencil in (0, 1):
        out = O.gg_step(eta, p, stencil=stencil)
        for i in range(4):
            others = sum(v[j] ** 2 for j in range(4) if j != i)
            exp = v[i] - dt * L * (2 * ga * v[i] * others + be * v[i] ** 3 - al * v[i])
            assert np.max(np.abs(out[i] - exp)) <= 1e-15
    # without the coupling term the values would differ by O(dt L): pinned
    i = 0
    no_coupling = v[i] - dt * L * (be * v[i] ** 3 - al * v[i])
    assert abs(out[i, 0, 0] - no_coupling) > 1e-4


def test_gg_zero_field_embedding(O):
    """A state of q fields with a (q+1)-th field identically 0 evolves like
    the q-field state (0^2 adds exactly nothing to the coupling sum), and the
    zero field stays exactly 0: bitwise."""
    rng = np.random.default_rng(3)
    p = O.gg_params_ref(1, 48, 48)
    eta2 = rng.uniform(0, 1, (2, 48, 48))
    eta3 = np.concatenate([eta2, np.zeros((1, 48, 48))])
    for stencil in (0, 1):
        a = O.gg_step(eta2, p, stencil=stencil)
        b = O.gg_step(eta3, p, stencil=stencil)
        assert np.array_equal(a, b[:2])
        assert not np.any(b[2])


# -------------------------------------------------------------- sintering
def sn_p(O):
    return O.sn_params_config(64, 64)


def test_sn_uniform_eta_reaction_at_partial_density(O):
    p = sn_p(O)
    dt, L, B = p[O.SN["DT"]], p[O.SN["L"]], p[O.SN["B"]]
    for c0, a1, a2 in ((0.4, 0.3, 0.5), (0.75, 0.6, 0.15), (0.2, 0.05, 0.0)):
        c, e1, e2 = (np.full((32, 32), v) for v in (c0, a1, a2))
        co, o1, o2 = O.sn_step(c, e1, e2, p)
        assert np.array_equal(co, c)  # uniform: lap mu = 0 exactly
        s2 = a1 * a1 + a2 * a2
        for a, o in ((a1, o1), (a2, o2)):
            d = B * (-12 * a * a * (2 - c0) + 12 * a * (1 - c0) + 12 * a * s2)
            exp = min(max(a - dt * L * d, 0.0), 1.0)
            assert np.max(np.abs(o - exp)) <= 1e-15
            # the (1 - c) term's sign matters at c != 1
            flip = a - dt * L * B * (-12 * a * a * (2 - c0) - 12 * a * (1 - c0) + 12 * a * s2)
            if a > 0:
                assert abs(o[0, 0] - flip) > 1e-6


def test_sn_cahn_hilliard_linearisation(O):
    """c = c0 + a sin(w x) with uniform etas: to first order in a,
    c' - c = dt M(c0) lambda (f''(c0) - kc/2 lambda) a sin(w x), with
    lambda = (2 cos w - 2)/dx^2, f''(c) = 2B + 2A(1 - 6c + 6c^2) and
    M = Dvol h + Dvap (1-h) + Dsurf c(1-c) + Dgb 2 eta1 eta2, h = c^3(10 -
    15c + 6c^2)."""
    p = sn_p(O)
    dx, dt, A, B, kc = (p[O.SN[k]] for k in ("DX", "DT", "A", "B", "KC"))
    dvol, dvap, dsur, dgb = (p[O.SN[k]] for k in ("DVOL", "DVAP", "DSURF", "DGB"))
    n, k, amp = 64, 5, 1e-6
    w = 2 * math.pi * k / n
    mode = np.sin(w * np.arange(n))[None, :] * np.ones((n, 1))
    for c0, a1, a2 in ((0.4, 0.2, 0.3), (0.7, 0.5, 0.45)):
        c = c0 + amp * mode
        e1, e2 = np.full((n, n), a1), np.full((n, n), a2)
        co, _, _ = O.sn_step(c, e1, e2, p)
        lam = (2 * math.cos(w) - 2) / (dx * dx)
        fpp = 2 * B + 2 * A * (1 - 6 * c0 + 6 * c0 * c0)
        h = c0 ** 3 * (10 - 15 * c0 + 6 * c0 * c0)
        M = dvol * h + dvap * (1 - h) + dsur * c0 * (1 - c0) + dgb * 2 * a1 * a2
        g = dt * M * lam * (fpp - 0.5 * kc * lam)
        big = np.abs(mode) > 0.5
        ratio = (co - c)[big] / (amp * mode[big])
        assert np.max(np.abs(ratio - g)) <= 1e-4 * abs(g)


def test_sn_allen_cahn_linearisation(O):
    """eta1 = e0 + a sin(w y), uniform c and eta2: to first order,
    eta1' - eta1 = g a sin(w y) with
    g = -dt L (d'(e0) - ke/2 lambda) - dt L s dc,
    d'(e) = B(-24 e (2-c) + 12 (1-c) + 12 (e^2 + eta2^2) + 24 e^2), where the
    eta update sees the new c: mu varies with eta1 (d mu/d eta1 = s =
    B(12 e^2 - 12 e)), so c moves by dc = dt M(c0) lambda s per unit mode
    amplitude, and d(reaction)/dc = s as well."""
    p = sn_p(O)
    dx, dt, L, B, ke = (p[O.SN[k]] for k in ("DX", "DT", "L", "B", "KE"))
    dvol, dvap, dsur, dgb = (p[O.SN[k]] for k in ("DVOL", "DVAP", "DSURF", "DGB"))
    n, k, amp = 64, 3, 1e-6
    w = 2 * math.pi * k / n
    mode = np.sin(w * np.arange(n))[:, None] * np.ones((1, n))
    for c0, e0, b0 in ((0.5, 0.4, 0.3), (0.9, 0.7, 0.1)):
        c, e2 = np.full((n, n), c0), np.full((n, n), b0)
        e1 = e0 + amp * mode
        _, o1, _ = O.sn_step(c, e1, e2, p)
        lam = (2 * math.cos(w) - 2) / (dx * dx)
        dprime = B * (-24 * e0 * (2 - c0) + 12 * (1 - c0) + 12 * (e0 * e0 + b0 * b0) + 24 * e0 * e0)
        sl = B * (12 * e0 * e0 - 12 * e0)
        h = c0 ** 3 * (10 - 15 * c0 + 6 * c0 * c0)
        M = dvol * h + dvap * (1 - h) + dsur * c0 * (1 - c0) + dgb * 2 * e0 * b0
        g = -dt * L * (dprime - 0.5 * ke * lam) - dt * L * sl * (dt * M * lam * sl)
        base = e0 - dt * L * B * (-12 * e0 * e0 * (2 - c0) + 12 * e0 * (1 - c0) + 12 * e0 * (e0 * e0 + b0 * b0))
        big = np.abs(mode) > 0.5
        ratio = ((o1 - base) - amp * mode)[big] / (amp * mode[big])
        assert np.max(np.abs(ratio - g)) <= 1e-4 * max(abs(g), 1e-12) + 1e-6

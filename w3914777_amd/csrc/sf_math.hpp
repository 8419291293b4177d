// sf_math.hpp — snowflake anisotropy without trigonometry, shared by the
// forward sweeps and the adjoint.
//
// For integer anisotropy n and theta = atan2(gy, gx):
//   cos(n theta) + i sin(n theta) = ((gx + i gy) / |g|)^n,
// evaluated by binary powering after an exact power-of-two rescale of
// (gx, gy), then rotated by n*theta0 with host-computed cos/sin.  At |g| = 0
// theta = atan2(0, 0) = 0 (cos = 1, sin = 0), as in the reference, which
// evaluates eps = epsb(1 + delta cos(aniso(theta - theta0))) and
// eps' = -epsb aniso delta sin(aniso(theta - theta0)) (snowflake/function.h:32-34).
#pragma once

#include "tile_common.hpp"

namespace amps {

__device__ inline double rcp_any(double x) { return rcp_nr(x); }
__device__ inline float rcp_any(float x) { return __builtin_amdgcn_rcpf(x); }

// (c, s) = (cos n theta, sin n theta) of the gradient direction.
template <typename R>
__device__ inline void sf_unit_power(R gx, R gy, int n, R& c, R& s) {
    R mx = fmax(fabs(gx), fabs(gy));
    c = R(1);
    s = R(0);
    if (mx != R(0)) {
        int e;
        frexp(mx, &e);
        R u = ldexp(gx, -e), v = ldexp(gy, -e);  // exact scaling, max(|u|,|v|) in [0.5, 1)
        R wr = R(1), wi = R(0), br = u, bi = v;
        int k = n;
        while (true) {
            if (k & 1) {
                R t = wr * br - wi * bi;
                wi = wr * bi + wi * br;
                wr = t;
            }
            k >>= 1;
            if (!k) break;
            R t = br * br - bi * bi;
            bi = R(2) * br * bi;
            br = t;
        }
        R r2 = u * u + v * v;
        R pw = R(1), b = r2;
        int h = n >> 1;
        while (h) {
            if (h & 1) pw *= b;
            h >>= 1;
            if (h) b *= b;
        }
        if (n & 1) pw *= sqrt(r2);
        R inv = rcp_any(pw);
        c = wr * inv;
        s = wi * inv;
    }
}

// (cos 6 theta, sin 6 theta): the unit vector z^2/|z|^2 = (cos 2 theta,
// sin 2 theta) of z = gx + i gy, cubed.  (gx, gy) is first scaled exactly by
// a power of two to max(|u|,|v|) in [0.5, 1), so the squares neither
// underflow nor overflow for any finite gradient, subnormal ones included (a
// reciprocal of a subnormal maximum would overflow).  A zero gradient gives
// (1, 0), as atan2(0, 0) = 0 does.
__device__ inline int frexp_exp(float x) { return __builtin_amdgcn_frexp_expf(x); }
__device__ inline int frexp_exp(double x) { return __builtin_amdgcn_frexp_exp(x); }
__device__ inline float ldexp_any(float x, int e) { return __builtin_amdgcn_ldexpf(x, e); }
__device__ inline double ldexp_any(double x, int e) { return __builtin_amdgcn_ldexp(x, e); }

template <typename R>
__device__ inline void sf_unit_power6(R gx, R gy, R& c, R& s) {
    const R mx = fmax(fabs(gx), fabs(gy));
    const int e = frexp_exp(mx);
    const R u = ldexp_any(gx, -e), v = ldexp_any(gy, -e);
    const R uu = u * u, vv = v * v, r2 = uu + vv;
    const bool nz = r2 > R(0);
    const R inv = rcp_any(nz ? r2 : R(1));
    const R c2 = nz ? (uu - vv) * inv : R(1), s2 = nz ? R(2) * u * v * inv : R(0);
    const R cc = c2 * c2, ss = s2 * s2;
    c = c2 * (cc - R(3) * ss);
    s = s2 * (R(3) * cc - ss);
}

template <typename R>
__device__ inline void sf_aniso(R gx, R gy, int n, R cth0, R sth0, R epsb, R epsb_delta, R m_epsb_n_delta,
                                R& eps, R& epsd) {
    R c, s;
    if (n == 6) sf_unit_power6<R>(gx, gy, c, s);  // the snowflake's six-fold case
    else sf_unit_power<R>(gx, gy, n, c, s);
    R ca = c * cth0 + s * sth0;  // cos(n (theta - theta0))
    R sa = s * cth0 - c * sth0;  // sin(n (theta - theta0))
    eps = epsb + epsb_delta * ca;
    epsd = m_epsb_n_delta * sa;
}

}  // namespace amps

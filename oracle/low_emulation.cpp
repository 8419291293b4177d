// low_emulation.cpp — CPU emulation of the engine's low-precision tile step.
//
// TEST INFRASTRUCTURE ONLY, like oracle.cpp: tests/ call it as the checker
// of the HIP engine's low tiles; the product never links it.
//
// The reference's low-precision arithmetic (half2 with CUDA intrinsics,
// snowflake/function.h:114-292,417-537, graingrowth/function.h:74-228) cannot
// be run here and is not what the engine computes.  The engine's low tiles
// store fp16 words h with a per-tile FP64 offset and power-of-two scale
// (value = off + h * scl) and compute in FP32 on deviations from their own
// offset.  This file restates that scheme in plain C++ (float for the FP32
// arithmetic, double where the engine uses double) from its definition:
//
//   decode    own cell: h * scl (float); neighbour cell of a low tile:
//             (float)(off_n - off_own) + h * scl_n; of an FP64 tile:
//             (float)(v - off_own)
//   update    the reference formulas (snowflake/function.h:5-60,
//             graingrowth/function.h:5-24) on the deviations, the bulk terms
//             on off + deviation; the six-fold (or integer-n) anisotropy as
//             ((gx + i gy)/|g|)^n after an exact power-of-two rescale
//   re-centre mid = (min + max)/2, half-range hw, scl = 2^(e-14) with
//             hw < 2^e (e >= -100; hw = 0 gives scl = 0), h = fp16_rne((r -
//             mid) 2^(14-e)), new offset = off + mid (double)
//
// With half2 set (the engine's AMPS_FLAG_HALF2 snowflake tiles) the six-fold
// unit vector with its rotation and the driving force's atan follow the
// engine's packed-half2 sequence instead, one fp16 rounding per operation.
//
// The engine fuses multiply-adds and uses hardware reciprocal and atan
// approximations of ~1 ulp, so results agree to within rounding of the FP32
// arithmetic: decoded values within one fp16 quantum (scl) of the tile.
// Tiles of type 2 are not touched here (the FP64 oracle covers them).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/amps.h"

namespace {

inline int bmap(int i, int n, int bc) {
    if (i >= 0 && i < n) return i;
    if (bc == AMPS_BC_PERIODIC) return i < 0 ? i + n : i - n;
    return i < 0 ? -i : 2 * (n - 1) - i;
}

// IEEE binary16 <-> binary32, round to nearest even.
float h2f(uint16_t h) {
    const uint32_t s = (uint32_t)(h >> 15) << 31, e = (h >> 10) & 0x1F, m = h & 0x3FF;
    if (e == 0) {
        const float v = std::ldexp((float)m, -24);
        return s ? -v : v;
    }
    uint32_t b;
    if (e == 31) b = s | 0x7F800000u | (m << 13);
    else b = s | ((e + 112) << 23) | (m << 13);
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

uint16_t f2h(float x) {
    uint32_t b;
    std::memcpy(&b, &x, 4);
    const uint16_t s = (uint16_t)((b >> 16) & 0x8000);
    const float a = std::fabs(x);
    if (std::isnan(x)) return s | 0x7E00;
    if (a >= 65520.0f) return s | 0x7C00;  // rounds to infinity
    if (a < std::ldexp(1.0f, -14)) {       // subnormal range: multiples of 2^-24
        const float q = std::nearbyint(a * 16777216.0f);  // exact scaling, RNE (default mode)
        return s | (uint16_t)q;
    }
    int e;
    const float fr = std::frexp(a, &e);  // a = fr 2^e, fr in [0.5, 1)
    const float mant = std::nearbyint(std::ldexp(fr, 11));  // 11 significant bits, RNE
    int E = e - 1;
    uint32_t M = (uint32_t)mant;  // in [1024, 2048]
    if (M == 2048) { M = 1024; E += 1; }
    if (E > 15) return s | 0x7C00;
    return s | (uint16_t)((E + 15) << 10) | (uint16_t)(M - 1024);
}

struct Code {
    float mid, inv, scl;
};

Code tile_code(float mn, float mx) {
    Code c;
    c.mid = 0.5f * (mn + mx);
    const float hw = 0.5f * (mx - mn);
    int ex = hw == 0.0f ? -100 : 0;
    if (hw > 0.0f && hw <= 3.0e38f) std::frexp(hw, &ex);
    if (ex < -100) ex = -100;
    c.scl = hw == 0.0f ? 0.0f : std::ldexp(1.0f, ex - 14);
    c.inv = std::ldexp(1.0f, 14 - ex);
    return c;
}

uint32_t lowbias32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
uint32_t noise_key(uint64_t seed, uint64_t step) {
    return (uint32_t)splitmix64(seed ^ splitmix64(step + 0x632BE59BD9B4E019ull));
}
float noise_u24(uint32_t key, uint32_t cell) { return (float)(lowbias32(key ^ cell) >> 8) * 0x1.0p-24f; }

// (cos n theta, sin n theta) of the gradient direction (theta = atan2(gy, gx),
// 0 for a zero gradient), from ((u + iv))^n / |u + iv|^n after scaling
// (gx, gy) by 2^-e so that max(|u|,|v|) is in [0.5, 1).
void unit_power(float gx, float gy, int n, float& c, float& s) {
    const float mx = std::fmax(std::fabs(gx), std::fabs(gy));
    c = 1.0f;
    s = 0.0f;
    if (mx == 0.0f) return;
    int e;
    std::frexp(mx, &e);
    const float u = std::ldexp(gx, -e), v = std::ldexp(gy, -e);
    if (n == 6) {
        const float uu = u * u, vv = v * v, r2 = uu + vv;
        if (!(r2 > 0.0f)) return;
        const float inv = 1.0f / r2;
        const float c2 = (uu - vv) * inv, s2 = 2.0f * u * v * inv;
        const float cc = c2 * c2, ss = s2 * s2;
        c = c2 * (cc - 3.0f * ss);
        s = s2 * (3.0f * cc - ss);
        return;
    }
    float wr = 1.0f, wi = 0.0f, br = u, bi = v;
    for (int k = n;;) {
        if (k & 1) {
            const float t = wr * br - wi * bi;
            wi = wr * bi + wi * br;
            wr = t;
        }
        k >>= 1;
        if (!k) break;
        const float t = br * br - bi * bi;
        bi = 2.0f * br * bi;
        br = t;
    }
    const float r2 = u * u + v * v;
    float pw = 1.0f, b = r2;
    for (int h = n >> 1; h;) {
        if (h & 1) pw *= b;
        h >>= 1;
        if (h) b *= b;
    }
    if (n & 1) pw *= std::sqrt(r2);
    c = wr / pw;
    s = wi / pw;
}

// ---- packed-half2 terms (the engine's AMPS_FLAG_HALF2 low tiles).  Every
// fp16 operation rounds once (to nearest even, subnormals kept, overflow to
// infinity).  Values are carried as doubles that are fp16 values: a product
// of two is exact in double, and the sums and fused multiply-adds below span
// fewer than 53 bits, so rounding the double result to fp16 is the one
// rounding of the fp16 operation.
double rh(double x) {
    if (x == 0.0 || !std::isfinite(x)) return x;
    const double a = std::fabs(x);
    int e;
    std::frexp(a, &e);                       // a = m 2^e, m in [0.5, 1)
    const int q = e - 11 < -24 ? -24 : e - 11;  // quantum: 11 significant bits, subnormals 2^-24
    double r = std::ldexp(std::nearbyint(std::ldexp(a, -q)), q);  // RNE (default mode)
    if (r > 65504.0) r = INFINITY;
    return std::copysign(r, x);
}

// (cos 6(theta - theta0), sin 6(theta - theta0)) from the rescaled gradient
// (u, v), as the engine's packed sequence: w = (u + iv)^2 = A + iB,
// (A (A^2 - 3B^2), B (3A^2 - B^2)) / (R2 R2 R2) with R2 = u^2 + v^2, then the
// rotation by the fp16 cos/sin of 6 theta0.
void aniso6_h2(float u, float v, float cth0, float sth0, float& ca, float& sa) {
    const double U = rh(u), V = rh(v);
    const double UU = rh(U * U), VV = rh(V * V);
    const double R2 = rh(UU + VV), A = rh(UU - VV);
    const double B = rh(rh(U + U) * V);
    const double AA = rh(A * A), BB = rh(B * B);
    const double Cn = rh(A * rh(std::fma(BB, -3.0, AA)));
    const double Sn = rh(B * rh(std::fma(AA, 3.0, -BB)));
    const double inv = rh(1.0 / rh(rh(R2 * R2) * R2));
    const double C = rh(Cn * inv), S = rh(Sn * inv);
    const double c0 = rh(cth0), s0 = rh(sth0);
    ca = (float)rh(std::fma(S, s0, rh(C * c0)));
    sa = (float)rh(std::fma(S, c0, -rh(C * s0)));
}

// atan in fp16: z = min(|x|, 1/|x|), p = z (c0 + c1 z^2 + c2 z^4 + c3 z^6),
// pi/2 - p where |x| > 1 (as (pi/2 - 2p) f + p with f = clamp(2048 |x| - 2048,
// 0, 1)), the sign of x.  x is the FP32 argument.
float atan_h2(float x) {
    const double c0 = rh(0.99926763f), c1 = rh(-0.32142954f), c2 = rh(0.14661292f), c3 = rh(-0.03913253f);
    const double X = rh(x), AX = std::fabs(X);
    const double Z = std::fmin(AX, rh(1.0 / AX));
    const double Z2 = rh(Z * Z);
    double p = rh(std::fma(Z2, c3, c2));
    p = rh(std::fma(Z2, p, c1));
    p = rh(std::fma(Z2, p, c0));
    p = rh(p * Z);
    const double f = std::fmin(std::fmax(rh(std::fma(AX, 2048.0, -2048.0)), 0.0), 1.0);
    const double r = rh(std::fma(f, rh(std::fma(p, -2.0, rh(1.5707963705062866))), p));
    return std::copysign((float)r, std::signbit(X) ? -1.0f : 1.0f);
}

}  // namespace

extern "C" {

// One NO_SKIP step of every low tile (type 1) of a snowflake grid.
//   types   [ty][tx] 1 low, 2 FP64
//   phi, T  FP64 values of the FP64 tiles ([ny][nx]; other cells unused)
//   w       fp16 words [2][ny][nx] (phi, T), meaningful on low tiles
//   off     [2][ty][tx] FP64 offsets, scl [2][ty][tx] scales
//   w_out, off_out, scl_out: the new words / codes of the low tiles (entries
//   of FP64 tiles are left as they are)
// noise_skip: the noise term is evaluated only where |phi (1 - phi)| >
// noise_skip (0: wherever it is non-zero, the NO_SKIP rule).
void oracle_sf_low_step(int nx, int ny, int tile, const double* p, int stencil, int bc, uint64_t seed, int64_t step,
                        const int32_t* types, const double* phi, const double* T, const uint16_t* w,
                        const double* off, const float* scl, float noise_skip, uint16_t* w_out, double* off_out,
                        float* scl_out, int half2) {
    const int TS = tile, H = 2, LW = TS + 2 * H, RW = TS + 2;
    const int tx = nx / TS, ty = ny / TS, nt = tx * ty;
    const size_t ncell = (size_t)nx * ny;
    // FP32 constants as the engine rounds them (double products then float)
    const double n_an = p[AMPS_SF_ANISO];
    const float inv_dxdy = (float)(1.0 / (p[AMPS_SF_DX] * p[AMPS_SF_DY]));
    const float inv2dx = (float)(1.0 / (2.0 * p[AMPS_SF_DX])), inv2dy = (float)(1.0 / (2.0 * p[AMPS_SF_DY]));
    const float dtt = (float)(p[AMPS_SF_DT] / p[AMPS_SF_TAU]), dt = (float)p[AMPS_SF_DT];
    const float kap = (float)p[AMPS_SF_KAPPA], aop = (float)(p[AMPS_SF_ALPHA] / p[AMPS_SF_PI]);
    const float gam = (float)p[AMPS_SF_GAMMA], epsb = (float)p[AMPS_SF_EPSB];
    const float epsbd = (float)(p[AMPS_SF_EPSB] * p[AMPS_SF_DELTA]);
    const float mend = (float)(-p[AMPS_SF_EPSB] * n_an * p[AMPS_SF_DELTA]);
    const float cth0 = (float)std::cos(n_an * p[AMPS_SF_THETA0]), sth0 = (float)std::sin(n_an * p[AMPS_SF_THETA0]);
    const float noise = (float)p[AMPS_SF_NOISE];
    const uint32_t key = noise_key(seed, (uint64_t)step);
    const int n = (int)n_an;
    const bool ref = stencil == AMPS_STENCIL_REF;
    std::memcpy(w_out, w, 2 * ncell * sizeof(uint16_t));
    std::memcpy(off_out, off, 2 * (size_t)nt * sizeof(double));
    std::memcpy(scl_out, scl, 2 * (size_t)nt * sizeof(float));
#pragma omp parallel for schedule(dynamic)
    for (int t = 0; t < nt; ++t) {
        if (types[t] != 1) continue;
        const int tyi = t / tx, txi = t - tyi * tx, ox = txi * TS, oy = tyi * TS;
        const double offp = off[t], offT = off[nt + t];
        std::vector<float> bp(LW * LW), bt(LW * LW), P(RW * RW), Q(RW * RW), E2(RW * RW);
        // box of deviations from the tile's own offsets
        for (int by = 0; by < LW; ++by)
            for (int bx = 0; bx < LW; ++bx) {
                const int gy = bmap(oy - H + by, ny, bc), gx = bmap(ox - H + bx, nx, bc);
                const int nt_ = (gy / TS) * tx + gx / TS;
                const size_t c = (size_t)gy * nx + gx;
                float vp, vt;
                if (nt_ == t) {
                    vp = h2f(w[c]) * scl[t];
                    vt = h2f(w[ncell + c]) * scl[nt + t];
                } else if (types[nt_] == 2) {
                    vp = (float)(phi[c] - offp);
                    vt = (float)(T[c] - offT);
                } else {
                    vp = (float)(off[nt_] - offp) + h2f(w[c]) * scl[nt_];
                    vt = (float)(off[nt + nt_] - offT) + h2f(w[ncell + c]) * scl[nt + nt_];
                }
                bp[by * LW + bx] = vp;
                bt[by * LW + bx] = vt;
            }
        // ring intermediates (zero-flux: a ring cell outside the domain takes
        // its mirror image's values)
        for (int ry = 0; ry < RW; ++ry)
            for (int rx = 0; rx < RW; ++rx) {
                int sy = ry, sx = rx;
                if (bc != AMPS_BC_PERIODIC) {
                    sy = bmap(oy - 1 + ry, ny, bc) - oy + 1;
                    sx = bmap(ox - 1 + rx, nx, bc) - ox + 1;
                }
                const int c = (sy + 1) * LW + sx + 1;
                const float gx = (bp[c + 1] - bp[c - 1]) * inv2dx;
                const float gy = (bp[c + LW] - bp[c - LW]) * inv2dy;
                float ca, sa;
                if (half2 && n == 6) {
                    const float mx = std::fmax(std::fabs(gx), std::fabs(gy));
                    int ex;
                    std::frexp(mx, &ex);
                    const float u = mx > 0.0f ? std::ldexp(gx, -ex) : 1.0f, v = mx > 0.0f ? std::ldexp(gy, -ex) : 0.0f;
                    aniso6_h2(u, v, cth0, sth0, ca, sa);
                } else {
                    float cn, sn;
                    unit_power(gx, gy, n, cn, sn);
                    ca = cn * cth0 + sn * sth0;
                    sa = sn * cth0 - cn * sth0;
                }
                const float eps = epsb + epsbd * ca, epsd = mend * sa;
                const float A = eps * epsd;
                P[ry * RW + rx] = A * gx;
                Q[ry * RW + rx] = A * gy;
                E2[ry * RW + rx] = eps * eps;
            }
        std::vector<float> rp(TS * TS), rt(TS * TS);
        float mnp = INFINITY, mxp = -INFINITY, mnt = INFINITY, mxt = -INFINITY;
        const float phi_off = (float)offp, teq_m = (float)(p[AMPS_SF_TEQ] - offT);
        for (int ly = 0; ly < TS; ++ly)
            for (int lx = 0; lx < TS; ++lx) {
                const int b = (ly + H) * LW + lx + H;
                auto lap = [&](const std::vector<float>& F) {
                    if (ref)
                        return (F[b - 1] + F[b - 2] + F[b + 1] + F[b + 2] + F[b - LW] + F[b - 2 * LW] + F[b + LW] +
                                F[b + 2 * LW] - 8.0f * F[b]) * inv_dxdy;
                    return (F[b - 1] + F[b + 1] + F[b - LW] + F[b + LW] - 4.0f * F[b]) * inv_dxdy;
                };
                const float lp = lap(bp), lt = lap(bt);
                const int r = (ly + 1) * RW + lx + 1;
                const float term1 = (P[r + RW] - P[r - RW]) * inv2dy;
                const float term2 = -(Q[r + 1] - Q[r - 1]) * inv2dx;
                const float flux = term1 + term2 + E2[r] * lp;
                const float ph = bp[b], tt = bt[b];
                const float pa = phi_off + ph;
                const float m = aop * (half2 ? atan_h2(gam * (teq_m - tt)) : std::atan(gam * (teq_m - tt)));
                const float pp = pa * (1.0f - pa);
                float br = flux + pp * (pa - 0.5f + m);
                if (noise != 0.0f && std::fabs(pp) > noise_skip) {
                    const float u = noise_u24(key, (uint32_t)((oy + ly) * nx + ox + lx));
                    br = br + noise * pa * (1.0f - pa) * (u - 0.5f);
                }
                const float phn = ph + dtt * br;
                const float tn = tt + dt * lt + kap * (phn - ph);
                rp[ly * TS + lx] = phn;
                rt[ly * TS + lx] = tn;
                mnp = std::fmin(mnp, phn); mxp = std::fmax(mxp, phn);
                mnt = std::fmin(mnt, tn); mxt = std::fmax(mxt, tn);
            }
        const Code cp = tile_code(mnp, mxp), ct = tile_code(mnt, mxt);
        for (int ly = 0; ly < TS; ++ly)
            for (int lx = 0; lx < TS; ++lx) {
                const size_t c = (size_t)(oy + ly) * nx + ox + lx;
                w_out[c] = f2h((rp[ly * TS + lx] - cp.mid) * cp.inv);
                w_out[ncell + c] = f2h((rt[ly * TS + lx] - ct.mid) * ct.inv);
            }
        off_out[t] = offp + (double)cp.mid;
        off_out[nt + t] = offT + (double)ct.mid;
        scl_out[t] = cp.scl;
        scl_out[nt + t] = ct.scl;
    }
}

// One step of the grain-growth low tiles (the 16x16 one-wave sweep's FP32
// path; each field has its own tile types): for every low tile of field f,
// the box of deviations from the tile's offset is decoded as above, then per
// cell, with eta = (float)off + dv and S the carried sum of squares (the
// decoded fields' squares summed in field order, FP64, rounded to float):
//   lap (Moore): fma(-9, dv, (cs[x-1] + cs[x]) + cs[x+1]) / (dx dy), cs the
//                column sums (u + m) + d of the 3x3 block
//   lap (5-point): fma(-4, dv, ((m[x-1] + m[x+1]) + u) + d) / (dx dy)
//   others = fma(-eta, eta, S), df = fma(2 gamma eta, others, fma(beta,
//   eta^2, -alpha) eta), r = fma(-dt L, fma(-kappa, lap, df), dv)
// (graingrowth/function.h:20-23; the engine writes these fused multiply-adds
// explicitly), re-centred as above.  A field whose codes
// are exactly 0 (offset 0, scale 0, low) on the tile and its 8 neighbours
// stays exactly 0.  Arrays are [q][ny][nx] (fields, words) and [q][ty][tx]
// (types, offsets, scales).
void oracle_gg_low_step(int nx, int ny, int q, int tile, const double* p, int stencil, int bc, const int32_t* types,
                        const double* v, const uint16_t* w, const double* off, const float* scl, uint16_t* w_out,
                        double* off_out, float* scl_out) {
    const int TS = tile, tx = nx / TS, ty = ny / TS, nt = tx * ty;
    const size_t ncell = (size_t)nx * ny;
    const float invf = (float)(1.0 / (p[AMPS_GG_DX] * p[AMPS_GG_DY])), dtlf = (float)(p[AMPS_GG_DT] * p[AMPS_GG_L]);
    const float kapf = (float)p[AMPS_GG_KAPPA], alf = (float)p[AMPS_GG_ALPHA], bef = (float)p[AMPS_GG_BETA];
    const float g2f = (float)(2.0 * p[AMPS_GG_GAMMA]);
    const bool moore = stencil == AMPS_STENCIL_REF;
    std::vector<double> S(ncell, 0.0);  // accumulated as the engine does, with fused multiply-adds
    for (int f = 0; f < q; ++f)
        for (size_t c = 0; c < ncell; ++c) S[c] = std::fma(v[f * ncell + c], v[f * ncell + c], S[c]);
    auto zero_code = [&](int f, int t) { return types[f * nt + t] == 1 && off[f * nt + t] == 0.0 && scl[f * nt + t] == 0.0f; };
    const int BW = TS + 2;
    std::vector<float> box((size_t)BW * BW), r((size_t)TS * TS);
    for (int f = 0; f < q; ++f)
        for (int t = 0; t < nt; ++t) {
            if (types[f * nt + t] != 1) continue;
            const int tyi = t / tx, txi = t % tx, ox = txi * TS, oy = tyi * TS;
            bool zall = true;
            for (int k = 0; k < 9 && zall; ++k) {
                const int gy = bmap(k / 3 == 0 ? oy - 1 : (k / 3 == 2 ? oy + TS : oy), ny, bc);
                const int gx = bmap(k % 3 == 0 ? ox - 1 : (k % 3 == 2 ? ox + TS : ox), nx, bc);
                zall = zero_code(f, (gy / TS) * tx + gx / TS);
            }
            if (zall) {
                for (int ly = 0; ly < TS; ++ly)
                    for (int lx = 0; lx < TS; ++lx) w_out[f * ncell + (size_t)(oy + ly) * nx + ox + lx] = 0;
                off_out[f * nt + t] = 0.0;
                scl_out[f * nt + t] = 0.0f;
                continue;
            }
            const double own = off[f * nt + t];
            const float sc = scl[f * nt + t];
            for (int by = -1; by <= TS; ++by)
                for (int bx = -1; bx <= TS; ++bx) {
                    const int gy = bmap(oy + by, ny, bc), gx = bmap(ox + bx, nx, bc);
                    const size_t c = (size_t)gy * nx + gx;
                    const int tn = (gy / TS) * tx + gx / TS;
                    float x;
                    if (tn == t) x = h2f(w[f * ncell + c]) * sc;
                    else if (types[f * nt + tn] == 2) x = (float)(v[f * ncell + c] - own);
                    else x = (float)(off[f * nt + tn] - own) + h2f(w[f * ncell + c]) * scl[f * nt + tn];
                    box[(size_t)(by + 1) * BW + bx + 1] = x;
                }
            float mn = INFINITY, mx = -INFINITY;
            const float offf = (float)own;
            for (int ly = 0; ly < TS; ++ly)
                for (int lx = 0; lx < TS; ++lx) {
                    const float* u = &box[(size_t)ly * BW + lx];
                    const float* m = u + BW;
                    const float* d = m + BW;
                    float lap;
                    if (moore) {
                        const float c0 = (u[0] + m[0]) + d[0], c1 = (u[1] + m[1]) + d[1], c2 = (u[2] + m[2]) + d[2];
                        lap = std::fmaf(-9.0f, m[1], (c0 + c1) + c2) * invf;
                    } else {
                        lap = std::fmaf(-4.0f, m[1], ((m[0] + m[2]) + u[1]) + d[1]) * invf;
                    }
                    const float dv = m[1], eta = offf + dv;
                    const float others = std::fmaf(-eta, eta, (float)S[(size_t)(oy + ly) * nx + ox + lx]);
                    const float tt = std::fmaf(bef, eta * eta, -alf);
                    const float df = std::fmaf(g2f * eta, others, tt * eta);
                    const float rn = std::fmaf(-dtlf, std::fmaf(-kapf, lap, df), dv);
                    r[(size_t)ly * TS + lx] = rn;
                    mn = std::fmin(mn, rn);
                    mx = std::fmax(mx, rn);
                }
            const Code cd = tile_code(mn, mx);
            for (int ly = 0; ly < TS; ++ly)
                for (int lx = 0; lx < TS; ++lx)
                    w_out[f * ncell + (size_t)(oy + ly) * nx + ox + lx] = f2h((r[(size_t)ly * TS + lx] - cd.mid) * cd.inv);
            off_out[f * nt + t] = own + (double)cd.mid;
            scl_out[f * nt + t] = cd.scl;
        }
}

// fp16 word -> double, for the tests' decoding of emulated tiles.
double oracle_h2d(uint16_t h) { return (double)h2f(h); }
double oracle_rh(double x) { return rh(x); }
void oracle_atan_h2(const float* x, float* r, int n) {
    for (int i = 0; i < n; ++i) r[i] = atan_h2(x[i]);
}
void oracle_aniso6_h2(const float* u, const float* v, float cth0, float sth0, float* ca, float* sa, int n) {
    for (int i = 0; i < n; ++i) aniso6_h2(u[i], v[i], cth0, sth0, ca[i], sa[i]);
}
uint16_t oracle_f2h(float x) { return f2h(x); }

}  // extern "C"

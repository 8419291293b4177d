// k_snowflake.hip — Kobayashi snowflake (phi, T) explicit-Euler step for gfx950.
//
// One fused sweep per step replaces the reference's kernel1/kernel2 pair and
// its six intermediate arrays (snowflake/function.h:5-60): a workgroup stages
// phi and T of its tile plus a 2-cell halo in LDS, evaluates the anisotropy
// (eps, eps') once per cell of the tile and its 1-cell ring, and writes only
// the new phi and T.  Algorithmic traffic: 2 fields x (8 B read + 8 B write)
// = 32 B per FP64 cell-update, 8 B per fp16 cell-update.
//
// Anisotropy without trigonometry: for integer aniso n,
//   cos(n theta) + i sin(n theta) = ((gx + i gy) / |g|)^n,   theta = atan2(gy, gx),
// then rotated by n*theta0 with host-computed cos/sin.  theta = atan2(0,0) = 0
// at |g| = 0 as in the reference.  The reference evaluates
// eps = epsb(1 + delta cos(aniso(theta - theta0))),
// eps' = -epsb aniso delta sin(aniso(theta - theta0)) (function.h:32-34).
//
// Tiles of type 2 run in FP64 from FP64 storage.  Tiles of type 1 hold fp16
// deviations from a per-tile FP64 offset (see tile_common.hpp), compute in
// FP32 on deviations from their own offset, and re-centre offset and scale
// after every step from the block min/max of the new values.  Halo cells are
// read from the storage their tile's type says is current; both candidate
// values are fetched in the first load round (with the neighbours' type,
// offset and scale), so a tile needs a single round of memory latency.
//
// Data movement: every thread owns R = TS*TS/256 horizontally adjacent cells
// of one row; it loads them with one vector load per field (FP64: 32 B,
// fp16: 8 B), keeps them in registers for the update and writes the results
// with one vector store per field.  The 2-cell halo frame is loaded element
// by element, each from the storage its tile's type says is current.  LDS
// rows are padded so that the update's neighbour reads are 16-byte aligned.
#include "kernels.hpp"
#include "tile_common.hpp"
#include "sf_math.hpp"

#include <type_traits>

namespace amps {

// Geometry of a snowflake tile of TS x TS cells updated by NT threads.
template <int TS, int NT = WG>
struct SfGeo {
    static constexpr int H = 2;
    static constexpr int LW = TS + 2 * H;          // box row stride (16-byte multiple for float and double)
    static constexpr int R = TS * TS / NT;         // cells per thread
    static constexpr int TPR = TS / R;             // threads per row
    static constexpr int RW = TS + 2;              // ring box rows/cols
    // Ring arrays: ring cell (ry, rx) at RB + ry*PW + rx - 1, so that the ring
    // cell of tile cell (ly, x) sits at RB + (ly+1)*PW + x: 16-byte aligned for
    // x % 4 == 0.  Rows do not overlap because RW - 1 < PW.
    static constexpr int PW = TS + 4;
    static constexpr int RB = 4;
    static constexpr int RSZ = RB + RW * PW;
    static constexpr int NH = LW * LW - TS * TS;   // halo frame elements
    static constexpr int NHT = (NH + NT - 1) / NT;
};

template <int TS, typename Rt>
struct SfLds {
    using G = SfGeo<TS>;
    Rt phi[G::LW * G::LW], T[G::LW * G::LW];
    Rt P[G::RSZ], Q[G::RSZ], E2[G::RSZ];
    NbField nbf[2];
    float red[8];
};

// Box coordinates of halo-frame element h (top band, bottom band, then the
// left and right bands row by row).
template <int TS>
__device__ inline void halo_coord(int h, int& ly, int& lx) {
    constexpr int H = SfGeo<TS>::H, LW = SfGeo<TS>::LW;
    if (h < H * LW) { ly = h / LW; lx = h - ly * LW; return; }
    h -= H * LW;
    if (h < H * LW) { ly = H + TS + h / LW; lx = h % LW; return; }
    h -= H * LW;
    ly = H + h / (2 * H);
    const int j = h % (2 * H);
    lx = j < H ? j : TS + j;
}

// Ring coordinate used for the intermediates of ring cell r (box-relative).
// Periodic: the box already holds the wrapped values, so r itself.  Zero
// flux: a ring cell outside the domain takes the intermediates of its mirror
// image inside the tile, as the oracle's index map does.
__device__ inline int ring_src(int r, int o, int n, int bc) {
    if (bc == AMPS_BC_PERIODIC) return r;
    int g = o - 1 + r;
    if (g < 0) g = -g;
    else if (g >= n) g = 2 * (n - 1) - g;
    return g - o + 1;
}

// Phase 1 on a box of R-typed values: P = eps eps' gx, Q = eps eps' gy, E2 = eps^2
// over the (TS+2)^2 ring box (stored as described at SfGeo::PW).
//
// fast: index the ring directly where no ring cell needs the zero-flux
// mirror (a tile off the domain edge, or periodic boundaries).
template <int TS, typename Rt, int NT = WG>
__device__ inline void sf_phase1(const Rt* phi, Rt* P, Rt* Q, Rt* E2, const InstDesc& d, int ox, int oy,
                                 Rt inv2dx, Rt inv2dy, Rt cth0, Rt sth0, Rt epsb, Rt epsbd, Rt mend,
                                 bool fast = false) {
    using G = SfGeo<TS>;
    const int n = (int)d.k[SFK_ANISO_N];
    const bool edge = d.bc != AMPS_BC_PERIODIC &&
                      (ox == 0 || oy == 0 || ox + TS == d.nx || oy + TS == d.ny);
    if (fast && !edge) {
        // box index (ry0 + 1) LW + rx0 + 1 = e + 2 ry0 + LW + 1, ring index
        // RB + ry0 PW + rx0 - 1 = e + 2 ry0 + RB - 1 (LW = PW = RW + 2)
        static_assert(G::LW == G::RW + 2 && G::PW == G::RW + 2, "ring and box rows differ by 2");
        for (int e = threadIdx.x; e < G::RW * G::RW; e += NT) {
            const int ry0 = e / G::RW;
            const int c = e + 2 * ry0 + G::LW + 1, o = e + 2 * ry0 + G::RB - 1;
            Rt gx = (phi[c + 1] - phi[c - 1]) * inv2dx;
            Rt gy = (phi[c + G::LW] - phi[c - G::LW]) * inv2dy;
            Rt eps, epsd;
            sf_aniso<Rt>(gx, gy, n, cth0, sth0, epsb, epsbd, mend, eps, epsd);
            Rt A = eps * epsd;
            P[o] = A * gx;
            Q[o] = A * gy;
            E2[o] = eps * eps;
        }
        return;
    }
    for (int e = threadIdx.x; e < G::RW * G::RW; e += NT) {
        const int ry0 = e / G::RW, rx0 = e - ry0 * G::RW;
        int ry = ry0, rx = rx0;
        if (edge) {
            ry = ring_src(ry, oy, d.ny, d.bc);
            rx = ring_src(rx, ox, d.nx, d.bc);
        }
        const int c = (ry + 1) * G::LW + rx + 1;
        Rt gx = (phi[c + 1] - phi[c - 1]) * inv2dx;
        Rt gy = (phi[c + G::LW] - phi[c - G::LW]) * inv2dy;
        Rt eps, epsd;
        sf_aniso<Rt>(gx, gy, n, cth0, sth0, epsb, epsbd, mend, eps, epsd);
        Rt A = eps * epsd;
        const int o = G::RB + ry0 * G::PW + rx0 - 1;
        P[o] = A * gx;
        Q[o] = A * gy;
        E2[o] = eps * eps;
    }
}

// Laplacian of the R cells (row ly, columns x0..x0+R-1) of a box field;
// also returns the cells' own values.  Reads are 16-byte aligned rows.
// Multiply-adds of the FP32 low path are written as fmaf: one rounding at
// every site, whatever the compiler would fuse, so that the paths that
// update the same tile (the flat sweep's and the full low update) give the
// same bits.
// T' = t + dt lap T + kappa dphi (FP32: fmaf(kappa, dphi, fmaf(dt, lap, t))).
template <typename Rt>
__device__ inline Rt sf_t_update(Rt t, Rt lt, Rt dt, Rt kap, Rt dphi) {
    if constexpr (std::is_same<Rt, float>::value) return fmaf(kap, dphi, fmaf(dt, lt, t));
    else return t + dt * lt + kap * dphi;
}

// A neighbour's fp16 word as a deviation from this tile's offset:
// (float)(its offset - ours) + word * its scale.  The conversion is kept a
// separate instruction (the empty asm pins its result in a VGPR), so every
// site rounds with the same v_cvt_f32_f16 + v_fma_f32 pair rather than a
// v_fma_mix_f32 at some sites and the pair at others.
__device__ inline float sf_decode_rel(double doff, unsigned short w, float scl) {
    float h = __half2float(__ushort_as_half(w));
    asm volatile("" : "+v"(h));
    return fmaf(h, scl, (float)doff);
}

template <int TS, typename Rt, int NT = WG>
__device__ inline void sf_lap(const Rt* F, int ly, int x0, bool ref, Rt inv_dxdy, Rt* own, Rt* lap) {
    using Gm = SfGeo<TS, NT>;
    constexpr int H = Gm::H, LW = Gm::LW, R = Gm::R;
    const int b = (ly + H) * LW + x0;
    Rt row[R + 4];
#pragma unroll
    for (int j = 0; j < R + 4; ++j) row[j] = F[b + j];
#pragma unroll
    for (int i = 0; i < R; ++i) own[i] = row[i + 2];
    if (ref) {
        Rt u1[R], u2[R], d1[R], d2[R];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            u1[j] = F[b - LW + H + j]; u2[j] = F[b - 2 * LW + H + j];
            d1[j] = F[b + LW + H + j]; d2[j] = F[b + 2 * LW + H + j];
        }
        // order of snowflake/function.h:27-28: xs1, xs2, xa1, xa2, ys1, ys2, ya1, ya2
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const Rt sum = row[i + 1] + row[i] + row[i + 3] + row[i + 4] + u1[i] + u2[i] + d1[i] + d2[i];
            if constexpr (std::is_same<Rt, float>::value) lap[i] = fmaf(-8.0f, row[i + 2], sum) * inv_dxdy;
            else lap[i] = (sum - Rt(8) * row[i + 2]) * inv_dxdy;
        }
    } else {
        Rt u1[R], d1[R];
#pragma unroll
        for (int j = 0; j < R; ++j) { u1[j] = F[b - LW + H + j]; d1[j] = F[b + LW + H + j]; }
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const Rt sum = row[i + 1] + row[i + 3] + u1[i] + d1[i];
            if constexpr (std::is_same<Rt, float>::value) lap[i] = fmaf(-4.0f, row[i + 2], sum) * inv_dxdy;
            else lap[i] = (sum - Rt(4) * row[i + 2]) * inv_dxdy;
        }
    }
}

// ------------------------------------------------------------ packed half2
// Pointwise terms of the low tiles in packed half2 (AMPS_FLAG_HALF2): each
// operation below is one v_pk_* instruction on two cells, rounded once to
// fp16 (round to nearest even); `fp contract(off)` keeps every written
// product and sum a rounding of its own, and the fused multiply-adds are the
// explicit __builtin_elementwise_fma calls.  Inputs arrive in FP32 and
// results leave in FP32.
typedef _Float16 sf_h2 __attribute__((ext_vector_type(2)));

__device__ inline sf_h2 h2_of(float a, float b) { return sf_h2{(_Float16)a, (_Float16)b}; }
__device__ inline sf_h2 h2_splat(float a) { return sf_h2{(_Float16)a, (_Float16)a}; }
__device__ inline sf_h2 h2_fma(sf_h2 a, sf_h2 b, sf_h2 c) { return __builtin_elementwise_fma(a, b, c); }

// (cos n(theta - theta0), sin n(theta - theta0)) for n = 6 of two ring cells
// from their exactly rescaled gradients (u, v) (max(|u|, |v|) in [0.5, 1), or
// u = 1, v = 0 for a zero gradient): w = (u + i v)^2 = A + i B, the unit
// vector w^3 / |w|^3 = (A (A^2 - 3 B^2), B (3 A^2 - B^2)) / R^3 with R = |w|,
// then the rotation by 6 theta0 (host cos/sin).
__device__ inline void aniso6_h2(float u0, float v0, float u1, float v1, float cth0, float sth0, float& ca0,
                                 float& sa0, float& ca1, float& sa1) {
#pragma clang fp contract(off)
    const sf_h2 U = h2_of(u0, u1), V = h2_of(v0, v1);
    const sf_h2 UU = U * U, VV = V * V;
    const sf_h2 R2 = UU + VV, A = UU - VV;
    const sf_h2 B = (U + U) * V;
    const sf_h2 AA = A * A, BB = B * B;
    const sf_h2 Cn = A * h2_fma(BB, h2_splat(-3.0f), AA);
    const sf_h2 Sn = B * h2_fma(AA, h2_splat(3.0f), -BB);
    const sf_h2 inv = h2_splat(1.0f) / ((R2 * R2) * R2);
    const sf_h2 C = Cn * inv, S = Sn * inv;
    const sf_h2 c0 = h2_splat(cth0), s0 = h2_splat(sth0);
    const sf_h2 CA = h2_fma(S, s0, C * c0);
    const sf_h2 SA = h2_fma(S, c0, -(C * s0));
    ca0 = (float)CA.x; ca1 = (float)CA.y;
    sa0 = (float)SA.x; sa1 = (float)SA.y;
}

// atan of two arguments: z = min(|x|, 1/|x|), p = z (c0 + c1 z^2 + c2 z^4 +
// c3 z^6) (least-squares fit of atan on [0, 1], 8.7e-5 in exact arithmetic;
// 1.5e-3 absolute with the fp16 roundings over |x| <= 20), pi/2 - p where
// |x| > 1 (f = clamp(2048 (|x| - 1), 0, 1) selects it: |x| > 1 in fp16 means
// |x| - 1 >= 2^-10), and the sign of x.
constexpr float SF_ATAN_H2_C0 = 0.99926763f, SF_ATAN_H2_C1 = -0.32142954f, SF_ATAN_H2_C2 = 0.14661292f,
                SF_ATAN_H2_C3 = -0.03913253f;

__device__ inline void atan_h2(float x0, float x1, float& r0, float& r1) {
#pragma clang fp contract(off)
    const sf_h2 X = h2_of(x0, x1);
    // |X| by its bits: the reciprocal then reads these fp16 values (an
    // elementwise abs lets the compiler re-round the FP32 product into it)
    const sf_h2 AX = __builtin_bit_cast(sf_h2, __builtin_bit_cast(unsigned, X) & 0x7fff7fffu);
    const sf_h2 Z = __builtin_elementwise_min(AX, h2_splat(1.0f) / AX);
    const sf_h2 Z2 = Z * Z;
    sf_h2 p = h2_fma(Z2, h2_splat(SF_ATAN_H2_C3), h2_splat(SF_ATAN_H2_C2));
    p = h2_fma(Z2, p, h2_splat(SF_ATAN_H2_C1));
    p = h2_fma(Z2, p, h2_splat(SF_ATAN_H2_C0));
    p = p * Z;
    const sf_h2 f = __builtin_elementwise_min(
        __builtin_elementwise_max(h2_fma(AX, h2_splat(2048.0f), h2_splat(-2048.0f)), h2_splat(0.0f)), h2_splat(1.0f));
    const sf_h2 r = h2_fma(f, h2_fma(p, h2_splat(-2.0f), h2_splat(1.5707963705062866f)), p);
    // r >= 0: the sign bits of x complete it
    const sf_h2 rs = __builtin_bit_cast(
        sf_h2, __builtin_bit_cast(unsigned, r) | (__builtin_bit_cast(unsigned, X) & 0x80008000u));
    r0 = (float)rs.x;
    r1 = (float)rs.y;
}

// Phase 1 of the low tiles with the six-fold unit vector and its rotation in
// packed half2: thread t pairs ring cells e and e + RW^2/2 (e = t + k WG), so
// that a wave's LDS reads and writes stay on consecutive words for both
// (pairing neighbouring cells of a row strides them by two words: 2-way bank
// conflicts); the gradients, the exact power-of-two rescale, eps, eps' and
// P, Q, E2 stay in FP32 as in sf_phase1.  An aniso other than 6 runs
// sf_phase1.
template <int TS>
__device__ inline void sf_phase1_h2(const float* phi, float* P, float* Q, float* E2, const InstDesc& d, int ox, int oy,
                                    float inv2dx, float inv2dy, float cth0, float sth0, float epsb, float epsbd,
                                    float mend) {
    using G = SfGeo<TS>;
    static_assert((G::RW * G::RW) % 2 == 0, "the ring splits into two halves");
    constexpr int HALF = G::RW * G::RW / 2;
    static_assert(G::LW == G::RW + 2 && G::PW == G::RW + 2, "ring and box rows differ by 2");
    if ((int)d.k[SFK_ANISO_N] != 6) {
        sf_phase1<TS, float>(phi, P, Q, E2, d, ox, oy, inv2dx, inv2dy, cth0, sth0, epsb, epsbd, mend, false);
        return;
    }
    const bool edge = d.bc != AMPS_BC_PERIODIC &&
                      (ox == 0 || oy == 0 || ox + TS == d.nx || oy + TS == d.ny);
    for (int e = threadIdx.x; e < HALF; e += WG) {
        int cb[2], ob[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int r = e + k * HALF;
            const int ry0 = r / G::RW, rx0 = r - ry0 * G::RW;
            // box index (ry0 + 1) LW + rx0 + 1 = r + 2 ry0 + LW + 1, ring index
            // RB + ry0 PW + rx0 - 1 = r + 2 ry0 + RB - 1
            cb[k] = edge ? (ring_src(ry0, oy, d.ny, d.bc) + 1) * G::LW + ring_src(rx0, ox, d.nx, d.bc) + 1
                         : r + 2 * ry0 + G::LW + 1;
            ob[k] = r + 2 * ry0 + G::RB - 1;
        }
        float gx[2], gy[2], u[2], v[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            gx[k] = (phi[cb[k] + 1] - phi[cb[k] - 1]) * inv2dx;
            gy[k] = (phi[cb[k] + G::LW] - phi[cb[k] - G::LW]) * inv2dy;
            const float mx = fmaxf(fabsf(gx[k]), fabsf(gy[k]));
            const int ex = frexp_exp(mx);
            u[k] = mx > 0.0f ? ldexp_any(gx[k], -ex) : 1.0f;  // zero gradient: theta = 0
            v[k] = ldexp_any(gy[k], -ex);
        }
        float ca[2], sa[2];
        aniso6_h2(u[0], v[0], u[1], v[1], cth0, sth0, ca[0], sa[0], ca[1], sa[1]);
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const float eps = epsb + epsbd * ca[k], epsd = mend * sa[k];
            const float A = eps * epsd;
            P[ob[k]] = A * gx[k];
            Q[ob[k]] = A * gy[k];
            E2[ob[k]] = eps * eps;
        }
    }
}

// Phase 2 for the R cells (row ly, columns x0 .. x0+R-1) of this thread:
// writes the new values to np/nt (and the old ones to pc/tc).  The noise
// term a*phi(1-phi)(u-1/2) is evaluated only where |phi(1-phi)| > noise_skip:
// 0 on the FP64 path (bitwise the same result), a bound far below the fp16
// resolution on the low path.
template <int TS, typename Rt, bool H2 = false, int NT = WG>
__device__ inline void sf_phase2(const Rt* F, const Rt* G, const Rt* P, const Rt* Q, const Rt* E2, int ly, int x0,
                                 Rt* pc, Rt* tc, Rt* np, Rt* nt, bool ref, Rt inv_dxdy, Rt inv2dx,
                                 Rt inv2dy, Rt dtt, Rt dt, Rt kap, Rt aop, Rt gam, Rt teq_minus_off, Rt phi_off,
                                 Rt noise, Rt noise_skip, uint32_t key, int gcell0) {
    using Gm = SfGeo<TS, NT>;
    constexpr int R = Gm::R, PW = Gm::PW, RB = Gm::RB;
    Rt lp[R], lt[R];
    sf_lap<TS, Rt, NT>(F, ly, x0, ref, inv_dxdy, pc, lp);
    sf_lap<TS, Rt, NT>(G, ly, x0, ref, inv_dxdy, tc, lt);
    // ring arrays: cell (ly, x) is ring (ly+1, x+1), stored at RB + (ly+1)*PW + x
    const int rb = RB + (ly + 1) * PW + x0;
    Rt flux[R];
    {
        Rt pu[R], pd[R], q[R + 2], e2[R];
#pragma unroll
        for (int j = 0; j < R; ++j) { pu[j] = P[rb - PW + j]; pd[j] = P[rb + PW + j]; e2[j] = E2[rb + j]; }
#pragma unroll
        for (int j = 0; j < R + 2; ++j) q[j] = Q[rb - 1 + j];
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const Rt term1 = (pd[i] - pu[i]) * inv2dy;
            const Rt term2 = -(q[i + 2] - q[i]) * inv2dx;
            flux[i] = term1 + term2 + e2[i] * lp[i];
        }
    }
    Rt mv[R];
    if constexpr (H2) {
        static_assert(R % 2 == 0, "half2 driving force pairs the cells of a thread");
#pragma unroll
        for (int i = 0; i < R; i += 2) {
            float r0, r1;
            atan_h2(gam * (teq_minus_off - tc[i]), gam * (teq_minus_off - tc[i + 1]), r0, r1);
            mv[i] = aop * r0;
            mv[i + 1] = aop * r1;
        }
    } else {
#pragma unroll
        for (int i = 0; i < R; ++i) mv[i] = aop * atan(gam * (teq_minus_off - tc[i]));
    }
#pragma unroll
    for (int i = 0; i < R; ++i) {
        const Rt ph = pc[i], t = tc[i];
        const Rt pa = phi_off + ph;
        const Rt m = mv[i];
        const Rt pp = pa * (Rt(1) - pa);
        Rt br = flux[i] + pp * (pa - Rt(0.5) + m);
        if (noise != Rt(0) && fabs(pp) > noise_skip) {
            const Rt u = (Rt)noise_u24(key, (uint32_t)(gcell0 + i));
            br = br + noise * pa * (Rt(1) - pa) * (u - Rt(0.5));
        }
        const Rt phn = ph + dtt * br;
        np[i] = phn;
        nt[i] = sf_t_update<Rt>(t, lt[i], dt, kap, phn - ph);
    }
}

// ------------------------------------------------------------- FP64 tile
// Halo frame of the tile (box coordinates and cell offsets) staged by this
// thread: element h = threadIdx.x + k*WG of the frame.
template <int TS, int NT = WG>
struct SfHalo {
    int off[SfGeo<TS, NT>::NHT];
    int box[SfGeo<TS, NT>::NHT];
    int reg[SfGeo<TS, NT>::NHT];
};

template <int TS, int NT = WG>
__device__ inline void sf_halo_map(SfHalo<TS, NT>& hm, const InstDesc& d, int ox, int oy) {
    using G = SfGeo<TS, NT>;
    constexpr int H = G::H;
    const bool interior = ox >= H && oy >= H && ox + TS + H <= d.nx && oy + TS + H <= d.ny;
#pragma unroll
    for (int k = 0; k < G::NHT; ++k) {
        const int h = threadIdx.x + k * NT;
        int bly = 0, blx = 0;
        if (h < G::NH) halo_coord<TS>(h, bly, blx);
        int gy = oy - H + bly, gx = ox - H + blx;
        if (!interior) { gy = bmap(gy, d.ny, d.bc); gx = bmap(gx, d.nx, d.bc); }
        hm.off[k] = h < G::NH ? gy * d.nx + gx : -1;
        hm.box[k] = bly * G::LW + blx;
        hm.reg[k] = box_region<TS, H>(bly, blx);
    }
}

// Halo elements of the frame as raw words, loaded by one sequence for both
// storages (addresses selected, no branches): a branch per storage makes
// the compiler wait for each group of loads before issuing the next.  An
// element of an FP64 tile is its 8 bytes, one of a low tile the aligned
// 4-word fp16 group holding it (grids are multiples of the tile, so the
// group lies inside its row).  The loads go straight into the registers
// they are decoded from, so the branch around a thread's missing element
// needs no wait at its join.
// Per-launch constants of a thread for the tiles of a one-instance grid that
// need no boundary map: its halo elements relative to the tile's first cell
// (box index, region; region -1: no element), its ring cells for phase 1
// and, for lanes 0..8, the offset of tile neighbour k = lane in the tile
// index.  Computing these per tile costs divisions and boundary maps that
// an interior tile does not need.
template <int TS>
struct SfTileConst {
    using G = SfGeo<TS>;
    static constexpr int NR = (G::RW * G::RW + WG - 1) / WG;
    int hrel[G::NHT], hbox[G::NHT], hreg[G::NHT];
    int nbrel;
};

template <int TS>
__device__ inline void sf_tile_const(SfTileConst<TS>& t, const InstDesc& d) {
    using G = SfGeo<TS>;
    constexpr int H = G::H;
#pragma unroll
    for (int k = 0; k < G::NHT; ++k) {
        const int h = threadIdx.x + k * WG;
        int bly = 0, blx = 0;
        if (h < G::NH) halo_coord<TS>(h, bly, blx);
        t.hrel[k] = (bly - H) * d.nx + (blx - H);
        t.hbox[k] = bly * G::LW + blx;
        t.hreg[k] = h < G::NH ? box_region<TS, H>(bly, blx) : -1;
    }
    const int kk = (threadIdx.x & 63) < 9 ? (int)(threadIdx.x & 63) : 8;
    t.nbrel = (kk / 3 - 1) * d.tx + (kk % 3 - 1);
}

// The halo map of tile (tile_x, tile_y): from the launch constants when the
// tile's halo needs no boundary map, else computed.
template <int TS>
__device__ inline void sf_halo_of(SfHalo<TS>& hm, const InstDesc& d, int tile_x, int tile_y,
                                  const SfTileConst<TS>* tc) {
    using G = SfGeo<TS>;
    constexpr int H = G::H;
    const int ox = tile_x * TS, oy = tile_y * TS;
    const bool interior = ox >= H && oy >= H && ox + TS + H <= d.nx && oy + TS + H <= d.ny;
    if (tc && interior) {
        const int org = oy * d.nx + ox;
#pragma unroll
        for (int k = 0; k < G::NHT; ++k) {
            hm.off[k] = tc->hreg[k] >= 0 ? org + tc->hrel[k] : -1;
            hm.box[k] = tc->hbox[k];
            hm.reg[k] = tc->hreg[k] >= 0 ? tc->hreg[k] : 4;
        }
    } else {
        sf_halo_map<TS>(hm, d, ox, oy);
    }
}

// phi = false: the T elements only (hp untouched).
template <int TS, bool ADAPT, int NT = WG>
__device__ inline void sf_halo_load(uint2* hp, uint2* ht, const StepArgs& a, const InstDesc& d,
                                    const SfHalo<TS, NT>& hm, unsigned hmask, bool phi = true) {
    const double* fphi = a.in.f64 + d.cell_base;
    const double* fT = a.in.f64 + a.total_cells + d.cell_base;
    const unsigned short* hphi = reinterpret_cast<const unsigned short*>(a.in.f16 + d.cell_base);
    const unsigned short* hT = reinterpret_cast<const unsigned short*>(a.in.f16 + a.total_cells + d.cell_base);
#pragma unroll
    for (int k = 0; k < SfGeo<TS, NT>::NHT; ++k) {
        const int off = hm.off[k];
        if (off < 0) continue;
        const bool h64 = !ADAPT || ((hmask >> hm.reg[k]) & 1u);
        if (phi)
            hp[k] = *(h64 ? reinterpret_cast<const uint2*>(fphi + off)
                          : reinterpret_cast<const uint2*>(hphi + (off & ~3)));
        ht[k] = *(h64 ? reinterpret_cast<const uint2*>(fT + off) : reinterpret_cast<const uint2*>(hT + (off & ~3)));
    }
}

template <int TS, bool ADAPT, int NT = WG>
__device__ void sf_tile_f64(SfLds<TS, double>& L, const StepArgs& a, const InstDesc& d, int tile_x, int tile_y,
                            unsigned hmask, long long step) {
    using G = SfGeo<TS, NT>;
    constexpr int H = G::H, LW = G::LW, R = G::R;
    const int ox = tile_x * TS, oy = tile_y * TS;
    const int ly = threadIdx.x / G::TPR, x0 = (threadIdx.x % G::TPR) * R;
    const int gcell0 = (oy + ly) * d.nx + ox + x0;
    const double* fphi = a.in.f64 + d.cell_base;
    const double* fT = a.in.f64 + a.total_cells + d.cell_base;
    // own cells (vector loads) and halo elements, all loads in flight together;
    // a halo element is read from the storage its tile's type says is current
    double pc[R], tc[R];
    if (R % 2 == 0) {
#pragma unroll
        for (int j = 0; j < R; j += 2) {
            const double2 u = *reinterpret_cast<const double2*>(fphi + gcell0 + j);
            const double2 v = *reinterpret_cast<const double2*>(fT + gcell0 + j);
            pc[j] = u.x; pc[j + 1] = u.y; tc[j] = v.x; tc[j + 1] = v.y;
        }
    } else {
#pragma unroll
        for (int j = 0; j < R; ++j) { pc[j] = fphi[gcell0 + j]; tc[j] = fT[gcell0 + j]; }
    }
    SfHalo<TS, NT> hm;
    sf_halo_map<TS, NT>(hm, d, ox, oy);
    // halo elements as raw words, one load sequence for both storages (see
    // sf_halo_load): FP64 bits or the aligned 4-word fp16 group
    uint2 hp[G::NHT], ht[G::NHT];
    sf_halo_load<TS, ADAPT, NT>(hp, ht, a, d, hm, hmask);
    if (ADAPT) {
        Nb nb;
        nb_setup(nb, d, tile_x, tile_y, TS);
        nbfield_fill2(&L.nbf[0], &L.nbf[1], a, a.in, nb, 0, 1, 0);
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
        L.phi[(ly + H) * LW + x0 + H + j] = pc[j];
        L.T[(ly + H) * LW + x0 + H + j] = tc[j];
    }
#pragma unroll
    for (int k = 0; k < G::NHT; ++k) {
        if (hm.off[k] < 0) continue;
        const int r = hm.reg[k];
        if (!ADAPT || ((hmask >> r) & 1u)) {
            L.phi[hm.box[k]] = raw_f64(hp[k].x, hp[k].y);
            L.T[hm.box[k]] = raw_f64(ht[k].x, ht[k].y);
        } else {
            const int off = hm.off[k];
            L.phi[hm.box[k]] =
                L.nbf[0].off[r] + (double)__half2float(__ushort_as_half((unsigned short)raw_word4(hp[k], off))) * (double)L.nbf[0].scl[r];
            L.T[hm.box[k]] =
                L.nbf[1].off[r] + (double)__half2float(__ushort_as_half((unsigned short)raw_word4(ht[k], off))) * (double)L.nbf[1].scl[r];
        }
    }
    __syncthreads();

    sf_phase1<TS, double, NT>(L.phi, L.P, L.Q, L.E2, d, ox, oy, d.k[SFK_INV_2DX], d.k[SFK_INV_2DY],
                          d.k[SFK_COS_N_THETA0], d.k[SFK_SIN_N_THETA0], d.k[SFK_EPSB], d.k[SFK_EPSB_DELTA],
                          d.k[SFK_MEPSB_ANISO_DELTA]);
    __syncthreads();

    double np[R], nt[R];
    sf_phase2<TS, double, false, NT>(L.phi, L.T, L.P, L.Q, L.E2, ly, x0, pc, tc, np, nt,
                          d.stencil == AMPS_STENCIL_REF, d.k[SFK_INV_DXDY], d.k[SFK_INV_2DX], d.k[SFK_INV_2DY],
                          d.k[SFK_DT_OVER_TAU], d.k[SFK_DT], d.k[SFK_KAPPA], d.k[SFK_ALPHA_OVER_PI], d.k[SFK_GAMMA],
                          d.k[SFK_TEQ], 0.0, d.k[SFK_NOISE], 0.0, noise_key(d.seed, (uint64_t)step), gcell0);
    double* ophi = a.out.f64 + d.cell_base;
    double* oT = a.out.f64 + a.total_cells + d.cell_base;
    if (R % 2 == 0) {
#pragma unroll
        for (int j = 0; j < R; j += 2) {
            *reinterpret_cast<double2*>(ophi + gcell0 + j) = make_double2(np[j], np[j + 1]);
            *reinterpret_cast<double2*>(oT + gcell0 + j) = make_double2(nt[j], nt[j + 1]);
        }
    } else {
#pragma unroll
        for (int j = 0; j < R; ++j) { ophi[gcell0 + j] = np[j]; oT[gcell0 + j] = nt[j]; }
    }
}

// ------------------------------------------------------------- low tile
// Bulk-phase hold: when every phi value of a low tile's box lies within
// SF_BULK_EPS of the same stable bulk value (0 = liquid or 1 = solid; both
// are stable fixed points of phi(1-phi)(phi-1/2+m) since |m| < alpha/2 < 1/2),
// phi is held and only T advances (T' = T + dt lap T): the omitted phi change
// is below ~SF_BULK_EPS in total.  The noise term is skipped where
// |phi(1-phi)| <= SF_NOISE_SKIP_LOW (its per-step contribution is then below
// 2e-9).
constexpr float SF_BULK_EPS = 1.0e-8f;
constexpr int SAME_PHI = 1, SAME_T = 2;  // same_flags bits: both buffers hold the same phi / T
constexpr float SF_NOISE_SKIP_LOW = 1.0e-6f;

// Unpack R (= 4) fp16 words of an 8-byte load.
__device__ inline float half_of(const uint2 w, int j) {
    const unsigned v = j < 2 ? w.x : w.y;
    return __half2float(__ushort_as_half((unsigned short)((j & 1) ? (v >> 16) : (v & 0xFFFFu))));
}

// Encode R new deviations into fp16 words.
template <int R>
__device__ inline void sf_encode(const float* r, const TileCode& tc, unsigned short* w) {
#pragma unroll
    for (int i = 0; i < R; ++i) w[i] = __half_as_ushort(__float2half_rn((r[i] - tc.mid) * tc.inv));
}

template <int R>
__device__ inline void sf_store_low(__half* dst16, int gcell0, const unsigned short* w) {
    if (R == 4) {
        *reinterpret_cast<uint2*>(dst16 + gcell0) =
            make_uint2((unsigned)w[0] | ((unsigned)w[1] << 16), (unsigned)w[2] | ((unsigned)w[3] << 16));
    } else {
#pragma unroll
        for (int i = 0; i < R; ++i) dst16[gcell0 + i] = __ushort_as_half(w[i]);
    }
}

// Everything a low tile needs from memory, fetched in one round: own fp16
// words, the halo elements (each from the storage its tile's type in `hmask`
// says is current) and the offsets/scales of the tile and its neighbours.
template <int TS>
struct LowFetch {
    using G = SfGeo<TS>;
    unsigned short op[G::R], ot[G::R];              // own fp16 words (phi, T)
    uint2 hp[G::NHT], ht[G::NHT];                   // halo elements: FP64 bits or the fp16 word's 4-word group
    double offp, offT;
    float sclp, sclT;
};

// tc (may be null): launch constants of a one-instance grid.
template <int TS>
__device__ inline void low_fetch(LowFetch<TS>& f, const StepArgs& a, const InstDesc& d, int w, unsigned hmask,
                                 const SfTileConst<TS>* tc = nullptr) {
    using G = SfGeo<TS>;
    constexpr int R = G::R;
    const int lt = w - d.tile_base;
    const int tile_y = lt / d.tx, tile_x = lt - tile_y * d.tx;
    const int ox = tile_x * TS, oy = tile_y * TS;
    const int ly = threadIdx.x / G::TPR, x0 = (threadIdx.x % G::TPR) * R;
    const int gcell0 = (oy + ly) * d.nx + ox + x0;
    const unsigned short* hphi = reinterpret_cast<const unsigned short*>(a.in.f16 + d.cell_base);
    const unsigned short* hT = reinterpret_cast<const unsigned short*>(a.in.f16 + a.total_cells + d.cell_base);
    if (R == 4) {
        const uint2 u = *reinterpret_cast<const uint2*>(hphi + gcell0);
        const uint2 v = *reinterpret_cast<const uint2*>(hT + gcell0);
        f.op[0] = u.x & 0xFFFFu; f.op[1 % R] = u.x >> 16; f.op[2 % R] = u.y & 0xFFFFu; f.op[3 % R] = u.y >> 16;
        f.ot[0] = v.x & 0xFFFFu; f.ot[1 % R] = v.x >> 16; f.ot[2 % R] = v.y & 0xFFFFu; f.ot[3 % R] = v.y >> 16;
    } else {
#pragma unroll
        for (int j = 0; j < R; ++j) { f.op[j] = hphi[gcell0 + j]; f.ot[j] = hT[gcell0 + j]; }
    }
    SfHalo<TS> hm;
    sf_halo_of<TS>(hm, d, tile_x, tile_y, tc);
    sf_halo_load<TS, true>(f.hp, f.ht, a, d, hm, hmask);
    f.offp = a.in.off[w];
    f.offT = a.in.off[a.total_tiles + w];
    f.sclp = a.in.scl[w];
    f.sclT = a.in.scl[a.total_tiles + w];
}

// LDS of the low sweep: the float boxes, the neighbour offset/scale table
// (two slots: the next tile's is written while the current tile computes),
// the per-wave bulk-hold flags and the min/max reduction slots.
template <int TS>
struct SfLowLds {
    using G = SfGeo<TS>;
    float phi[G::LW * G::LW], T[G::LW * G::LW];
    float P[G::RSZ], Q[G::RSZ], E2[G::RSZ];
    double nbo[2][2][9];  // [slot][field][neighbour]
    float nbs[2][2][9];
    unsigned hflag[WG / 64];
    float red[4 * WG / 64];
};

// Offsets and scales of neighbour tile k = threadIdx.x (threads 0..8) of
// tile w, both fields.
struct NbCode {
    double o0, o1;
    float s0, s1;
};

// Tile index of neighbour k = min(lane, 8) of tile w: w + the launch
// constant for a tile off the tile grid's border (tc given), else mapped.
template <int TS>
__device__ inline int nb_tile_of(const InstDesc& d, int w, const SfTileConst<TS>* tc) {
    const int lt = w - d.tile_base;
    const int tile_y = lt / d.tx, tile_x = lt - tile_y * d.tx;
    if (tc && tile_x > 0 && tile_y > 0 && tile_x < d.tx - 1 && tile_y < d.ty - 1) return w + tc->nbrel;
    Nb nb;
    nb_setup(nb, d, tile_x, tile_y, TS);
    const int l = threadIdx.x & 63;
    return nb_tile(nb, l < 9 ? l : 8);
}

template <int TS>
__device__ inline void nb_code_fetch(NbCode& c, const StepArgs& a, const InstDesc& d, int w,
                                     const SfTileConst<TS>* tc = nullptr) {
    // every thread loads (threads 9.. repeat neighbour 8, cache hits): no
    // branch around the loads, so their wait lands where they are used
    // (threads 0..8 are lanes 0..8 of wave 0)
    const int g = nb_tile_of<TS>(d, w, tc);
    c.o0 = a.in.off[g];
    c.o1 = a.in.off[a.total_tiles + g];
    c.s0 = a.in.scl[g];
    c.s1 = a.in.scl[a.total_tiles + g];
}

template <int TS>
__device__ inline void nb_code_store(SfLowLds<TS>& L, int slot, const NbCode& c) {
    if (threadIdx.x < 9) {
        L.nbo[slot][0][threadIdx.x] = c.o0;
        L.nbo[slot][1][threadIdx.x] = c.o1;
        L.nbs[slot][0][threadIdx.x] = c.s0;
        L.nbs[slot][1][threadIdx.x] = c.s1;
    }
}

// How flat a low tile's neighbourhood is, from the codes of its 9 tiles
// (lane k of every wave holds tile k, k = 4 the tile itself); uniform over
// the workgroup.  0: none; 1: the phi box decodes to exact zeros (the 9
// tiles low, phi scale 0, one phi offset); 2: and that offset is 0 or 1 as
// a float; 3: and the T box decodes to exact zeros too (T scale 0, one T
// offset).
__device__ inline int low_flat_level(const NbCode& c, unsigned hmask) {
    if (hmask != 0u) return 0;
    const int l = threadIdx.x & 63;
    const double op = __shfl(c.o0, 4), oT = __shfl(c.o1, 4);
    const bool fp = l >= 9 || (c.s0 == 0.0f && c.o0 == op);
    const bool fT = l >= 9 || (c.s1 == 0.0f && c.o1 == oT);
    if (__ballot(!fp) != 0ull) return 0;
    const float po = (float)op;
    if (po != 0.0f && po != 1.0f) return 1;
    return __ballot(!fT) == 0ull ? 3 : 2;
}

// A level-3 tile: the update of a level-2 tile (sf_tile_low) with the T
// Laplacian, T and phi all signed zeros gives T' = (+-0) + kap (+0 - ph) =
// +0 and phi' = +0 in every cell, so the new words and codes follow from
// the tile's offsets alone (those of tile_code(+0, +0), as the min/max of
// the new values would give): no word is read.
// NT threads cover the cells of the WG threads of sf_tile_low (virtual
// thread v = threadIdx.x + k NT).
template <int TS, int NT = WG>
__device__ inline void sf_tile_zero(const StepArgs& a, const InstDesc& d, int gtile, double offp, double offT) {
    using G = SfGeo<TS>;
    constexpr int R = G::R;
    static_assert(WG % NT == 0, "whole passes");
    const int lt = gtile - d.tile_base;
    const int tile_y = lt / d.tx, tile_x = lt - tile_y * d.tx;
    const TileCode z = tile_code(0.0f, 0.0f);
    float zr[R];
#pragma unroll
    for (int i = 0; i < R; ++i) zr[i] = 0.0f;
    unsigned short w[R];
    sf_encode<R>(zr, z, w);
#pragma unroll
    for (int k = 0; k < WG / NT; ++k) {
        const int v = threadIdx.x + k * NT;
        const int ly = v / G::TPR, x0 = (v % G::TPR) * R;
        const int gcell0 = (tile_y * TS + ly) * d.nx + tile_x * TS + x0;
        sf_store_low<R>(a.out.f16 + d.cell_base, gcell0, w);
        sf_store_low<R>(a.out.f16 + a.total_cells + d.cell_base, gcell0, w);
    }
    if (threadIdx.x == 0) {
        a.out.off[gtile] = offp + (double)z.mid;
        a.out.scl[gtile] = z.scl;
        a.out.off[a.total_tiles + gtile] = offT + (double)z.mid;
        a.out.scl[a.total_tiles + gtile] = z.scl;
        a.same_flags[gtile] = 0;
    }
}

// T' of a cell of a level-2 tile (sf_flat_kernel): the update of sf_phase2
// with phi' = +0.
__device__ inline float sf_t_flat(float t, float lt, float dt, float kap, float ph) {
    const float phn = 0.0f;
    return sf_t_update<float>(t, lt, dt, kap, phn - ph);
}

// Phase 1 of a low tile whose phi box decodes to exact zeros (the tile and
// its 8 neighbours low, phi scale 0, every offset the tile's own): each ring
// gradient is exactly +0, so each ring cell takes the zero-gradient values
// (theta = 0), computed once with the operations of the loop (bitwise the
// values it would write, signs of zero included).  Exact, not a skip.
template <int TS, bool H2>
__device__ inline void sf_phase1_flat(float* P, float* Q, float* E2, const InstDesc& d, float cth0, float sth0,
                                      float epsb, float epsbd, float mend) {
    using G = SfGeo<TS>;
    const int n = (int)d.k[SFK_ANISO_N];
    float eps, epsd;
    if (H2 && n == 6) {
        float ca, sa, ca1, sa1;
        aniso6_h2(1.0f, 0.0f, 1.0f, 0.0f, cth0, sth0, ca, sa, ca1, sa1);
        eps = epsb + epsbd * ca;
        epsd = mend * sa;
    } else {
        sf_aniso<float>(0.0f, 0.0f, n, cth0, sth0, epsb, epsbd, mend, eps, epsd);
    }
    const float A = eps * epsd;
    const float pq = A * 0.0f, e2 = eps * eps;
    for (int e = threadIdx.x; e < G::RW * G::RW; e += WG) {
        const int ry0 = e / G::RW;
        const int o = e + 2 * ry0 + G::RB - 1;  // RB + ry0 PW + rx0 - 1 (PW = RW + 2)
        P[o] = pq;
        Q[o] = pq;
        E2[o] = e2;
    }
}

// One low tile.  Barriers: A (box staged, bulk-hold flags), B (phase 1
// done; interface tiles only) and C (min/max of the new values).  The
// neighbour table of the next tile (`nxt`) is stored into slot `slot ^ 1`
// after A.  There is no barrier at the end of a tile: a thread that starts
// staging the next tile has passed C, so every read of this tile's boxes is
// done; the next tile's writes of the flags, the reduction slots and P/Q/E2
// all come after its own barrier A, which every thread reaches only after
// its last read of this tile's values.
// NOSKIP: no bulk hold and no noise skip (every term of the update is
// evaluated, as the reference's type-1 branch does); flat: the tile's
// low_flat_level (0 or 1: level 1 takes sf_phase1_flat).
template <int TS, bool NOSKIP, bool H2 = false>
__device__ void sf_tile_low(SfLowLds<TS>& L, int slot, const NbCode& nxt, const StepArgs& a, const InstDesc& d,
                            int gtile, const LowFetch<TS>& f, unsigned hmask, long long step,
                            const SfTileConst<TS>* tcon = nullptr, int flat = 0) {
    using G = SfGeo<TS>;
    constexpr int H = G::H, LW = G::LW, R = G::R;
    const int lt = gtile - d.tile_base;
    const int tile_y = lt / d.tx, tile_x = lt - tile_y * d.tx;
    const int ox = tile_x * TS, oy = tile_y * TS;
    const int ly = threadIdx.x / G::TPR, x0 = (threadIdx.x % G::TPR) * R;
    const int gcell0 = (oy + ly) * d.nx + ox + x0;
    const double offp = f.offp, offT = f.offT;
    const float sclp = f.sclp, sclT = f.sclT;
    float pc[R], tc[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        pc[j] = __half2float(__ushort_as_half(f.op[j])) * sclp;
        tc[j] = __half2float(__ushort_as_half(f.ot[j])) * sclT;
    }
    SfHalo<TS> hm;
    sf_halo_of<TS>(hm, d, tile_x, tile_y, tcon);
    const float c0 = (float)offp, c1 = (float)(1.0 - offp);
    bool near0 = true, near1 = true;
#pragma unroll
    for (int j = 0; j < R; ++j) {
        L.phi[(ly + H) * LW + x0 + H + j] = pc[j];
        L.T[(ly + H) * LW + x0 + H + j] = tc[j];
        if (!NOSKIP) {
            near0 = near0 && fabsf(c0 + pc[j]) <= SF_BULK_EPS;
            near1 = near1 && fabsf(c1 - pc[j]) <= SF_BULK_EPS;
        }
    }
#pragma unroll
    for (int k = 0; k < G::NHT; ++k) {
        if (hm.off[k] < 0) continue;
        const int r = hm.reg[k];
        float vp;
        if ((hmask >> r) & 1u) {
            vp = (float)(raw_f64(f.hp[k].x, f.hp[k].y) - offp);
            L.T[hm.box[k]] = (float)(raw_f64(f.ht[k].x, f.ht[k].y) - offT);
        } else {
            const unsigned short wp = (unsigned short)raw_word4(f.hp[k], hm.off[k]);
            const unsigned short wt = (unsigned short)raw_word4(f.ht[k], hm.off[k]);
            vp = sf_decode_rel(L.nbo[slot][0][r] - offp, wp, L.nbs[slot][0][r]);
            L.T[hm.box[k]] =
                sf_decode_rel(L.nbo[slot][1][r] - offT, wt, L.nbs[slot][1][r]);
        }
        L.phi[hm.box[k]] = vp;
        if (!NOSKIP) {
            near0 = near0 && fabsf(c0 + vp) <= SF_BULK_EPS;
            near1 = near1 && fabsf(c1 - vp) <= SF_BULK_EPS;
        }
    }
    if (!NOSKIP) {
        const unsigned long long b0 = __ballot(near0), b1 = __ballot(near1);
        if ((threadIdx.x & 63) == 0) L.hflag[threadIdx.x >> 6] = (b0 == ~0ull ? 1u : 0u) | (b1 == ~0ull ? 2u : 0u);
    }
    __syncthreads();  // A
    bool hold = false;
    if (!NOSKIP) {
        unsigned fl = L.hflag[0];
#pragma unroll
        for (int i = 1; i < WG / 64; ++i) fl &= L.hflag[i];
        hold = fl != 0u;
    }
    nb_code_store<TS>(L, slot ^ 1, nxt);
    __half* dphi = a.out.f16 + d.cell_base;
    __half* dT = a.out.f16 + a.total_cells + d.cell_base;
    if (hold) {
        // phi held (same fp16 words, offset and scale); T' = T + dt lap T
        float lt[R];
        sf_lap<TS, float>(L.T, ly, x0, d.stencil == AMPS_STENCIL_REF, d.kf[SFK_INV_DXDY], tc, lt);
        float mnt = INFINITY, mxt = -INFINITY, dmn = 0.0f, dmx = 0.0f;
#pragma unroll
        for (int i = 0; i < R; ++i) {
            tc[i] = tc[i] + d.kf[SFK_DT] * lt[i];
            mnt = fminf(mnt, tc[i]);
            mxt = fmaxf(mxt, tc[i]);
        }
        block_minmax2(mnt, mxt, dmn, dmx, L.red);  // C
        const TileCode ct = tile_code(mnt, mxt);
        const double offTn = offT + (double)ct.mid;
        unsigned short w[R];
        sf_encode<R>(tc, ct, w);
        sf_store_low<R>(dT, gcell0, w);
        sf_store_low<R>(dphi, gcell0, f.op);
        if (threadIdx.x == 0) {
            a.out.off[gtile] = offp;
            a.out.scl[gtile] = sclp;
            a.out.off[a.total_tiles + gtile] = offTn;
            a.out.scl[a.total_tiles + gtile] = ct.scl;
            a.same_flags[gtile] = SAME_PHI;  // phi copied: both buffers hold it
        }
        return;
    }

    // a phi box of exact zeros needs no ring gradients (sf_phase1_flat)
    if (NOSKIP && flat == 1)
        sf_phase1_flat<TS, H2>(L.P, L.Q, L.E2, d, d.kf[SFK_COS_N_THETA0], d.kf[SFK_SIN_N_THETA0], d.kf[SFK_EPSB],
                               d.kf[SFK_EPSB_DELTA], d.kf[SFK_MEPSB_ANISO_DELTA]);
    else if constexpr (H2)
        sf_phase1_h2<TS>(L.phi, L.P, L.Q, L.E2, d, ox, oy, d.kf[SFK_INV_2DX], d.kf[SFK_INV_2DY],
                         d.kf[SFK_COS_N_THETA0], d.kf[SFK_SIN_N_THETA0], d.kf[SFK_EPSB], d.kf[SFK_EPSB_DELTA],
                         d.kf[SFK_MEPSB_ANISO_DELTA]);
    else
        sf_phase1<TS, float>(L.phi, L.P, L.Q, L.E2, d, ox, oy, d.kf[SFK_INV_2DX], d.kf[SFK_INV_2DY],
                             d.kf[SFK_COS_N_THETA0], d.kf[SFK_SIN_N_THETA0], d.kf[SFK_EPSB], d.kf[SFK_EPSB_DELTA],
                             d.kf[SFK_MEPSB_ANISO_DELTA], tcon != nullptr);
    __syncthreads();  // B

    float rp[R], rt[R];
    sf_phase2<TS, float, H2>(L.phi, L.T, L.P, L.Q, L.E2, ly, x0, pc, tc, rp, rt,
                         d.stencil == AMPS_STENCIL_REF, d.kf[SFK_INV_DXDY], d.kf[SFK_INV_2DX], d.kf[SFK_INV_2DY],
                         d.kf[SFK_DT_OVER_TAU], d.kf[SFK_DT], d.kf[SFK_KAPPA], d.kf[SFK_ALPHA_OVER_PI],
                         d.kf[SFK_GAMMA], (float)(d.k[SFK_TEQ] - offT), (float)offp, d.kf[SFK_NOISE],
                         NOSKIP ? 0.0f : SF_NOISE_SKIP_LOW, noise_key(d.seed, (uint64_t)step), gcell0);
    float mnp = INFINITY, mxp = -INFINITY, mnt = INFINITY, mxt = -INFINITY;
#pragma unroll
    for (int i = 0; i < R; ++i) {
        mnp = fminf(mnp, rp[i]); mxp = fmaxf(mxp, rp[i]);
        mnt = fminf(mnt, rt[i]); mxt = fmaxf(mxt, rt[i]);
    }
    block_minmax2(mnp, mxp, mnt, mxt, L.red);  // C
    const TileCode cp = tile_code(mnp, mxp);
    const TileCode ct = tile_code(mnt, mxt);
    const double offpn = offp + (double)cp.mid, offTn = offT + (double)ct.mid;
    unsigned short w[R];
    sf_encode<R>(rp, cp, w);
    sf_store_low<R>(dphi, gcell0, w);
    sf_encode<R>(rt, ct, w);
    sf_store_low<R>(dT, gcell0, w);
    if (threadIdx.x == 0) {
        a.out.off[gtile] = offpn;
        a.out.scl[gtile] = cp.scl;
        a.out.off[a.total_tiles + gtile] = offTn;
        a.out.scl[a.total_tiles + gtile] = ct.scl;
        a.same_flags[gtile] = 0;
    }
}

// ------------------------------------------------------------- hold sweep
// Low tiles in the bulk-hold state by their tile codes: the tile and its 8
// neighbours are low and each one's phi code bounds every phi value within
// SF_BULK_EPS of 0 (or each one's within SF_BULK_EPS of 1): |phi - off| <=
// 2^14 scl since |h| <= 2^14, so |off| + 2^14 scl <= eps implies |phi| <= eps
// for every cell.  This implies the per-value test of the full low sweep, so
// a tile held here would be held there too.  Such a tile keeps phi (same
// words, offset and scale) and advances only T' = T + dt lap T, with the
// arithmetic of the full sweep's hold path.  same_flags[tile] bit 0 records
// that both buffers already hold that phi (a previous hold step copied it),
// so the phi words are neither read nor written again.  A held tile is also
// T-flat when the codes of the tile and of its four edge neighbours bound
// every T of its Laplacian box within SF_T_FLAT of the tile's offset
// (|off_n - off| + 2^14 scl_n <= SF_T_FLAT): its T is held as well (the
// update would change it by less than dt * 8 SF_T_FLAT / dx^2, ~SF_T_FLAT on
// the workload grids).  Far from the crystal T is flat down to float
// subnormals; holding it keeps those tiles free of work.  A flat tile is
// copied once into the other buffer (bit 1) and skipped outright after that:
// no loads beyond its codes, no stores.  Tiles that fail the hold test are
// appended to the mixed sweep's list, which updates them with the full low
// step beside the FP64 tiles.
//
// One wave per tile, no LDS barrier: lane l owns R = TS*TS/64 cells of row
// l / TPR; the T halo is the cross of depth H (no corners: the Laplacian
// needs none), every element from a low neighbour.
constexpr double SF_H_BOUND = 16384.0;
constexpr double SF_T_FLAT = 1.0e-12;

template <int TS, int H>
struct HoldGeo {
    static constexpr int R = TS * TS / 64;     // cells per lane
    static constexpr int TPR = TS / R;         // lanes per row
    static constexpr int C0 = 4;               // box column of tile column 0 (16-byte aligned)
    // box row stride: a multiple of 4 floats with LW/4 odd, so the 16-byte
    // reads and writes of a quarter wave (4 rows x 2 row halves) fall on
    // distinct bank groups
    static constexpr int LW = TS + 12;
    static_assert((LW / 4) % 2 == 1, "odd number of 16-byte groups per row");
    static constexpr int ROWS = TS + 2 * H;
    static constexpr int NX = 4 * TS * H;      // halo cross elements
    static constexpr int NXL = NX / 64;        // per lane
    static_assert(NX % 64 == 0, "halo cross must fill whole waves");
};

struct HoldMeta {
    int type;
    double op, ot;
    float sp, st;
};

// Bulk-hold bits of one tile's phi code: bit 0 every value within
// SF_BULK_EPS of 0, bit 1 within SF_BULK_EPS of 1 (FP64 tiles: neither).  A
// tile is held when the AND over itself and its 8 neighbours is non-zero.
__device__ inline unsigned hold_bits(int type, double op, float sp) {
    if (type != 1) return 0u;
    const double bnd = SF_H_BOUND * (double)sp, eps = (double)SF_BULK_EPS;
    return (fabs(op) + bnd <= eps ? 1u : 0u) | (fabs(1.0 - op) + bnd <= eps ? 2u : 0u);
}

__device__ inline double readlane_d(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xFFFFFFFFll), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ inline float readlane_f(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }

// R fp16 words of this lane's cells (R = 16: two 16-byte loads; R = 4: one 8-byte load).
template <int R>
__device__ inline void load_words(const unsigned short* p, unsigned* wd) {
    if constexpr (R == 16) {
        const uint4 u = reinterpret_cast<const uint4*>(p)[0], v = reinterpret_cast<const uint4*>(p)[1];
        wd[0] = u.x; wd[1] = u.y; wd[2] = u.z; wd[3] = u.w; wd[4] = v.x; wd[5] = v.y; wd[6] = v.z; wd[7] = v.w;
    } else {
        static_assert(R == 4, "16x16 or 32x32 tiles");
        const uint2 u = *reinterpret_cast<const uint2*>(p);
        wd[0] = u.x; wd[1] = u.y;
    }
}
template <int R>
__device__ inline void store_words(unsigned short* p, const unsigned* wd) {
    if constexpr (R == 16) {
        reinterpret_cast<uint4*>(p)[0] = make_uint4(wd[0], wd[1], wd[2], wd[3]);
        reinterpret_cast<uint4*>(p)[1] = make_uint4(wd[4], wd[5], wd[6], wd[7]);
    } else {
        *reinterpret_cast<uint2*>(p) = make_uint2(wd[0], wd[1]);
    }
}
__device__ inline float word_half(const unsigned* wd, int j) {
    const unsigned v = wd[j >> 1];
    return __half2float(__ushort_as_half((unsigned short)((j & 1) ? (v >> 16) : (v & 0xFFFFu))));
}

// A held, T-flat tile: both fields keep their words, offsets and scales.
// Copies whichever field the other buffer does not hold yet.
template <int TS, int R>
__device__ void sf_flat_copy(const StepArgs& a, const InstDesc& d, int w, int same) {
    constexpr int TPR = TS / R;
    const int lane = threadIdx.x & 63;
    const int lt = w - d.tile_base;
    const int tile_y = lt / d.tx, tile_x = lt - tile_y * d.tx;
    const int gcell0 = (tile_y * TS + lane / TPR) * d.nx + tile_x * TS + (lane % TPR) * R;
#pragma unroll
    for (int f = 0; f < 2; ++f) {
        if (same & (f == 0 ? SAME_PHI : SAME_T)) continue;
        unsigned wd[R / 2];
        load_words<R>(reinterpret_cast<const unsigned short*>(a.in.f16 + (size_t)f * a.total_cells + d.cell_base) + gcell0, wd);
        store_words<R>(reinterpret_cast<unsigned short*>(a.out.f16 + (size_t)f * a.total_cells + d.cell_base) + gcell0, wd);
        if (lane == 0) {
            a.out.off[(size_t)f * a.total_tiles + w] = a.in.off[(size_t)f * a.total_tiles + w];
            a.out.scl[(size_t)f * a.total_tiles + w] = a.in.scl[(size_t)f * a.total_tiles + w];
        }
    }
    if (lane == 0) a.same_flags[w] = SAME_PHI | SAME_T;
}

template <int TS, int H>
__device__ void sf_hold_tile(float* Tb, const StepArgs& a, const InstDesc& d, int w, const HoldMeta& m, int lb,
                             bool same) {
    using G = HoldGeo<TS, H>;
    constexpr int R = G::R, LW = G::LW, C0 = G::C0;
    const int lane = threadIdx.x & 63;
    const int lt = w - d.tile_base;
    const int tile_y = lt / d.tx, tile_x = lt - tile_y * d.tx;
    const int ox = tile_x * TS, oy = tile_y * TS;
    const int ly = lane / G::TPR, x0 = (lane % G::TPR) * R;
    const int gcell0 = (oy + ly) * d.nx + ox + x0;
    const unsigned short* iT = reinterpret_cast<const unsigned short*>(a.in.f16 + a.total_cells + d.cell_base);
    const unsigned short* iP = reinterpret_cast<const unsigned short*>(a.in.f16 + d.cell_base);
    unsigned tw[R / 2], pw[R / 2];
    load_words<R>(iT + gcell0, tw);
    // halo cross: band 0 rows above, 1 rows below, 2 columns left, 3 columns right
    const bool interior = ox >= H && oy >= H && ox + TS + H <= d.nx && oy + TS + H <= d.ny;
    unsigned short hw[G::NXL];
    int hbox[G::NXL], hband[G::NXL];
#pragma unroll
    for (int k = 0; k < G::NXL; ++k) {
        const int e = lane + 64 * k, band = e / (TS * H), i = e - band * (TS * H);
        int gy, gx, by, bx;
        if (band < 2) {
            const int r = i / TS, c = i - r * TS;
            gy = band == 0 ? oy - H + r : oy + TS + r;
            by = band == 0 ? r : H + TS + r;
            gx = ox + c;
            bx = C0 + c;
        } else {
            const int r = i / H, c = i - r * H;
            gy = oy + r;
            by = H + r;
            gx = band == 2 ? ox - H + c : ox + TS + c;
            bx = band == 2 ? C0 - H + c : C0 + TS + c;
        }
        if (!interior) { gy = bmap(gy, d.ny, d.bc); gx = bmap(gx, d.nx, d.bc); }
        hw[k] = iT[gy * d.nx + gx];
        hbox[k] = by * LW + bx;
        hband[k] = band;
    }
    if (!same) load_words<R>(iP + gcell0, pw);  // issued after the T loads: they do not wait on it
    // codes: own tile (neighbour 4) and the edge neighbours 1, 7, 3, 5
    const double offT = readlane_d(m.ot, lb + 4), offP = readlane_d(m.op, lb + 4);
    const float sclT = readlane_f(m.st, lb + 4), sclP = readlane_f(m.sp, lb + 4);
    const float o1 = (float)(readlane_d(m.ot, lb + 1) - offT), o7 = (float)(readlane_d(m.ot, lb + 7) - offT);
    const float o3 = (float)(readlane_d(m.ot, lb + 3) - offT), o5 = (float)(readlane_d(m.ot, lb + 5) - offT);
    const float s1 = readlane_f(m.st, lb + 1), s7 = readlane_f(m.st, lb + 7), s3 = readlane_f(m.st, lb + 3),
                s5 = readlane_f(m.st, lb + 5);
    float tc[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        tc[j] = word_half(tw, j) * sclT;
        Tb[(ly + H) * LW + C0 + x0 + j] = tc[j];
    }
#pragma unroll
    for (int k = 0; k < G::NXL; ++k) {
        const int bd = hband[k];
        const float o = bd == 0 ? o1 : (bd == 1 ? o7 : (bd == 2 ? o3 : o5));
        const float sc = bd == 0 ? s1 : (bd == 1 ? s7 : (bd == 2 ? s3 : s5));
        Tb[hbox[k]] = o + __half2float(__ushort_as_half(hw[k])) * sc;
    }
    __syncthreads();  // one wave: orders the LDS writes before the reads
    // T' = T + dt lap T, same operand order as the full sweep's Laplacian;
    // every box read is an aligned 16-byte read: the own row as columns
    // x0-4 .. x0+R+3, the rows above and below as columns x0 .. x0+R-1
    constexpr int LW4 = LW / 4, Q = R / 4;
    const float4* B4 = reinterpret_cast<const float4*>(Tb);
    const int b4 = ((ly + H) * LW + C0 + x0) / 4;
    float row[R + 8];
#pragma unroll
    for (int q = 0; q < Q + 2; ++q) {
        float4 v = B4[b4 - 1 + q];
        // keep the whole 16-byte read (only part of the first and last
        // groups is used; split into 4-byte reads they would conflict)
        asm volatile("" : "+v"(v.x), "+v"(v.y), "+v"(v.z), "+v"(v.w));
        row[4 * q] = v.x; row[4 * q + 1] = v.y; row[4 * q + 2] = v.z; row[4 * q + 3] = v.w;
    }
    const bool ref = H == 2 && d.stencil == AMPS_STENCIL_REF;
    const float inv = d.kf[SFK_INV_DXDY], dt = d.kf[SFK_DT];
    float mn = INFINITY, mx = -INFINITY;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const float4 u1 = B4[b4 - LW4 + q], d1 = B4[b4 + LW4 + q];
        float4 u2 = u1, d2 = d1;
        if (ref) { u2 = B4[b4 - 2 * LW4 + q]; d2 = B4[b4 + 2 * LW4 + q]; }
        const float U1[4] = {u1.x, u1.y, u1.z, u1.w}, D1[4] = {d1.x, d1.y, d1.z, d1.w};
        const float U2[4] = {u2.x, u2.y, u2.z, u2.w}, D2[4] = {d2.x, d2.y, d2.z, d2.w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int j = 4 * q + i;
            const float* c = row + 4 + j;  // c[0]: this cell
            float l;
            if (ref)
                l = (c[-1] + c[-2] + c[1] + c[2] + U1[i] + U2[i] + D1[i] + D2[i] - 8.0f * c[0]) * inv;
            else
                l = (c[-1] + c[1] + U1[i] + D1[i] - 4.0f * c[0]) * inv;
            tc[j] = tc[j] + dt * l;
            mn = fminf(mn, tc[j]);
            mx = fmaxf(mx, tc[j]);
        }
    }
    wave_minmax(mn, mx);
    const TileCode ct = tile_code(mn, mx);
    unsigned nw[R / 2];
#pragma unroll
    for (int j = 0; j < R; j += 2) {
        const unsigned lo = __half_as_ushort(__float2half_rn((tc[j] - ct.mid) * ct.inv));
        const unsigned hi = __half_as_ushort(__float2half_rn((tc[j + 1] - ct.mid) * ct.inv));
        nw[j >> 1] = lo | (hi << 16);
    }
    unsigned short* oT = reinterpret_cast<unsigned short*>(a.out.f16 + a.total_cells + d.cell_base);
    unsigned short* oP = reinterpret_cast<unsigned short*>(a.out.f16 + d.cell_base);
    store_words<R>(oT + gcell0, nw);
    if (!same) store_words<R>(oP + gcell0, pw);
    if (lane == 0) {
        a.out.off[a.total_tiles + w] = offT + (double)ct.mid;
        a.out.scl[a.total_tiles + w] = ct.scl;
        if (!same) {
            a.out.off[w] = offP;
            a.out.scl[w] = sclP;
        }
        a.same_flags[w] = SAME_PHI;  // phi is now in both buffers
    }
    __syncthreads();  // the box is rewritten by the next tile
}

// Each wave takes a group of HOLD_GT list positions g, g + ngr, g + 2 ngr,
// ... (ngr groups: the tiles of a group lie far apart, so the active tiles,
// which cluster in the list, spread over the waves) and reads the codes of
// all of them in one round: lane 9 j + k holds neighbour k of the group's
// tile j.  Holding, flatness and the same_flags of the group are decided
// together; only the tiles that need work touch their fields.
// 4 tiles x 9 neighbours (36 lanes): more, shorter groups run in parallel
// (4096^2 snowflake: 7.2 us a launch with 7 tiles a group, 6.7 with 3 or 4,
// 7.4 with 2)
constexpr int HOLD_GT = 4;

template <int TS, int H, bool MULTI>
__global__ __launch_bounds__(64) void sf_hold_kernel(StepArgs a, const InstDesc* __restrict__ inst,
                                                     const int* __restrict__ tinst, const int* __restrict__ wl,
                                                     const int* __restrict__ wlc) {
    using G = HoldGeo<TS, H>;
    __shared__ __align__(16) float Tb[G::ROWS * G::LW];
    const long long step = *a.step_ctr + a.step_add;
    // the next step appends to the other parity's segments: clear them (the
    // kernel that read them, the previous step's mixed sweep, has finished)
    if (blockIdx.x == 0 && threadIdx.x < 8) a.full_cnt[(int)((step + 1) & 1) * 8 + threadIdx.x] = 0;
    int* fcnt = a.full_cnt + (int)(step & 1) * 8;
    const int n = *wlc;
    const int ngr = (n + HOLD_GT - 1) / HOLD_GT;
    const int lane = threadIdx.x & 63, tj = lane / 9, tk = lane - 9 * tj;
    for (int g = blockIdx.x; g < ngr; g += gridDim.x) {
        // tile of this lane's group slot (lanes past the last slot repeat it)
        const int jj = tj < HOLD_GT ? tj : HOLD_GT - 1;
        const int pj = g + jj * ngr;
        const int pc = pj < n ? pj : g;  // an absent slot repeats slot 0 (loaded, never used)
        const int w = wl[pc];
        const InstDesc& dj = inst[MULTI ? tinst[w] : 0];
        HoldMeta m;
        {
            const int lt = w - dj.tile_base;
            const int tile_y = lt / dj.tx, tile_x = lt - tile_y * dj.tx;
            Nb nb;
            nb_setup(nb, dj, tile_x, tile_y, TS);
            const int gt = nb_tile(nb, tk < 9 ? tk : 8);
            m.type = a.type[gt];
            m.op = a.in.off[gt];
            m.ot = a.in.off[a.total_tiles + gt];
            m.sp = a.in.scl[gt];
            m.st = a.in.scl[a.total_tiles + gt];
        }
        const int sf = a.same_flags[w];
        const unsigned hbits = tj < HOLD_GT ? hold_bits(m.type, m.op, m.sp) : 0u;
        const unsigned long long n0 = __ballot(hbits & 1u), n1 = __ballot(hbits & 2u);
        // T-flat: |off_T - own off_T| + 2^14 scl_T <= SF_T_FLAT on slots 1, 3, 4, 5, 7 of the tile
        const double own_ot = __shfl(m.ot, 9 * jj + 4, 64);
        const bool fe = tj < HOLD_GT && (tk == 1 || tk == 3 || tk == 4 || tk == 5 || tk == 7);
        const unsigned long long fl = __ballot(fe && fabs(m.ot - own_ot) + SF_H_BOUND * (double)m.st <= SF_T_FLAT);
        for (int j = 0; j < HOLD_GT; ++j) {
            const int p = g + j * ngr;
            if (p >= n) break;
            const unsigned long long msk = 0x1FFull << (9 * j), fmsk = 0xBAull << (9 * j);
            const int wt = __builtin_amdgcn_readlane(w, 9 * j);
            const InstDesc& d = inst[MULTI ? tinst[wt] : 0];
            if ((n0 & msk) == msk || (n1 & msk) == msk) {
                const int same = __builtin_amdgcn_readlane(sf, 9 * j);
                if ((fl & fmsk) != fmsk) sf_hold_tile<TS, H>(Tb, a, d, wt, m, 9 * j, (same & SAME_PHI) != 0);
                else if (same != (SAME_PHI | SAME_T)) sf_flat_copy<TS, G::R>(a, d, wt, same);
            } else if (lane == 0) {
                // for the mixed sweep; segment = list position mod 8 (at most
                // ceil(n / 8) entries each; 8 counters spread the atomics)
                const int seg = p & 7;
                a.full_seg[seg * a.full_cap + atomicAdd(fcnt + seg, 1)] = wt;
            }
        }
    }
}

// ---------------------------------------------------------------- kernels
// Sweeps over device-built tile work-lists (rebuilt after every
// classification): the FP64 list and the low-precision list.
//
// The read-only metadata (instance descriptors, tile -> instance map, the
// work-list and its count) arrive as separate restrict-qualified arguments:
// the compiler then reads them through the scalar cache instead of a chain of
// vector loads and waits per tile, and with MULTI = false (one instance) the
// descriptor is loaded once per workgroup.
// Type of neighbour tile (lane % 64) of tile w, for lanes 0..8 of every wave
// (2 elsewhere); hmask_of() turns it into the wave-uniform 9-bit mask of FP64
// neighbours.
template <int TS>
__device__ inline int nb_type_lane(const StepArgs& a, const InstDesc& d, int w, const SfTileConst<TS>* tc = nullptr) {
    const int lane = threadIdx.x & 63;
    if (a.type == nullptr) return 2;
    // every lane loads (lanes 9.. repeat neighbour 8): no branch around the
    // load, so its wait lands where the type is used
    const int t = a.type[nb_tile_of<TS>(d, w, tc)];
    return lane < 9 ? t : 2;
}

__device__ inline unsigned hmask_of(int t) {
    return (unsigned)(__ballot((threadIdx.x & 63) < 9 && t == 2) & 0x1FFull);
}

// PURE_FP64 sweep: one workgroup per tile (the grid is the work count).
// With no loop over tiles nothing stays live from one tile to the next (a
// persistent loop keeps the per-tile constants live, spilling SGPRs).  The
// LDS boxes (50.5 KB) allow 3 workgroups per CU; 32^2 tiles run with NT =
// 512 threads, 2 cells each (78 VGPRs against 114 at 256 threads and 4
// cells), so 24 waves per CU share the tile chains' latency instead of 12.
template <int TS, bool MULTI, int NT = WG>
__global__ __launch_bounds__(NT) void sf_f64_tile_kernel(StepArgs a, const InstDesc* __restrict__ inst,
                                                         const int* __restrict__ tinst) {
    __shared__ __align__(16) SfLds<TS, double> L;
    const long long step = *a.step_ctr + a.step_add;
    const int b = blockIdx.x;
    if (b >= a.nwork) return;
    const int w = xcd_remap(b, a.nwork);
    const InstDesc& d = inst[MULTI ? tinst[w] : 0];
    const int lt = w - d.tile_base;
    const int tile_y = lt / d.tx, tile_x = lt - tile_y * d.tx;
    sf_tile_f64<TS, false, NT>(L, a, d, tile_x, tile_y, 0x1FFu, step);
}

// Mixed sweep of the adaptive step: the FP64 tiles (FP64 work-list) and the
// low tiles that failed the hold sweep's bulk-hold test (its 8 segments of
// this step's parity) are the items of one persistent launch, so the two
// latency-bound tile chains run side by side.  Items [0, nh) are FP64 tiles
// (work-list entry i, stored XCD-remapped by the work-list kernel), the rest
// low tiles; the LDS is a union of the two layouts, with a barrier after
// every item.
template <int TS>
union SfMixLds {
    SfLds<TS, double> h;
    SfLowLds<TS> l;
};

//
// NOSKIP: the low items are the tiles the flat sweep (sf_flat_kernel)
// appended to this step parity's NSK_SEGS segments, levels 0 and 1 of
// low_flat_level, each updated by the full low path (H2: its half2 form).
template <int TS, bool MULTI, bool NOSKIP, bool H2 = false>
__global__ __launch_bounds__(256, 3) void sf_mixed_kernel(StepArgs a, const InstDesc* __restrict__ inst,
                                                      const int* __restrict__ tinst, const int* __restrict__ wl,
                                                      const int* __restrict__ wlc, const int* __restrict__ fseg,
                                                      const int* __restrict__ fcnt2) {
    __shared__ __align__(16) SfMixLds<TS> U;
    // Everything the first item needs is read in one round: both parities'
    // segment counts (no wait on the step counter first) and work-list entry
    // blockIdx.x (entries sit at XCD-remapped positions, so item i is entry i;
    // the index and value are clamped in case there is no such FP64 item).
    const int G = gridDim.x, T = a.total_tiles;
    const int wfirst = wl[blockIdx.x < (unsigned)T ? (int)blockIdx.x : T - 1];
    int c0[8], c1[8];
    const int lane = threadIdx.x & 63;
    if (NOSKIP) {
        // lane k: segment k's count of either parity
        c0[0] = fcnt2[lane * NSK_CNT_STRIDE];
        c1[0] = fcnt2[(NSK_SEGS + lane) * NSK_CNT_STRIDE];
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) { c0[k] = fcnt2[k]; c1[k] = fcnt2[8 + k]; }
    }
    const int nh = *wlc;
    const long long step = *a.step_ctr + a.step_add;
    const bool odd = (step & 1) != 0;
    int cnt[8], nf = 0, sincl = 0;
    if (NOSKIP) {
        // inclusive prefix of the segment counts over the lanes: item j of
        // the concatenated segments lies in the first segment whose prefix
        // exceeds j
        sincl = odd ? c1[0] : c0[0];
#pragma unroll
        for (int dlt = 1; dlt < 64; dlt <<= 1) {
            const int t = __shfl_up(sincl, dlt);
            if (lane >= dlt) sincl += t;
        }
        nf = __shfl(sincl, 63);
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) { cnt[k] = odd ? c1[k] : c0[k]; nf += cnt[k]; }
    }
    const int n = nh + nf;
    asm volatile("" ::"s"(wfirst));  // the entry's load stays in the counts' round (not sunk into the loop)
    if (blockIdx.x == 0 && threadIdx.x == 0) {  // tile counts of the step (one writer per launch)
        a.stats[2] += (unsigned long long)nh;
        // NOSKIP: the flat sweep counts every low tile; these are the ones it listed
        a.stats[NOSKIP ? 4 : 3] += (unsigned long long)nf;
    }
    for (int i = blockIdx.x; i < n; i += G) {
        if (i < nh) {
            const int w = i == (int)blockIdx.x ? wfirst : wl[i];
            const InstDesc& d = inst[MULTI ? tinst[w] : 0];
            const int t = nb_type_lane<TS>(a, d, w);
            const int lt = w - d.tile_base;
            const int tile_y = lt / d.tx, tile_x = lt - tile_y * d.tx;
            sf_tile_f64<TS, true>(U.h, a, d, tile_x, tile_y, hmask_of(t), step);
        } else {
            int w;
            if (NOSKIP) {
                const int j = i - nh;
                const int seg = __ffsll((long long)__ballot(sincl > j)) - 1;
                const int excl = __shfl(sincl, seg) - __shfl(odd ? c1[0] : c0[0], seg);
                w = fseg[seg * nsk_cap(a.total_tiles) + (j - excl)];
            } else {
                // entry j of the concatenated segments
                int j = i - nh, seg = 0;
#pragma unroll
                for (int k = 0; k < 7; ++k)
                    if (seg == k && j >= cnt[k]) { j -= cnt[k]; seg = k + 1; }
                w = fseg[seg * a.full_cap + j];
            }
            const InstDesc& d = inst[MULTI ? tinst[w] : 0];
            const int t = nb_type_lane<TS>(a, d, w);
            NbCode c0;
            nb_code_fetch<TS>(c0, a, d, w);
            const unsigned hmask = hmask_of(t);
            LowFetch<TS> f;
            low_fetch<TS>(f, a, d, w, hmask);
            nb_code_store<TS>(U.l, 0, c0);
            __syncthreads();  // neighbour table visible
            sf_tile_low<TS, NOSKIP, H2>(U.l, 0, c0, a, d, w, f, hmask, step, nullptr,
                                        NOSKIP ? low_flat_level(c0, hmask) : 0);
        }
        __syncthreads();  // the LDS union is rewritten by the next item
    }
}

// ------------------------------------------------------------- flat sweep
// First sweep of the NO_SKIP adaptive step: one wave per tile of the low
// work-list.  Workgroup b takes entry b; the list places runs of XCD_CHUNK
// consecutive tiles on one XCD (a tile's left and right halo lines are then
// mostly in its own XCD's L2) and deals the runs round-robin, so the costly
// tiles near the crystal spread evenly over the XCDs (one contiguous run
// per XCD left the XCDs holding the crystal's rows working 1.5x longer).
//
// Tiles of low_flat_level 3 and 2 are finished here, with the values the
// full low update (sf_tile_low) gives them.  On a level-2 tile (the phi box
// decodes to exact zeros at phi 0 or 1) pa(1 - pa) is +-0 in every cell;
// the flux is +0 (differences of equal P and Q, a Laplacian of zeros),
// pa(1 - pa)(pa - 1/2 + m) and the noise term (not taken: |pp| = 0) are
// +-0, so br = +0 and phi' = ph + dtt br = +0, and T' is the update of
// sf_phase2 with phi' - ph = +0 - ph = +0 for either sign of ph (so the
// phi words need not be read: sf_t_flat).  A level-3 tile adds a T box of
// exact zeros: sf_tile_zero.  A tile of level 0 or 1 is appended to this
// step parity's list (NSK_SEGS segments) for the mixed sweep, which updates
// it beside the FP64 tiles.
//
// A level-2 tile needs only its T box in LDS (5 KB at tile 32): about 30
// of these workgroups fit a CU against 6 of the full low tile (25 KB), so
// that many more tiles' load rounds overlap.  The wave covers a tile as
// the WG threads of sf_tile_low would, in WG/64 passes (virtual thread
// v = lane + 64 k); neighbour codes come from lanes 0..8 by cross-lane
// reads, and the tile's min/max is a wave reduction.
constexpr int FLAT_NT = 64;

template <int TS, bool MULTI>
__global__ __launch_bounds__(FLAT_NT) void sf_flat_kernel(StepArgs a, const InstDesc* __restrict__ inst,
                                                          const int* __restrict__ tinst, const int* __restrict__ wl,
                                                          const int* __restrict__ wlc) {
    using G = SfGeo<TS>;
    using GF = SfGeo<TS, FLAT_NT>;
    constexpr int H = G::H, LW = G::LW, R = G::R, NP = WG / FLAT_NT;
    __shared__ __align__(16) float Tb[LW * LW];
    const int b = blockIdx.x;
    const int w = wl[b];
    const int n = *wlc;
    const long long step = *a.step_ctr + a.step_add;
    if (b == 0) {
        if (threadIdx.x == 0) a.stats[3] += (unsigned long long)n;  // low tiles of the step (one writer)
        // the next step's counters (lane k: segment k): their last reader, the
        // previous step's mixed sweep, has finished
        static_assert(FLAT_NT == NSK_SEGS, "one lane per segment");
        a.full_cnt[((int)((step + 1) & 1) * NSK_SEGS + (int)threadIdx.x) * NSK_CNT_STRIDE] = 0;
    }
    if (b >= n) return;
    const InstDesc& d = inst[MULTI ? tinst[w] : 0];
    const int tt = nb_type_lane<TS>(a, d, w);
    NbCode c;
    nb_code_fetch<TS>(c, a, d, w);
    const unsigned hm = hmask_of(tt);
    const int flat = low_flat_level(c, hm);
    const double offp = __shfl(c.o0, 4), offT = __shfl(c.o1, 4);  // lane 4: the tile itself
    if (flat == 3) {
        sf_tile_zero<TS, FLAT_NT>(a, d, w, offp, offT);
        return;
    }
    if (flat < 2) {
        // segment = list position mod NSK_SEGS: at most nsk_cap entries each,
        // and the appends spread over NSK_SEGS counters on their own lines
        if (threadIdx.x == 0) {
            const int seg = b & (NSK_SEGS - 1);
            const int j = atomicAdd(a.full_cnt + ((int)(step & 1) * NSK_SEGS + seg) * NSK_CNT_STRIDE, 1);
            a.full_seg[seg * nsk_cap(a.total_tiles) + j] = w;
        }
        return;
    }
    // level 2: the T words of the tile and of its halo (no FP64 neighbour:
    // every element from the fp16 storage)
    const int lt = w - d.tile_base;
    const int tile_y = lt / d.tx, tile_x = lt - tile_y * d.tx;
    const int ox = tile_x * TS, oy = tile_y * TS;
    const unsigned short* hT = reinterpret_cast<const unsigned short*>(a.in.f16 + a.total_cells + d.cell_base);
    unsigned short ot[NP][R];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const int v = threadIdx.x + k * FLAT_NT;
        const int gcell0 = (oy + v / G::TPR) * d.nx + ox + (v % G::TPR) * R;
        if (R == 4) {
            const uint2 u = *reinterpret_cast<const uint2*>(hT + gcell0);
            ot[k][0] = u.x & 0xFFFFu; ot[k][1 % R] = u.x >> 16; ot[k][2 % R] = u.y & 0xFFFFu; ot[k][3 % R] = u.y >> 16;
        } else {
#pragma unroll
            for (int j = 0; j < R; ++j) ot[k][j] = hT[gcell0 + j];
        }
    }
    SfHalo<TS, FLAT_NT> hmap;
    sf_halo_map<TS, FLAT_NT>(hmap, d, ox, oy);
    uint2 hp[GF::NHT], ht[GF::NHT];
    sf_halo_load<TS, true, FLAT_NT>(hp, ht, a, d, hmap, 0u, false);
    const float sclT = __shfl(c.s1, 4);
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const int v = threadIdx.x + k * FLAT_NT;
        const int ly = v / G::TPR, x0 = (v % G::TPR) * R;
#pragma unroll
        for (int j = 0; j < R; ++j)
            Tb[(ly + H) * LW + x0 + H + j] = __half2float(__ushort_as_half(ot[k][j])) * sclT;
    }
#pragma unroll
    for (int k = 0; k < GF::NHT; ++k) {
        const int r = hmap.reg[k];
        const double o1 = __shfl(c.o1, r);  // every lane takes part in the cross-lane reads
        const float s1 = __shfl(c.s1, r);
        if (hmap.off[k] < 0) continue;
        const unsigned short wt = (unsigned short)raw_word4(ht[k], hmap.off[k]);
        Tb[hmap.box[k]] = sf_decode_rel(o1 - offT, wt, s1);
    }
    __syncthreads();  // the T box is in LDS
    const float dt = d.kf[SFK_DT], kap = d.kf[SFK_KAPPA];
    float rt[NP][R];
    float mnt = INFINITY, mxt = -INFINITY;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const int v = threadIdx.x + k * FLAT_NT;
        const int ly = v / G::TPR, x0 = (v % G::TPR) * R;
        float tc[R], lp[R];
        sf_lap<TS, float>(Tb, ly, x0, d.stencil == AMPS_STENCIL_REF, d.kf[SFK_INV_DXDY], tc, lp);
#pragma unroll
        for (int i = 0; i < R; ++i) {
            rt[k][i] = sf_t_flat(tc[i], lp[i], dt, kap, 0.0f);
            mnt = fminf(mnt, rt[k][i]);
            mxt = fmaxf(mxt, rt[k][i]);
        }
    }
    wave_minmax(mnt, mxt);
    const TileCode cp = tile_code(0.0f, 0.0f), ct = tile_code(mnt, mxt);
    float zr[R];
#pragma unroll
    for (int i = 0; i < R; ++i) zr[i] = 0.0f;
    unsigned short wz[R];
    sf_encode<R>(zr, cp, wz);
    __half* dphi = a.out.f16 + d.cell_base;
    __half* dT = a.out.f16 + a.total_cells + d.cell_base;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const int v = threadIdx.x + k * FLAT_NT;
        const int gcell0 = (oy + v / G::TPR) * d.nx + ox + (v % G::TPR) * R;
        unsigned short wt[R];
        sf_encode<R>(rt[k], ct, wt);
        sf_store_low<R>(dphi, gcell0, wz);
        sf_store_low<R>(dT, gcell0, wt);
    }
    if (threadIdx.x == 0) {
        a.out.off[w] = offp + (double)cp.mid;
        a.out.scl[w] = cp.scl;
        a.out.off[a.total_tiles + w] = offT + (double)ct.mid;
        a.out.scl[a.total_tiles + w] = ct.scl;
        a.same_flags[w] = 0;
    }
}

// ------------------------------------------------------------ host launch
// Persistent grid of kernel K: resident workgroups per CU (occupancy query,
// cached per kernel) x CUs, at most one workgroup per work item.
template <auto K, int BS = WG>
static int resident_grid(int nwork, int cap) {
    static int per_cu = 0, cus = 0;
    if (per_cu == 0) {
        int v = 0, dev = 0, c = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, K, BS, 0) != hipSuccess || v <= 0) v = 2;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0)
            c = 256;
        cus = c;
        per_cu = v;
    }
    long long g = (long long)per_cu * cus;
    if (cap > 0 && cap < g) g = cap;
    return (int)(g < nwork ? g : nwork);
}

template <int TS, bool MULTI>
static void sf_launch(const StepArgs& a, bool adaptive, hipStream_t s, const hipEvent_t* ev) {
    dim3 blk(WG);
    if (adaptive && (a.flags & AMPS_FLAG_NO_SKIP)) {
        // flat sweep over the low work-list (one wave per tile; it lists the
        // tiles of low_flat_level 0 and 1), then the FP64 tiles and those
        // low tiles in one launch; HALF2 is validated to tile 32 (R = 4
        // cells per thread, in pairs)
        constexpr auto kf = sf_flat_kernel<TS, MULTI>;
        launch_sweep(kf, dim3(a.total_tiles), dim3(FLAT_NT), s, ev, a, a.inst, a.tile_inst, a.wl_low,
                     a.wl_low_count);
        if constexpr (TS == 32) {
            if (a.flags & AMPS_FLAG_HALF2) {
                constexpr auto km = sf_mixed_kernel<TS, MULTI, true, true>;
                launch_sweep(km, dim3(resident_grid<km>(a.total_tiles, a.grid_cap)), blk, s, ev ? ev + 2 : nullptr,
                             a, a.inst, a.tile_inst, a.worklist, a.wl_count, (const int*)a.full_seg,
                             (const int*)a.full_cnt);
                return;
            }
        }
        constexpr auto km = sf_mixed_kernel<TS, MULTI, true, false>;
        launch_sweep(km, dim3(resident_grid<km>(a.total_tiles, a.grid_cap)), blk, s, ev ? ev + 2 : nullptr, a,
                     a.inst, a.tile_inst, a.worklist, a.wl_count, (const int*)a.full_seg, (const int*)a.full_cnt);
    } else if (adaptive) {
        // bulk-hold sweep first (it lists the low tiles it does not hold),
        // then FP64 tiles and those low tiles in one launch
        constexpr auto km = sf_mixed_kernel<TS, MULTI, false>;
        if (a.t_halo == 2) {
            constexpr auto kb = sf_hold_kernel<TS, 2, MULTI>;
            launch_sweep(kb, dim3(resident_grid<kb, 64>(a.total_tiles, 0)), dim3(64), s, ev, a, a.inst, a.tile_inst,
                         a.wl_low, a.wl_low_count);
        } else {
            constexpr auto kb = sf_hold_kernel<TS, 1, MULTI>;
            launch_sweep(kb, dim3(resident_grid<kb, 64>(a.total_tiles, 0)), dim3(64), s, ev, a, a.inst, a.tile_inst,
                         a.wl_low, a.wl_low_count);
        }
        launch_sweep(km, dim3(resident_grid<km>(a.total_tiles, a.grid_cap)), blk, s, ev ? ev + 2 : nullptr, a, a.inst,
                     a.tile_inst, a.worklist, a.wl_count, (const int*)a.full_seg, (const int*)a.full_cnt);
    } else {
        // 32^2 tiles: 512 threads, 2 cells each (see sf_f64_tile_kernel)
        constexpr int NT = TS == 32 ? 512 : WG;
        launch_sweep(sf_f64_tile_kernel<TS, MULTI, NT>, dim3(a.nwork), dim3(NT), s, ev, a, a.inst, a.tile_inst);
    }
}

int launch_sf_step(const StepArgs& a, int tile, bool adaptive, hipStream_t s, const hipEvent_t* ev) {
    const bool multi = a.ninst > 1;
    if (tile == 16) {
        if (multi) sf_launch<16, true>(a, adaptive, s, ev);
        else sf_launch<16, false>(a, adaptive, s, ev);
    } else if (tile == 32) {
        if (multi) sf_launch<32, true>(a, adaptive, s, ev);
        else sf_launch<32, false>(a, adaptive, s, ev);
    } else {
        return AMPS_E_UNSUPPORTED;
    }
    return hipGetLastError() == hipSuccess ? AMPS_OK : AMPS_E_HIP;
}

}  // namespace amps

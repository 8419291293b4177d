// k_adjoint.hip — reverse-time adjoint of one FP64 explicit-Euler step of the
// snowflake and sintering models.
//
// Given the state u before a step, the state u' after it and the cotangent
// lam' = dJ/du', the kernels produce lam = dJ/du (the transposed Jacobian of
// the step applied to lam') and add the step's contribution to dJ/dp for the
// differentiable parameters.  They transpose exactly the arithmetic of the
// forward sweeps (same stencils, same anisotropy evaluation, same noise):
// every forward read of a neighbour through the boundary map bmap becomes, in
// the adjoint, a gather over the pre-images {j : bmap(j + d) = i}, so zero-flux
// mirrors and periodic wrap transpose without special cases.
//
// Each kernel is one workgroup per tile, one thread per cell of the tile
// (TS*TS/256 cells per thread), and writes full-grid work arrays read by the
// next kernel.  Parameter partial sums are reduced per workgroup and added to
// a per-tile accumulator slot, so the sums are deterministic.
//
// Snowflake (forward: snowflake/function.h:5-60 restated in k_snowflake.hip):
//   mu      = dt/tau (lam_phi' + K lam_T')                   cotangent of the bracket
//   lamP_e  = 1/(2dy) (sum mu over rows whose +y read is e  - ... -y read is e)
//   lamQ_e  = -1/(2dx) (same along x)
//   lamE2_e = mu_e lap(phi)_e
//   -> lam(eps), lam(eps'), lam(theta) -> lam(gx), lam(gy) at every cell
//   lam_phi = lam_phi' + mu (dR/dphi + dN/dphi) + div^T(lam g) + lap^T(E2 mu)
//   lam_T   = lam_T'  + mu phi(1-phi) dm/dT + dt lap^T(lam_T')
// Sintering (forward: sintering/function.h:51-159, order of main.cu:136-146):
//   the clamps pass the cotangent where 0 <= v < 1 and stop it elsewhere;
//   the eta updates read the NEW c, so their cotangent flows into c' first.
#include "kernels.hpp"
#include "tile_common.hpp"
#include "sf_math.hpp"

namespace amps {

// Calls f(j) for every j in [0, n) with bmap(j + d, n, bc) == i.
template <class F>
__device__ inline void for_preimages(int i, int d, int n, int bc, F&& f) {
    if (bc == AMPS_BC_PERIODIC) {
        int j = i - d;
        if (j < 0) j += n;
        else if (j >= n) j -= n;
        f(j);
        return;
    }
    int j = i - d;
    if (j >= 0 && j < n) f(j);
    if (i > 0) {  // left mirror: j + d < 0, -(j + d) == i
        j = -i - d;
        if (j >= 0 && j < n && j + d < 0) f(j);
    }
    if (i < n - 1) {  // right mirror: j + d >= n, 2(n-1) - (j + d) == i
        j = 2 * (n - 1) - i - d;
        if (j >= 0 && j < n && j + d >= n) f(j);
    }
}

// Sum over pre-images along x (row y fixed) of a[row][j].
__device__ inline double gather_x(const double* a, int x, int y, int d, const InstDesc& D) {
    double s = 0.0;
    const double* row = a + (size_t)y * D.nx;
    for_preimages(x, d, D.nx, D.bc, [&](int j) { s += row[j]; });
    return s;
}
__device__ inline double gather_y(const double* a, int x, int y, int d, const InstDesc& D) {
    double s = 0.0;
    for_preimages(y, d, D.ny, D.bc, [&](int j) { s += a[(size_t)j * D.nx + x]; });
    return s;
}

// Transposed Laplacian of the field a (instance view) at (x, y), without the
// 1/(dx dy) factor: sum over the stencil's neighbour offsets of the
// pre-image values minus centre weight times a(x, y).
__device__ inline double lapT(const double* a, int x, int y, bool ref, const InstDesc& D) {
    double s = gather_x(a, x, y, -1, D) + gather_x(a, x, y, 1, D) + gather_y(a, x, y, -1, D) +
               gather_y(a, x, y, 1, D);
    double w0 = 4.0;
    if (ref) {
        s += gather_x(a, x, y, -2, D) + gather_x(a, x, y, 2, D) + gather_y(a, x, y, -2, D) + gather_y(a, x, y, 2, D);
        w0 = 8.0;
    }
    return s - w0 * a[(size_t)y * D.nx + x];
}

// Forward Laplacian (same operand order as the sweeps) of field f at (x, y).
__device__ inline double lap_fwd(const double* f, int x, int y, bool ref, double inv, const InstDesc& D) {
    const int nx = D.nx, ny = D.ny, bc = D.bc;
    const int xs1 = bmap(x - 1, nx, bc), xa1 = bmap(x + 1, nx, bc);
    const int ys1 = bmap(y - 1, ny, bc), ya1 = bmap(y + 1, ny, bc);
    const double* r = f + (size_t)y * nx;
    const double c = r[x];
    if (ref) {
        const int xs2 = bmap(x - 2, nx, bc), xa2 = bmap(x + 2, nx, bc);
        const int ys2 = bmap(y - 2, ny, bc), ya2 = bmap(y + 2, ny, bc);
        return (r[xs1] + r[xs2] + r[xa1] + r[xa2] + f[(size_t)ys1 * nx + x] + f[(size_t)ys2 * nx + x] +
                f[(size_t)ya1 * nx + x] + f[(size_t)ya2 * nx + x] - 8.0 * c) * inv;
    }
    return (r[xs1] + r[xa1] + f[(size_t)ys1 * nx + x] + f[(size_t)ya1 * nx + x] - 4.0 * c) * inv;
}

// Workgroup sum of NG partials; thread 0 adds them to acc[g][tile].
template <int NG>
__device__ inline void block_accumulate(double (&v)[NG], double* acc, int total_tiles, int tile, double* red) {
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        double x = v[g];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if ((threadIdx.x & 63) == 0) red[g * 4 + (threadIdx.x >> 6)] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int g = 0; g < NG; ++g)
            acc[(size_t)g * total_tiles + tile] += (red[g * 4] + red[g * 4 + 1]) + (red[g * 4 + 2] + red[g * 4 + 3]);
    }
}

// Cell (x, y) of instance D handled by this thread for item i of the tile.
template <int TS>
__device__ inline void tile_cell(const InstDesc& D, int w, int i, int& x, int& y) {
    const int lt = w - D.tile_base;
    const int ty = lt / D.tx, tx = lt - ty * D.tx;
    const int e = threadIdx.x + i * WG;
    y = ty * TS + e / TS;
    x = tx * TS + e % TS;
}

// ============================================================== snowflake
// gradient slots: 0 delta, 1 K, 2 epsb, 3 tau
constexpr int SF_NG = 4;

// Kernel 1: mu = dt/tau (lam_phi' + K lam_T'); partials of K and tau.
template <int TS>
__global__ __launch_bounds__(256) void sf_adj1_kernel(AdjArgs a) {
    constexpr int PER = TS * TS / WG;
    __shared__ double red[4 * SF_NG];
    const int w = blockIdx.x;
    const InstDesc& D = a.inst[a.tile_inst[w]];
    const size_t TC = a.total_cells;
    const double dtt = D.k[SFK_DT_OVER_TAU], K = D.k[SFK_KAPPA], tau = D.p[AMPS_SF_TAU];
    double g[SF_NG] = {0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < PER; ++i) {
        int x, y;
        tile_cell<TS>(D, w, i, x, y);
        const size_t c = D.cell_base + (size_t)y * D.nx + x;
        const double lp = a.lam_n[c], lt = a.lam_n[TC + c];
        const double big = lp + K * lt;
        a.w0[c] = dtt * big;
        const double dphi = a.un[c] - a.u[c];
        g[1] += lt * dphi;
        g[3] -= big * dphi / tau;
    }
    block_accumulate<SF_NG>(g, a.gacc, a.total_tiles, w, red);
}

// Kernel 2: lam(gx), lam(gy) and E2*mu at every cell; partials of delta, epsb.
template <int TS>
__global__ __launch_bounds__(256) void sf_adj2_kernel(AdjArgs a) {
    constexpr int PER = TS * TS / WG;
    __shared__ double red[4 * SF_NG];
    const int w = blockIdx.x;
    const InstDesc& D = a.inst[a.tile_inst[w]];
    const double* phi = a.u + D.cell_base;
    const double* mu = a.w0 + D.cell_base;
    const bool ref = D.stencil == AMPS_STENCIL_REF;
    const int n = (int)D.k[SFK_ANISO_N];
    const double inv2dx = D.k[SFK_INV_2DX], inv2dy = D.k[SFK_INV_2DY], inv = D.k[SFK_INV_DXDY];
    const double cth0 = D.k[SFK_COS_N_THETA0], sth0 = D.k[SFK_SIN_N_THETA0];
    const double epsb = D.k[SFK_EPSB], epsbd = D.k[SFK_EPSB_DELTA], mend = D.k[SFK_MEPSB_ANISO_DELTA];
    const double delta = D.p[AMPS_SF_DELTA];
    double g[SF_NG] = {0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < PER; ++i) {
        int x, y;
        tile_cell<TS>(D, w, i, x, y);
        const size_t cl = (size_t)y * D.nx + x;
        const int xs1 = bmap(x - 1, D.nx, D.bc), xa1 = bmap(x + 1, D.nx, D.bc);
        const int ys1 = bmap(y - 1, D.ny, D.bc), ya1 = bmap(y + 1, D.ny, D.bc);
        const double gx = (phi[(size_t)y * D.nx + xa1] - phi[(size_t)y * D.nx + xs1]) * inv2dx;
        const double gy = (phi[(size_t)ya1 * D.nx + x] - phi[(size_t)ys1 * D.nx + x]) * inv2dy;
        double c, s;
        sf_unit_power<double>(gx, gy, n, c, s);
        const double ca = c * cth0 + s * sth0, sa = s * cth0 - c * sth0;
        const double eps = epsb + epsbd * ca, epsd = mend * sa;
        const double A = eps * epsd;
        const double lamP = inv2dy * (gather_y(mu, x, y, 1, D) - gather_y(mu, x, y, -1, D));
        const double lamQ = -inv2dx * (gather_x(mu, x, y, 1, D) - gather_x(mu, x, y, -1, D));
        const double mue = mu[cl];
        const double lamE2 = mue * lap_fwd(phi, x, y, ref, inv, D);
        const double lamA = lamP * gx + lamQ * gy;
        const double lame = 2.0 * eps * lamE2 + epsd * lamA;
        const double lamd = eps * lamA;
        const double lamth = lame * epsd + lamd * ((double)n * mend * ca);
        // d theta / d(gx, gy) = (-gy, gx) / |g|^2, evaluated on the rescaled
        // gradient; the cotangent is multiplied in before the power-of-two
        // scale 2^-e, which overflows on its own for subnormal gradients (a
        // zero cotangent then gives 0, not 0 * inf)
        double thx = 0.0, thy = 0.0;
        const double mx = fmax(fabs(gx), fabs(gy));
        if (mx != 0.0) {
            int e;
            frexp(mx, &e);
            const double u = ldexp(gx, -e), v = ldexp(gy, -e);
            const double ir2 = 1.0 / (u * u + v * v);
            thx = ldexp(lamth * (-v * ir2), -e);
            thy = ldexp(lamth * (u * ir2), -e);
        }
        a.w1[D.cell_base + cl] = A * lamP + thx;
        a.w2[D.cell_base + cl] = A * lamQ + thy;
        a.w3[D.cell_base + cl] = eps * eps * mue;
        g[0] += lame * (epsb * ca) + lamd * (-epsb * (double)n * sa);
        g[2] += lame * (1.0 + delta * ca) + lamd * (-(double)n * delta * sa);
    }
    block_accumulate<SF_NG>(g, a.gacc, a.total_tiles, w, red);
}

// Kernel 3: lam_phi, lam_T of the state before the step.
template <int TS>
__global__ __launch_bounds__(256) void sf_adj3_kernel(AdjArgs a) {
    constexpr int PER = TS * TS / WG;
    const int w = blockIdx.x;
    const InstDesc& D = a.inst[a.tile_inst[w]];
    const size_t TC = a.total_cells;
    const double* phi = a.u + D.cell_base;
    const double* T = a.u + TC + D.cell_base;
    const double* lamT1 = a.lam_n + TC + D.cell_base;
    const double* mu = a.w0 + D.cell_base;
    const double* lgx = a.w1 + D.cell_base;
    const double* lgy = a.w2 + D.cell_base;
    const double* z = a.w3 + D.cell_base;
    const bool ref = D.stencil == AMPS_STENCIL_REF;
    const double inv2dx = D.k[SFK_INV_2DX], inv2dy = D.k[SFK_INV_2DY], inv = D.k[SFK_INV_DXDY];
    const double dt = D.k[SFK_DT], aop = D.k[SFK_ALPHA_OVER_PI], gam = D.k[SFK_GAMMA], teq = D.k[SFK_TEQ];
    const double noise = D.k[SFK_NOISE];
    const uint32_t key = noise_key(D.seed, (uint64_t)(*a.step_ctr + a.step_add));
    for (int i = 0; i < PER; ++i) {
        int x, y;
        tile_cell<TS>(D, w, i, x, y);
        const size_t cl = (size_t)y * D.nx + x;
        const double p = phi[cl], t = T[cl], m_u = mu[cl];
        const double arg = gam * (teq - t);
        const double m = aop * atan(arg);
        const double pp = p * (1.0 - p);
        double dR = (1.0 - 2.0 * p) * (p - 0.5 + m) + pp;
        if (noise != 0.0) {
            const double u = (double)noise_u24(key, (uint32_t)cl);
            dR += noise * (1.0 - 2.0 * p) * (u - 0.5);
        }
        double lp = a.lam_n[D.cell_base + cl] + m_u * dR;
        lp += inv2dx * (gather_x(lgx, x, y, 1, D) - gather_x(lgx, x, y, -1, D));
        lp += inv2dy * (gather_y(lgy, x, y, 1, D) - gather_y(lgy, x, y, -1, D));
        lp += inv * lapT(z, x, y, ref, D);
        const double dmdT = aop * (-gam) / (1.0 + arg * arg);
        const double lt = lamT1[cl] + m_u * pp * dmdT + dt * inv * lapT(lamT1, x, y, ref, D);
        a.lam[D.cell_base + cl] = lp;
        a.lam[TC + D.cell_base + cl] = lt;
    }
}

// ============================================================== sintering
// gradient slots: 0 Dsurf, 1 Dgb, 2 kappa_c, 3 kappa_eta, 4 L, 5 Dvol
constexpr int SN_NG = 6;

__device__ inline double sn_mu(const double* c, const double* e1, const double* e2, int x, int y,
                               const InstDesc& D) {
    const size_t cl = (size_t)y * D.nx + x;
    const double cv = c[cl], a1 = e1[cl], a2 = e2[cl];
    const double con_lap = lap_fwd(c, x, y, false, D.k[SNK_INV_DXDY], D);
    const double sum2 = a1 * a1 + a2 * a2, sum3 = a1 * a1 * a1 + a2 * a2 * a2, om = 1.0 - cv;
    const double B = D.k[SNK_B], A2 = D.k[SNK_A2];
    const double df = B * (2.0 * cv + 4.0 * sum3 - 6.0 * sum2) - A2 * (cv * cv) * om + A2 * cv * (om * om);
    return df - D.k[SNK_HALF_KC] * con_lap;
}

// Kernel 0: mu of the state before the step (w0).
template <int TS>
__global__ __launch_bounds__(256) void sn_adj0_kernel(AdjArgs a) {
    constexpr int PER = TS * TS / WG;
    const int w = blockIdx.x;
    const InstDesc& D = a.inst[a.tile_inst[w]];
    const size_t TC = a.total_cells;
    const double* c = a.u + D.cell_base;
    const double* e1 = a.u + TC + D.cell_base;
    const double* e2 = a.u + 2 * TC + D.cell_base;
    for (int i = 0; i < PER; ++i) {
        int x, y;
        tile_cell<TS>(D, w, i, x, y);
        a.w0[D.cell_base + (size_t)y * D.nx + x] = sn_mu(c, e1, e2, x, y, D);
    }
}

// Kernel 1: clamp masks, the cotangent kappa of c's pre-clamp update, nu_i of
// the etas', Y = kappa dt M (for lap^T into mu); partials of Dsurf, Dgb, Dvol,
// kappa_eta, L.
template <int TS>
__global__ __launch_bounds__(256) void sn_adj1_kernel(AdjArgs a) {
    constexpr int PER = TS * TS / WG;
    __shared__ double red[4 * SN_NG];
    const int w = blockIdx.x;
    const InstDesc& D = a.inst[a.tile_inst[w]];
    const size_t TC = a.total_cells;
    const double* c = a.u + D.cell_base;
    const double* e[2] = {a.u + TC + D.cell_base, a.u + 2 * TC + D.cell_base};
    const double* mu = a.w0 + D.cell_base;
    const double inv = D.k[SNK_INV_DXDY], dt = D.k[SNK_DT], B = D.k[SNK_B];
    const double dvol = D.k[SNK_DVOL], dvap = D.k[SNK_DVAP], dsur = D.k[SNK_DSURF], dgb = D.k[SNK_DGB];
    const double dtl = D.k[SNK_DT_L], hke = D.k[SNK_HALF_KE];
    double g[SN_NG] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < PER; ++i) {
        int x, y;
        tile_cell<TS>(D, w, i, x, y);
        const size_t cl = (size_t)y * D.nx + x;
        const double cv = c[cl], a1 = e[0][cl], a2 = e[1][cl];
        const double mu_lap = lap_fwd(mu, x, y, false, inv, D);
        const double h = cv * cv * cv * (10.0 - 15.0 * cv + 6.0 * (cv * cv));
        const double mob = dvol * h + dvap * (1.0 - h) + dsur * cv * (1.0 - cv) + dgb * (a1 * a2 * 2.0);
        const double v = cv + dt * mob * mu_lap;
        const double cn = v >= 1.0 ? 1.0 : (v < 0.0 ? 0.0 : v);
        const double gc = (v >= 0.0 && v < 1.0) ? 1.0 : 0.0;
        const double sum2 = a1 * a1 + a2 * a2;
        double lcn = a.lam_n[D.cell_base + cl];
        double nu[2];
        for (int k = 0; k < 2; ++k) {
            const double ak = k == 0 ? a1 : a2;
            const double l = lap_fwd(e[k], x, y, false, inv, D);
            const double dk = B * (-12.0 * (ak * ak) * (2.0 - cn) + 12.0 * ak * (1.0 - cn) + 12.0 * ak * sum2);
            const double wk = ak - dtl * (dk - hke * l);
            const double gk = (wk >= 0.0 && wk < 1.0) ? 1.0 : 0.0;
            nu[k] = a.lam_n[(size_t)(1 + k) * TC + D.cell_base + cl] * gk;
            lcn += nu[k] * (-dtl) * B * (12.0 * ak * ak - 12.0 * ak);
            g[3] += nu[k] * dtl * 0.5 * l;                 // d w_k / d kappa_eta
            g[4] += nu[k] * (-dt) * (dk - hke * l);        // d w_k / d L
        }
        const double kap = lcn * gc;
        a.w1[D.cell_base + cl] = kap * dt * mob;
        a.w2[D.cell_base + cl] = kap;
        a.w3[D.cell_base + cl] = nu[0];
        a.w4[D.cell_base + cl] = nu[1];
        const double q = kap * dt * mu_lap;
        g[0] += q * cv * (1.0 - cv);
        g[1] += q * (a1 * a2 * 2.0);
        g[5] += q * h;
    }
    block_accumulate<SN_NG>(g, a.gacc, a.total_tiles, w, red);
}

// Kernel 2: lam_mu = lap^T(Y) (w5); partial of kappa_c.
template <int TS>
__global__ __launch_bounds__(256) void sn_adj2_kernel(AdjArgs a) {
    constexpr int PER = TS * TS / WG;
    __shared__ double red[4 * SN_NG];
    const int w = blockIdx.x;
    const InstDesc& D = a.inst[a.tile_inst[w]];
    const double* c = a.u + D.cell_base;
    const double* Y = a.w1 + D.cell_base;
    const double inv = D.k[SNK_INV_DXDY];
    double g[SN_NG] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < PER; ++i) {
        int x, y;
        tile_cell<TS>(D, w, i, x, y);
        const size_t cl = (size_t)y * D.nx + x;
        const double lmu = inv * lapT(Y, x, y, false, D);
        a.w5[D.cell_base + cl] = lmu;
        g[2] += lmu * (-0.5) * lap_fwd(c, x, y, false, inv, D);
    }
    block_accumulate<SN_NG>(g, a.gacc, a.total_tiles, w, red);
}

// Kernel 3: lam_c, lam_eta1, lam_eta2 of the state before the step.
template <int TS>
__global__ __launch_bounds__(256) void sn_adj3_kernel(AdjArgs a) {
    constexpr int PER = TS * TS / WG;
    const int w = blockIdx.x;
    const InstDesc& D = a.inst[a.tile_inst[w]];
    const size_t TC = a.total_cells;
    const double* c = a.u + D.cell_base;
    const double* e1 = a.u + TC + D.cell_base;
    const double* e2 = a.u + 2 * TC + D.cell_base;
    const double* mu = a.w0 + D.cell_base;
    const double* kapa = a.w2 + D.cell_base;
    const double* nua[2] = {a.w3 + D.cell_base, a.w4 + D.cell_base};
    const double* lmua = a.w5 + D.cell_base;
    const double inv = D.k[SNK_INV_DXDY], dt = D.k[SNK_DT], B = D.k[SNK_B], A2 = D.k[SNK_A2];
    const double dvol = D.k[SNK_DVOL], dvap = D.k[SNK_DVAP], dsur = D.k[SNK_DSURF], dgb = D.k[SNK_DGB];
    const double dtl = D.k[SNK_DT_L], hke = D.k[SNK_HALF_KE], hkc = D.k[SNK_HALF_KC];
    for (int i = 0; i < PER; ++i) {
        int x, y;
        tile_cell<TS>(D, w, i, x, y);
        const size_t cl = (size_t)y * D.nx + x;
        const double cv = c[cl], a1 = e1[cl], a2 = e2[cl], om = 1.0 - cv;
        const double cn = a.un[D.cell_base + cl];
        const double mu_lap = lap_fwd(mu, x, y, false, inv, D);
        const double kap = kapa[cl], lmu = lmua[cl];
        const double h = cv * cv * cv * (10.0 - 15.0 * cv + 6.0 * (cv * cv));
        const double hp = 3.0 * cv * cv * (10.0 - 15.0 * cv + 6.0 * (cv * cv)) + cv * cv * cv * (-15.0 + 12.0 * cv);
        (void)h;
        const double dMdc = (dvol - dvap) * hp + dsur * (1.0 - 2.0 * cv);
        const double dfdc = 2.0 * B - A2 * (2.0 * cv * om - cv * cv) + A2 * (om * om - 2.0 * cv * om);
        const double q = kap * dt * mu_lap;
        const double lc = kap + q * dMdc + lmu * dfdc - hkc * inv * lapT(lmua, x, y, false, D);
        a.lam[D.cell_base + cl] = lc;
        const double sum2 = a1 * a1 + a2 * a2;
        for (int k = 0; k < 2; ++k) {
            const double ak = k == 0 ? a1 : a2, aj = k == 0 ? a2 : a1;
            const double nuk = nua[k][cl], nuj = nua[1 - k][cl];
            const double ddk = B * (-24.0 * ak * (2.0 - cn) + 12.0 * (1.0 - cn) + 12.0 * sum2 + 24.0 * ak * ak);
            double la = nuk * (1.0 - dtl * ddk) + nuj * (-dtl) * B * 24.0 * aj * ak;
            la += q * dgb * 2.0 * aj + lmu * B * (12.0 * ak * ak - 12.0 * ak);
            la += dtl * hke * inv * lapT(nua[k], x, y, false, D);
            a.lam[(size_t)(1 + k) * TC + D.cell_base + cl] = la;
        }
    }
}

// ------------------------------------------------------------ host launch
int launch_adjoint_step(const AdjArgs& a, int model, int tile, hipStream_t s) {
    const dim3 g(a.total_tiles), b(WG);
    if (tile != 16 && tile != 32) return AMPS_E_UNSUPPORTED;
    if (model == AMPS_SNOWFLAKE) {
        if (tile == 16) {
            hipLaunchKernelGGL(sf_adj1_kernel<16>, g, b, 0, s, a);
            hipLaunchKernelGGL(sf_adj2_kernel<16>, g, b, 0, s, a);
            hipLaunchKernelGGL(sf_adj3_kernel<16>, g, b, 0, s, a);
        } else {
            hipLaunchKernelGGL(sf_adj1_kernel<32>, g, b, 0, s, a);
            hipLaunchKernelGGL(sf_adj2_kernel<32>, g, b, 0, s, a);
            hipLaunchKernelGGL(sf_adj3_kernel<32>, g, b, 0, s, a);
        }
    } else if (model == AMPS_SINTERING) {
        if (tile == 16) {
            hipLaunchKernelGGL(sn_adj0_kernel<16>, g, b, 0, s, a);
            hipLaunchKernelGGL(sn_adj1_kernel<16>, g, b, 0, s, a);
            hipLaunchKernelGGL(sn_adj2_kernel<16>, g, b, 0, s, a);
            hipLaunchKernelGGL(sn_adj3_kernel<16>, g, b, 0, s, a);
        } else {
            hipLaunchKernelGGL(sn_adj0_kernel<32>, g, b, 0, s, a);
            hipLaunchKernelGGL(sn_adj1_kernel<32>, g, b, 0, s, a);
            hipLaunchKernelGGL(sn_adj2_kernel<32>, g, b, 0, s, a);
            hipLaunchKernelGGL(sn_adj3_kernel<32>, g, b, 0, s, a);
        }
    } else {
        return AMPS_E_UNSUPPORTED;
    }
    return hipGetLastError() == hipSuccess ? AMPS_OK : AMPS_E_HIP;
}

}  // namespace amps

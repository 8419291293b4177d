// k_sintering.hip — Wang solid-state sintering (c, eta_1, eta_2), one fused
// sweep per step.
//
// Reference: con1_pure -> con2_pure -> phi1_pure -> phi2_pure with a device
// synchronisation after each (sintering/function.h:51-159, main.cu:136-146):
//   mu   = B(2c + 4 sum eta^3 - 6 sum eta^2) - 2A c^2 (1-c) + 2A c (1-c)^2 - kc/2 lap c
//   M    = Dvol h + Dvap (1-h) + Dsurf c(1-c) + Dgb 2 eta1 eta2,  h = c^3 (10 - 15c + 6c^2)
//   c'   = clamp01(c + dt M lap mu)
//   eta' = clamp01(eta - dt L (B(-12 eta^2 (2-c') + 12 eta (1-c') + 12 eta sum eta^2) - ke/2 lap eta))
// with 5-point Laplacians, mu from the old state and the eta updates using
// the NEW c and the OLD etas.  Here a workgroup stages c with a 2-cell halo
// and the etas with a 1-cell halo, computes mu on the tile and its ring in
// LDS (FP64) and applies all three updates: 48 B per FP64 cell-update.
//
// c, mu and lap mu are always FP64 (the chemical-potential divergence is the
// precision-critical term).  In low tiles the etas are stored in fp16 with a
// per-tile offset and scale: 16 + 2 x 4 = 24 B per cell-update.
#include "kernels.hpp"
#include "tile_common.hpp"

namespace amps {

template <int TS, bool CONS>
struct SnLds {
    static constexpr int LWC = TS + 4, LWE = TS + 2, RW = TS + 2;
    double c[LWC * LWC];
    double e1[LWE * LWE], e2[LWE * LWE];
    double mu[RW * RW];
    double mob[CONS ? RW * RW : 1];  // divergence form: mobility on the tile and its ring
    NbField nbf[2];
    double red[4];
    float redf[4 * WG / 64];  // eta bounds of a low tile (block_minmax2)
};

// Mobility M = Dvol h + Dvap (1 - h) + Dsurf c(1-c) + Dgb 2 eta1 eta2,
// h = c^3 (10 - 15c + 6c^2) (sintering/function.h:97-99).
__device__ inline double sn_mobility(double cv, double a1, double a2, double dvol, double dvap, double dsur,
                                     double dgb) {
    const double h = cv * cv * cv * (10.0 - 15.0 * cv + 6.0 * (cv * cv));
    return dvol * h + dvap * (1.0 - h) + dsur * cv * (1.0 - cv) + dgb * (a1 * a2 * 2.0);
}

__device__ inline int ring_src_sn(int r, int o, int n, int bc) {
    if (bc == AMPS_BC_PERIODIC) return r;
    int g = o - 1 + r;
    if (g < 0) g = -g;
    else if (g >= n) g = 2 * (n - 1) - g;
    return g - o + 1;
}

__device__ inline double clamp01_ref(double v) {
    // sintering/function.h:100-101: >= 1 -> 1, < 0 -> 0
    if (v >= 1.0) return 1.0;
    if (v < 0.0) return 0.0;
    return v;
}

// Read-only metadata (descriptors, tile -> instance map, tile types) arrive
// as restrict-qualified arguments so that the per-tile reads go through the
// scalar cache; with MULTI = false the descriptor is loop-invariant.
//
// CONS: the Cahn-Hilliard update in divergence form, c' = clamp01(c + dt
// div(M grad mu)), with M also evaluated on the ring and face fluxes
// F = 0.5 (M_left + M_right) (mu_right - mu_left); a face shared by two cells
// (in two tiles too) gives both the bitwise same flux, so the sum of c is
// conserved up to the clamp and the final roundings.
template <int TS, bool ADAPT, bool MULTI, bool CONS>
__global__ __launch_bounds__(256) void sn_step_kernel(StepArgs a, const InstDesc* __restrict__ inst,
                                                      const int* __restrict__ tinst, const int* __restrict__ types) {
    constexpr int LWC = TS + 4, LWE = TS + 2, RW = TS + 2, PER = TS * TS / WG;
    __shared__ SnLds<TS, CONS> L;
    for (int b = blockIdx.x; b < a.nwork; b += gridDim.x) {
        const int w = xcd_remap(b, a.nwork);
        const InstDesc& d = inst[MULTI ? tinst[w] : 0];
        const int lt = w - d.tile_base;
        const int tile_y = lt / d.tx, tile_x = lt - tile_y * d.tx;
        const int ox = tile_x * TS, oy = tile_y * TS;
        Nb nb;
        nb_setup(nb, d, tile_x, tile_y, TS);
        const int ty = ADAPT ? types[w] : 2;
        BoxMap<TS, 2> mc;
        BoxMap<TS, 1> me;
        box_map<TS, 2>(mc, d, tile_x, tile_y);
        box_map<TS, 1>(me, d, tile_x, tile_y);
        const double* fc = a.in.f64 + d.cell_base;
        const double* f1 = a.in.f64 + a.total_cells + d.cell_base;
        const double* f2 = a.in.f64 + 2 * a.total_cells + d.cell_base;
        if (ADAPT) {
            // one memory round: c (always FP64), the etas of the own tile from
            // its current storage, the 1-cell halo ring from both storages
            // (selected by the neighbour table loaded alongside)
            constexpr int NC = BoxMap<TS, 2>::N, NE = BoxMap<TS, 1>::N;
            nbfield_fill2(&L.nbf[0], &L.nbf[1], a, a.in, nb, 1, 2, 0);
            const __half* h1 = a.in.f16 + a.total_cells + d.cell_base;
            const __half* h2 = a.in.f16 + 2 * a.total_cells + d.cell_base;
            double vc[NC], v1[NE], v2[NE];
            __half q1[NE], q2[NE];
#pragma unroll
            for (int k = 0; k < NC; ++k) if (mc.off[k] >= 0) vc[k] = fc[mc.off[k]];
#pragma unroll
            for (int k = 0; k < NE; ++k) {
                if (me.off[k] < 0) continue;
                const bool own = me.reg[k] == 4;
                if (!own || ty == 2) { v1[k] = f1[me.off[k]]; v2[k] = f2[me.off[k]]; }
                if (!own || ty != 2) { q1[k] = h1[me.off[k]]; q2[k] = h2[me.off[k]]; }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < NC; ++k) if (mc.off[k] >= 0) L.c[threadIdx.x + k * WG] = vc[k];
#pragma unroll
            for (int k = 0; k < NE; ++k) {
                if (me.off[k] < 0) continue;
                const int r = me.reg[k];
                if (L.nbf[0].type[r] == 2) {
                    L.e1[threadIdx.x + k * WG] = v1[k];
                    L.e2[threadIdx.x + k * WG] = v2[k];
                } else {
                    L.e1[threadIdx.x + k * WG] = L.nbf[0].off[r] + (double)__half2float(q1[k]) * (double)L.nbf[0].scl[r];
                    L.e2[threadIdx.x + k * WG] = L.nbf[1].off[r] + (double)__half2float(q2[k]) * (double)L.nbf[1].scl[r];
                }
            }
        } else {
            // all three fields' loads in flight together
            constexpr int NC = BoxMap<TS, 2>::N, NE = BoxMap<TS, 1>::N;
            double vc[NC], v1[NE], v2[NE];
#pragma unroll
            for (int k = 0; k < NC; ++k) if (mc.off[k] >= 0) vc[k] = fc[mc.off[k]];
#pragma unroll
            for (int k = 0; k < NE; ++k) if (me.off[k] >= 0) { v1[k] = f1[me.off[k]]; v2[k] = f2[me.off[k]]; }
#pragma unroll
            for (int k = 0; k < NC; ++k) if (mc.off[k] >= 0) L.c[threadIdx.x + k * WG] = vc[k];
#pragma unroll
            for (int k = 0; k < NE; ++k) if (me.off[k] >= 0) { L.e1[threadIdx.x + k * WG] = v1[k]; L.e2[threadIdx.x + k * WG] = v2[k]; }
        }
        __syncthreads();

        const double inv = d.k[SNK_INV_DXDY], A2 = d.k[SNK_A2], B = d.k[SNK_B], hkc = d.k[SNK_HALF_KC];
        const bool edge = d.bc != AMPS_BC_PERIODIC && (ox == 0 || oy == 0 || ox + TS == d.nx || oy + TS == d.ny);
        // phase 1: mu on the tile and its ring (sintering/function.h:67-72)
        for (int e = threadIdx.x; e < RW * RW; e += WG) {
            int ry = e / RW, rx = e - ry * RW;
            int sy = edge ? ring_src_sn(ry, oy, d.ny, d.bc) : ry, sx = edge ? ring_src_sn(rx, ox, d.nx, d.bc) : rx;
            int cc = (sy + 1) * LWC + sx + 1;
            int ce = sy * LWE + sx;
            double cv = L.c[cc];
            double con_lap = (L.c[cc - 1] + L.c[cc + 1] + L.c[cc - LWC] + L.c[cc + LWC] - 4.0 * cv) * inv;
            double a1 = L.e1[ce], a2 = L.e2[ce];
            double sum2 = a1 * a1 + a2 * a2;
            double sum3 = a1 * a1 * a1 + a2 * a2 * a2;
            double om = 1.0 - cv;
            double df = B * (2.0 * cv + 4.0 * sum3 - 6.0 * sum2) - A2 * (cv * cv) * om + A2 * cv * (om * om);
            L.mu[e] = df - hkc * con_lap;
            if (CONS)
                L.mob[e] = sn_mobility(cv, a1, a2, d.k[SNK_DVOL], d.k[SNK_DVAP], d.k[SNK_DSURF], d.k[SNK_DGB]);
        }
        __syncthreads();

        // phase 2: c (con2_pure), then eta1, eta2 (phi1_pure, phi2_pure)
        const double dt = d.k[SNK_DT], dvol = d.k[SNK_DVOL], dvap = d.k[SNK_DVAP], dsur = d.k[SNK_DSURF];
        const double dgb = d.k[SNK_DGB], dtl = d.k[SNK_DT_L], hke = d.k[SNK_HALF_KE];
        double n1[PER], n2[PER];
        size_t gcell[PER];
        double* oc = a.out.f64 + d.cell_base;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            int e = threadIdx.x + i * WG;
            int ly = e / TS, lx = e - ly * TS;
            gcell[i] = (size_t)(oy + ly) * d.nx + ox + lx;
            int rc = (ly + 1) * RW + lx + 1;
            double cv = L.c[(ly + 2) * LWC + lx + 2];
            int ce = (ly + 1) * LWE + lx + 1;
            double a1 = L.e1[ce], a2 = L.e2[ce];
            double cn;
            if (CONS) {
                // faces (left a, right b): 0.5 (M_a + M_b) (mu_b - mu_a)
                const double* M = L.mob;
                const double* U = L.mu;
                const double fe = 0.5 * (M[rc] + M[rc + 1]) * (U[rc + 1] - U[rc]);
                const double fw = 0.5 * (M[rc - 1] + M[rc]) * (U[rc] - U[rc - 1]);
                const double fn = 0.5 * (M[rc] + M[rc + RW]) * (U[rc + RW] - U[rc]);
                const double fs = 0.5 * (M[rc - RW] + M[rc]) * (U[rc] - U[rc - RW]);
                cn = clamp01_ref(cv + dt * (((fe - fw) + (fn - fs)) * inv));
            } else {
                double mu_lap = (L.mu[rc - 1] + L.mu[rc + 1] + L.mu[rc - RW] + L.mu[rc + RW] - 4.0 * L.mu[rc]) * inv;
                double mob = sn_mobility(cv, a1, a2, dvol, dvap, dsur, dgb);
                cn = clamp01_ref(cv + dt * mob * mu_lap);
            }
            oc[gcell[i]] = cn;
            double sum2 = a1 * a1 + a2 * a2;
            double l1 = (L.e1[ce - 1] + L.e1[ce + 1] + L.e1[ce - LWE] + L.e1[ce + LWE] - 4.0 * a1) * inv;
            double l2 = (L.e2[ce - 1] + L.e2[ce + 1] + L.e2[ce - LWE] + L.e2[ce + LWE] - 4.0 * a2) * inv;
            double d1 = B * (-12.0 * (a1 * a1) * (2.0 - cn) + 12.0 * a1 * (1.0 - cn) + 12.0 * a1 * sum2);
            double d2 = B * (-12.0 * (a2 * a2) * (2.0 - cn) + 12.0 * a2 * (1.0 - cn) + 12.0 * a2 * sum2);
            n1[i] = clamp01_ref(a1 - dtl * (d1 - hke * l1));
            n2[i] = clamp01_ref(a2 - dtl * (d2 - hke * l2));
        }
        if (ty == 2) {
            double* o1 = a.out.f64 + a.total_cells + d.cell_base;
            double* o2 = a.out.f64 + 2 * a.total_cells + d.cell_base;
#pragma unroll
            for (int i = 0; i < PER; ++i) { o1[gcell[i]] = n1[i]; o2[gcell[i]] = n2[i]; }
        } else {
            // encode both etas of the tile (offset = mid-range, power-of-two
            // scale).  The bounds are the values' min and max rounded outward to
            // float (a range that covers every value), all four reduced over the
            // workgroup with one barrier.
            float lo0 = INFINITY, hi0 = -INFINITY, lo1 = INFINITY, hi1 = -INFINITY;
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                lo0 = fminf(lo0, __double2float_rd(n1[i]));
                hi0 = fmaxf(hi0, __double2float_ru(n1[i]));
                lo1 = fminf(lo1, __double2float_rd(n2[i]));
                hi1 = fmaxf(hi1, __double2float_ru(n2[i]));
            }
            block_minmax2(lo0, hi0, lo1, hi1, L.redf);
            for (int k = 0; k < 2; ++k) {
                const double* v = k == 0 ? n1 : n2;
                const double mn = k == 0 ? lo0 : lo1, mx = k == 0 ? hi0 : hi1;
                double off = 0.5 * (mn + mx), hw = 0.5 * (mx - mn);
                int ex = hw == 0.0 ? -100 : 0;  // a uniform tile takes the smallest scale
                if (hw > 0.0) frexp(hw, &ex);
                if (ex < -100) ex = -100;
                const float invs = ldexpf(1.0f, 14 - ex);
                __half* o = a.out.f16 + (size_t)(1 + k) * a.total_cells + d.cell_base;
#pragma unroll
                for (int i = 0; i < PER; ++i) o[gcell[i]] = __float2half_rn((float)(v[i] - off) * invs);
                if (threadIdx.x == 0) {
                    a.out.off[(size_t)(1 + k) * a.total_tiles + w] = off;
                    a.out.scl[(size_t)(1 + k) * a.total_tiles + w] = hw == 0.0 ? 0.0f : ldexpf(1.0f, ex - 14);
                }
            }
        }
        __syncthreads();
    }
}

template <int TS, bool ADAPT, bool CONS>
static void sn_launch2(const StepArgs& a, int grid, hipStream_t s, const hipEvent_t* ev) {
    if (a.ninst > 1)
        launch_sweep(sn_step_kernel<TS, ADAPT, true, CONS>, dim3(grid), dim3(WG), s, ev, a, a.inst, a.tile_inst, a.type);
    else
        launch_sweep(sn_step_kernel<TS, ADAPT, false, CONS>, dim3(grid), dim3(WG), s, ev, a, a.inst, a.tile_inst, a.type);
}

template <int TS, bool ADAPT>
static void sn_launch(const StepArgs& a, int grid, hipStream_t s, const hipEvent_t* ev) {
    if (a.flags & AMPS_FLAG_CH_CONSERVATIVE) sn_launch2<TS, ADAPT, true>(a, grid, s, ev);
    else sn_launch2<TS, ADAPT, false>(a, grid, s, ev);
}

int launch_sn_step(const StepArgs& a, int tile, bool adaptive, int grid, hipStream_t s, const hipEvent_t* ev) {
    if (tile == 16) {
        if (adaptive) sn_launch<16, true>(a, grid, s, ev);
        else sn_launch<16, false>(a, grid, s, ev);
    } else if (tile == 32) {
        if (adaptive) sn_launch<32, true>(a, grid, s, ev);
        else sn_launch<32, false>(a, grid, s, ev);
    } else {
        return AMPS_E_UNSUPPORTED;
    }
    return hipGetLastError() == hipSuccess ? AMPS_OK : AMPS_E_HIP;
}

}  // namespace amps

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

// Ring element r of the box with halo H around a TS x TS tile (box width
// LW = TS + 2H): the H top rows, the H bottom rows, then the H left and H
// right cells of each of the TS middle rows.
template <int TS, int H>
__device__ inline void sn_ring_coord(int r, int& ly, int& lx) {
    constexpr int LW = TS + 2 * H;
    if (r < 2 * H * LW) {
        const int row = r / LW;
        lx = r - row * LW;
        ly = row < H ? row : TS + row;
    } else {
        const int q = r - 2 * H * LW, row = q / (2 * H), j = q - row * (2 * H);
        ly = H + row;
        lx = j < H ? j : TS + j;
    }
}

// Global cell offset (within the instance) of box element (ly, lx) of the
// box with halo H whose tile starts at (ox, oy), boundary-mapped.
template <int H>
__device__ inline int sn_box_cell(const InstDesc& d, int ox, int oy, int ly, int lx) {
    return bmap(oy - H + ly, d.ny, d.bc) * d.nx + bmap(ox - H + lx, d.nx, d.bc);
}

// Phase 1: mu (and with CONS the mobility) on the tile and its ring
// (sintering/function.h:67-72), from the staged c and eta boxes.
template <int TS, bool CONS>
__device__ inline void sn_phase1(SnLds<TS, CONS>& L, const InstDesc& d, int ox, int oy) {
    constexpr int LWC = TS + 4, LWE = TS + 2, RW = TS + 2;
    const double inv = d.k[SNK_INV_DXDY], A2 = d.k[SNK_A2], B = d.k[SNK_B], hkc = d.k[SNK_HALF_KC];
    const bool edge = d.bc != AMPS_BC_PERIODIC && (ox == 0 || oy == 0 || ox + TS == d.nx || oy + TS == d.ny);
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
        if (CONS) L.mob[e] = sn_mobility(cv, a1, a2, d.k[SNK_DVOL], d.k[SNK_DVAP], d.k[SNK_DSURF], d.k[SNK_DGB]);
    }
}

// Phase 2 for tile cell (ly, lx): the new c (con2_pure), then eta1, eta2
// from the new c and the old etas (phi1_pure, phi2_pure).
template <int TS, bool CONS>
__device__ inline void sn_phase2_cell(const SnLds<TS, CONS>& L, const InstDesc& d, int ly, int lx, double& cn,
                                      double& n1, double& n2) {
    constexpr int LWC = TS + 4, LWE = TS + 2, RW = TS + 2;
    const double inv = d.k[SNK_INV_DXDY], B = d.k[SNK_B];
    const double dt = d.k[SNK_DT], dtl = d.k[SNK_DT_L], hke = d.k[SNK_HALF_KE];
    const int rc = (ly + 1) * RW + lx + 1;
    const double cv = L.c[(ly + 2) * LWC + lx + 2];
    const int ce = (ly + 1) * LWE + lx + 1;
    const double a1 = L.e1[ce], a2 = L.e2[ce];
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
        const double mu_lap = (L.mu[rc - 1] + L.mu[rc + 1] + L.mu[rc - RW] + L.mu[rc + RW] - 4.0 * L.mu[rc]) * inv;
        const double mob = sn_mobility(cv, a1, a2, d.k[SNK_DVOL], d.k[SNK_DVAP], d.k[SNK_DSURF], d.k[SNK_DGB]);
        cn = clamp01_ref(cv + dt * mob * mu_lap);
    }
    const double sum2 = a1 * a1 + a2 * a2;
    const double l1 = (L.e1[ce - 1] + L.e1[ce + 1] + L.e1[ce - LWE] + L.e1[ce + LWE] - 4.0 * a1) * inv;
    const double l2 = (L.e2[ce - 1] + L.e2[ce + 1] + L.e2[ce - LWE] + L.e2[ce + LWE] - 4.0 * a2) * inv;
    const double d1 = B * (-12.0 * (a1 * a1) * (2.0 - cn) + 12.0 * a1 * (1.0 - cn) + 12.0 * a1 * sum2);
    const double d2 = B * (-12.0 * (a2 * a2) * (2.0 - cn) + 12.0 * a2 * (1.0 - cn) + 12.0 * a2 * sum2);
    n1 = clamp01_ref(a1 - dtl * (d1 - hke * l1));
    n2 = clamp01_ref(a2 - dtl * (d2 - hke * l2));
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
//
// FP64 sweep (every tile FP64): box elements e = threadIdx.x + k*WG of the
// c box (halo 2) and the eta boxes (halo 1), all three fields' loads in
// flight together.
template <int TS, bool MULTI, bool CONS>
__global__ __launch_bounds__(256) void sn_f64_kernel(StepArgs a, const InstDesc* __restrict__ inst,
                                                     const int* __restrict__ tinst) {
    constexpr int PER = TS * TS / WG;
    __shared__ SnLds<TS, CONS> L;
    for (int b = blockIdx.x; b < a.nwork; b += gridDim.x) {
        const int w = xcd_remap(b, a.nwork);
        const InstDesc& d = inst[MULTI ? tinst[w] : 0];
        const int lt = w - d.tile_base;
        const int tile_y = lt / d.tx, tile_x = lt - tile_y * d.tx;
        const int ox = tile_x * TS, oy = tile_y * TS;
        BoxMap<TS, 2> mc;
        BoxMap<TS, 1> me;
        box_map<TS, 2>(mc, d, tile_x, tile_y);
        box_map<TS, 1>(me, d, tile_x, tile_y);
        const double* fc = a.in.f64 + d.cell_base;
        const double* f1 = a.in.f64 + a.total_cells + d.cell_base;
        const double* f2 = a.in.f64 + 2 * a.total_cells + d.cell_base;
        {
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
        sn_phase1<TS, CONS>(L, d, ox, oy);
        __syncthreads();
        double* oc = a.out.f64 + d.cell_base;
        double* o1 = a.out.f64 + a.total_cells + d.cell_base;
        double* o2 = a.out.f64 + 2 * a.total_cells + d.cell_base;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = threadIdx.x + i * WG;
            const int ly = e / TS, lx = e - ly * TS;
            const size_t g = (size_t)(oy + ly) * d.nx + ox + lx;
            double cn, n1, n2;
            sn_phase2_cell<TS, CONS>(L, d, ly, lx, cn, n1, n2);
            oc[g] = cn;
            o1[g] = n1;
            o2[g] = n2;
        }
        __syncthreads();
    }
}

// Adaptive sweep, one workgroup per tile (the grid is the work count: with
// no loop over tiles nothing stays live from one tile to the next, and the
// kernel neither spills nor waits on a tile it has finished).  c is FP64 on
// every tile; the etas of a low tile are fp16 with the tile's offset and
// scale.  Staging, one memory round with no branch around a load (a branch
// join would wait for the loads inside it):
// - a thread's R = TS*TS/WG own cells are adjacent in one tile row: c with
//   vector loads; the etas from the address the tile's type selects (R
//   doubles, or the aligned fp16 group holding the R words);
// - the rings (c: 2 cells, etas: 1 cell) are one element per thread (threads
//   past a ring repeat element 0 and store nothing); a ring eta's neighbour
//   tile type, offsets and scales are loaded first, and the eta itself from
//   the storage that type selects once it arrives (the own cells and the c
//   ring are in flight meanwhile), then decoded in registers.
// One barrier then separates staging from phase 1, as in the FP64 sweep.
template <int TS, bool MULTI, bool CONS>
__global__ __launch_bounds__(256) void sn_amp_kernel(StepArgs a, const InstDesc* __restrict__ inst,
                                                     const int* __restrict__ tinst, const int* __restrict__ types) {
    constexpr int LWC = TS + 4, LWE = TS + 2;
    constexpr int R = TS * TS / WG, TPR = TS / R, PER = R;
    constexpr int RNC = LWC * LWC - TS * TS, RNE = LWE * LWE - TS * TS, KC = (RNC + WG - 1) / WG;
    static_assert(RNE <= WG, "one eta ring element per thread");
    static_assert(R == 1 || R == 4, "strips of 1 or 4 cells");
    __shared__ SnLds<TS, CONS> L;
    const int b = blockIdx.x;
    if (b >= a.nwork) return;
    const int w = xcd_remap(b, a.nwork);
    const InstDesc& d = inst[MULTI ? tinst[w] : 0];
    const int lt = w - d.tile_base;
    const int tile_y = lt / d.tx, tile_x = lt - tile_y * d.tx;
    const int ox = tile_x * TS, oy = tile_y * TS;
    const int ty = types[w];
    const int sly = threadIdx.x / TPR, sx0 = (threadIdx.x % TPR) * R;
    const int g0 = (oy + sly) * d.nx + ox + sx0;  // first own cell of the strip
    const size_t T = (size_t)a.total_tiles;
    const double* fc = a.in.f64 + d.cell_base;
    const double* f1 = a.in.f64 + a.total_cells + d.cell_base;
    const double* f2 = a.in.f64 + 2 * a.total_cells + d.cell_base;
    const unsigned short* h1 = reinterpret_cast<const unsigned short*>(a.in.f16 + a.total_cells + d.cell_base);
    const unsigned short* h2 = reinterpret_cast<const unsigned short*>(a.in.f16 + 2 * a.total_cells + d.cell_base);
    // ---- the ring eta element's neighbour tile: its type, offsets and scales
    // first (the ring eta load below waits for the type alone: loads return
    // in issue order), then everything that does not depend on them
    const bool edge = ox < 2 || oy < 2 || ox + TS + 2 > d.nx || oy + TS + 2 > d.ny;
    int reb, regi, nty;
    double no1, no2;
    float ns1, ns2;
    {
        int bly = 0, blx = 0;
        sn_ring_coord<TS, 1>(threadIdx.x < RNE ? (int)threadIdx.x : 0, bly, blx);
        reb = threadIdx.x < RNE ? bly * LWE + blx : -1;
        regi = edge ? sn_box_cell<1>(d, ox, oy, bly, blx) : (oy - 1 + bly) * d.nx + ox - 1 + blx;
        Nb nb;
        nb_setup(nb, d, tile_x, tile_y, TS);
        const int g = nb_tile(nb, box_region<TS, 1>(bly, blx));
        nty = types[g];
        no1 = a.in.off[T + g];
        no2 = a.in.off[2 * T + g];
        ns1 = a.in.scl[T + g];
        ns2 = a.in.scl[2 * T + g];
    }
    // ---- own cells
    double oc[R];
    if (R == 4) {
        const double2 u0 = *reinterpret_cast<const double2*>(fc + g0);
        const double2 u1 = *reinterpret_cast<const double2*>(fc + g0 + 2);
        oc[0] = u0.x; oc[1 % R] = u0.y; oc[2 % R] = u1.x; oc[3 % R] = u1.y;
    } else {
        oc[0] = fc[g0];
    }
    const bool own64 = ty == 2;
    uint4 p1[2], p2[2];
    if (R == 4) {
        // FP64: R doubles in two 16-byte loads; fp16: the aligned 8-word group,
        // loaded twice
        const uint4* a1 = own64 ? reinterpret_cast<const uint4*>(f1 + g0) : reinterpret_cast<const uint4*>(h1 + (g0 & ~7));
        const uint4* a2 = own64 ? reinterpret_cast<const uint4*>(f2 + g0) : reinterpret_cast<const uint4*>(h2 + (g0 & ~7));
        p1[0] = a1[0];
        p1[1] = a1[own64 ? 1 : 0];
        p2[0] = a2[0];
        p2[1] = a2[own64 ? 1 : 0];
    } else {
        const uint2 u1 = *(own64 ? reinterpret_cast<const uint2*>(f1 + g0) : reinterpret_cast<const uint2*>(h1 + (g0 & ~3)));
        const uint2 u2 = *(own64 ? reinterpret_cast<const uint2*>(f2 + g0) : reinterpret_cast<const uint2*>(h2 + (g0 & ~3)));
        p1[0] = p1[1] = make_uint4(u1.x, u1.y, 0u, 0u);
        p2[0] = p2[1] = make_uint4(u2.x, u2.y, 0u, 0u);
    }
    const double oo1 = a.in.off[T + w], oo2 = a.in.off[2 * T + w];
    const float os1 = a.in.scl[T + w], os2 = a.in.scl[2 * T + w];
    // ---- c ring (boxes crossing the grid edge need the boundary map)
    double rc[KC];
    int rcb[KC];
#pragma unroll
    for (int k = 0; k < KC; ++k) {
        const int r = threadIdx.x + k * WG;
        int bly = 0, blx = 0;
        sn_ring_coord<TS, 2>(r < RNC ? r : 0, bly, blx);
        rcb[k] = r < RNC ? bly * LWC + blx : -1;
        const int gi = edge ? sn_box_cell<2>(d, ox, oy, bly, blx) : (oy - 2 + bly) * d.nx + ox - 2 + blx;
        rc[k] = fc[gi];
    }
    // ---- eta ring: only the storage the neighbour's type says is current
    // (the other copy is stale, and a ring column reads a whole memory line
    // per element)
    const bool n64 = nty == 2;
    const uint2 r1 = *(n64 ? reinterpret_cast<const uint2*>(f1 + regi) : reinterpret_cast<const uint2*>(h1 + (regi & ~3)));
    const uint2 r2 = *(n64 ? reinterpret_cast<const uint2*>(f2 + regi) : reinterpret_cast<const uint2*>(h2 + (regi & ~3)));
    // ---- box staging
    {
        double* cp = &L.c[(sly + 2) * LWC + sx0 + 2];
#pragma unroll
        for (int j = 0; j < R; ++j) cp[j] = oc[j];
#pragma unroll
        for (int k = 0; k < KC; ++k)
            if (rcb[k] >= 0) L.c[rcb[k]] = rc[k];
        double* q1 = &L.e1[(sly + 1) * LWE + sx0 + 1];
        double* q2 = &L.e2[(sly + 1) * LWE + sx0 + 1];
        if (own64) {
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const uint4 u1 = p1[j / 2], u2 = p2[j / 2];
                q1[j] = (j & 1) ? raw_f64(u1.z, u1.w) : raw_f64(u1.x, u1.y);
                q2[j] = (j & 1) ? raw_f64(u2.z, u2.w) : raw_f64(u2.x, u2.y);
            }
        } else {
            const uint2 w1 = R == 4 ? raw_words8(p1[0], g0) : make_uint2(p1[0].x, p1[0].y);
            const uint2 w2 = R == 4 ? raw_words8(p2[0], g0) : make_uint2(p2[0].x, p2[0].y);
#pragma unroll
            for (int j = 0; j < R; ++j) {
                q1[j] = oo1 + (double)__half2float(__ushort_as_half((unsigned short)raw_word4(w1, g0 + j))) * (double)os1;
                q2[j] = oo2 + (double)__half2float(__ushort_as_half((unsigned short)raw_word4(w2, g0 + j))) * (double)os2;
            }
        }
        if (reb >= 0) {
            if (n64) {
                L.e1[reb] = raw_f64(r1.x, r1.y);
                L.e2[reb] = raw_f64(r2.x, r2.y);
            } else {
                L.e1[reb] = no1 + (double)__half2float(__ushort_as_half((unsigned short)raw_word4(r1, regi))) * (double)ns1;
                L.e2[reb] = no2 + (double)__half2float(__ushort_as_half((unsigned short)raw_word4(r2, regi))) * (double)ns2;
            }
        }
    }
    __syncthreads();
    sn_phase1<TS, CONS>(L, d, ox, oy);
    __syncthreads();
    double nc[PER], n1[PER], n2[PER];
    int gc[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int e = threadIdx.x + i * WG;
        const int ly = e / TS, lx = e - ly * TS;
        gc[i] = (oy + ly) * d.nx + ox + lx;
        sn_phase2_cell<TS, CONS>(L, d, ly, lx, nc[i], n1[i], n2[i]);
    }
    double* oc_ = a.out.f64 + d.cell_base;
#pragma unroll
    for (int i = 0; i < PER; ++i) oc_[gc[i]] = nc[i];
    if (own64) {
        double* o1 = a.out.f64 + a.total_cells + d.cell_base;
        double* o2 = a.out.f64 + 2 * a.total_cells + d.cell_base;
#pragma unroll
        for (int i = 0; i < PER; ++i) { o1[gc[i]] = n1[i]; o2[gc[i]] = n2[i]; }
    } else {
        // encode both etas of the tile (offset = mid-range, power-of-two
        // scale).  The bounds are the values' min and max rounded outward to
        // float (a range that covers every value), all four reduced over the
        // workgroup with one barrier.
        double m1 = n1[0], x1 = n1[0], m2 = n2[0], x2 = n2[0];
#pragma unroll
        for (int i = 1; i < PER; ++i) {
            m1 = fmin(m1, n1[i]); x1 = fmax(x1, n1[i]);
            m2 = fmin(m2, n2[i]); x2 = fmax(x2, n2[i]);
        }
        float lo0 = __double2float_rd(m1), hi0 = __double2float_ru(x1);
        float lo1 = __double2float_rd(m2), hi1 = __double2float_ru(x2);
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
            for (int i = 0; i < PER; ++i) o[gc[i]] = __float2half_rn((float)(v[i] - off) * invs);
            if (threadIdx.x == 0) {
                a.out.off[(size_t)(1 + k) * a.total_tiles + w] = off;
                a.out.scl[(size_t)(1 + k) * a.total_tiles + w] = hw == 0.0 ? 0.0f : ldexpf(1.0f, ex - 14);
            }
        }
    }
}

template <int TS, bool ADAPT, bool CONS>
static void sn_launch2(const StepArgs& a, int grid, hipStream_t s, const hipEvent_t* ev) {
    if (ADAPT) {
        // one workgroup per tile
        if (a.ninst > 1)
            launch_sweep(sn_amp_kernel<TS, true, CONS>, dim3(a.nwork), dim3(WG), s, ev, a, a.inst, a.tile_inst, a.type);
        else
            launch_sweep(sn_amp_kernel<TS, false, CONS>, dim3(a.nwork), dim3(WG), s, ev, a, a.inst, a.tile_inst, a.type);
    } else {
        if (a.ninst > 1)
            launch_sweep(sn_f64_kernel<TS, true, CONS>, dim3(grid), dim3(WG), s, ev, a, a.inst, a.tile_inst);
        else
            launch_sweep(sn_f64_kernel<TS, false, CONS>, dim3(grid), dim3(WG), s, ev, a, a.inst, a.tile_inst);
    }
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

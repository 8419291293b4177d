// k_graingrowth.hip — Fan-Chen grain growth with q order parameters, one
// fused sweep per step for all fields.
//
// Reference (q = 2): kernel1_pure, graingrowth/function.h:5-24, launched once
// per field per step (purehigh.cu:31-32):
//   lap   = (sum of the 8 Moore neighbours - 8 eta) / dxdy
//   df    = 2 eta sum_{j!=i} eta_j^2 + eta^3 - eta          (alpha=beta=gamma=1)
//   eta' = eta - dt L (df - kappa lap)
// Here every field of a tile is advanced by one wave.  The per-cell sum of
// squares S = sum_j eta_j^2 of the current state is carried in an FP64 side
// array written by the previous sweep, so sum_{j!=i} = S - eta_i^2 costs O(1)
// per field instead of O(q): traffic per cell-update is 16 B per FP64 field
// (4 B per fp16 field) + 16 B for S.
//
// Tile types are per (field, tile): a field that is ~0 across a tile (a grain
// that is absent there) stays fp16 while the few fields forming the local
// grain boundaries run in FP64.
//
// Work mapping (16x16 tiles): a workgroup is ONE wave and owns one tile; it
// walks the q fields of the tile with no workgroup barrier (a one-wave
// barrier is free), each lane updating a column strip of R = TS*TS/64 cells.
// Per field the wave stages the (TS+2)^2 box in LDS (8 KB per wave with the
// neighbour tables), so many independent waves per CU hide each other's load
// latency; a low-precision field's re-centring needs only a wave reduction.
// 32x32 tiles use a 256-thread workgroup per tile with the same arithmetic.
#include "kernels.hpp"
#include "tile_common.hpp"

namespace amps {

constexpr int GW = 64;   // one wave per workgroup

// One-wave workgroup, 16x16 tile: lane l owns the row strip of 4 cells
// (row l/4, columns 4(l%4) .. +3).  LDS rows are padded so that each lane's
// strip and its neighbours are read with aligned 16-byte loads.
struct GgWaveLds {
    static constexpr int TS = 16, LWD = 20, C0D = 2, LWF = 24, C0F = 4;
    union {
        double d[(TS + 2) * LWD];
        float f[(TS + 2) * LWF];
    };
    // precision state of the 3x3 tiles around the tile, per field
    int ntype[AMPS_MAX_FIELDS][9];
    double noff[AMPS_MAX_FIELDS][9];
    float nscl[AMPS_MAX_FIELDS][9];
};

// Moore or 5-point Laplacians of a row strip of 4 cells: u, m, d hold the 6
// values of columns x0-1 .. x0+4 of the rows above, at and below the strip;
// operand order as the reference's (graingrowth/function.h:20).  Written
// with constant indices so that the rows stay in registers.
template <typename T>
__device__ inline void gg_lap_row4(const T (&u)[6], const T (&m)[6], const T (&d)[6], bool moore, T inv, T (&lap)[4]) {
    if (moore) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
            lap[i] = (u[i] + u[i + 1] + u[i + 2] + m[i] + m[i + 2] + d[i] + d[i + 1] + d[i + 2] - T(8) * m[i + 1]) * inv;
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) lap[i] = (m[i] + m[i + 2] + u[i + 1] + d[i + 1] - T(4) * m[i + 1]) * inv;
    }
}

// Column strip form (32x32 workgroup kernel): cv[] holds rows ly0-1 .. ly0+R
// of the column; side columns are read from LDS.
template <typename T>
__device__ inline T gg_lap(const T* B, int c, int LW, const T* cv, int i, bool moore, T inv) {
    if (moore)
        return (B[c - LW - 1] + cv[i] + B[c - LW + 1] + B[c - 1] + B[c + 1] + B[c + LW - 1] + cv[i + 2] +
                B[c + LW + 1] - T(8) * cv[i + 1]) * inv;
    return (B[c - 1] + B[c + 1] + cv[i] + cv[i + 2] - T(4) * cv[i + 1]) * inv;
}

// Low-tile update of one cell, FP32 on the deviation dv from the tile's
// offset (eta = off + dv), written with explicit fused multiply-adds so
// that its rounding does not depend on the compiler's contraction (the
// oracle's low-tile emulation restates it operation by operation):
//   others = S - eta^2, df = 2 gamma eta others + (beta eta^2 - alpha) eta,
//   r = dv - dt L (df - kappa lap)   (graingrowth/function.h:20-23).
__device__ inline float gg_low_update(float dv, float lap, float offf, float Sf, float g2f, float bef, float alf,
                                      float kapf, float dtlf) {
    const float eta = offf + dv;
    const float others = fmaf(-eta, eta, Sf);
    const float t = fmaf(bef, eta * eta, -alf);
    const float df = fmaf(g2f * eta, others, t * eta);
    return fmaf(-dtlf, fmaf(-kapf, lap, df), dv);
}

// Low-tile Laplacians of the 4 cells of a row strip (FP32), with the centre
// term fused: Moore as the 3x3 column sums minus 9 dv, 5-point as the four
// neighbours minus 4 dv.
__device__ inline void gg_lap_row4_low(const float (&u)[6], const float (&m)[6], const float (&d)[6], bool moore,
                                       float inv, float (&lap)[4]) {
    if (moore) {
        float cs[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) cs[j] = (u[j] + m[j]) + d[j];
#pragma unroll
        for (int i = 0; i < 4; ++i) lap[i] = fmaf(-9.0f, m[i + 1], (cs[i] + cs[i + 1]) + cs[i + 2]) * inv;
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) lap[i] = fmaf(-4.0f, m[i + 1], ((m[i] + m[i + 2]) + u[i + 1]) + d[i + 1]) * inv;
    }
}

// Halo element h (0..67) of a 16x16 tile: box position (by, bx) with
// by, bx in [-1, 16], top row, bottom row, left column, right column.
__device__ inline void gg_halo16(int h, int& by, int& bx) {
    if (h < 18) { by = -1; bx = h - 1; }
    else if (h < 36) { by = 16; bx = h - 19; }
    else if (h < 52) { by = h - 36; bx = -1; }
    else { by = h - 52; bx = 16; }
}

// Field data of one field for the lane, loaded by one sequence of loads for
// both precisions (addresses selected, no branches), so that the loads of the
// next field stay in flight while the current one is computed: own strip as
// 4 FP64 values (a: cells 0-1, b: cells 2-3) or, for fp16, the aligned
// 8-word group holding the 4 own words (loaded twice); each halo element as
// its FP64 bits or the aligned 4-word group holding its fp16 word.  Grids
// are multiples of the tile, so every group lies inside its row.  A lane
// with one halo element reads its own cell for the second (a cache hit; one
// shared address for every such lane would load a single memory channel).
struct GgRaw {
    uint4 a, b;
    uint2 halo[2];
};

template <bool ADAPT>
__device__ inline void gg_wave_load(GgRaw& r, const StepArgs& a, const InstDesc& d, int f, int ty, int cell0,
                                    const int* hoff, const int* hreg, int nh, const GgWaveLds& L) {
    const double* f64 = a.in.f64 + (size_t)f * a.total_cells + d.cell_base;
    const unsigned short* f16 =
        reinterpret_cast<const unsigned short*>(a.in.f16 + (size_t)f * a.total_cells + d.cell_base);
    const bool own64 = !ADAPT || ty == 2;
    const uint4* po = own64 ? reinterpret_cast<const uint4*>(f64 + cell0) : reinterpret_cast<const uint4*>(f16 + (cell0 & ~7));
    r.a = po[0];
    r.b = po[own64 ? 1 : 0];
#pragma unroll
    for (int k = 0; k < 2; ++k) {  // a lane with one halo element loads its own cell0 again (unused)
        const bool h64 = !ADAPT || L.ntype[f][hreg[k]] == 2;
        r.halo[k] = *(h64 ? reinterpret_cast<const uint2*>(f64 + hoff[k])
                          : reinterpret_cast<const uint2*>(f16 + (hoff[k] & ~3)));
    }
}

__device__ inline float half_word(unsigned w) { return __half2float(__ushort_as_half((unsigned short)w)); }

// Held field-tiles (adaptive step): a low field-tile whose tile code and the
// codes of all 8 neighbour tiles bound every value of the field within
// GG_FLAT of 0 (|off| + 2^14 scl <= GG_FLAT) is an absent grain, decaying
// towards 0 by ~2.5% a step on the workload grids; it is set to exactly 0
// (uniform: words 0, offset 0, scale 0), a change below GG_FLAT, and stays
// there (0 is a fixed point of the update).  Buffer zmask bit f of a tile
// records "field f is exactly 0 here": a field whose bit is set on all 9
// tiles needs no neighbour codes at all, and the monitor's value of such a
// field is exactly 0.  A field
// bounded within GG_FLAT of 1 on the 9 tiles, on a tile where every other
// field is held near 0, is a grain interior: eta = 1, S - eta^2 = 0 is a
// fixed point of the update when alpha = beta, so it is held as well.
// same_flags[f][tile] records that both buffers already hold it, so its
// loads and stores are skipped; its share of S is then 0, or off^2 near 1
// (|v^2 - off^2| <= 3 GG_FLAT).  The held set of a tile is decided once from
// the neighbour tables; the next-field prefetch skips held fields.
constexpr double GG_FLAT = 1.0e-10;

// Read-only metadata (descriptors, tile -> instance map, tile types) arrive
// as restrict-qualified arguments so that the per-tile reads go through the
// scalar cache; with MULTI = false the descriptor is loop-invariant.
// ONE: the grid covers the work (one tile per workgroup); the loop then has
// a trip count of one that the compiler can see, so nothing is kept live
// across tiles (a loop of unknown trip count keeps every per-tile constant
// in registers and spills SGPRs into VGPR lanes).
// One-wave mapping of 16x16 tiles: per field one vector load of the lane's
// own strip and at most two halo elements (offsets computed once per tile),
// the next field's loads in flight while the current field is computed.
template <int TS, bool ADAPT, bool MULTI, bool ONE>
__global__ __launch_bounds__(64, 4) void gg_wave_kernel(StepArgs a, const InstDesc* __restrict__ inst,
                                                     const int* __restrict__ tinst, const int* __restrict__ types) {
    static_assert(TS == 16, "the one-wave mapping covers 16x16 tiles");
    using Lw = GgWaveLds;
    constexpr int LWD = Lw::LWD, C0D = Lw::C0D, LWF = Lw::LWF, C0F = Lw::C0F;
    __shared__ __align__(16) GgWaveLds L;
    const int lane = threadIdx.x;
    const int ly = lane >> 2, x0 = (lane & 3) << 2;
    const int nh = lane + 64 < 68 ? 2 : 1;
    for (int b = blockIdx.x, it = 0; b < a.nwork && (!ONE || it == 0); b += gridDim.x, ++it) {
        const int w = xcd_remap(b, a.nwork);
        const InstDesc& d = inst[MULTI ? tinst[w] : 0];
        const int lt = w - d.tile_base;
        const int tile_y = lt / d.tx, tile_x = lt - tile_y * d.tx;
        const int ox = tile_x * TS, oy = tile_y * TS;
        const bool interior = ox >= 1 && oy >= 1 && ox + TS + 1 <= d.nx && oy + TS + 1 <= d.ny;
        const double inv = d.k[GGK_INV_DXDY], dtl = d.k[GGK_DT_L], kap = d.k[GGK_KAPPA];
        const double al = d.k[GGK_ALPHA], be = d.k[GGK_BETA], g2 = d.k[GGK_2GAMMA];
        const bool moore = d.stencil == AMPS_STENCIL_REF;
        const int cell0 = (oy + ly) * d.nx + ox + x0;
        // halo elements of the lane: cell offsets, regions, LDS rows/columns
        int hoff[2] = {cell0, cell0}, hreg[2] = {4, 4}, hby[2] = {0, 0}, hbx[2] = {0, 0};
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            if (k < nh) {
                int by, bx;
                gg_halo16(lane + 64 * k, by, bx);
                int gy = oy + by, gx = ox + bx;
                if (!interior) { gy = bmap(gy, d.ny, d.bc); gx = bmap(gx, d.nx, d.bc); }
                hoff[k] = gy * d.nx + gx;
                hreg[k] = (by < 0 ? 0 : (by >= TS ? 6 : 3)) + (bx < 0 ? 0 : (bx >= TS ? 2 : 1));
                hby[k] = by + 1;
                hbx[k] = bx;
            }
        }
        // carried sum of squares and the same-data flags: independent of the
        // tables below, so their loads go first
        double S[4], Sn[4] = {0.0, 0.0, 0.0, 0.0};
        {
            const double2 u = *reinterpret_cast<const double2*>(a.aux_in + d.cell_base + cell0);
            const double2 v = *reinterpret_cast<const double2*>(a.aux_in + d.cell_base + cell0 + 2);
            S[0] = u.x; S[1] = u.y; S[2] = v.x; S[3] = v.y;
        }
        int sfl = 0;
        if (ADAPT && lane < a.q) sfl = a.same_flags[(size_t)lane * a.total_tiles + w];
        unsigned long long zall = 0ull;  // fields exactly 0 on all 9 tiles
        if (ADAPT) {
            Nb nb;
            nb_setup(nb, d, tile_x, tile_y, TS);
            const unsigned long long zm = a.in.zmask[nb_tile(nb, lane < 9 ? lane : 8)];
            zall = ~0ull;
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                const unsigned lo = __builtin_amdgcn_readlane((unsigned)(zm & 0xFFFFFFFFull), k);
                const unsigned hi = __builtin_amdgcn_readlane((unsigned)(zm >> 32), k);
                zall &= ((unsigned long long)hi << 32) | lo;
            }
            // neighbour tables of the fields not zero on all 9 tiles: the loads
            // of NB entries per lane in flight together, then their LDS stores
            constexpr int NE = (AMPS_MAX_FIELDS * 9 + GW - 1) / GW, NB = 3;
            const unsigned qe = (unsigned)a.q * 9u;
#pragma unroll 1
            for (int i0 = 0; i0 < NE; i0 += NB) {
                int tv[NB];
                double ov[NB];
                float sv[NB];
                unsigned ent[NB];
#pragma unroll
                for (int i = 0; i < NB; ++i) {
                    const unsigned e = (unsigned)(lane + (i0 + i) * GW), f = e / 9u, k = e - f * 9u;
                    ent[i] = e < qe && !((zall >> f) & 1ull) ? f * 9u + k : ~0u;
                    tv[i] = 0;
                    ov[i] = 0.0;
                    sv[i] = 0.0f;
                    if (ent[i] != ~0u) {
                        const unsigned idx = f * (unsigned)a.total_tiles + (unsigned)nb_tile(nb, (int)k);
                        tv[i] = types[idx];
                        ov[i] = a.in.off[idx];
                        sv[i] = a.in.scl[idx];
                    }
                }
#pragma unroll
                for (int i = 0; i < NB; ++i) {
                    if (ent[i] != ~0u) {
                        (&L.ntype[0][0])[ent[i]] = tv[i];
                        (&L.noff[0][0])[ent[i]] = ov[i];
                        (&L.nscl[0][0])[ent[i]] = sv[i];
                    }
                }
            }
        }
        __syncthreads();  // neighbour tables visible
        // held fields (bit f) and those both buffers already hold
        unsigned long long held = 0ull, skip = 0ull, samem = 0ull, near0 = 0ull, zout = 0ull;
        if (ADAPT) {
            bool h0 = lane < a.q, h1 = lane < a.q;
            if (lane < a.q && ((zall >> lane) & 1ull)) {
                h1 = false;  // zero on all 9 tiles (codes not loaded): the update keeps it exactly 0
            } else if (lane < a.q && (a.flags & AMPS_FLAG_NO_SKIP)) {
                h0 = h1 = false;  // no flush to 0 and no interior hold
            } else if (lane < a.q) {
#pragma unroll
                for (int k = 0; k < 9; ++k) {
                    const bool low = L.ntype[lane][k] == 1;
                    const double o = L.noff[lane][k], bnd = 16384.0 * (double)L.nscl[lane][k];
                    h0 = h0 && low && fabs(o) + bnd <= GG_FLAT;
                    h1 = h1 && low && fabs(1.0 - o) + bnd <= GG_FLAT;
                }
            }
            near0 = __ballot(h0);
            const unsigned long long near1 = __ballot(h1);
            const unsigned long long all = (1ull << a.q) - 1ull;
            held = near0;
            // a field near 1 is held when every other field of the tile is held near 0
            // (eta = 1 is the bulk value only when alpha = beta)
            if (al == be && near1 && __popcll(near1) == 1 && (near0 | near1) == all && (near0 & near1) == 0ull)
                held |= near1;
            samem = __ballot(sfl != 0);
            skip = held & samem;
        }
        const unsigned long long live = ~held;  // fields that are updated
        const int f0 = (int)__builtin_ctzll(live | (1ull << 63));
        GgRaw raw;
        if (f0 < a.q) gg_wave_load<ADAPT>(raw, a, d, f0, ADAPT ? L.ntype[f0][4] : 2, cell0, hoff, hreg, nh, L);
        for (int f = 0; f < a.q; ++f) {
            if (ADAPT && ((held >> f) & 1ull) && ((near0 >> f) & 1ull)) {
                // absent grain: exactly 0 from now on
                zout |= 1ull << f;
                if ((skip >> f) & 1ull) continue;  // both buffers hold the zeros
                *reinterpret_cast<uint2*>(a.out.f16 + (size_t)f * a.total_cells + d.cell_base + cell0) = make_uint2(0u, 0u);
                if (lane == 0) {
                    const bool in_zero = ((zall >> f) & 1ull) || (L.noff[f][4] == 0.0 && L.nscl[f][4] == 0.0f);
                    a.out.off[(size_t)f * a.total_tiles + w] = 0.0;
                    a.out.scl[(size_t)f * a.total_tiles + w] = 0.0f;
                    a.same_flags[(size_t)f * a.total_tiles + w] = in_zero ? 1 : 0;  // equal once the input is 0 too
                }
                continue;
            }
            if (ADAPT && ((held >> f) & 1ull)) {
                const double off = L.noff[f][4];
                if ((skip >> f) & 1ull) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) Sn[i] += off * off;
                    continue;
                }
                // first held step: copy into the other buffer (words, codes)
                const float sc = L.nscl[f][4];
                const __half* in16 = a.in.f16 + (size_t)f * a.total_cells + d.cell_base;
                const uint2 wv = *reinterpret_cast<const uint2*>(in16 + cell0);
                *reinterpret_cast<uint2*>(a.out.f16 + (size_t)f * a.total_cells + d.cell_base + cell0) = wv;
                const unsigned hv[4] = {wv.x & 0xFFFFu, wv.x >> 16, wv.y & 0xFFFFu, wv.y >> 16};
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const double v = off + (double)half_word(hv[i]) * (double)sc;
                    Sn[i] += v * v;
                }
                if (lane == 0) {
                    a.out.off[(size_t)f * a.total_tiles + w] = off;
                    a.out.scl[(size_t)f * a.total_tiles + w] = sc;
                    a.same_flags[(size_t)f * a.total_tiles + w] = 1;
                }
                continue;
            }
            const int ty = ADAPT ? L.ntype[f][4] : 2;
            // stage field f (own strip + halo) into the LDS box
            if (ty == 2) {
                double* row = L.d + (ly + 1) * LWD + C0D + x0;
                *reinterpret_cast<double2*>(row) = make_double2(raw_f64(raw.a.x, raw.a.y), raw_f64(raw.a.z, raw.a.w));
                *reinterpret_cast<double2*>(row + 2) = make_double2(raw_f64(raw.b.x, raw.b.y), raw_f64(raw.b.z, raw.b.w));
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    if (k < nh) {
                        const int r = hreg[k];
                        L.d[hby[k] * LWD + C0D + hbx[k]] =
                            (!ADAPT || L.ntype[f][r] == 2)
                                ? raw_f64(raw.halo[k].x, raw.halo[k].y)
                                : L.noff[f][r] + (double)half_word(raw_word4(raw.halo[k], hoff[k])) * (double)L.nscl[f][r];
                    }
                }
            } else {
                const double own_off = L.noff[f][4];
                const float sc = L.nscl[f][4];
                float* row = L.f + (ly + 1) * LWF + C0F + x0;
                const uint2 o16 = raw_words8(raw.a, x0);
                *reinterpret_cast<float4*>(row) =
                    make_float4(half_word(o16.x & 0xFFFFu) * sc, half_word(o16.x >> 16) * sc,
                                half_word(o16.y & 0xFFFFu) * sc, half_word(o16.y >> 16) * sc);
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    if (k < nh) {
                        const int r = hreg[k];
                        float x;
                        if (L.ntype[f][r] == 2) x = (float)(raw_f64(raw.halo[k].x, raw.halo[k].y) - own_off);
                        else x = (float)(L.noff[f][r] - own_off) + half_word(raw_word4(raw.halo[k], hoff[k])) * L.nscl[f][r];
                        L.f[hby[k] * LWF + C0F + hbx[k]] = x;
                    }
                }
            }
            {
                // next updated field's loads in flight while this one is computed
                const int fn = f + 1 < a.q ? (int)__builtin_ctzll((live >> (f + 1) << (f + 1)) | (1ull << 63)) : a.q;
                const int fl = fn < a.q ? fn : f;  // past the last field: loads of this field again (unused)
                gg_wave_load<ADAPT>(raw, a, d, fl, ADAPT ? L.ntype[fl][4] : 2, cell0, hoff, hreg, nh, L);
            }
            __syncthreads();
            if (ty == 2) {
                double ru[6], rm[6], rd[6];
                {
                    const double* p = L.d + ly * LWD + C0D + x0 - 1;
#define GG_ROW6D(R, P)                                                                   \
    do {                                                                                 \
        R[0] = (P)[0];                                                                   \
        const double2 m1_ = *reinterpret_cast<const double2*>((P) + 1);                  \
        const double2 m2_ = *reinterpret_cast<const double2*>((P) + 3);                  \
        R[1] = m1_.x; R[2] = m1_.y; R[3] = m2_.x; R[4] = m2_.y;                          \
        R[5] = (P)[5];                                                                   \
    } while (0)
                    GG_ROW6D(ru, p);
                    GG_ROW6D(rm, p + LWD);
                    GG_ROW6D(rd, p + 2 * LWD);
#undef GG_ROW6D
                }
                double lapv[4], en[4];
                gg_lap_row4<double>(ru, rm, rd, moore, inv, lapv);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const double eta = rm[i + 1];
                    const double lap = lapv[i];
                    const double others = S[i] - eta * eta;
                    const double df = g2 * eta * others + be * (eta * eta * eta) - al * eta;
                    en[i] = eta - dtl * (df - kap * lap);
                    Sn[i] += en[i] * en[i];
                }
                double* out = a.out.f64 + (size_t)f * a.total_cells + d.cell_base + cell0;
                *reinterpret_cast<double2*>(out) = make_double2(en[0], en[1]);
                *reinterpret_cast<double2*>(out + 2) = make_double2(en[2], en[3]);
            } else {
                const float invf = d.kf[GGK_INV_DXDY], dtlf = d.kf[GGK_DT_L], kapf = d.kf[GGK_KAPPA];
                const float alf = d.kf[GGK_ALPHA], bef = d.kf[GGK_BETA], g2f = d.kf[GGK_2GAMMA];
                const double own_off = L.noff[f][4];
                const float offf = (float)own_off;
                float ru[6], rm[6], rd[6];
                {
                    const float* p = L.f + ly * LWF + C0F + x0;
#define GG_ROW6F(R, P)                                                                   \
    do {                                                                                 \
        R[0] = (P)[-1];                                                                  \
        const float4 m_ = *reinterpret_cast<const float4*>(P);                           \
        R[1] = m_.x; R[2] = m_.y; R[3] = m_.z; R[4] = m_.w;                              \
        R[5] = (P)[4];                                                                   \
    } while (0)
                    GG_ROW6F(ru, p);
                    GG_ROW6F(rm, p + LWF);
                    GG_ROW6F(rd, p + 2 * LWF);
#undef GG_ROW6F
                }
                float lapv[4], rn[4];
                gg_lap_row4_low(ru, rm, rd, moore, invf, lapv);
                float mn = INFINITY, mx = -INFINITY;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    rn[i] = gg_low_update(rm[i + 1], lapv[i], offf, (float)S[i], g2f, bef, alf, kapf, dtlf);
                    mn = fminf(mn, rn[i]);
                    mx = fmaxf(mx, rn[i]);
                }
                wave_minmax(mn, mx);
                const TileCode tc = tile_code(mn, mx);
                const double noff = own_off + (double)tc.mid;
                unsigned short hw[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const __half h = __float2half_rn((rn[i] - tc.mid) * tc.inv);
                    hw[i] = __half_as_ushort(h);
                    const double v = noff + (double)__half2float(h) * (double)tc.scl;
                    Sn[i] += v * v;
                }
                __half* out = a.out.f16 + (size_t)f * a.total_cells + d.cell_base;
                *reinterpret_cast<uint2*>(out + cell0) =
                    make_uint2((unsigned)hw[0] | ((unsigned)hw[1] << 16), (unsigned)hw[2] | ((unsigned)hw[3] << 16));
                if (lane == 0) {
                    a.out.off[(size_t)f * a.total_tiles + w] = noff;
                    a.out.scl[(size_t)f * a.total_tiles + w] = tc.scl;
                    if (ADAPT && ((samem >> f) & 1ull)) a.same_flags[(size_t)f * a.total_tiles + w] = 0;
                }
                if (tc.scl == 0.0f && noff == 0.0) zout |= 1ull << f;
            }
            __syncthreads();  // the box is rewritten for the next field
        }
        double* so = a.aux_out + d.cell_base + cell0;
        *reinterpret_cast<double2*>(so) = make_double2(Sn[0], Sn[1]);
        *reinterpret_cast<double2*>(so + 2) = make_double2(Sn[2], Sn[3]);
        if (ADAPT && lane == 0) a.out.zmask[w] = zout;
    }
}

template <int TS>
struct GgBlockLds {
    static constexpr int H = 1, LW = TS + 2;
    union {
        double d[LW * LW];
        float f[LW * LW];
    };
    // precision state of the 3x3 tiles around the workgroup's tile, per field
    int ntype[AMPS_MAX_FIELDS][9];
    double noff[AMPS_MAX_FIELDS][9];
    float nscl[AMPS_MAX_FIELDS][9];
    float red[16];  // block_minmax of up to 8 waves
};

// Raw box values of field f: FP64 bits, or the fp16 word where the element's
// tile is low (the type table in LDS says which storage is current).
template <int TS, bool ADAPT, int NT = WG>
__device__ inline void gg_load_raw(unsigned long long* raw, const StepArgs& a, const InstDesc& d, int f,
                                   const BoxMap<TS, 1, NT>& m, const GgBlockLds<TS>& L) {
    constexpr int N = BoxMap<TS, 1, NT>::N;
    const double* f64 = a.in.f64 + (size_t)f * a.total_cells + d.cell_base;
    const unsigned short* f16 =
        reinterpret_cast<const unsigned short*>(a.in.f16 + (size_t)f * a.total_cells + d.cell_base);
#pragma unroll
    for (int k = 0; k < N; ++k) {
        if (m.off[k] >= 0) {
            if (!ADAPT || L.ntype[f][m.reg[k]] == 2) raw[k] = (unsigned long long)__double_as_longlong(f64[m.off[k]]);
            else raw[k] = f16[m.off[k]];
        }
    }
}

// Workgroup-per-tile variant for 32x32 tiles (a one-wave mapping of a 32x32
// tile needs more registers than a SIMD can give 4 waves).  Per tile: one memory round brings the neighbour precision table of every
// field, the carried sum of squares and field 0's box; afterwards the box of
// field f+1 is loaded into registers while field f is computed, so each field
// costs one overlapped load round.  NT threads own column strips of
// TS*TS/NT cells; the FP64 sweep runs NT = 512 (2 cells per thread, 54
// VGPRs against 82 at 256 and 4 cells): twice the waves per CU share each
// tile's load rounds.
template <int TS, bool ADAPT, bool MULTI, bool ONE, int NT = WG>
__global__ __launch_bounds__(NT, 4) void gg_block_kernel(StepArgs a, const InstDesc* __restrict__ inst,
                                                      const int* __restrict__ tinst, const int* __restrict__ types) {
    constexpr int H = 1, LW = TS + 2, R = TS * TS / NT, N = BoxMap<TS, 1, NT>::N;
    __shared__ GgBlockLds<TS> L;
    for (int b = blockIdx.x, it = 0; b < a.nwork && (!ONE || it == 0); b += gridDim.x, ++it) {
        const int w = xcd_remap(b, a.nwork);
        const InstDesc& d = inst[MULTI ? tinst[w] : 0];
        const int lt = w - d.tile_base;
        const int tile_y = lt / d.tx, tile_x = lt - tile_y * d.tx;
        const int ox = tile_x * TS, oy = tile_y * TS;
        BoxMap<TS, H, NT> m;
        box_map<TS, H, NT>(m, d, tile_x, tile_y);
        const double inv = d.k[GGK_INV_DXDY], dtl = d.k[GGK_DT_L], kap = d.k[GGK_KAPPA];
        const double al = d.k[GGK_ALPHA], be = d.k[GGK_BETA], g2 = d.k[GGK_2GAMMA];
        const bool moore = d.stencil == AMPS_STENCIL_REF;
        const int lx = threadIdx.x % TS, ly0 = (threadIdx.x / TS) * R;

        if (ADAPT) {
            Nb nb;
            nb_setup(nb, d, tile_x, tile_y, TS);
            for (int e = threadIdx.x; e < a.q * 9; e += NT) {
                const int f = e / 9, k = e - f * 9;
                const int g = nb_tile(nb, k);
                L.ntype[f][k] = types[(size_t)f * a.total_tiles + g];
                L.noff[f][k] = a.in.off[(size_t)f * a.total_tiles + g];
                L.nscl[f][k] = a.in.scl[(size_t)f * a.total_tiles + g];
            }
            __syncthreads();
        }
        double S[R], Sn[R];
        int gcell[R];
#pragma unroll
        for (int i = 0; i < R; ++i) {
            gcell[i] = (oy + ly0 + i) * d.nx + ox + lx;
            S[i] = a.aux_in[d.cell_base + gcell[i]];
            Sn[i] = 0.0;
        }
        unsigned long long raw[N];
        gg_load_raw<TS, ADAPT, NT>(raw, a, d, 0, m, L);
        for (int f = 0; f < a.q; ++f) {
            const int ty = ADAPT ? L.ntype[f][4] : 2;
            // stage field f from the raw values, then start loading field f+1
            if (ty == 2) {
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    if (m.off[k] < 0) continue;
                    const int r = m.reg[k];
                    L.d[threadIdx.x + k * NT] =
                        (!ADAPT || L.ntype[f][r] == 2)
                            ? __longlong_as_double((long long)raw[k])
                            : L.noff[f][r] + (double)__half2float(__ushort_as_half((unsigned short)raw[k])) *
                                                 (double)L.nscl[f][r];
                }
            } else {
                const double own_off = L.noff[f][4];
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    if (m.off[k] < 0) continue;
                    const int r = m.reg[k];
                    float x;
                    if (L.ntype[f][r] == 2) x = (float)(__longlong_as_double((long long)raw[k]) - own_off);
                    else if (r == 4) x = __half2float(__ushort_as_half((unsigned short)raw[k])) * L.nscl[f][4];
                    else
                        x = (float)(L.noff[f][r] - own_off) +
                            __half2float(__ushort_as_half((unsigned short)raw[k])) * L.nscl[f][r];
                    L.f[threadIdx.x + k * NT] = x;
                }
            }
            if (f + 1 < a.q) gg_load_raw<TS, ADAPT, NT>(raw, a, d, f + 1, m, L);
            __syncthreads();
            if (ty == 2) {
                double* out = a.out.f64 + (size_t)f * a.total_cells + d.cell_base;
                double cv[R + 2];
#pragma unroll
                for (int j = 0; j < R + 2; ++j) cv[j] = L.d[(ly0 + j) * LW + lx + H];
#pragma unroll
                for (int i = 0; i < R; ++i) {
                    const int c = (ly0 + i + H) * LW + lx + H;
                    const double eta = cv[i + 1];
                    const double lap = gg_lap<double>(L.d, c, LW, cv, i, moore, inv);
                    const double others = S[i] - eta * eta;
                    const double df = g2 * eta * others + be * (eta * eta * eta) - al * eta;
                    const double en = eta - dtl * (df - kap * lap);
                    out[gcell[i]] = en;
                    Sn[i] += en * en;
                }
            } else {
                const float invf = d.kf[GGK_INV_DXDY], dtlf = d.kf[GGK_DT_L], kapf = d.kf[GGK_KAPPA];
                const float alf = d.kf[GGK_ALPHA], bef = d.kf[GGK_BETA], g2f = d.kf[GGK_2GAMMA];
                const double own_off = L.noff[f][4];
                const float offf = (float)own_off;
                float cv[R + 2];
#pragma unroll
                for (int j = 0; j < R + 2; ++j) cv[j] = L.f[(ly0 + j) * LW + lx + H];
                float r[R];
                float mn = INFINITY, mx = -INFINITY;
#pragma unroll
                for (int i = 0; i < R; ++i) {
                    const int c = (ly0 + i + H) * LW + lx + H;
                    const float dv = cv[i + 1];
                    const float lap = gg_lap<float>(L.f, c, LW, cv, i, moore, invf);
                    const float rn = gg_low_update(dv, lap, offf, (float)S[i], g2f, bef, alf, kapf, dtlf);
                    r[i] = rn;
                    mn = fminf(mn, rn);
                    mx = fmaxf(mx, rn);
                }
                block_minmax<NT / 64>(mn, mx, L.red);
                const TileCode tc = tile_code(mn, mx);
                const double noff = own_off + (double)tc.mid;
                __half* out = a.out.f16 + (size_t)f * a.total_cells + d.cell_base;
#pragma unroll
                for (int i = 0; i < R; ++i) {
                    const __half h = __float2half_rn((r[i] - tc.mid) * tc.inv);
                    out[gcell[i]] = h;
                    const double v = noff + (double)__half2float(h) * (double)tc.scl;
                    Sn[i] += v * v;
                }
                if (threadIdx.x == 0) {
                    a.out.off[(size_t)f * a.total_tiles + w] = noff;
                    a.out.scl[(size_t)f * a.total_tiles + w] = tc.scl;
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int i = 0; i < R; ++i) a.aux_out[d.cell_base + gcell[i]] = Sn[i];
    }
}

// S = sum_i eta_i^2 from the current storage of buffer `in` (low tiles
// decoded), written to aux_in.  One workgroup per tile.  With type_old given,
// only tiles where some field changed precision are refreshed (a conversion
// re-rounds the stored values; elsewhere S already matches them).
template <int TS>
__global__ __launch_bounds__(256) void gg_aux_kernel(StepArgs a, const int* __restrict__ type_old,
                                                     const int* __restrict__ types) {
    constexpr int PER = TS * TS / WG;
    __shared__ int chg[WG], todo[WG];
    __shared__ int ntodo;
    // chunks of WG / q tiles: thread (tile j, field f) checks one type pair,
    // the tiles with a changed field (every tile without type_old) are
    // compacted and recomputed one by one
    const int tpc = WG / a.q > 0 ? WG / a.q : 1;
    for (int c0 = blockIdx.x * tpc; c0 < a.total_tiles; c0 += gridDim.x * tpc) {
        if ((int)threadIdx.x < tpc) chg[threadIdx.x] = type_old ? 0 : 1;
        __syncthreads();
        if (type_old) {
            const int j = (int)threadIdx.x / a.q, f = (int)threadIdx.x - j * a.q, w = c0 + j;
            if (j < tpc && w < a.total_tiles &&
                type_old[(size_t)f * a.total_tiles + w] != types[(size_t)f * a.total_tiles + w])
                chg[j] = 1;  // every writer stores 1
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int k = 0;
            for (int i = 0; i < tpc; ++i)
                if (chg[i] && c0 + i < a.total_tiles) todo[k++] = c0 + i;
            ntodo = k;
        }
        __syncthreads();
        const int nt = ntodo;
        for (int it = 0; it < nt; ++it) {
        const int w = todo[it];
        const InstDesc& d = a.inst[a.tile_inst[w]];
        const int lt = w - d.tile_base;
        const int tile_y = lt / d.tx, tile_x = lt - tile_y * d.tx;
        double S[PER];
        size_t gc[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            int e = threadIdx.x + i * WG;
            int ly = e / TS, lx = e - ly * TS;
            gc[i] = d.cell_base + (size_t)(tile_y * TS + ly) * d.nx + tile_x * TS + lx;
            S[i] = 0.0;
        }
        for (int f = 0; f < a.q; ++f) {
            const int ty = a.type ? a.type[(size_t)f * a.total_tiles + w] : 2;
            const size_t fb = (size_t)f * a.total_cells;
            if (ty == 2) {
#pragma unroll
                for (int i = 0; i < PER; ++i) { double v = a.in.f64[fb + gc[i]]; S[i] += v * v; }
            } else {
                const double off = a.in.off[(size_t)f * a.total_tiles + w];
                const double scl = a.in.scl[(size_t)f * a.total_tiles + w];
#pragma unroll
                for (int i = 0; i < PER; ++i) {
                    double v = off + (double)__half2float(a.in.f16[fb + gc[i]]) * scl;
                    S[i] += v * v;
                }
            }
        }
#pragma unroll
        for (int i = 0; i < PER; ++i) a.aux_in[gc[i]] = S[i];
        }
        __syncthreads();  // todo and ntodo are rewritten by the next chunk
    }
}

template <int TS, bool ADAPT, bool ONE>
static void gg_launch3(const StepArgs& a, int grid, hipStream_t s, const hipEvent_t* ev) {
    if constexpr (TS == 16) {
        if (a.ninst > 1)
            launch_sweep(gg_wave_kernel<TS, ADAPT, true, ONE>, dim3(grid), dim3(GW), s, ev, a, a.inst, a.tile_inst, a.type);
        else
            launch_sweep(gg_wave_kernel<TS, ADAPT, false, ONE>, dim3(grid), dim3(GW), s, ev, a, a.inst, a.tile_inst, a.type);
    } else {
        // FP64 sweep: 512 threads, 2 cells each (see gg_block_kernel)
        if constexpr (!ADAPT) {
            if (a.ninst > 1)
                launch_sweep(gg_block_kernel<TS, ADAPT, true, ONE, 512>, dim3(grid), dim3(512), s, ev, a, a.inst, a.tile_inst, a.type);
            else
                launch_sweep(gg_block_kernel<TS, ADAPT, false, ONE, 512>, dim3(grid), dim3(512), s, ev, a, a.inst, a.tile_inst, a.type);
            return;
        }
        if (a.ninst > 1)
            launch_sweep(gg_block_kernel<TS, ADAPT, true, ONE>, dim3(grid), dim3(WG), s, ev, a, a.inst, a.tile_inst, a.type);
        else
            launch_sweep(gg_block_kernel<TS, ADAPT, false, ONE>, dim3(grid), dim3(WG), s, ev, a, a.inst, a.tile_inst, a.type);
    }
}

template <int TS, bool ADAPT>
static void gg_launch(const StepArgs& a, int grid, hipStream_t s, const hipEvent_t* ev) {
    if (grid >= a.nwork) gg_launch3<TS, ADAPT, true>(a, grid, s, ev);
    else gg_launch3<TS, ADAPT, false>(a, grid, s, ev);
}

int launch_gg_step(const StepArgs& a, int tile, bool adaptive, int grid, hipStream_t s, const hipEvent_t* ev) {
    if (tile == 16) {
        if (adaptive) gg_launch<16, true>(a, grid, s, ev);
        else gg_launch<16, false>(a, grid, s, ev);
    } else if (tile == 32) {
        if (adaptive) gg_launch<32, true>(a, grid, s, ev);
        else gg_launch<32, false>(a, grid, s, ev);
    } else {
        return AMPS_E_UNSUPPORTED;
    }
    return hipGetLastError() == hipSuccess ? AMPS_OK : AMPS_E_HIP;
}

int launch_gg_aux(const StepArgs& a, int tile, const int* type_old, hipStream_t s) {
    const int tpc = WG / a.q > 0 ? WG / a.q : 1;
    int grid = (a.total_tiles + tpc - 1) / tpc;
    if (grid > 65536) grid = 65536;
    if (grid < 1) grid = 1;
    if (tile == 16) hipLaunchKernelGGL((gg_aux_kernel<16>), dim3(grid), dim3(WG), 0, s, a, type_old, a.type);
    else if (tile == 32) hipLaunchKernelGGL((gg_aux_kernel<32>), dim3(grid), dim3(WG), 0, s, a, type_old, a.type);
    else return AMPS_E_UNSUPPORTED;
    return hipGetLastError() == hipSuccess ? AMPS_OK : AMPS_E_HIP;
}

}  // namespace amps

// k_graingrowth.hip — Fan-Chen grain growth with q order parameters, one
// fused sweep per step for all fields.
//
// Reference (q = 2): kernel1_pure, graingrowth/function.h:5-24, launched once
// per field per step (purehigh.cu:31-32):
//   lap   = (sum of the 8 Moore neighbours - 8 eta) / dxdy
//   df    = 2 eta sum_{j!=i} eta_j^2 + eta^3 - eta          (alpha=beta=gamma=1)
//   eta' = eta - dt L (df - kappa lap)
// Here every field of a tile is advanced by one workgroup.  The per-cell sum
// of squares S = sum_j eta_j^2 of the current state is carried in an FP64
// side array written by the previous sweep, so sum_{j!=i} = S - eta_i^2 costs
// O(1) per field instead of O(q): traffic per cell-update is 16 B per FP64
// field (4 B per fp16 field) + 16 B for S.
//
// Tile types are per (field, tile): a field that is ~0 across a tile (a grain
// that is absent there) stays fp16 while the few fields forming the local
// grain boundaries run in FP64.
#include "kernels.hpp"
#include "tile_common.hpp"

namespace amps {

template <int TS>
struct GgLds {
    static constexpr int H = 1, LW = TS + 2;
    union {
        double d[LW * LW];
        float f[LW * LW];
    };
    // precision state of the 3x3 tiles around the workgroup's tile, per field
    int ntype[AMPS_MAX_FIELDS][9];
    double noff[AMPS_MAX_FIELDS][9];
    float nscl[AMPS_MAX_FIELDS][9];
    float red[8];
};

// Moore or 5-point Laplacian of the cell at strip position i (column values
// cv[] hold rows ly0-1 .. ly0+R of the column; side columns read from LDS).
template <typename T>
__device__ inline T gg_lap(const T* B, int c, int LW, const T* cv, int i, bool moore, T inv) {
    if (moore)
        return (B[c - LW - 1] + cv[i] + B[c - LW + 1] + B[c - 1] + B[c + 1] + B[c + LW - 1] + cv[i + 2] +
                B[c + LW + 1] - T(8) * cv[i + 1]) * inv;
    return (B[c - 1] + B[c + 1] + cv[i] + cv[i + 2] - T(4) * cv[i + 1]) * inv;
}

// Raw box values of field f: FP64 bits, or the fp16 word where the element's
// tile is low (the type table in LDS says which storage is current).
template <int TS, bool ADAPT>
__device__ inline void gg_load_raw(unsigned long long* raw, const StepArgs& a, const InstDesc& d, int f,
                                   const BoxMap<TS, 1>& m, const GgLds<TS>& L) {
    constexpr int N = BoxMap<TS, 1>::N;
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

// Read-only metadata (descriptors, tile -> instance map, tile types) arrive
// as restrict-qualified arguments so that the per-tile reads go through the
// scalar cache; with MULTI = false the descriptor is loop-invariant.
//
// Per tile: one memory round brings the neighbour precision table of every
// field, the carried sum of squares and field 0's box; afterwards the box of
// field f+1 is loaded into registers while field f is computed, so each field
// costs one overlapped load round.
template <int TS, bool ADAPT, bool MULTI>
__global__ __launch_bounds__(256, 4) void gg_step_kernel(StepArgs a, const InstDesc* __restrict__ inst,
                                                      const int* __restrict__ tinst, const int* __restrict__ types) {
    constexpr int H = 1, LW = TS + 2, R = TS * TS / WG, N = BoxMap<TS, 1>::N;
    __shared__ GgLds<TS> L;
    for (int b = blockIdx.x; b < a.nwork; b += gridDim.x) {
        const int w = xcd_remap(b, a.nwork);
        const InstDesc& d = inst[MULTI ? tinst[w] : 0];
        const int lt = w - d.tile_base;
        const int tile_y = lt / d.tx, tile_x = lt - tile_y * d.tx;
        const int ox = tile_x * TS, oy = tile_y * TS;
        BoxMap<TS, H> m;
        box_map<TS, H>(m, d, tile_x, tile_y);
        const double inv = d.k[GGK_INV_DXDY], dtl = d.k[GGK_DT_L], kap = d.k[GGK_KAPPA];
        const double al = d.k[GGK_ALPHA], be = d.k[GGK_BETA], g2 = d.k[GGK_2GAMMA];
        const bool moore = d.stencil == AMPS_STENCIL_REF;
        const int lx = threadIdx.x % TS, ly0 = (threadIdx.x / TS) * R;

        if (ADAPT) {
            Nb nb;
            nb_setup(nb, d, tile_x, tile_y, TS);
            for (int e = threadIdx.x; e < a.q * 9; e += WG) {
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
        gg_load_raw<TS, ADAPT>(raw, a, d, 0, m, L);
        for (int f = 0; f < a.q; ++f) {
            const int ty = ADAPT ? L.ntype[f][4] : 2;
            // stage field f from the raw values, then start loading field f+1
            if (ty == 2) {
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    if (m.off[k] < 0) continue;
                    const int r = m.reg[k];
                    L.d[threadIdx.x + k * WG] =
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
                    L.f[threadIdx.x + k * WG] = x;
                }
            }
            if (f + 1 < a.q) gg_load_raw<TS, ADAPT>(raw, a, d, f + 1, m, L);
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
                    const float eta = offf + dv;
                    const float others = (float)S[i] - eta * eta;
                    const float df = g2f * eta * others + bef * (eta * eta * eta) - alf * eta;
                    const float rn = dv - dtlf * (df - kapf * lap);
                    r[i] = rn;
                    mn = fminf(mn, rn);
                    mx = fmaxf(mx, rn);
                }
                block_minmax(mn, mx, L.red);
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
// decoded), written to aux_in.  One workgroup per tile.
template <int TS>
__global__ __launch_bounds__(256) void gg_aux_kernel(StepArgs a) {
    constexpr int PER = TS * TS / WG;
    for (int w = blockIdx.x; w < a.total_tiles; w += gridDim.x) {
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
}

template <int TS, bool ADAPT>
static void gg_launch(const StepArgs& a, int grid, hipStream_t s) {
    if (a.ninst > 1)
        hipLaunchKernelGGL((gg_step_kernel<TS, ADAPT, true>), dim3(grid), dim3(WG), 0, s, a, a.inst, a.tile_inst, a.type);
    else
        hipLaunchKernelGGL((gg_step_kernel<TS, ADAPT, false>), dim3(grid), dim3(WG), 0, s, a, a.inst, a.tile_inst, a.type);
}

int launch_gg_step(const StepArgs& a, int tile, bool adaptive, int grid, hipStream_t s) {
    if (tile == 16) {
        if (adaptive) gg_launch<16, true>(a, grid, s);
        else gg_launch<16, false>(a, grid, s);
    } else if (tile == 32) {
        if (adaptive) gg_launch<32, true>(a, grid, s);
        else gg_launch<32, false>(a, grid, s);
    } else {
        return AMPS_E_UNSUPPORTED;
    }
    return hipGetLastError() == hipSuccess ? AMPS_OK : AMPS_E_HIP;
}

int launch_gg_aux(const StepArgs& a, int tile, hipStream_t s) {
    int grid = a.total_tiles > 65536 ? 65536 : a.total_tiles;
    if (tile == 16) hipLaunchKernelGGL((gg_aux_kernel<16>), dim3(grid), dim3(WG), 0, s, a);
    else if (tile == 32) hipLaunchKernelGGL((gg_aux_kernel<32>), dim3(grid), dim3(WG), 0, s, a);
    else return AMPS_E_UNSUPPORTED;
    return hipGetLastError() == hipSuccess ? AMPS_OK : AMPS_E_HIP;
}

}  // namespace amps

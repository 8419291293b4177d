// k_common.hip — model-independent kernels: field initialisation, the
// per-tile precision monitor, tile classification and FP64 <-> fp16 tile
// conversion, plus decode/encode used by get/set of fields.
#include "kernels.hpp"
#include "tile_common.hpp"

namespace amps {

// ================================================================== init
// snowflake/tools.h:24-33 (disc, T = 0); graingrowth/tools.h:24-38 (circle
// grain) or seeded U[-amp, amp]; sintering/main.cu:45-60 (1-based discs) or
// tanh discs plus seeded noise.  One thread per cell, every field of one
// instance; writes FP64 storage of buffer `b`.
__global__ void init_kernel(const InstDesc* inst_all, int inst, int model, int init, int q, Buf b,
                            long long total_cells) {
    const InstDesc& d = inst_all[inst];
    const long long n = (long long)d.nx * d.ny;
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n;
         c += (long long)gridDim.x * blockDim.x) {
        const int y = (int)(c / d.nx), x = (int)(c - (long long)y * d.nx);
        double* f = b.f64 + d.cell_base + c;
        if (model == AMPS_SNOWFLAKE) {
            double dx = x - d.p[AMPS_SF_CX], dy = y - d.p[AMPS_SF_CY];
            f[0] = (dx * dx + dy * dy < d.p[AMPS_SF_R2]) ? 1.0 : 0.0;
            f[total_cells] = 0.0;
        } else if (model == AMPS_GRAINGROWTH) {
            if (init == AMPS_INIT_REF) {
                double dx = x - d.p[AMPS_GG_CX], dy = y - d.p[AMPS_GG_CY], R = d.p[AMPS_GG_R];
                bool in = dx * dx + dy * dy <= R * R;
                for (int i = 0; i < q; ++i) f[i * total_cells] = 0.0;
                f[0] = in ? 0.0 : 1.0;
                f[total_cells] = in ? 1.0 : 0.0;
            } else {
                double amp = d.p[AMPS_GG_AMP];
                for (int i = 0; i < q; ++i) {
                    uint64_t key = rng_key(d.seed, STREAM_GG_INIT + (uint64_t)i);
                    f[i * total_cells] = amp * (2.0 * rng_u01_keyed(key, (uint64_t)c) - 1.0);
                }
            }
        } else {
            const double R1 = d.p[AMPS_SN_R1], R2 = d.p[AMPS_SN_R2];
            const double X1 = d.p[AMPS_SN_X1], Y1 = d.p[AMPS_SN_Y1], X2 = d.p[AMPS_SN_X2], Y2 = d.p[AMPS_SN_Y2];
            double cv, e1, e2;
            if (init == AMPS_INIT_REF) {
                const int xx = x + 1, yy = y + 1;  // reference loops are 1-based
                cv = e1 = e2 = 0.0;
                if ((xx - X1) * (xx - X1) + (yy - Y1) * (yy - Y1) <= R1 * R1) { cv = 1.0; e1 = 1.0; }
                if ((xx - X2) * (xx - X2) + (yy - Y2) * (yy - Y2) <= R2 * R2) { cv = 1.0; e2 = 1.0; e1 = 0.0; }
            } else {
                const double W = d.p[AMPS_SN_WIDTH], a = d.p[AMPS_SN_NOISE];
                double d1 = sqrt((x - X1) * (x - X1) + (y - Y1) * (y - Y1));
                double d2 = sqrt((x - X2) * (x - X2) + (y - Y2) * (y - Y2));
                e1 = 0.5 * (1.0 - tanh((d1 - R1) / W));
                e2 = 0.5 * (1.0 - tanh((d2 - R2) / W));
                uint64_t key = rng_key(d.seed, STREAM_SN_INIT);
                cv = e1 + e2 + a * (2.0 * rng_u01_keyed(key, (uint64_t)c) - 1.0);
                cv = cv < 0.0 ? 0.0 : (cv > 1.0 ? 1.0 : cv);
            }
            f[0] = cv;
            f[total_cells] = e1;
            f[2 * total_cells] = e2;
        }
    }
}

// ============================================================ fp16 codec
// Direct double -> binary16 rounding to nearest even: round to odd into
// float (24 bits), then to nearest even into half (11 bits); with at least
// two extra bits the two-step result equals one correct rounding.
__device__ inline __half f64_to_f16_rne(double x) {
    float f = (float)x;
    double back = (double)f;
    if (back != x && !isnan(x)) {
        uint32_t bits = __float_as_uint(f);
        if ((bits & 1u) == 0u) bits += (fabs(back) < fabs(x)) ? 1u : (uint32_t)-1;
        f = __uint_as_float(bits);
    }
    return __float2half_rn(f);
}

// Encode the FP64 values of tile `gtile`, field f of buffer b into fp16
// (offset = mid-range, power-of-two scale).  Whole workgroup; barriers inside.
template <int TS>
__device__ void encode_tile(const Buf& b, long long total_cells, int total_tiles, const InstDesc& d, int gtile,
                            int field, double* red) {
    const int lt = gtile - d.tile_base, tyi = lt / d.tx, txi = lt - tyi * d.tx;
    const double* f64 = b.f64 + (size_t)field * total_cells + d.cell_base;
    __half* f16 = b.f16 + (size_t)field * total_cells + d.cell_base;
    double mn = INFINITY, mx = -INFINITY;
    for (int e = threadIdx.x; e < TS * TS; e += WG) {
        int ly = e / TS, lx = e - ly * TS;
        double v = f64[(size_t)(tyi * TS + ly) * d.nx + txi * TS + lx];
        mn = fmin(mn, v);
        mx = fmax(mx, v);
    }
    mx = block_max_d(mx, red);
    __syncthreads();
    mn = -block_max_d(-mn, red);
    double off = 0.5 * (mn + mx), hw = 0.5 * (mx - mn);
    int ex = hw == 0.0 ? -100 : 0;  // a uniform tile takes the smallest scale
    if (hw > 0.0 && hw <= 1e300) frexp(hw, &ex);
    if (ex < -100) ex = -100;
    if (ex > 100) ex = 100;
    const double inv = ldexp(1.0, 14 - ex), scl = ldexp(1.0, ex - 14);
    double* w64 = const_cast<double*>(f64);
    for (int e = threadIdx.x; e < TS * TS; e += WG) {
        int ly = e / TS, lx = e - ly * TS;
        size_t c = (size_t)(tyi * TS + ly) * d.nx + txi * TS + lx;
        const __half h = f64_to_f16_rne((f64[c] - off) * inv);
        f16[c] = h;
        // FP64 storage keeps the decoded value: neighbours read halos there
        w64[c] = off + (double)__half2float(h) * scl;
    }
    if (threadIdx.x == 0) {
        b.off[(size_t)field * total_tiles + gtile] = off;
        b.scl[(size_t)field * total_tiles + gtile] = hw == 0.0 ? 0.0f : ldexpf(1.0f, ex - 14);  // 0: uniform
    }
}

template <int TS>
__device__ void decode_tile(const Buf& b, long long total_cells, int total_tiles, const InstDesc& d, int gtile,
                            int field, double* dst_f64 /* field base incl. instance */) {
    const int lt = gtile - d.tile_base, tyi = lt / d.tx, txi = lt - tyi * d.tx;
    const __half* f16 = b.f16 + (size_t)field * total_cells + d.cell_base;
    const double off = b.off[(size_t)field * total_tiles + gtile];
    const double scl = b.scl[(size_t)field * total_tiles + gtile];
    for (int e = threadIdx.x; e < TS * TS; e += WG) {
        int ly = e / TS, lx = e - ly * TS;
        size_t c = (size_t)(tyi * TS + ly) * d.nx + txi * TS + lx;
        dst_f64[c] = off + (double)__half2float(f16[c]) * scl;
    }
}

// Convert tiles whose type changed (old -> new) for every field of the group:
// 2 -> 1 encodes FP64 into fp16, 1 -> 2 decodes into FP64.  force_encode /
// force_decode convert every tile whose new type is 1 / every type-1 tile.
// One workgroup per (group, tile).
template <int TS>
__global__ __launch_bounds__(256) void convert_kernel(const InstDesc* __restrict__ inst,
                                                       const int* __restrict__ tile_inst, Buf b,
                                                       long long total_cells, int total_tiles,
                                                       const int* __restrict__ type_old,
                                                       const int* __restrict__ type, GroupMap gm, int mode) {
    __shared__ double red[4];
    __shared__ int todo[WG];
    __shared__ int ntodo;
    const int ngw = gm.ngroups * total_tiles;
    // chunks of WG items: every thread decides one item, the items that
    // convert are compacted and converted one by one by the workgroup
    for (int c0 = blockIdx.x * WG; c0 < ngw; c0 += gridDim.x * WG) {
        const int w = c0 + (int)threadIdx.x;
        int op = 0;  // 1 encode, 2 decode
        if (w < ngw) {
            const int tn = type[w];
            const int to = type_old ? type_old[w] : tn;
            if (mode == CONVERT_CHANGED) op = (to == 2 && tn == 1) ? 1 : ((to == 1 && tn == 2) ? 2 : 0);
            else if (mode == CONVERT_ENCODE_LOW) op = tn == 1 ? 1 : 0;
            else if (mode == CONVERT_DECODE_LOW) op = tn == 1 ? 2 : 0;
        }
        if (threadIdx.x == 0) ntodo = 0;
        __syncthreads();
        if (op) todo[atomicAdd(&ntodo, 1)] = w * 2 + (op - 1);  // LDS atomic: order does not matter
        __syncthreads();
        const int nt = ntodo;
        for (int i = 0; i < nt; ++i) {
            const int it = todo[i], wi = it >> 1;
            const bool enc = (it & 1) == 0;
            const int g = wi / total_tiles, gtile = wi - g * total_tiles;
            const InstDesc& d = inst[tile_inst[gtile]];
            for (int f = 0; f < gm.nfields; ++f) {
                if (gm.fg[f] != g) continue;
                if (enc) encode_tile<TS>(b, total_cells, total_tiles, d, gtile, f, red);
                else decode_tile<TS>(b, total_cells, total_tiles, d, gtile, f, b.f64 + (size_t)f * total_cells + d.cell_base);
                __syncthreads();
            }
        }
        __syncthreads();  // todo and ntodo are rewritten by the next chunk
    }
}

// ================================================================ monitor
// Sample lattices of the reference (tile 16): {0,5,10,15} for Monitor1
// (snowflake/tools.h:186), {2,6,10,14} for Monitor2 (tools.h:203); for
// tile t: {0,(t-1)/3,2(t-1)/3,t-1} and {t/8,3t/8,5t/8,7t/8}.
__device__ inline int sample_off(int tile, int kind, int i) {
    if (kind == AMPS_MON_SPATIAL) return i == 0 ? 0 : (i == 1 ? (tile - 1) / 3 : (i == 2 ? 2 * (tile - 1) / 3 : tile - 1));
    return (2 * i + 1) * tile / 8;
}

__device__ inline double read_decoded(const Buf& b, long long total_cells, int total_tiles, const InstDesc& d,
                                      int field, int gtile, int ty, int y, int x) {
    size_t c = (size_t)field * total_cells + d.cell_base + (size_t)y * d.nx + x;
    if (ty == 2) return b.f64[c];
    return b.off[(size_t)field * total_tiles + gtile] + (double)__half2float(b.f16[c]) *
                                                         (double)b.scl[(size_t)field * total_tiles + gtile];
}

// Monitor1 / Monitor2 on the 4x4 sample lattice, one thread per (group,
// tile).  Monitor1: get_max_diff1 (snowflake/tools.h:177-193), min starts
// at 1 and max at 0.  Monitor2: get_max_diff2 (tools.h:194-211) against the
// samples kept from the previous pass, which are then refreshed (the
// reference refreshes phi_last right after the monitor, function.h:106-113).
__global__ void monitor_samples_kernel(const InstDesc* inst, const int* tile_inst, Buf b, long long total_cells,
                                       int total_tiles, const int* type, int ngroups_eval, int field0,
                                       int field_stride, int type_groups, int tile, int kind, double* md,
                                       double* last, unsigned long long* __restrict__ reset) {
    if (reset && blockIdx.x == 0 && threadIdx.x == 0) *reset = 0ull;  // counter the classification adds to
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= ngroups_eval * total_tiles) return;
    const int g = w / total_tiles, gtile = w - g * total_tiles;
    const int field = field0 + g * field_stride;
    const InstDesc& d = inst[tile_inst[gtile]];
    const int lt = gtile - d.tile_base, tyi = lt / d.tx, txi = lt - tyi * d.tx;
    const int tgroup = type_groups > 1 ? (g < type_groups ? g : 0) : 0;
    const int ty = type ? type[(size_t)tgroup * total_tiles + gtile] : 2;
    if (kind == AMPS_MON_SPATIAL) {
        double mn = 1.0, mx = 0.0;
        for (int j = 0; j < 4; ++j)
            for (int i = 0; i < 4; ++i) {
                double v = read_decoded(b, total_cells, total_tiles, d, field, gtile, ty,
                                        tyi * tile + sample_off(tile, kind, j), txi * tile + sample_off(tile, kind, i));
                if (v < mn) mn = v;
                if (v > mx) mx = v;
            }
        md[w] = mx - mn;
    } else {
        double m = 0.0;
        double* lw = last + (size_t)w * 16;
        for (int j = 0; j < 4; ++j)
            for (int i = 0; i < 4; ++i) {
                double v = read_decoded(b, total_cells, total_tiles, d, field, gtile, ty,
                                        tyi * tile + sample_off(tile, kind, j), txi * tile + sample_off(tile, kind, i));
                double df = fabs(lw[j * 4 + i] - v);
                if (df > m) m = df;
                lw[j * 4 + i] = v;
            }
        md[w] = m;
    }
}

// Monitor2 history at initialisation: the complement of the field
// (phi_old = phi == 0 ? 1 : 0, snowflake/end2end_mix.cu:44-46).
__global__ void monitor_last_init_kernel(const InstDesc* inst, const int* tile_inst, Buf b, long long total_cells,
                                         int total_tiles, int ngroups_eval, int field0, int field_stride, int tile,
                                         double* last) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= ngroups_eval * total_tiles) return;
    const int g = w / total_tiles, gtile = w - g * total_tiles;
    const int field = field0 + g * field_stride;
    const InstDesc& d = inst[tile_inst[gtile]];
    const int lt = gtile - d.tile_base, tyi = lt / d.tx, txi = lt - tyi * d.tx;
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i) {
            double v = read_decoded(b, total_cells, total_tiles, d, field, gtile, 2,
                                    tyi * tile + sample_off(tile, AMPS_MON_TEMPORAL, j),
                                    txi * tile + sample_off(tile, AMPS_MON_TEMPORAL, i));
            last[(size_t)w * 16 + j * 4 + i] = v == 0.0 ? 1.0 : 0.0;
        }
}

// Gradient monitor: max over the tile of (0.5*(f[x+1]-f[x-1]))^2 +
// (0.5*(f[y+1]-f[y-1])*dx/dy)^2, neighbours through the boundary map.
// Contraction into FMA is off so the value matches a CPU evaluation bit for
// bit.  One workgroup per (group, tile).
template <int TS>
__global__ __launch_bounds__(256) void monitor_grad_kernel(StepArgs a, int ngroups_eval, int field0,
                                                           int field_stride, int type_groups, double* md, unsigned long long* __restrict__ reset) {
    if (reset && blockIdx.x == 0 && threadIdx.x == 0) *reset = 0ull;  // counter the classification adds to
    constexpr int H = 1, LW = TS + 2;
    __shared__ double box[LW * LW];
    __shared__ NbField nbf;
    __shared__ double red[4];
    const int ngw = ngroups_eval * a.total_tiles;
    for (int w = blockIdx.x; w < ngw; w += gridDim.x) {
        const int g = w / a.total_tiles, gtile = w - g * a.total_tiles;
        const int field = field0 + g * field_stride;
        const int tgroup = type_groups > 1 ? g : 0;
        const InstDesc& d = a.inst[a.tile_inst[gtile]];
        const int lt = gtile - d.tile_base, tyi = lt / d.tx, txi = lt - tyi * d.tx;
        Nb nb;
        nb_setup(nb, d, txi, tyi, TS);
        const bool uni = (a.type == nullptr);
        if (!uni) {
            nbfield_fill(&nbf, a, a.in, nb, field, tgroup);
            __syncthreads();
        }
        BoxMap<TS, H> bm;
        box_map<TS, H>(bm, d, txi, tyi);
        const double* f64 = a.in.f64 + (size_t)field * a.total_cells + d.cell_base;
        if (uni) box_load1_f64<TS, H>(box, f64, bm);
        else box_load_mixed_f64<TS, H>(box, f64, a.in.f16 + (size_t)field * a.total_cells + d.cell_base, &nbf, bm);
        __syncthreads();
        const double r = d.p[0] / d.p[1];  // dx / dy (slots 0 and 1 of every model)
        double m = 0.0;
        for (int e = threadIdx.x; e < TS * TS; e += WG) {
#pragma clang fp contract(off)
            int ly = e / TS, lx = e - ly * TS;
            int c = (ly + 1) * LW + lx + 1;
            double hx = 0.5 * (box[c + 1] - box[c - 1]);
            double hy = 0.5 * (box[c + LW] - box[c - LW]) * r;
            double qv = hx * hx + hy * hy;
            if (qv > m) m = qv;
        }
        m = block_max_d(m, red);
        if (threadIdx.x == 0) md[w] = m;
        __syncthreads();
    }
}

// The same monitor for 16x16 tiles with one wave per tile evaluating every
// monitored field: lane l owns the row strip (row l/4, columns 4(l%4)..+3),
// loaded with one vector load, plus at most two halo elements whose offsets
// are computed once per tile; the neighbour precision table of every field
// is read once.  Values are decoded and the quotient evaluated exactly as
// above, so the monitor values are identical.
struct MonWaveLds {
    static constexpr int LWD = 22, C0 = 2;  // row stride (even, for 16-byte reads; 11 16-byte groups) and first own column
    double box[18 * LWD];
    int ntype[AMPS_MAX_FIELDS][9];
    double noff[AMPS_MAX_FIELDS][9];
    float nscl[AMPS_MAX_FIELDS][9];
};

__global__ __launch_bounds__(64) void monitor_grad_wave16(StepArgs a, const InstDesc* __restrict__ inst,
                                                          const int* __restrict__ tinst,
                                                          const int* __restrict__ types, int ngroups_eval, int field0,
                                                          int field_stride, int type_groups, double* __restrict__ md, unsigned long long* __restrict__ reset) {
    if (reset && blockIdx.x == 0 && threadIdx.x == 0) *reset = 0ull;  // counter the classification adds to
    constexpr int TS = 16, LWD = MonWaveLds::LWD, C0 = MonWaveLds::C0;
    static_assert(LWD % 2 == 0 && LWD >= C0 + TS + 2, "16-byte reads of columns x0-2 .. x0+5");
    __shared__ __align__(16) MonWaveLds L;
    const int lane = threadIdx.x;
    const int ly = lane >> 2, x0 = (lane & 3) << 2;
    const int nh = lane + 64 < 68 ? 2 : 1;
    const bool uni = types == nullptr;
    for (int w = blockIdx.x; w < a.total_tiles; w += gridDim.x) {
        const InstDesc& d = inst[tinst[w]];
        const int lt = w - d.tile_base, tyi = lt / d.tx, txi = lt - tyi * d.tx;
        const int ox = txi * TS, oy = tyi * TS;
        const bool interior = ox >= 1 && oy >= 1 && ox + TS + 1 <= d.nx && oy + TS + 1 <= d.ny;
        const int cell0 = (oy + ly) * d.nx + ox + x0;
        int hoff[2] = {cell0, cell0}, hreg[2] = {4, 4}, hidx[2] = {0, 0};  // no element: own cell (unused)
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            if (k < nh) {
                const int h = lane + 64 * k;
                int by, bx;
                if (h < 18) { by = -1; bx = h - 1; }
                else if (h < 36) { by = 16; bx = h - 19; }
                else if (h < 52) { by = h - 36; bx = -1; }
                else { by = h - 52; bx = 16; }
                int gy = oy + by, gx = ox + bx;
                if (!interior) { gy = bmap(gy, d.ny, d.bc); gx = bmap(gx, d.nx, d.bc); }
                hoff[k] = gy * d.nx + gx;
                hreg[k] = (by < 0 ? 0 : (by >= TS ? 6 : 3)) + (bx < 0 ? 0 : (bx >= TS ? 2 : 1));
                hidx[k] = (by + 1) * LWD + C0 + bx;
            }
        }
        // fields exactly 0 on the tile and its edge neighbours (grain growth's
        // zero masks): every central difference is 0, so the value is 0
        unsigned long long z5 = 0ull;
        if (!uni && a.in.zmask) {
            Nb nb;
            nb_setup(nb, d, txi, tyi, TS);
            const unsigned long long zm = a.in.zmask[nb_tile(nb, lane < 9 ? lane : 8)];
            z5 = ~0ull;
            const int kk[5] = {1, 3, 4, 5, 7};  // edge neighbours and the tile
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                const unsigned lo = __builtin_amdgcn_readlane((unsigned)(zm & 0xFFFFFFFFull), kk[i]);
                const unsigned hi = __builtin_amdgcn_readlane((unsigned)(zm >> 32), kk[i]);
                z5 &= ((unsigned long long)hi << 32) | lo;
            }
            if (field_stride != 1 || field0 != 0) z5 = 0ull;  // bits are fields; groups must be fields 0..
            if (lane < ngroups_eval && ((z5 >> lane) & 1ull)) md[(size_t)lane * a.total_tiles + w] = 0.0;
        }
        if (!uni) {
            Nb nb;
            nb_setup(nb, d, txi, tyi, TS);
            // the loads of NB entries per lane in flight together, then their
            // LDS stores (fields whose value is already written are skipped)
            constexpr int NE = (AMPS_MAX_FIELDS * 9 + 63) / 64, NB = 3;
            const unsigned ge = (unsigned)ngroups_eval * 9u;
#pragma unroll 1
            for (int i0 = 0; i0 < NE; i0 += NB) {
                int tv[NB];
                double ov[NB];
                float sv[NB];
                unsigned ent[NB];
#pragma unroll
                for (int i = 0; i < NB; ++i) {
                    const unsigned e = (unsigned)(lane + (i0 + i) * 64), g = e / 9u, k = e - g * 9u;
                    ent[i] = e < ge && !((z5 >> g) & 1ull) ? e : ~0u;
                    tv[i] = 0;
                    ov[i] = 0.0;
                    sv[i] = 0.0f;
                    if (ent[i] != ~0u) {
                        const unsigned field = (unsigned)(field0 + (int)g * field_stride);
                        const unsigned tgroup = type_groups > 1 ? g : 0u;
                        const unsigned gt = (unsigned)nb_tile(nb, (int)k);
                        tv[i] = types[tgroup * (unsigned)a.total_tiles + gt];
                        ov[i] = a.in.off[field * (unsigned)a.total_tiles + gt];
                        sv[i] = a.in.scl[field * (unsigned)a.total_tiles + gt];
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
        __syncthreads();
        const double r = d.p[0] / d.p[1];  // dx / dy (slots 0 and 1 of every model)
        // raw loads of field g, one sequence for both precisions (addresses
        // selected, no branches) so that they stay in flight while the previous
        // field is evaluated: own strip as 4 FP64 values (ra: cells 0-1, rb:
        // cells 2-3) or the aligned 8-word fp16 group holding the 4 own words;
        // halo elements as FP64 bits or the aligned 4-word fp16 group (grids
        // are multiples of the tile, so every group lies inside its row; a lane
        // with one halo element loads its own cell again, unused)
        uint4 ra = make_uint4(0u, 0u, 0u, 0u), rb = ra;
        uint2 rh[2] = {make_uint2(0u, 0u), make_uint2(0u, 0u)};
        auto load = [&](int g) {
            const int field = field0 + g * field_stride;
            const double* f64 = a.in.f64 + (size_t)field * a.total_cells + d.cell_base;
            const unsigned short* f16 =
                reinterpret_cast<const unsigned short*>(a.in.f16 + (size_t)field * a.total_cells + d.cell_base);
            const bool own64 = uni || L.ntype[g][4] == 2;
            const uint4* po = own64 ? reinterpret_cast<const uint4*>(f64 + cell0)
                                    : reinterpret_cast<const uint4*>(f16 + (cell0 & ~7));
            ra = po[0];
            rb = po[own64 ? 1 : 0];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const bool h64 = uni || L.ntype[g][hreg[k]] == 2;
                rh[k] = *(h64 ? reinterpret_cast<const uint2*>(f64 + hoff[k])
                              : reinterpret_cast<const uint2*>(f16 + (hoff[k] & ~3)));
            }
        };
        const unsigned long long live = ~z5;
        const int g0 = (int)__builtin_ctzll(live | (1ull << 63));
        if (g0 < ngroups_eval) load(g0);
        for (int g = 0; g < ngroups_eval; ++g) {
            if ((z5 >> g) & 1ull) continue;
            double own[4];
            if (uni || L.ntype[g][4] == 2) {
                own[0] = raw_f64(ra.x, ra.y);
                own[1] = raw_f64(ra.z, ra.w);
                own[2] = raw_f64(rb.x, rb.y);
                own[3] = raw_f64(rb.z, rb.w);
            } else {
                const double o = L.noff[g][4], sc = (double)L.nscl[g][4];
                const uint2 ownh = raw_words8(ra, x0);
                own[0] = o + (double)__half2float(__ushort_as_half((unsigned short)(ownh.x & 0xFFFFu))) * sc;
                own[1] = o + (double)__half2float(__ushort_as_half((unsigned short)(ownh.x >> 16))) * sc;
                own[2] = o + (double)__half2float(__ushort_as_half((unsigned short)(ownh.y & 0xFFFFu))) * sc;
                own[3] = o + (double)__half2float(__ushort_as_half((unsigned short)(ownh.y >> 16))) * sc;
            }
            double hv[2] = {0.0, 0.0};
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                if (k < nh) {
                    const int rg = hreg[k];
                    if (uni || L.ntype[g][rg] == 2) hv[k] = raw_f64(rh[k].x, rh[k].y);
                    else
                        hv[k] = L.noff[g][rg] + (double)__half2float(__ushort_as_half((unsigned short)raw_word4(rh[k], hoff[k]))) *
                                                    (double)L.nscl[g][rg];
                }
            }
            double* row = L.box + (ly + 1) * LWD + C0 + x0;
            *reinterpret_cast<double2*>(row) = make_double2(own[0], own[1]);
            *reinterpret_cast<double2*>(row + 2) = make_double2(own[2], own[3]);
#pragma unroll
            for (int k = 0; k < 2; ++k)
                if (k < nh) L.box[hidx[k]] = hv[k];
            {
                // next evaluated field in flight during this one
                const int gn = g + 1 < ngroups_eval ? (int)__builtin_ctzll((live >> (g + 1) << (g + 1)) | (1ull << 63))
                                                    : ngroups_eval;
                load(gn < ngroups_eval ? gn : g);  // past the last field: this field again (unused)
            }
            __syncthreads();
            double m = 0.0;
            {
#pragma clang fp contract(off)
                // 16-byte reads: columns x0-2 .. x0+5 of the row, x0 .. x0+3 above and below
                const double* c = L.box + (ly + 1) * LWD + C0 + x0;
                const double2 m0 = *reinterpret_cast<const double2*>(c - 2);
                const double2 m1 = *reinterpret_cast<const double2*>(c);
                const double2 m2 = *reinterpret_cast<const double2*>(c + 2);
                const double2 m3 = *reinterpret_cast<const double2*>(c + 4);
                const double2 u0 = *reinterpret_cast<const double2*>(c - LWD);
                const double2 u1 = *reinterpret_cast<const double2*>(c - LWD + 2);
                const double2 d0 = *reinterpret_cast<const double2*>(c + LWD);
                const double2 d1 = *reinterpret_cast<const double2*>(c + LWD + 2);
                const double row[6] = {m0.y, m1.x, m1.y, m2.x, m2.y, m3.x};
                const double up[4] = {u0.x, u0.y, u1.x, u1.y}, dn[4] = {d0.x, d0.y, d1.x, d1.y};
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const double hx = 0.5 * (row[i + 2] - row[i]);
                    const double hy = 0.5 * (dn[i] - up[i]) * r;
                    const double qv = hx * hx + hy * hy;
                    if (qv > m) m = qv;
                }
            }
            m = wave_max_d(m);
            if (lane == 0) md[(size_t)g * a.total_tiles + w] = m;
            __syncthreads();
        }
    }
}

// get_type (snowflake/tools.h:212-230): type_old <- type; neighbour max over
// self and the 4 non-periodic von-Neumann neighbours with fmax (CUDA's max);
// < thr -> 1, >= thr -> 2, unordered keeps the old type.  One thread per
// (group, tile); with broadcast every group takes group 0's monitor value.
constexpr int CL_PT = 1;  // (group, tile) items per thread: one dependent load chain each

__global__ __launch_bounds__(1024) void classify_kernel(const InstDesc* __restrict__ inst,
                                                        const int* __restrict__ tile_inst, int total_tiles,
                                                        int ngroups, int broadcast, int single, const double* md,
                                                        double thr, int dilate, int* type_old, int* type,
                                                        unsigned long long* stats) {
    __shared__ unsigned cnt[16];
    unsigned nhi = 0;
    const int n = ngroups * total_tiles;
#pragma unroll
    for (int k = 0; k < CL_PT; ++k) {
        const int w = (blockIdx.x * CL_PT + k) * blockDim.x + threadIdx.x;
        if (w >= n) continue;
        const int g = w / total_tiles, gtile = w - g * total_tiles;
        const InstDesc& d = inst[single ? 0 : tile_inst[gtile]];  // one instance: no tile -> instance load
        const int lt = gtile - d.tile_base, uy = lt / d.tx, ux = lt - uy * d.tx;
        const double* m0 = md + (size_t)(broadcast ? 0 : g) * total_tiles;
        double m = m0[gtile];
        if (uy < d.ty - 1) m = fmax(m, m0[gtile + d.tx]);
        if (uy > 0) m = fmax(m, m0[gtile - d.tx]);
        if (ux < d.tx - 1) m = fmax(m, m0[gtile + 1]);
        if (ux > 0) m = fmax(m, m0[gtile - 1]);
        // dilation of the FP64 set: the maximum over the tiles within
        // Chebyshev distance `dilate` (non-periodic, inside the instance)
        if (dilate > 0) {
            const int y0 = max(uy - dilate, 0), y1 = min(uy + dilate, d.ty - 1);
            const int x0 = max(ux - dilate, 0), x1 = min(ux + dilate, d.tx - 1);
            for (int yy = y0; yy <= y1; ++yy)
                for (int xx = x0; xx <= x1; ++xx) m = fmax(m, m0[d.tile_base + yy * d.tx + xx]);
        }
        int t = type[w];
        type_old[w] = t;
        if (m < thr) t = 1;
        else if (m >= thr) t = 2;
        type[w] = t;
        nhi += t == 2 ? 1u : 0u;
    }
    // FP64 count of the block: wave sums, then one atomic per counter per
    // block (per-wave atomics on one address serialise)
    const unsigned ws = (unsigned)__ockl_wfred_add_i32((int)nhi);
    if ((threadIdx.x & 63) == 0) cnt[threadIdx.x >> 6] = ws;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) tot += cnt[i];
        if (tot) {
            atomicAdd(&stats[0], tot);
            atomicAdd(&stats[1], tot);
        }
    }
}

// Clears same_flags[f][tile] where the type group of field f changed type on
// the tile (the monitor pass converted it: its two buffers no longer agree).
__global__ void clear_changed_kernel(const int* __restrict__ type_old, const int* __restrict__ type, GroupMap gm,
                                     int nfields, int total_tiles, int* __restrict__ same_flags,
                                     unsigned long long* __restrict__ zmask) {
    const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (e >= (long long)nfields * total_tiles) return;
    const int f = (int)(e / total_tiles), t = (int)(e - (long long)f * total_tiles);
    const int g = gm.fg[f];
    if (g < 0) return;
    const size_t k = (size_t)g * total_tiles + t;
    if (type_old[k] != type[k]) {
        same_flags[e] = 0;
        if (zmask) atomicAnd(zmask + t, ~(1ull << f));
    }
}

// Work-lists of FP64 (type 2) and low (type 1) tiles (group 0 types).  The
// FP64 entry of rank r in tile order sits at xcd_unmap(r, n), so the
// workgroup that reads entry b (dealt to XCD b mod 8) gets rank
// xcd_remap(b, n): each XCD a contiguous run of tiles, with no count needed
// to find the entry.  The low entry of rank r sits at xcd_chunk_unmap(r,
// n): runs of XCD_CHUNK tiles per XCD, the runs dealt round-robin (the low
// tiles' costs vary across the grid).  Workgroup j owns tiles
// [1024 j, 1024 j + 1024): it counts the entries of all earlier and later
// tiles itself (vector loads, a block sum), so no workgroup waits on
// another, then writes its own entries with wave ballots.  The last
// workgroup records the totals.  Also clears same_flags (when given) of
// every tile whose type changed (type_old given: the tiles a monitor pass
// converted) or of every tile (type_old null: fields or types were set from
// the host): the two buffers of such a tile can no longer be assumed equal.
constexpr int WL_TILES = 1024;

__global__ __launch_bounds__(1024) void worklist_kernel(const int* __restrict__ type, const int* __restrict__ type_old,
                                                        int total_tiles, int* __restrict__ wl_hi,
                                                        int* __restrict__ wl_lo, int* __restrict__ counts,
                                                        int* __restrict__ same_flags) {
    __shared__ int ws[4][16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int t0 = blockIdx.x * WL_TILES;
    int bh = 0, bl = 0;
    for (int t = tid * 4; t < t0; t += WL_TILES * 4) {
        const int4 v = *reinterpret_cast<const int4*>(type + t);
        bh += (v.x == 2) + (v.y == 2) + (v.z == 2) + (v.w == 2);
        bl += (v.x == 1) + (v.y == 1) + (v.z == 1) + (v.w == 1);
    }
    // FP64 and low tiles after this block's chunk: the totals place the entries
    int ah = 0, al = 0;
    for (int u = t0 + WL_TILES + tid; u < total_tiles; u += WL_TILES) {
        const int tu = type[u];
        ah += tu == 2;
        al += tu == 1;
    }
    const int t = t0 + tid;
    const int ty = t < total_tiles ? type[t] : 0;
    const bool hi = ty == 2, lo = ty == 1;
    const unsigned long long mh = __ballot(hi), ml = __ballot(lo);
    bh = __ockl_wfred_add_i32(bh);
    bl = __ockl_wfred_add_i32(bl);
    ah = __ockl_wfred_add_i32(ah);
    al = __ockl_wfred_add_i32(al);
    if (lane == 0) {
        ws[0][wave] = bh;
        ws[1][wave] = bl;
        ws[2][wave] = ah;
        ws[3][wave] = al;
    }
    __syncthreads();
    int oh = 0, ol = 0, nafter = 0, lafter = 0;
#pragma unroll
    for (int w = 0; w < 16; ++w) { oh += ws[0][w]; ol += ws[1][w]; nafter += ws[2][w]; lafter += ws[3][w]; }
    __syncthreads();
    if (lane == 0) {
        ws[0][wave] = __popcll(mh);
        ws[1][wave] = __popcll(ml);
    }
    __syncthreads();
    int th = oh, tl = ol;
#pragma unroll
    for (int w = 0; w < 16; ++w) {
        const int x = ws[0][w], y = ws[1][w];
        if (w < wave) { oh += x; ol += y; }
        th += x;
        tl += y;
    }
    const unsigned long long below = (1ull << lane) - 1ull;
    // entry of rank r (tile order) at the position the consumers' item r maps to
    const int nhi = th + nafter, nlo = tl + lafter;
    if (hi) wl_hi[xcd_unmap(oh + __popcll(mh & below), nhi)] = t;
    if (lo) wl_lo[xcd_chunk_unmap(ol + __popcll(ml & below), nlo)] = t;
    if (same_flags && t < total_tiles && (!type_old || type_old[t] != ty)) same_flags[t] = 0;
    if (tid == 0 && t0 + WL_TILES >= total_tiles) { counts[0] = th; counts[1] = tl; }
}

// Decode every low tile of field f into a dense FP64 array (instance order),
// for amps_get_field.  FP64 tiles are copied.
template <int TS>
__global__ __launch_bounds__(256) void gather_kernel(const InstDesc* inst, Buf b, long long total_cells,
                                                     int total_tiles, const int* type, int group, int instid,
                                                     int field, double* dst) {
    const InstDesc& d = inst[instid];
    const int ntile = d.tx * d.ty;
    for (int lt = blockIdx.x; lt < ntile; lt += gridDim.x) {
        const int gtile = d.tile_base + lt;
        const int ty = (type && group >= 0) ? type[(size_t)group * total_tiles + gtile] : 2;
        const int tyi = lt / d.tx, txi = lt - tyi * d.tx;
        if (ty == 1) {
            decode_tile<TS>(b, total_cells, total_tiles, d, gtile, field, dst);
        } else {
            const double* src = b.f64 + (size_t)field * total_cells + d.cell_base;
            for (int e = threadIdx.x; e < TS * TS; e += WG) {
                int ly = e / TS, lx = e - ly * TS;
                size_t c = (size_t)(tyi * TS + ly) * d.nx + txi * TS + lx;
                dst[c] = src[c];
            }
        }
    }
}

// Count non-finite values of a dense array (amps_get_field check).
__global__ void nonfinite_kernel(const double* v, long long n, unsigned long long* cnt) {
    unsigned long long local = 0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        local += isfinite(v[i]) ? 0 : 1;
    if (local) atomicAdd(cnt, local);
}

__global__ void fill_u64_kernel(unsigned long long* p, int n, unsigned long long v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void advance_step_kernel(long long* step_ctr, int n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step_ctr += n;
}

// ============================================================ host launch
static int hip_ok() { return hipGetLastError() == hipSuccess ? AMPS_OK : AMPS_E_HIP; }

int launch_init(const InstDesc* inst_dev, int inst, long long ncells, int model, int init, int q, const Buf& b,
                long long total_cells, hipStream_t s) {
    int grid = (int)((ncells + 255) / 256);
    if (grid > 65536) grid = 65536;
    hipLaunchKernelGGL(init_kernel, dim3(grid), dim3(256), 0, s, inst_dev, inst, model, init, q, b, total_cells);
    return hip_ok();
}

int launch_convert(const InstDesc* inst, const int* tile_inst, const Buf& b, long long total_cells, int total_tiles,
                   const int* type_old, const int* type, const GroupMap& gm, int mode, int tile, hipStream_t s) {
    const int items = gm.ngroups * total_tiles;
    int grid = (items + WG - 1) / WG;
    if (grid > 65536) grid = 65536;
    if (grid < 1) grid = 1;
    if (tile == 16) hipLaunchKernelGGL((convert_kernel<16>), dim3(grid), dim3(256), 0, s, inst, tile_inst, b, total_cells, total_tiles, type_old, type, gm, mode);
    else if (tile == 32) hipLaunchKernelGGL((convert_kernel<32>), dim3(grid), dim3(256), 0, s, inst, tile_inst, b, total_cells, total_tiles, type_old, type, gm, mode);
    else if (tile == 64) hipLaunchKernelGGL((convert_kernel<64>), dim3(grid), dim3(256), 0, s, inst, tile_inst, b, total_cells, total_tiles, type_old, type, gm, mode);
    else return AMPS_E_UNSUPPORTED;
    return hip_ok();
}

// Gradient monitor for 32x32 tiles and one monitored field, one wave per
// tile over a persistent grid: lane l owns the 16 cells of row l/2, columns
// 16(l%2) .. +15, and two of the 128 halo-cross elements (rows above and
// below, columns left and right: the central differences need no corners).
// The codes of the tile and its 8 neighbours are read one tile ahead.  A low
// tile that is uniform (scale 0), with edge neighbours uniform at the same
// offset, has every difference exactly 0: its value 0 is written without
// reading the field.  Otherwise values are decoded and the quotient
// evaluated exactly as in monitor_grad_kernel, so the values are identical.
// Box rows are 38 doubles: an odd number of 16-byte groups, so the 16-byte
// accesses of 8 rows x 2 row halves fall on distinct bank groups.
struct MonMeta {
    int type;
    double off;
    float scl;
};

__device__ inline void mon_meta_fetch(MonMeta& m, const StepArgs& a, const InstDesc& d, const int* types, int field,
                                      int w) {
    const int lane = threadIdx.x & 63;
    const int lt = w - d.tile_base, tyi = lt / d.tx, txi = lt - tyi * d.tx;
    Nb nb;
    nb_setup(nb, d, txi, tyi, 32);
    const int g = nb_tile(nb, lane < 9 ? lane : 8);  // every lane loads: no branch around the loads
    // without a type map the loads read the first descriptor (a valid
    // address) and the values are replaced: no branch around the loads
    const bool has = types != nullptr;
    const size_t idx = (size_t)field * a.total_tiles + g;
    const int tv = *(has ? types + g : reinterpret_cast<const int*>(a.inst));
    const double ov = *(has ? a.in.off + idx : reinterpret_cast<const double*>(a.inst));
    const float sv = *(has ? a.in.scl + idx : reinterpret_cast<const float*>(a.inst));
    m.type = has ? tv : 2;
    m.off = has ? ov : 0.0;
    m.scl = has ? sv : 0.0f;
}

__device__ inline double mon_rl_d(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xFFFFFFFFll), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__global__ __launch_bounds__(64) void monitor_grad_wave32(StepArgs a, const InstDesc* __restrict__ inst,
                                                          const int* __restrict__ tinst,
                                                          const int* __restrict__ types, int field0,
                                                          double* __restrict__ md, unsigned long long* __restrict__ reset) {
    if (reset && blockIdx.x == 0 && threadIdx.x == 0) *reset = 0ull;  // counter the classification adds to
    constexpr int TS = 32, LWD = 38, C0 = 2;
    __shared__ __align__(16) double box[(TS + 2) * LWD];
    const int lane = threadIdx.x;
    const int ly = lane >> 1, x0 = (lane & 1) << 4;
    const int n = a.total_tiles, NG = gridDim.x;
    int w = blockIdx.x;
    if (w >= n) return;
    // the codes are fetched one tile ahead, alternating between two sets
    // (ma, mb) so that no copy of a set in flight is needed: a copy waits for
    // its loads, and a wait on the vector-memory counter also waits for
    // every load issued after them
    auto tile = [&](const MonMeta& mc, int w) {
        const InstDesc& d = inst[tinst[w]];
        // uniform low tile between uniform equal edge neighbours: value 0
        const double own_off = mon_rl_d(mc.off, 4);
        const bool edge = lane == 1 || lane == 3 || lane == 4 || lane == 5 || lane == 7;
        const unsigned long long flat = __ballot(edge && mc.type == 1 && mc.scl == 0.0f && mc.off == own_off);
        if ((flat & 0xBAull) == 0xBAull) {
            if (lane == 0) md[w] = 0.0;
            return;
        }
        const int lt = w - d.tile_base, tyi = lt / d.tx, txi = lt - tyi * d.tx;
        const int ox = txi * TS, oy = tyi * TS;
        const bool interior = ox >= 1 && oy >= 1 && ox + TS + 1 <= d.nx && oy + TS + 1 <= d.ny;
        const double* f64 = a.in.f64 + (size_t)field0 * a.total_cells + d.cell_base;
        const unsigned short* f16 =
            reinterpret_cast<const unsigned short*>(a.in.f16 + (size_t)field0 * a.total_cells + d.cell_base);
        const int cell0 = (oy + ly) * d.nx + ox + x0;
        const int town = __builtin_amdgcn_readlane(mc.type, 4);
        // loads: own strip as 8 x 16 bytes, FP64 values or (fp16) the 2 x 16
        // bytes of words read 4 times over, one load sequence for both storages
        // (a branch per storage made the compiler wait for the next tile's
        // codes before the FP64 loads), and two halo elements
        uint4 ro[8];
        {
            const uint4* po = town == 2 ? reinterpret_cast<const uint4*>(f64 + cell0)
                                        : reinterpret_cast<const uint4*>(f16 + cell0);
#pragma unroll
            for (int q = 0; q < 8; ++q) ro[q] = po[town == 2 ? q : (q & 1)];
        }
        double own[16];
        // halo elements as raw words by one load sequence for both storages:
        // FP64 bits or the aligned 4-word fp16 group holding the word
        uint2 hraw[2];
        int hbox[2], hreg[2], hoff[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int e = lane + 64 * k, band = e >> 5, i = e & 31;
            int by, bx;
            if (band == 0) { by = -1; bx = i; }
            else if (band == 1) { by = TS; bx = i; }
            else if (band == 2) { by = i; bx = -1; }
            else { by = i; bx = TS; }
            int gy = oy + by, gx = ox + bx;
            if (!interior) { gy = bmap(gy, d.ny, d.bc); gx = bmap(gx, d.nx, d.bc); }
            hreg[k] = band == 0 ? 1 : (band == 1 ? 7 : (band == 2 ? 3 : 5));
            hbox[k] = (by + 1) * LWD + C0 + bx;
            const int off = gy * d.nx + gx;
            hoff[k] = off;
            const int rt = __shfl(mc.type, hreg[k], 64);  // this element's region type
            hraw[k] = *(rt == 2 ? reinterpret_cast<const uint2*>(f64 + off)
                                : reinterpret_cast<const uint2*>(f16 + (off & ~3)));
        }
        if (town == 2) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                own[2 * q] = raw_f64(ro[q].x, ro[q].y);
                own[2 * q + 1] = raw_f64(ro[q].z, ro[q].w);
            }
        } else {
            const double o = own_off, sc = (double)__shfl(mc.scl, 4, 64);
            const unsigned wd[8] = {ro[0].x, ro[0].y, ro[0].z, ro[0].w, ro[1].x, ro[1].y, ro[1].z, ro[1].w};
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const unsigned v = wd[j >> 1];
                own[j] = o + (double)__half2float(__ushort_as_half((unsigned short)((j & 1) ? (v >> 16) : (v & 0xFFFFu)))) * sc;
            }
        }
        double* row = box + (ly + 1) * LWD + C0 + x0;
#pragma unroll
        for (int q = 0; q < 8; ++q) *reinterpret_cast<double2*>(row + 2 * q) = make_double2(own[2 * q], own[2 * q + 1]);
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            // shuffles outside the branch: a lane shuffles in 0 from an inactive lane
            const int rg = hreg[k];
            const int rt = __shfl(mc.type, rg, 64);
            const double ro = __shfl(mc.off, rg, 64);
            const float rs = __shfl(mc.scl, rg, 64);
            box[hbox[k]] = rt == 2 ? raw_f64(hraw[k].x, hraw[k].y)
                                   : ro + (double)__half2float(__ushort_as_half((unsigned short)raw_word4(hraw[k], hoff[k]))) *
                                              (double)rs;
        }
        __syncthreads();  // one wave: orders the box writes before the reads
        const double r = d.p[0] / d.p[1];  // dx / dy (slots 0 and 1 of every model)
        double m = 0.0;
        {
#pragma clang fp contract(off)
            const double* c = box + (ly + 1) * LWD + C0 + x0;
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const double hx = 0.5 * (c[i + 1] - c[i - 1]);
                const double hy = 0.5 * (c[i + LWD] - c[i - LWD]) * r;
                const double qv = hx * hx + hy * hy;
                if (qv > m) m = qv;
            }
        }
        m = wave_max_d(m);
        if (lane == 0) md[w] = m;
        __syncthreads();  // the box is rewritten by the next tile
    };
    MonMeta ma, mb;
    mon_meta_fetch(ma, a, inst[tinst[w]], types, field0, w);
    for (;;) {
        int wn = w + NG < n ? w + NG : w;
        mon_meta_fetch(mb, a, inst[tinst[wn]], types, field0, wn);
        tile(ma, w);
        w += NG;
        if (w >= n) break;
        wn = w + NG < n ? w + NG : w;
        mon_meta_fetch(ma, a, inst[tinst[wn]], types, field0, wn);
        tile(mb, w);
        w += NG;
        if (w >= n) break;
    }
}

int launch_monitor(const StepArgs& a, int kind, int tile, int ngroups_eval, int field0, int field_stride,
                   int type_groups, double* md, double* last, unsigned long long* reset, hipStream_t s) {
    const int n = ngroups_eval * a.total_tiles;
    if (kind == AMPS_MON_GRADIENT) {
        int grid = n > 65536 ? 65536 : n;
        if (tile == 16) {
            const int g16 = a.total_tiles > 262144 ? 262144 : a.total_tiles;
            hipLaunchKernelGGL(monitor_grad_wave16, dim3(g16), dim3(64), 0, s, a, a.inst, a.tile_inst, a.type,
                               ngroups_eval, field0, field_stride, type_groups, md, reset);
        }
        else if (tile == 32 && ngroups_eval == 1) {
            static int per_cu = 0, cus = 0;
            if (per_cu == 0) {
                int v = 0, dev = 0, c = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, monitor_grad_wave32, 64, 0) != hipSuccess || v <= 0) v = 8;
                if (hipGetDevice(&dev) != hipSuccess ||
                    hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0)
                    c = 256;
                per_cu = v;
                cus = c;
            }
            const int g32 = per_cu * cus < a.total_tiles ? per_cu * cus : a.total_tiles;
            hipLaunchKernelGGL(monitor_grad_wave32, dim3(g32 > 0 ? g32 : 1), dim3(64), 0, s, a, a.inst, a.tile_inst,
                               a.type, field0, md, reset);
        }
        else if (tile == 32) hipLaunchKernelGGL((monitor_grad_kernel<32>), dim3(grid), dim3(256), 0, s, a, ngroups_eval, field0, field_stride, type_groups, md, reset);
        else if (tile == 64) hipLaunchKernelGGL((monitor_grad_kernel<64>), dim3(grid), dim3(256), 0, s, a, ngroups_eval, field0, field_stride, type_groups, md, reset);
        else return AMPS_E_UNSUPPORTED;
    } else {
        hipLaunchKernelGGL(monitor_samples_kernel, dim3((n + 255) / 256), dim3(256), 0, s, a.inst, a.tile_inst, a.in,
                           a.total_cells, a.total_tiles, a.type, ngroups_eval, field0, field_stride, type_groups, tile,
                           kind, md, last, reset);
    }
    return hip_ok();
}

int launch_monitor_last_init(const StepArgs& a, int tile, int ngroups_eval, int field0, int field_stride,
                             double* last, hipStream_t s) {
    const int n = ngroups_eval * a.total_tiles;
    hipLaunchKernelGGL(monitor_last_init_kernel, dim3((n + 255) / 256), dim3(256), 0, s, a.inst, a.tile_inst, a.in,
                       a.total_cells, a.total_tiles, ngroups_eval, field0, field_stride, tile, last);
    return hip_ok();
}

int launch_classify(const InstDesc* inst, const int* tile_inst, int total_tiles, int ngroups, int broadcast,
                    int single, const double* md, double thr, int dilate, int* type_old, int* type,
                    unsigned long long* stats, hipStream_t s) {
    const int n = ngroups * total_tiles;
    const int per = 1024 * CL_PT;
    hipLaunchKernelGGL(classify_kernel, dim3(n > 0 ? (n + per - 1) / per : 1), dim3(1024), 0, s, inst, tile_inst,
                       total_tiles, ngroups, broadcast, single, md, thr, dilate, type_old, type, stats);
    return hip_ok();
}

int launch_gather(const InstDesc* inst, const Buf& b, long long total_cells, int total_tiles, const int* type,
                  int group, int instid, int ntiles_inst, int field, double* dst, int tile, hipStream_t s) {
    int grid = ntiles_inst > 65536 ? 65536 : ntiles_inst;
    if (tile == 16) hipLaunchKernelGGL((gather_kernel<16>), dim3(grid), dim3(256), 0, s, inst, b, total_cells, total_tiles, type, group, instid, field, dst);
    else if (tile == 32) hipLaunchKernelGGL((gather_kernel<32>), dim3(grid), dim3(256), 0, s, inst, b, total_cells, total_tiles, type, group, instid, field, dst);
    else if (tile == 64) hipLaunchKernelGGL((gather_kernel<64>), dim3(grid), dim3(256), 0, s, inst, b, total_cells, total_tiles, type, group, instid, field, dst);
    else return AMPS_E_UNSUPPORTED;
    return hip_ok();
}

int launch_clear_changed(const int* type_old, const int* type, const GroupMap& gm, int nfields, int total_tiles,
                         int* same_flags, unsigned long long* zmask, hipStream_t s) {
    const long long n = (long long)nfields * total_tiles;
    if (n <= 0) return AMPS_OK;
    hipLaunchKernelGGL(clear_changed_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, type_old, type, gm,
                       nfields, total_tiles, same_flags, zmask);
    return hip_ok();
}

int launch_worklist(const int* type, const int* type_old, int total_tiles, int* wl_hi, int* wl_lo, int* counts,
                    int* same_flags, hipStream_t s) {
    const int grid = total_tiles > 0 ? (total_tiles + WL_TILES - 1) / WL_TILES : 1;
    hipLaunchKernelGGL(worklist_kernel, dim3(grid), dim3(1024), 0, s, type, type_old, total_tiles, wl_hi, wl_lo, counts,
                       same_flags);
    return hip_ok();
}

int launch_nonfinite(const double* v, long long n, unsigned long long* cnt, hipStream_t s) {
    long long g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(nonfinite_kernel, dim3((int)g), dim3(256), 0, s, v, n, cnt);
    return hip_ok();
}

int launch_fill_u64(unsigned long long* p, int n, unsigned long long v, hipStream_t s) {
    hipLaunchKernelGGL(fill_u64_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
    return hip_ok();
}

int launch_advance_step(long long* step_ctr, int n, hipStream_t s) {
    hipLaunchKernelGGL(advance_step_kernel, dim3(1), dim3(64), 0, s, step_ctr, n);
    return hip_ok();
}

}  // namespace amps

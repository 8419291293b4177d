// tile_common.hpp — device building blocks shared by the model kernels:
// neighbour-tile table, precision-aware halo loads into LDS, block
// reductions and the per-tile fp16 encoder.
//
// A workgroup (256 threads) owns one square tile of TS x TS cells and stages
// the tile plus an H-cell halo ("box", LW = TS + 2H per side) in LDS.  Cells
// of a halo may belong to neighbour tiles held in the other precision; the
// loaders read each neighbour from the storage its type says is current:
//   FP64 tile: f64[cell]
//   low tile : off[tile] + f16[cell] * scl[tile]
#pragma once

#include <hip/hip_ext.h>

#include "amps_internal.hpp"

namespace amps {

constexpr int WG = 256;

// Launch a sweep kernel.  With ev (two timing events), hipExtLaunchKernel
// binds both events to the dispatch itself, so their elapsed time is the
// kernel's own execution time (no dispatch gap); without, a plain launch.
template <typename K, typename... A>
inline void launch_sweep(K k, dim3 grid, dim3 blk, hipStream_t s, const hipEvent_t* ev, A... args) {
    if (ev) hipExtLaunchKernelGGL(k, grid, blk, 0, s, ev[0], ev[1], 0, args...);
    else hipLaunchKernelGGL(k, grid, blk, 0, s, args...);
}

// Type of (tile, group); group < 0 means the field is always FP64.
__device__ inline int tile_type(const StepArgs& a, int group, int gtile) {
    if (a.type == nullptr || group < 0) return 2;
    return a.type[(size_t)group * a.total_tiles + gtile];
}

// The tiles around the workgroup's tile, in box order k = (dy+1)*3 + (dx+1).
// For zero-flux edges the out-of-domain direction maps back into the own tile
// (mirror), for periodic edges to the wrapped tile.
struct Nb {
    const InstDesc* d;
    int tile_x, tile_y, TS;
};

__device__ inline void nb_setup(Nb& nb, const InstDesc& d, int tile_x, int tile_y, int TS) {
    nb.d = &d;
    nb.tile_x = tile_x;
    nb.tile_y = tile_y;
    nb.TS = TS;
}

__device__ inline int nb_tile(const Nb& nb, int k) {
    const InstDesc& d = *nb.d;
    const int TS = nb.TS, ox = nb.tile_x * TS, oy = nb.tile_y * TS;
    const int dy = k / 3 - 1, dx = k % 3 - 1;
    const int gx = bmap(dx < 0 ? ox - 1 : (dx > 0 ? ox + TS : ox), d.nx, d.bc);
    const int gy = bmap(dy < 0 ? oy - 1 : (dy > 0 ? oy + TS : oy), d.ny, d.bc);
    return d.tile_base + (gy / TS) * d.tx + gx / TS;
}

// Box region (0..8) of box coordinates (ly, lx).
template <int TS, int H>
__device__ inline int box_region(int ly, int lx) {
    int rx = lx < H ? 0 : (lx < H + TS ? 1 : 2);
    int ry = ly < H ? 0 : (ly < H + TS ? 1 : 2);
    return ry * 3 + rx;
}

// Per-workgroup cache of the neighbour tiles' precision state for one field.
struct NbField {
    int type[9];
    double off[9];
    float scl[9];
};

// Threads 0..8 record the type, offset and scale of neighbour k = threadIdx.x
// (offset and scale are read whatever the type, so no load waits on another).
__device__ inline void nbfield_fill(NbField* s, const StepArgs& a, const Buf& b, const Nb& nb, int field,
                                    int group) {
    int t = threadIdx.x;
    if (t < 9) {
        int g = nb_tile(nb, t);
        int ty = tile_type(a, group, g);
        double off = b.off ? b.off[(size_t)field * a.total_tiles + g] : 0.0;
        float scl = b.scl ? b.scl[(size_t)field * a.total_tiles + g] : 0.0f;
        s->type[t] = ty;
        s->off[t] = off;
        s->scl[t] = scl;
    }
}

// Two fields of one type group: every load of both tables is issued before
// the first LDS store, so the two tables cost one round of loads.
__device__ inline void nbfield_fill2(NbField* s0, NbField* s1, const StepArgs& a, const Buf& b, const Nb& nb,
                                     int field0, int field1, int group) {
    int t = threadIdx.x;
    if (t < 9) {
        int g = nb_tile(nb, t);
        int ty = tile_type(a, group, g);
        double off0 = b.off ? b.off[(size_t)field0 * a.total_tiles + g] : 0.0;
        double off1 = b.off ? b.off[(size_t)field1 * a.total_tiles + g] : 0.0;
        float scl0 = b.scl ? b.scl[(size_t)field0 * a.total_tiles + g] : 0.0f;
        float scl1 = b.scl ? b.scl[(size_t)field1 * a.total_tiles + g] : 0.0f;
        s0->type[t] = ty;
        s0->off[t] = off0;
        s0->scl[t] = scl0;
        s1->type[t] = ty;
        s1->off[t] = off1;
        s1->scl[t] = scl1;
    }
}

// ------------------------------------------------------- batched box load
// Offsets (inside one instance's field array) of the box elements a thread
// stages: element e = threadIdx.x + k*WG.  Computed once per tile and shared
// by every field of the tile; interior tiles skip the boundary map.
template <int TS, int H, int NT = WG>
struct BoxMap {
    static constexpr int LW = TS + 2 * H, NE = LW * LW, N = (NE + NT - 1) / NT;
    int off[N];
    int reg[N];   // 3x3 region of the element (neighbour table index)
};

template <int TS, int H, int NT = WG>
__device__ inline void box_map(BoxMap<TS, H, NT>& m, const InstDesc& d, int tile_x, int tile_y) {
    constexpr int LW = TS + 2 * H, NE = LW * LW, N = BoxMap<TS, H, NT>::N;
    const int ox = tile_x * TS - H, oy = tile_y * TS - H;
    const bool interior = ox >= 0 && oy >= 0 && ox + LW <= d.nx && oy + LW <= d.ny;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const int e = threadIdx.x + k * NT;
        const int ly = e / LW, lx = e - ly * LW;
        int gy = oy + ly, gx = ox + lx;
        if (!interior) {
            gy = bmap(gy, d.ny, d.bc);
            gx = bmap(gx, d.nx, d.bc);
        }
        m.off[k] = (e < NE) ? gy * d.nx + gx : -1;
        m.reg[k] = box_region<TS, H>(ly, lx);
    }
}

// Stage two FP64 fields whose every tile is FP64.  All loads are issued
// before the first LDS store so their latencies overlap.
template <int TS, int H>
__device__ inline void box_load2_f64(double* b0, double* b1, const double* f0, const double* f1,
                                     const BoxMap<TS, H>& m) {
    constexpr int N = BoxMap<TS, H>::N;
    double v0[N], v1[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        if (m.off[k] >= 0) {
            v0[k] = f0[m.off[k]];
            v1[k] = f1[m.off[k]];
        }
    }
#pragma unroll
    for (int k = 0; k < N; ++k) {
        if (m.off[k] >= 0) {
            b0[threadIdx.x + k * WG] = v0[k];
            b1[threadIdx.x + k * WG] = v1[k];
        }
    }
}

template <int TS, int H>
__device__ inline void box_load1_f64(double* b0, const double* f0, const BoxMap<TS, H>& m) {
    constexpr int N = BoxMap<TS, H>::N;
    double v0[N];
#pragma unroll
    for (int k = 0; k < N; ++k)
        if (m.off[k] >= 0) v0[k] = f0[m.off[k]];
#pragma unroll
    for (int k = 0; k < N; ++k)
        if (m.off[k] >= 0) b0[threadIdx.x + k * WG] = v0[k];
}

// Stage one field as FP64 values when neighbour tiles may be low precision
// (decoded exactly: off + h * scl).
template <int TS, int H>
__device__ inline void box_load_mixed_f64(double* b0, const double* f64, const __half* f16, const NbField* nbf,
                                          const BoxMap<TS, H>& m) {
    constexpr int N = BoxMap<TS, H>::N;
    double v[N];
    __half h[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        if (m.off[k] >= 0) {
            if (nbf->type[m.reg[k]] == 2) v[k] = f64[m.off[k]];
            else h[k] = f16[m.off[k]];
        }
    }
#pragma unroll
    for (int k = 0; k < N; ++k) {
        if (m.off[k] >= 0) {
            const int r = m.reg[k];
            b0[threadIdx.x + k * WG] = nbf->type[r] == 2 ? v[k] : nbf->off[r] + (double)__half2float(h[k]) * (double)nbf->scl[r];
        }
    }
}

// Stage one field as FP32 deviations from own_off (the workgroup's tile is low).
template <int TS, int H>
__device__ inline void box_load_dev32(float* b0, const double* f64, const __half* f16, const NbField* nbf,
                                      double own_off, const BoxMap<TS, H>& m) {
    constexpr int N = BoxMap<TS, H>::N;
    double v[N];
    __half h[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        if (m.off[k] >= 0) {
            if (nbf->type[m.reg[k]] == 2) v[k] = f64[m.off[k]];
            else h[k] = f16[m.off[k]];
        }
    }
#pragma unroll
    for (int k = 0; k < N; ++k) {
        if (m.off[k] >= 0) {
            const int r = m.reg[k];
            float x;
            if (nbf->type[r] == 2) x = (float)(v[k] - own_off);
            else if (r == 4) x = __half2float(h[k]) * nbf->scl[4];
            else x = (float)(nbf->off[r] - own_off) + __half2float(h[k]) * nbf->scl[r];
            b0[threadIdx.x + k * WG] = x;
        }
    }
}

// 1/x to ~1 ulp: hardware reciprocal estimate refined by two Newton steps.
__device__ inline double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    r = fma(fma(-x, r, 1.0), r, r);
    return r;
}

// ----------------------------------------------------------- reductions
// Wave-wide maximum of doubles (DPP, result in every lane) from the device library.
extern "C" __device__ double __ockl_wfred_max_f64(double);

__device__ inline double wave_max_d(double v) { return __ockl_wfred_max_f64(v); }

// Order-preserving map between float bit patterns and signed integers (an
// involution): non-negative floats keep their bits, negative ones have their
// magnitude bits flipped, so integer order is float order (-0 < +0).
__device__ inline int f32_key(float v) {
    const int b = __float_as_int(v);
    return b ^ ((b >> 31) & 0x7FFFFFFF);
}
__device__ inline float key_f32(int k) { return __int_as_float(k ^ ((k >> 31) & 0x7FFFFFFF)); }

// Raw field data loaded as plain words (one load sequence for FP64 and fp16
// tiles): an FP64 value from its two 32-bit halves as stored (low first); the
// fp16 word at position p & 3 of an aligned 4-word group; the 4 words at
// position p & 4 of an aligned 8-word group.
__device__ inline double raw_f64(unsigned lo, unsigned hi) { return __hiloint2double((int)hi, (int)lo); }

__device__ inline unsigned raw_word4(uint2 g, int p) {
    const unsigned long long q = ((unsigned long long)g.y << 32) | g.x;
    return (unsigned)(q >> ((p & 3) << 4)) & 0xFFFFu;
}

__device__ inline uint2 raw_words8(uint4 g, int p) {
    const bool up = (p & 4) != 0;
    return make_uint2(up ? g.z : g.x, up ? g.w : g.y);
}

// Min and max of non-NaN floats over a full wave (all 64 lanes active),
// returned in every lane.  Integer keys let each step of the row_shr DPP
// ladder be one integer min/max with the shifted operand (lanes shifted in
// from outside the 16-lane row read the identity); lane 15 of each row then
// holds the row's extreme, and the four rows are combined from scalar reads.
__device__ inline void wave_minmax(float& mn, float& mx) {
    int a = f32_key(mn), b = f32_key(mx);
#define AMPS_DPP_STEP(CTRL)                                                                  \
    a = min(a, __builtin_amdgcn_update_dpp(0x7FFFFFFF, a, CTRL, 0xF, 0xF, false));          \
    b = max(b, __builtin_amdgcn_update_dpp((int)0x80000000, b, CTRL, 0xF, 0xF, false))
    AMPS_DPP_STEP(0x111);  // row_shr:1
    AMPS_DPP_STEP(0x112);  // row_shr:2
    AMPS_DPP_STEP(0x114);  // row_shr:4
    AMPS_DPP_STEP(0x118);  // row_shr:8
#undef AMPS_DPP_STEP
    const int a0 = __builtin_amdgcn_readlane(a, 15), a1 = __builtin_amdgcn_readlane(a, 31);
    const int a2 = __builtin_amdgcn_readlane(a, 47), a3 = __builtin_amdgcn_readlane(a, 63);
    const int b0 = __builtin_amdgcn_readlane(b, 15), b1 = __builtin_amdgcn_readlane(b, 31);
    const int b2 = __builtin_amdgcn_readlane(b, 47), b3 = __builtin_amdgcn_readlane(b, 63);
    mn = key_f32(min(min(a0, a1), min(a2, a3)));
    mx = key_f32(max(max(b0, b1), max(b2, b3)));
}

// Min and max over the workgroup; red must hold 2*WG/64 floats.  Contains
// barriers: every thread must call it.
// NWAVES waves per workgroup; red holds 2 * NWAVES floats.
template <int NWAVES = WG / 64>
__device__ inline void block_minmax(float& mn, float& mx, float* red) {
    wave_minmax(mn, mx);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) { red[w] = mn; red[NWAVES + w] = mx; }
    __syncthreads();
    mn = red[0];
    mx = red[NWAVES];
#pragma unroll
    for (int i = 1; i < NWAVES; ++i) {
        mn = fminf(mn, red[i]);
        mx = fmaxf(mx, red[NWAVES + i]);
    }
}

// Min and max of two values over the workgroup with a single barrier; red
// must hold 4*WG/64 floats.  No leading barrier: the caller guarantees that
// every thread has passed a barrier since it last read red.
__device__ inline void block_minmax2(float& mn0, float& mx0, float& mn1, float& mx1, float* red) {
    static_assert(WG == 256, "four waves per workgroup");
    wave_minmax(mn0, mx0);
    wave_minmax(mn1, mx1);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) *reinterpret_cast<float4*>(red + 4 * w) = make_float4(mn0, mx0, mn1, mx1);
    __syncthreads();
    const float4 r0 = reinterpret_cast<const float4*>(red)[0], r1 = reinterpret_cast<const float4*>(red)[1];
    const float4 r2 = reinterpret_cast<const float4*>(red)[2], r3 = reinterpret_cast<const float4*>(red)[3];
    mn0 = fminf(fminf(r0.x, r1.x), fminf(r2.x, r3.x));
    mx0 = fmaxf(fmaxf(r0.y, r1.y), fmaxf(r2.y, r3.y));
    mn1 = fminf(fminf(r0.z, r1.z), fminf(r2.z, r3.z));
    mx1 = fmaxf(fmaxf(r0.w, r1.w), fmaxf(r2.w, r3.w));
}

__device__ inline double block_max_d(double v, double* red) {
    v = wave_max_d(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    return fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// ------------------------------------------------------------ fp16 codec
// A low tile stores h = (value - off) / scl in fp16 (round to nearest even),
// where off is the tile's mid-range and scl = 2^(e-14) for a half-range
// below 2^e, so |h| < 2^14 (fp16 max 65504) and the scaling is exact.  e is
// clamped to >= -100.  A uniform tile (half-range 0, every h = 0) stores
// scale 0: its values decode to off exactly, |off| + 2^14 scl bounds them
// tightly, and scl == 0 tells a reader that the tile is uniform.
struct TileCode {
    float mid;   // added to the old offset
    float inv;   // 1/scl
    float scl;
};

__device__ inline TileCode tile_code(float mn, float mx) {
    TileCode t;
    t.mid = 0.5f * (mn + mx);
    float hw = 0.5f * (mx - mn);
    int ex = hw == 0.0f ? -100 : 0;  // a uniform tile takes the smallest scale
    if (hw > 0.0f && hw <= 3.0e38f) frexpf(hw, &ex);
    if (ex < -100) ex = -100;
    t.scl = hw == 0.0f ? 0.0f : ldexpf(1.0f, ex - 14);  // uniform tile: scale 0 marks it
    t.inv = ldexpf(1.0f, 14 - ex);
    return t;
}

}  // namespace amps

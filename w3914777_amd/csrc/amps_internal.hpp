// amps_internal.hpp — definitions shared by the host runtime and the gfx950
// kernels of the adaptive mixed-precision stencil engine.
//
// Device memory layout (one context, possibly many instances):
//   f64[buf][field]   FP64 fields, every instance's cells concatenated
//                     (instance i starts at cell_base[i]); row-major [ny][nx]
//   f16[buf][field]   fp16 payload of the same cells, meaningful only inside
//                     low-precision tiles; value = off + h * scl
//   off/scl[buf][field][tile]  per-tile FP64 offset and power-of-two scale
//   type[group][tile] 1 = low (fp16 storage, FP32 arithmetic), 2 = FP64
//   md[group][tile]   last monitor values; last[group][tile][16] Monitor2 history
// buf is the ping-pong index of the explicit-Euler step.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "../../include/amps.h"

#define AMPS_MAX_FIELDS 40   // grain growth q <= 38 plus headroom
#define AMPS_MAX_INST 256

namespace amps {

// Per-instance constants, built on the host, read by every kernel through a
// device pointer (scalar loads).
struct InstDesc {
    int nx, ny;          // grid
    int tx, ty;          // tiles per row / column
    int tile_base;       // first global tile index of this instance
    int bc, stencil;
    int pad0;
    long long cell_base; // first cell of this instance inside each field array
    unsigned long long seed; // counter-RNG seed of this instance
    double p[AMPS_NPARAMS];
    // derived constants (filled per model, see host_derive_constants)
    double k[24];
    float kf[24];
};

// Constants derived per model (index into InstDesc::k / kf).
enum {
    // snowflake
    SFK_INV_DXDY = 0, SFK_INV_2DX, SFK_INV_2DY, SFK_DT_OVER_TAU, SFK_DT, SFK_KAPPA, SFK_ALPHA_OVER_PI,
    SFK_GAMMA, SFK_TEQ, SFK_EPSB, SFK_EPSB_DELTA, SFK_MEPSB_ANISO_DELTA, SFK_COS_N_THETA0,
    SFK_SIN_N_THETA0, SFK_NOISE, SFK_ANISO_N,
    // grain growth
    GGK_INV_DXDY = 0, GGK_DT_L, GGK_KAPPA, GGK_ALPHA, GGK_BETA, GGK_2GAMMA,
    // sintering
    SNK_INV_DXDY = 0, SNK_DT, SNK_A2, SNK_B, SNK_HALF_KC, SNK_DVOL, SNK_DVAP, SNK_DSURF, SNK_DGB,
    SNK_DT_L, SNK_HALF_KE
};

// ----------------------------------------------------------- index mapping
// Periodic wrap (reference: snowflake/function.h:13-18) or zero-flux mirror
// f(-k) = f(k).  Valid for |overhang| < n.
__host__ __device__ inline int bmap(int i, int n, int bc) {
    if (i >= 0 && i < n) return i;
    if (bc == AMPS_BC_PERIODIC) return i < 0 ? i + n : i - n;
    return i < 0 ? -i : 2 * (n - 1) - i;
}

// ------------------------------------------------------------- counter RNG
// SplitMix64 finaliser keyed by (seed, stream, index); bit-identical to the
// host implementation so seeded noise is reproducible on any device.
__host__ __device__ inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t rng_key(uint64_t seed, uint64_t stream) {
    return splitmix64(seed ^ splitmix64(stream + 0x632BE59BD9B4E019ull));
}
__host__ __device__ inline double rng_u01_keyed(uint64_t key, uint64_t idx) {
    return (double)(splitmix64(key ^ idx) >> 11) * 0x1.0p-53;
}
// Per-step interface noise: a 32-bit key per (seed, step) from SplitMix64 and
// one 32-bit mix per cell (Wellons' "lowbias32" finaliser), 24 bits of which
// give u in [0, 1) exactly representable in FP32 and FP64.
__host__ __device__ inline uint32_t lowbias32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
__host__ __device__ inline uint32_t noise_key(uint64_t seed, uint64_t step) { return (uint32_t)rng_key(seed, step); }
__host__ __device__ inline float noise_u24(uint32_t key, uint32_t cell) {
    return (float)(lowbias32(key ^ cell) >> 8) * 0x1.0p-24f;
}
static const uint64_t STREAM_GG_INIT = 1ull << 40;
static const uint64_t STREAM_SN_INIT = 2ull << 40;

// ------------------------------------------------------------ field views
// One buffer generation of a context.  Field f of every instance lives at
// f64 + f * total_cells (likewise f16); per-tile arrays at off + f * total_tiles.
struct Buf {
    double* f64;   // [nfields][total_cells]
    __half* f16;   // [nfields][total_cells]
    double* off;   // [nfields][total_tiles]  low tiles: value = off + h * scl
    float* scl;    // [nfields][total_tiles]  power of two
    unsigned long long* zmask;  // grain growth: [total_tiles] bit f = field f is low and exactly 0
                                // on the tile (a clear bit only means "not known")
};

struct StepArgs {
    const InstDesc* inst;       // [ninst]
    const int* tile_inst;       // [total_tiles] -> instance
    const int* worklist;        // [nwork] global tile indices (nullptr: identity)
    const int* wl_count;        // device count of worklist entries (nullptr: nwork)
    const int* wl_low;          // low-precision tile work-list (snowflake adaptive)
    const int* wl_low_count;
    int* same_flags;            // [total_tiles] snowflake low tiles: bit 0 both buffers hold the same phi,
                                // bit 1 the same T (words, offset and scale)
    int* full_seg;              // snowflake low tiles listed for the mixed sweep this step: [8][full_cap]
                                // (bulk-hold sweep) or [NSK_SEGS][nsk_cap] (NO_SKIP flat sweep)
    int* full_cnt;              // entries per segment by step parity: [2][8], or NO_SKIP
                                // [2][NSK_SEGS] at a stride of NSK_CNT_STRIDE ints
    int full_cap;
    unsigned long long* stats;  // [2] FP64 tiles, [3] low tiles given the full update, [4] NO_SKIP: low
                                // tiles the flat sweep listed (summed over steps)
    const int* type;            // [group][total_tiles] (nullptr: all FP64)
    long long total_cells;
    int total_tiles;
    int nwork;
    int nfields;
    int q;                      // grain growth order parameters
    const long long* step_ctr;  // device step counter (steps before this launch)
    int step_add;               // added to *step_ctr
    int ninst;                  // instances in the context (1: descriptor is loop-invariant)
    int grid_cap;               // sweep workgroups at most (0: no cap)
    int t_halo;                 // snowflake: halo depth of the T Laplacian (2: reference stencil in some instance)
    int flags;                  // AMPS_FLAG_* of the run (NO_SKIP: no hold, flush or skip on low tiles)
    Buf in, out;
    double* aux_in;             // model auxiliary per-cell array (grain growth: sum of squares)
    double* aux_out;
};

// NO_SKIP list of the snowflake's flat sweep: NSK_SEGS segments (list
// position mod NSK_SEGS) of nsk_cap(total_tiles) entries, their counters
// NSK_CNT_STRIDE ints (256 B) apart so that the appends of one launch spread
// over as many cache lines (one line's atomics serialise).
constexpr int NSK_SEGS = 64, NSK_CNT_STRIDE = 64;
__host__ __device__ inline int nsk_cap(int total_tiles) { return (total_tiles + NSK_SEGS - 1) / NSK_SEGS; }

// Records the message returned by amps_last_error() for the calling thread.
void set_error_text(const char* msg);

// XCD-aware work mapping: the hardware deals workgroups round-robin over the 8
// XCDs; give each XCD a contiguous run of tiles so that halo re-reads hit the
// XCD's own L2.  Bijective on [0, n).
__device__ inline int xcd_remap(int b, int n) {
    int per = n >> 3;
    if (b < (per << 3)) return (b & 7) * per + (b >> 3);
    return b;
}

// Chunked, balanced variant for items of uneven cost: chunks of XCD_CHUNK
// consecutive items, chunk j to XCD j mod 8, so that neighbouring items
// share an XCD's L2 while every XCD samples the whole range.  The
// workgroup b (dealt to XCD b mod 8, its (b / 8)-th there) gets item
// xcd_chunk_remap(b, n); bijective on [0, n).
constexpr int XCD_CHUNK = 12;
__device__ inline int xcd_chunk_remap(int b, int n) {
    const int full = n / (8 * XCD_CHUNK) * (8 * XCD_CHUNK);
    if (b >= full) return b;
    const int x = b & 7, q = b >> 3;
    return ((q / XCD_CHUNK) * 8 + x) * XCD_CHUNK + q % XCD_CHUNK;
}

// Inverse of xcd_chunk_remap: the b with xcd_chunk_remap(b, n) == r.
__device__ inline int xcd_chunk_unmap(int r, int n) {
    const int full = n / (8 * XCD_CHUNK) * (8 * XCD_CHUNK);
    if (r >= full) return r;
    const int j = r / XCD_CHUNK, o = r - j * XCD_CHUNK;
    return ((j >> 3) * XCD_CHUNK + o) * 8 + (j & 7);
}

// Inverse of xcd_remap: the b with xcd_remap(b, n) == r.
__device__ inline int xcd_unmap(int r, int n) {
    int per = n >> 3;
    if (r < (per << 3)) return (r % per) * 8 + r / per;
    return r;
}

}  // namespace amps

/*
 * amps.h — C ABI of the MI355X adaptive mixed-precision phase-field stencil engine.
 *
 * One context (amps_ctx) owns every device buffer, stream and graph of one run
 * (or one batched ensemble of runs).  Fields cross this boundary as FP64,
 * row-major [ny][nx], whatever precision a tile is held in on the device.
 * Every call returns AMPS_OK (0) or a negative AMPS_E_* status; the library
 * never exits or aborts.  One host thread may drive a context at a time;
 * distinct contexts are independent.
 *
 * The reference (AMPStencil) has no library interface: its "boundary" is a set
 * of __global__ kernels on raw managed pointers driven by per-application
 * main()s, with compile-time parameters.  Each entry point below names the
 * reference code it replaces.
 */
#ifndef AMPS_H
#define AMPS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AMPS_ABI_VERSION 2

/* ---- status codes (replace CHECK_ERROR/exit(EXIT_FAILURE),
 *      snowflake/paras/total.h:48-65) ---- */
enum {
    AMPS_OK = 0,
    AMPS_E_ARG = -1,         /* invalid argument or configuration            */
    AMPS_E_NOMEM = -2,       /* device or host allocation failed             */
    AMPS_E_HIP = -3,         /* HIP runtime error                            */
    AMPS_E_PARSE = -4,       /* paras/ file could not be parsed              */
    AMPS_E_NONFINITE = -5,   /* NaN/Inf found in a field                     */
    AMPS_E_STATE = -6,       /* call not valid in the context's state        */
    AMPS_E_UNSUPPORTED = -7, /* combination not implemented                  */
    AMPS_E_IO = -8           /* file could not be read or written            */
};

/* ---- enumerations ---- */
typedef enum { AMPS_SNOWFLAKE = 1, AMPS_GRAINGROWTH = 2, AMPS_SINTERING = 3 } amps_model;

/* PURE_FP64  = the reference's PURE build with highprecision=double
 * PURE_HALF2 = every tile low precision (reference purelow.cu)
 * ADAPTIVE   = per-tile monitor + conversion (reference AMSTENCIL)
 * STATIC_MAP = caller-supplied tile map, never re-evaluated (GRAM-style) */
typedef enum { AMPS_PURE_FP64 = 0, AMPS_PURE_HALF2 = 1, AMPS_ADAPTIVE = 2, AMPS_STATIC_MAP = 3 } amps_mode;

/* SPATIAL  = Monitor1, max-min of a 4x4 sample lattice (snowflake/tools.h:177-193)
 * TEMPORAL = Monitor2, max |f - f_last| on a 4x4 lattice (snowflake/tools.h:194-211)
 * GRADIENT = max over the tile of |grad f|^2 dx^2 (central differences);
 *            compared against threshold^2.  No reference counterpart. */
typedef enum { AMPS_MON_NONE = 0, AMPS_MON_SPATIAL = 1, AMPS_MON_TEMPORAL = 2, AMPS_MON_GRADIENT = 3 } amps_monitor;

/* REF  = the reference operator of each model: radius-2 axis cross for the
 *        snowflake (snowflake/function.h:27-28), Moore-8 for grain growth
 *        (graingrowth/function.h:20), 5-point for sintering (sintering/function.h:67)
 * LAP5 = 5-point Laplacian for every model */
typedef enum { AMPS_STENCIL_REF = 0, AMPS_STENCIL_LAP5 = 1 } amps_stencil;

/* PERIODIC = the reference's wrap (snowflake/function.h:13-18)
 * NEUMANN  = zero flux, mirror ghost cells f(-k) = f(k) */
typedef enum { AMPS_BC_PERIODIC = 0, AMPS_BC_NEUMANN = 1 } amps_bc;

/* REF    = the reference's initial state (snowflake/tools.h:24-33,
 *          graingrowth/tools.h:24-38, sintering/main.cu:45-60)
 * CONFIG = seeded variants: snowflake disc (same rule, centre/radius from
 *          params), grain growth eta_i ~ U[-amp, amp], sintering tanh discs
 *          plus noise */
typedef enum { AMPS_INIT_REF = 0, AMPS_INIT_CONFIG = 1 } amps_init_kind;

/* ---- parameter slots of amps_config.params ---- */
/* snowflake (snowflake/paras/para1.h:4-15, total.h:47) */
enum {
    AMPS_SF_DX = 0, AMPS_SF_DY, AMPS_SF_DT, AMPS_SF_TAU, AMPS_SF_EPSB, AMPS_SF_KAPPA,
    AMPS_SF_DELTA, AMPS_SF_ANISO, AMPS_SF_ALPHA, AMPS_SF_GAMMA, AMPS_SF_TEQ,
    AMPS_SF_THETA0, AMPS_SF_PI, AMPS_SF_NOISE, AMPS_SF_R2, AMPS_SF_CX, AMPS_SF_CY
};
/* grain growth (graingrowth/paras/para1.h:4-6) */
enum {
    AMPS_GG_DX = 0, AMPS_GG_DY, AMPS_GG_DT, AMPS_GG_L, AMPS_GG_KAPPA, AMPS_GG_ALPHA,
    AMPS_GG_BETA, AMPS_GG_GAMMA, AMPS_GG_R, AMPS_GG_CX, AMPS_GG_CY, AMPS_GG_AMP
};
/* sintering (sintering/paras/para1.h:5-7, sintering/function.h:72) */
enum {
    AMPS_SN_DX = 0, AMPS_SN_DY, AMPS_SN_DT, AMPS_SN_A, AMPS_SN_B, AMPS_SN_KC, AMPS_SN_KE,
    AMPS_SN_L, AMPS_SN_DVOL, AMPS_SN_DVAP, AMPS_SN_DSURF, AMPS_SN_DGB, AMPS_SN_R1,
    AMPS_SN_R2, AMPS_SN_X1, AMPS_SN_Y1, AMPS_SN_X2, AMPS_SN_Y2, AMPS_SN_WIDTH, AMPS_SN_NOISE
};
#define AMPS_NPARAMS 32

/* amps_config.flags bits.
 * NO_SKIP   : low-precision tiles run the full update every step.  Without it
 *             the engine skips work it bounds as negligible: snowflake tiles
 *             whose phi is bulk (within 1e-8 of 0 or 1 on the tile and its 8
 *             neighbours) advance only T, and T is held where it is flat within
 *             1e-12; the low-tile noise term is skipped where |phi(1-phi)| <=
 *             1e-6; grain-growth fields within 1e-10 of 0 on 3x3 tiles are set
 *             to exactly 0, and grain interiors are held.  The reference updates
 *             every type-1 tile every step (snowflake/function.h:376-604,
 *             graingrowth/function.h:33-277): NO_SKIP is that schedule.
 *             (Under NO_SKIP a snowflake low tile whose 3x3 neighbourhood
 *             decodes to exact zeros at phi 0 or 1 gets phi' = +0 and T' from
 *             the T update alone: the values the full update gives, without
 *             the terms that are signed zeros there.)
 * CH_CONSERVATIVE : sintering Cahn-Hilliard update in divergence form,
 *             c' = c + dt div(M grad mu) with face mobilities
 *             M_{i+1/2} = (M_i + M_{i+1})/2, in FP64 (mass-conserving up to
 *             the [0,1] clamp).  Without it, the reference's c' = c + dt M lap mu
 *             (sintering/function.h:95-101).
 * HALF2     : snowflake low tiles evaluate their pointwise terms in packed
 *             half2 (v_pk_fma_f16) on the offset-and-scaled fields: the
 *             six-fold unit vector (cos 6 theta, sin 6 theta) of each ring
 *             cell and its rotation by 6 theta0 (two ring cells per packed
 *             operation), and the atan of the driving force (two cells per
 *             operation, a degree-7 odd polynomial on min(|x|, 1/|x|)).
 *             Gradients, Laplacians, flux sums, eps, eps' and the update
 *             stay in FP32.  Needs NO_SKIP and tile 32; an aniso other than 6
 *             keeps the FP32 anisotropy.  The reference's low branch is half2
 *             throughout (snowflake/function.h:417-537). */
enum { AMPS_FLAG_NO_SKIP = 1, AMPS_FLAG_CH_CONSERVATIVE = 2, AMPS_FLAG_HALF2 = 4 };

/* Field indices: snowflake 0 = phi, 1 = T; grain growth 0..q-1 = eta_i;
 * sintering 0 = c, 1 = eta_1, 2 = eta_2. */

typedef struct {
    int32_t model;          /* amps_model                                          */
    int32_t mode;           /* amps_mode                                           */
    int32_t monitor;        /* amps_monitor (ADAPTIVE only)                        */
    int32_t nx, ny;         /* grid; multiples of tile                             */
    int32_t q;              /* grain growth: order parameters (>=2); others ignore */
    int32_t tile;           /* square tile edge: 16, 32 or 64                      */
    int32_t monitor_every;  /* steps between monitor passes (reference: 10)        */
    int32_t monitor_field;  /* field the monitor reads; -1 (grain growth only):
                               each field gets its own per-tile type             */
    int32_t stencil;        /* amps_stencil                                        */
    int32_t bc;             /* amps_bc                                             */
    int32_t init;           /* amps_init_kind                                      */
    int32_t wg_per_cu;      /* sweep workgroups per CU (sets tiles per workgroup);
                               0 = as many as fit (occupancy maximum).  The
                               adaptive sintering sweep and the PURE_FP64
                               snowflake sweep run one workgroup per tile and
                               ignore it                                      */
    int32_t flags;          /* AMPS_FLAG_* bits; 0 = defaults                      */
    int32_t dilate;         /* ADAPTIVE: the FP64 set also takes every tile within
                               this Chebyshev distance (in tiles) of a tile whose
                               monitor value reaches the threshold; 0 = only the
                               reference's von-Neumann neighbour rule             */
    double threshold;       /* monitor threshold                                   */
    uint64_t seed;          /* counter-based RNG seed (noise, random init)         */
    double params[AMPS_NPARAMS];
} amps_config;

typedef struct {
    int64_t steps_done;      /* steps advanced since amps_init                      */
    int64_t monitor_passes;  /* monitor passes run                                  */
    int32_t ntiles;          /* tiles per instance (largest instance for batches)  */
    int32_t type_groups;     /* tile-type maps per instance (grain growth: q)      */
    int32_t nfields;         /* fields per instance                                 */
    int32_t ninstances;
    double hi_frac_last;     /* fraction of FP64 tiles after the last monitor pass */
    double hi_frac_mean;     /* mean over monitor passes                            */
    double cells;            /* grid points per step, summed over instances        */
    int64_t fp64_tile_updates;     /* snowflake low-precision steps since amps_init: FP64
                                      tiles updated, summed over steps                */
    int64_t low_full_tile_updates; /* ... and low tiles given the full update (without
                                      NO_SKIP: those outside the bulk-hold state)     */
    int64_t low_listed_tile_updates; /* NO_SKIP: low tiles the flat sweep left to the
                                        mixed sweep (their neighbourhood not flat)    */
} amps_info;

/* ---- configuration ---- */

/* Fill cfg with the reference's para1 values of `model`: 512x512, tile 16,
 * reference stencil/BC/init, PURE_FP64, monitor TEMPORAL every 10 steps with
 * threshold 0.99 (snowflake/paras/total.h:19-23, para1.h). */
int amps_default_config(int model, amps_config* cfg);

/* Parse a parameter file in the reference's header syntax
 * (`const <type> a=1.0,b=a*2;`, `typedef double highprecision;`,
 * `int timesteps=500;`, `#define Monitor1`) and override the values of cfg
 * it names.  Replaces the compile-time paras/paraK.h + total.h configuration and
 * the line rewriting of the harness (snowflake/end2end_mix.py:23-35).
 * *timesteps (if not NULL) receives `timesteps` when the file sets it. */
int amps_load_paras(const char* path, amps_config* cfg, int* timesteps);

/* ---- lifetime ---- */

typedef struct amps_ctx amps_ctx;

/* Allocate a context for one grid.  Replaces the cudaMallocManaged block of
 * every driver (e.g. snowflake/end2end_mix.cu:4-31, sintering/main.cu:6-44). */
int amps_create(const amps_config* cfg, amps_ctx** out);

/* Allocate one context for an ensemble of n grids of the same model and mode
 * (sizes and parameters may differ).  Each step advances every instance in one
 * launch over a tile work-list.  No reference counterpart (the reference runs
 * one compile-time grid per process). */
int amps_create_batch(const amps_config* cfgs, int n, amps_ctx** out);

/* Free every resource.  Replaces the cudaFree block (snowflake/end2end_mix.cu:123-145). */
void amps_destroy(amps_ctx* ctx);

/* ---- running ---- */

/* Write the initial fields of every instance (cfg.init) and reset the tile
 * map: ADAPTIVE starts all-FP64 (the reference's complement init of the
 * Monitor2 history makes every tile high at step 0, snowflake/end2end_mix.cu:42-47),
 * PURE_HALF2 all-low, STATIC_MAP all-FP64 until amps_set_tile_types.
 * Replaces dataprepare_high/half2 (snowflake/tools.h:24-58),
 * graingrowth/tools.h:24-81 and the host loop sintering/main.cu:45-66. */
int amps_init(amps_ctx* ctx);

/* Enqueue nsteps explicit-Euler steps (asynchronous).  In ADAPTIVE mode the
 * monitor, tile classification and conversion run before every step whose
 * index is a multiple of monitor_every.  Replaces the host step loops
 * snowflake/purehigh.cu:23-34, snowflake/end2end_mix.cu:80-105,
 * graingrowth/end2end_mix.cu:71-98, sintering/main.cu:162-206. */
int amps_step(amps_ctx* ctx, int nsteps);

/* Block until queued work is done.  elapsed_ms (if not NULL) receives the GPU
 * time of the steps enqueued since the previous amps_sync, measured with HIP
 * events around them (the reference's cudaEvent timing, snowflake/purehigh.cu:24-46). */
int amps_sync(amps_ctx* ctx, double* elapsed_ms);

/* Run one monitor + classification + conversion pass now, on the current
 * fields (ADAPTIVE only).  Replaces get_max_diff1/2 + get_type +
 * data_sychro_duringcomputation (snowflake/tools.h:177-246). */
int amps_monitor_now(amps_ctx* ctx);

/* ---- state exchange (FP64, row-major [ny][nx]) ---- */

/* Copy field `field` of instance `inst` to dst (n == nx*ny).  Low-precision
 * tiles are decoded exactly.  Replaces data_sychro_aftercomputation +
 * writetocsv (snowflake/tools.h:147-174, 5-23). */
int amps_get_field(amps_ctx* ctx, int inst, int field, double* dst, size_t n);

/* Overwrite field `field` of instance `inst` from src (n == nx*ny).  Tiles
 * currently held in low precision are re-encoded.  No reference counterpart
 * (the reference only initialises). */
int amps_set_field(amps_ctx* ctx, int inst, int field, const double* src, size_t n);

/* Tile map of instance `inst`: 1 = low precision, 2 = FP64 (the reference's
 * encoding, snowflake/end2end_mix.cu:23).  n == ntiles * type_groups; grain
 * growth keeps one map per order parameter ([q][tiles_y][tiles_x]). */
int amps_get_tile_types(amps_ctx* ctx, int inst, int32_t* dst, size_t n);

/* Impose a tile map (converting tiles whose type changes).  With STATIC_MAP
 * the map is never re-evaluated; this is how the GRAM comparators
 * (snowflake/function.h:72-79) are expressed. */
int amps_set_tile_types(amps_ctx* ctx, int inst, const int32_t* src, size_t n);

/* Per-tile monitor values of the last pass (max_diff of get_max_diff1/2,
 * snowflake/tools.h:192,210; squared gradient norm for GRADIENT).
 * n == ntiles * type_groups. */
int amps_get_monitor_values(amps_ctx* ctx, int inst, double* dst, size_t n);

/* Raw low-precision state of field `field` of instance `inst`: the fp16
 * words of every cell ([ny][nx], meaningful on tiles whose type is 1) and
 * each tile's FP64 offset and power-of-two scale ([tiles_y][tiles_x]); a
 * low tile's value is off + h * scl.  n == nx*ny, ntiles == tiles of the
 * instance.  Contexts with low-precision storage only.  No reference
 * counterpart (the reference keeps half2 arrays in managed memory that its
 * drivers read directly). */
int amps_get_raw_field(amps_ctx* ctx, int inst, int field, uint16_t* words, size_t n, double* off, float* scl,
                       size_t ntiles);

int amps_get_info(amps_ctx* ctx, amps_info* info);

/* ---- measurement ---- */

/* When enabled, every sweep kernel launched by amps_step carries a pair of
 * HIP events bound to its own dispatch (hipExtLaunchKernel), so the summed
 * durations are the kernels' execution times without dispatch gaps; graph
 * replay is suspended meanwhile, and the counters below restart from zero.  Replaces the per-kernel cudaEvent
 * pairs of the reference's timing-decomposition drivers
 * (snowflake/monitor_independent_withoutmonitor.cu:85-127). */
int amps_set_kernel_timing(amps_ctx* ctx, int enable);

/* Number of timed sweep-kernel launches and their summed duration (ms). */
int amps_get_kernel_time(amps_ctx* ctx, int64_t* launches, double* total_ms);

/* The same split by kernel.  The snowflake's low-precision step runs two:
 * first a sweep over the low tiles (AMPS_SWEEP_LOW), then one kernel that
 * updates the FP64 tiles together with the low tiles that sweep listed
 * (AMPS_SWEEP_FP64).  The first sweep is the bulk-hold sweep (it lists the
 * tiles outside the bulk-hold state) or, with AMPS_FLAG_NO_SKIP, the flat
 * sweep (it finishes the tiles whose neighbourhood decodes to exact zeros
 * at phi 0 or 1 and lists the others).  For every other step
 * AMPS_SWEEP_FP64 is the whole sweep and AMPS_SWEEP_LOW is zero.  Replaces the per-kernel
 * cudaEvent pairs around kernel1_mix/mix_kernel2
 * (snowflake/monitor_independent_withoutmonitor.cu:85-127). */
enum { AMPS_SWEEP_ALL = 0, AMPS_SWEEP_FP64 = 1, AMPS_SWEEP_LOW = 2 };
int amps_get_sweep_time(amps_ctx* ctx, int which, int64_t* launches, double* total_ms);

/* ---- stepping without allocation ---- */

/* One PURE_FP64 snowflake step of one grid (cfg: model SNOWFLAKE, mode
 * PURE_FP64, validated like amps_create) from d_in to d_out, both
 * caller-owned device buffers of [2][ny][nx] doubles (phi, then T) that must
 * not overlap; step is the step index (it keys the noise, like the step count
 * of a context since amps_init).  The work is enqueued on `stream` (a
 * hipStream_t; NULL: the null stream) and the call returns without waiting.
 * No reference counterpart: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  It allocates
 * nothing: the descriptor and step counter the step kernels read go to a
 * fixed 4 KiB device array of the library, one per GPU, written by a kernel
 * at each call; grids are limited only as in amps_create.  Calls that
 * overlap on one GPU (other streams or host threads) race on that array.
 * The result equals amps_step on a PURE_FP64 context of the same config. */
int amps_sf_step_fp64_noalloc(const amps_config* cfg, const double* d_in, double* d_out, int64_t step,
                              void* stream);

/* ---- adjoint (FP64 contexts; snowflake and sintering) ---- */

/* Parameters a gradient can be taken with respect to. */
enum {
    AMPS_GRAD_SF_DELTA = 1, AMPS_GRAD_SF_KAPPA = 2, AMPS_GRAD_SF_EPSB = 3, AMPS_GRAD_SF_TAU = 4,
    AMPS_GRAD_SN_DSURF = 101, AMPS_GRAD_SN_DGB = 102, AMPS_GRAD_SN_KC = 103,
    AMPS_GRAD_SN_KE = 104, AMPS_GRAD_SN_L = 105, AMPS_GRAD_SN_DVOL = 106
};

/* Start recording: the current state becomes checkpoint 0 and every later
 * amps_step keeps a checkpoint each ckpt_every steps.  No reference counterpart. */
int amps_tape_begin(amps_ctx* ctx, int ckpt_every);

/* Reverse sweep over the recorded steps.  dJ_du: cotangent of the final
 * fields, [ninstances][nfields][ny][nx] concatenated (instances of a batch in
 * order).  dJ_dp receives [ninstances][np] parameter gradients; dJ_du0 (may be
 * NULL) receives the cotangent of the recorded initial fields, same layout as
 * dJ_du.  Each segment between checkpoints is recomputed forward on the GPU. */
int amps_adjoint(amps_ctx* ctx, const double* dJ_du, const int32_t* param_ids, int np,
                 double* dJ_dp, double* dJ_du0);

/* ---- autotuning ---- */

/* A tiling of a run: tile edge and sweep workgroups per CU (0 = occupancy
 * maximum), with the time per step measured for it. */
typedef struct {
    int32_t tile;
    int32_t wg_per_cu;
    double us_per_step;
} amps_tuning;

/* Time `steps` explicit-Euler steps (after a short warm-up) of every candidate
 * tiling of the run cfgs[0..n-1] -- tile 16 or 32 (where the grids allow) x
 * workgroups per CU in {occupancy maximum, 1, 2, 3, 4} -- each on a scratch
 * context started from the run's initial state, and return the fastest in
 * *best.  cfgs' own tile and wg_per_cu are ignored; apply the result by
 * setting them in the configs given to amps_create.  Results are cached for
 * the process per (model, mode, monitor, grids, q, steps).  No reference
 * counterpart: the reference hard-codes 16x16 tiles and 8x16-thread blocks
 * (snowflake/paras/total.h:21, end2end_mix.cu:76-77). */
int amps_autotune(const amps_config* cfgs, int n, int steps, amps_tuning* best);

/* ---- diagnostics ---- */
const char* amps_status_string(int status);
/* Message of the most recent failure on the calling thread. */
const char* amps_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* AMPS_H */

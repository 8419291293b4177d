// api.cpp — host runtime of the C ABI declared in include/amps.h.
//
// A context holds every instance of a run (one grid, or an ensemble of
// uneven grids of one model), all device buffers, one HIP stream and the
// captured step graphs.  amps_step enqueues work without host
// synchronisation: the monitor cadence, tile classification, conversion and
// the explicit-Euler sweeps are all device work, replayed from hipGraphs in
// blocks of one monitor period.
#include <algorithm>
#include <map>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "kernels.hpp"

using namespace amps;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define HIPCHK(expr)                                                                                        \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess)                                                                               \
            return fail(AMPS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));                    \
    } while (0)

#define RCCHK(expr)                                                                                         \
    do {                                                                                                    \
        int r_ = (expr);                                                                                    \
        if (r_ != AMPS_OK) return r_ == AMPS_E_HIP ? fail(r_, std::string(#expr) + ": " +                  \
                                                        hipGetErrorString(hipGetLastError()))               \
                                                  : r_;                                                     \
    } while (0)

int nfields_of(const amps_config& c) {
    if (c.model == AMPS_SNOWFLAKE) return 2;
    if (c.model == AMPS_GRAINGROWTH) return c.q;
    return 3;
}

}  // namespace

namespace amps {
void set_error_text(const char* msg) { g_last_error = msg ? msg : ""; }
}  // namespace amps

// ints of the list counters: the larger of the two layouts (bulk-hold [2][8],
// NO_SKIP [2][NSK_SEGS] at a stride of NSK_CNT_STRIDE)
constexpr size_t FULL_CNT_INTS = 2 * (size_t)NSK_SEGS * NSK_CNT_STRIDE;

struct amps_ctx {
    int model = 0, mode = 0, monitor = 0, tile = 16, every = 10, mfield = 0;
    int nfields = 0, ngroups = 1, q = 0, ninst = 0;
    int grid_cap = 0;                      // sweep workgroups at most (wg_per_cu x CUs; 0: no cap)
    int flags = 0;                         // AMPS_FLAG_* of the run
    int dilate = 0;                        // FP64-set dilation radius in tiles (ADAPTIVE)
    double thr_eff = 0;
    std::vector<amps_config> cfg;
    std::vector<InstDesc> desc;
    std::vector<int> tile_inst;
    long long total_cells = 0;
    int total_tiles = 0;
    long long max_inst_cells = 0;
    GroupMap gm{};

    InstDesc* d_desc = nullptr;
    int* d_tile_inst = nullptr;
    Buf buf[2]{};
    int cur = 0;
    int* d_type = nullptr;
    int* d_type_old = nullptr;
    double* d_md = nullptr;
    double* d_last = nullptr;
    unsigned long long* d_stats = nullptr;  // [0] sum of FP64 tiles over passes, [1] last pass,
                                            // [2], [3], [4] snowflake sweep tile counts (summed)
    long long* d_step = nullptr;
    double* d_aux[2] = {nullptr, nullptr};
    double* d_scratch = nullptr;           // max instance cells (get/set)
    int* d_wl = nullptr;                   // [2][total_tiles]: FP64 and low tile work-lists
    int* d_wl_count = nullptr;             // [2]
    int* d_same_flags = nullptr;           // snowflake [total_tiles], grain growth [q][total_tiles]:
                                           // tile data identical in both buffers
    size_t same_flags_n = 0;
    int* d_full_seg = nullptr;             // snowflake: low tiles listed for the mixed sweep (StepArgs::full_seg)
    int* d_full_cnt = nullptr;             // their counts (StepArgs::full_cnt)
    int full_cap = 0;

    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timing_open = false;

    long long steps_done = 0;
    long long monitor_passes = 0;
    int pending_adv = 0;
    bool initialised = false;

    // graphs: one per starting buffer parity
    bool use_graph = true;
    int graph_period = 0;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};

    // adjoint tape
    int tape_every = 0;
    bool taping = false;
    long long tape_step0 = 0;
    std::vector<double*> tape;              // device checkpoints (all fields)
    std::vector<long long> tape_steps;

    // sweep-kernel timing (events around each step-kernel launch)
    bool ktiming = false;
    std::vector<hipEvent_t> kev;  // quads: start and end of a step's first and second sweep kernel
    std::vector<unsigned char> kev_pairs;  // per timed step: 1 or 2 sweep kernels (pairs recorded)
    int kev_used = 0;
    long long klaunches = 0;
    double kms = 0.0;                      // all sweep kernels
    double kms_part[2] = {0.0, 0.0};       // snowflake adaptive: FP64 sweep, low sweep

    std::vector<void*> allocs;
};

int tape_push(amps_ctx* c);

namespace {

int dmalloc(amps_ctx* c, void** p, size_t bytes) {
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(AMPS_E_NOMEM, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    c->allocs.push_back(*p);
    HIPCHK(hipMemset(*p, 0, bytes));
    return AMPS_OK;
}

bool finite_params(const amps_config& c) {
    for (int i = 0; i < AMPS_NPARAMS; ++i)
        if (!std::isfinite(c.params[i])) return false;
    return std::isfinite(c.threshold);
}

int validate(const amps_config& c) {
    if (c.model < 1 || c.model > 3) return fail(AMPS_E_ARG, "model must be 1 (snowflake), 2 (grain growth) or 3 (sintering)");
    if (c.mode < 0 || c.mode > 3) return fail(AMPS_E_ARG, "mode out of range");
    if (c.tile != 16 && c.tile != 32) return fail(AMPS_E_UNSUPPORTED, "tile must be 16 or 32");
    if (c.nx < c.tile || c.ny < c.tile || c.nx % c.tile || c.ny % c.tile)
        return fail(AMPS_E_ARG, "nx and ny must be positive multiples of tile");
    if ((long long)c.nx * c.ny > (1ll << 31)) return fail(AMPS_E_ARG, "grid too large (> 2^31 cells)");
    if (c.stencil != AMPS_STENCIL_REF && c.stencil != AMPS_STENCIL_LAP5) return fail(AMPS_E_ARG, "stencil out of range");
    if (c.bc != AMPS_BC_PERIODIC && c.bc != AMPS_BC_NEUMANN) return fail(AMPS_E_ARG, "bc out of range");
    if (c.init != AMPS_INIT_REF && c.init != AMPS_INIT_CONFIG) return fail(AMPS_E_ARG, "init out of range");
    if (!finite_params(c)) return fail(AMPS_E_ARG, "non-finite parameter");
    if (c.wg_per_cu < 0 || c.wg_per_cu > 64) return fail(AMPS_E_ARG, "wg_per_cu must be in [0, 64]");
    if (c.flags & ~(AMPS_FLAG_NO_SKIP | AMPS_FLAG_CH_CONSERVATIVE | AMPS_FLAG_HALF2))
        return fail(AMPS_E_ARG, "unknown flags bit");
    if ((c.flags & AMPS_FLAG_HALF2) &&
        (c.model != AMPS_SNOWFLAKE || c.tile != 32 || !(c.flags & AMPS_FLAG_NO_SKIP)))
        return fail(AMPS_E_UNSUPPORTED, "HALF2 applies to the snowflake with tile 32 and NO_SKIP");
    if ((c.flags & AMPS_FLAG_CH_CONSERVATIVE) && c.model != AMPS_SINTERING)
        return fail(AMPS_E_ARG, "CH_CONSERVATIVE applies to the sintering model only");
    if (c.dilate < 0 || c.dilate > 16) return fail(AMPS_E_ARG, "dilate must be in [0, 16]");
    if (c.mode == AMPS_ADAPTIVE) {
        if (c.monitor < 1 || c.monitor > 3) return fail(AMPS_E_ARG, "ADAPTIVE needs monitor 1, 2 or 3");
        if (c.monitor_every < 1) return fail(AMPS_E_ARG, "monitor_every must be >= 1");
    }
    const double* p = c.params;
    if (p[0] <= 0 || p[1] <= 0 || p[2] <= 0) return fail(AMPS_E_ARG, "dx, dy and dt must be positive");
    if (c.model == AMPS_SNOWFLAKE) {
        double n = p[AMPS_SF_ANISO];
        if (n < 1 || n > 12 || n != std::floor(n)) return fail(AMPS_E_UNSUPPORTED, "snowflake aniso must be an integer in [1, 12]");
        if (p[AMPS_SF_TAU] <= 0) return fail(AMPS_E_ARG, "tau must be positive");
        // explicit Euler on the T diffusion: spectral radius of the cross-2
        // operator is 12.5/(dx dy) (5-point: 8/(dx dy)); require dt*rho < 2
        double rho = (c.stencil == AMPS_STENCIL_REF ? 12.5 : 8.0) / (p[AMPS_SF_DX] * p[AMPS_SF_DY]);
        if (p[AMPS_SF_DT] * rho >= 2.0) return fail(AMPS_E_ARG, "dt violates the explicit-Euler bound of the T diffusion");
        if (c.monitor_field != 0 && c.mode == AMPS_ADAPTIVE && (c.monitor_field < 0 || c.monitor_field > 1))
            return fail(AMPS_E_ARG, "snowflake monitor_field must be 0 (phi) or 1 (T)");
    } else if (c.model == AMPS_GRAINGROWTH) {
        if (c.q < 2 || c.q > 38) return fail(AMPS_E_ARG, "grain growth q must be in [2, 38]");
        double rho = (c.stencil == AMPS_STENCIL_REF ? 12.0 : 8.0) / (p[AMPS_GG_DX] * p[AMPS_GG_DY]);
        if (p[AMPS_GG_DT] * p[AMPS_GG_L] * p[AMPS_GG_KAPPA] * rho >= 2.0)
            return fail(AMPS_E_ARG, "dt*L*kappa violates the explicit-Euler bound");
        if (c.mode == AMPS_ADAPTIVE && (c.monitor_field < -1 || c.monitor_field >= c.q))
            return fail(AMPS_E_ARG, "grain growth monitor_field must be -1 or a field index");
    } else {
        if (c.mode == AMPS_ADAPTIVE && (c.monitor_field < 0 || c.monitor_field > 2))
            return fail(AMPS_E_ARG, "sintering monitor_field must be 0, 1 or 2");
    }
    return AMPS_OK;
}

void derive(InstDesc& d, const amps_config& c) {
    const double* p = c.params;
    for (int i = 0; i < AMPS_NPARAMS; ++i) d.p[i] = p[i];
    for (int i = 0; i < 24; ++i) d.k[i] = 0;
    if (c.model == AMPS_SNOWFLAKE) {
        const double n = p[AMPS_SF_ANISO];
        d.k[SFK_INV_DXDY] = 1.0 / (p[AMPS_SF_DX] * p[AMPS_SF_DY]);
        d.k[SFK_INV_2DX] = 1.0 / (2.0 * p[AMPS_SF_DX]);
        d.k[SFK_INV_2DY] = 1.0 / (2.0 * p[AMPS_SF_DY]);
        d.k[SFK_DT_OVER_TAU] = p[AMPS_SF_DT] / p[AMPS_SF_TAU];
        d.k[SFK_DT] = p[AMPS_SF_DT];
        d.k[SFK_KAPPA] = p[AMPS_SF_KAPPA];
        d.k[SFK_ALPHA_OVER_PI] = p[AMPS_SF_ALPHA] / p[AMPS_SF_PI];
        d.k[SFK_GAMMA] = p[AMPS_SF_GAMMA];
        d.k[SFK_TEQ] = p[AMPS_SF_TEQ];
        d.k[SFK_EPSB] = p[AMPS_SF_EPSB];
        d.k[SFK_EPSB_DELTA] = p[AMPS_SF_EPSB] * p[AMPS_SF_DELTA];
        d.k[SFK_MEPSB_ANISO_DELTA] = -p[AMPS_SF_EPSB] * n * p[AMPS_SF_DELTA];
        d.k[SFK_COS_N_THETA0] = std::cos(n * p[AMPS_SF_THETA0]);
        d.k[SFK_SIN_N_THETA0] = std::sin(n * p[AMPS_SF_THETA0]);
        d.k[SFK_NOISE] = p[AMPS_SF_NOISE];
        d.k[SFK_ANISO_N] = n;
    } else if (c.model == AMPS_GRAINGROWTH) {
        d.k[GGK_INV_DXDY] = 1.0 / (p[AMPS_GG_DX] * p[AMPS_GG_DY]);
        d.k[GGK_DT_L] = p[AMPS_GG_DT] * p[AMPS_GG_L];
        d.k[GGK_KAPPA] = p[AMPS_GG_KAPPA];
        d.k[GGK_ALPHA] = p[AMPS_GG_ALPHA];
        d.k[GGK_BETA] = p[AMPS_GG_BETA];
        d.k[GGK_2GAMMA] = 2.0 * p[AMPS_GG_GAMMA];
    } else {
        d.k[SNK_INV_DXDY] = 1.0 / (p[AMPS_SN_DX] * p[AMPS_SN_DY]);
        d.k[SNK_DT] = p[AMPS_SN_DT];
        d.k[SNK_A2] = 2.0 * p[AMPS_SN_A];
        d.k[SNK_B] = p[AMPS_SN_B];
        d.k[SNK_HALF_KC] = 0.5 * p[AMPS_SN_KC];
        d.k[SNK_DVOL] = p[AMPS_SN_DVOL];
        d.k[SNK_DVAP] = p[AMPS_SN_DVAP];
        d.k[SNK_DSURF] = p[AMPS_SN_DSURF];
        d.k[SNK_DGB] = p[AMPS_SN_DGB];
        d.k[SNK_DT_L] = p[AMPS_SN_DT] * p[AMPS_SN_L];
        d.k[SNK_HALF_KE] = 0.5 * p[AMPS_SN_KE];
    }
    for (int i = 0; i < 24; ++i) d.kf[i] = (float)d.k[i];
}

// Descriptor of one instance of a validated config (tile base and cell base 0).
void fill_desc(InstDesc& d, const amps_config& c) {
    std::memset(&d, 0, sizeof(d));
    d.nx = c.nx;
    d.ny = c.ny;
    d.tx = c.nx / c.tile;
    d.ty = c.ny / c.tile;
    d.bc = c.bc;
    d.stencil = c.stencil;
    d.seed = c.seed;
    derive(d, c);
}

bool has_low(const amps_ctx* c) { return c->mode != AMPS_PURE_FP64; }

// The snowflake's low-precision step runs two sweep kernels: the bulk-hold
// sweep over the low tiles, then the mixed sweep (FP64 tiles and the low
// tiles it did not hold); with NO_SKIP the FP64 sweep, then the full low
// sweep.
bool two_sweeps(const amps_ctx* c) { return c->model == AMPS_SNOWFLAKE && has_low(c); }
bool low_sweep_first(const amps_ctx* c) { return two_sweeps(c); }

StepArgs step_args(amps_ctx* c, int from) {
    StepArgs a{};
    a.inst = c->d_desc;
    a.tile_inst = c->d_tile_inst;
    a.worklist = nullptr;
    a.wl_count = nullptr;
    a.wl_low = nullptr;
    a.wl_low_count = nullptr;
    if (c->model == AMPS_SNOWFLAKE && has_low(c)) {
        a.worklist = c->d_wl;
        a.wl_count = c->d_wl_count;
        a.wl_low = c->d_wl + c->total_tiles;
        a.wl_low_count = c->d_wl_count + 1;
        a.same_flags = c->d_same_flags;
        a.full_seg = c->d_full_seg;
        a.full_cnt = c->d_full_cnt;
        a.full_cap = c->full_cap;
        a.stats = c->d_stats;
    }
    if (c->model == AMPS_GRAINGROWTH && has_low(c)) a.same_flags = c->d_same_flags;
    a.type = has_low(c) ? c->d_type : nullptr;
    a.total_cells = c->total_cells;
    a.total_tiles = c->total_tiles;
    a.nwork = c->total_tiles;
    a.nfields = c->nfields;
    a.q = c->q;
    a.step_ctr = c->d_step;
    a.step_add = 0;
    a.ninst = c->ninst;
    a.t_halo = 1;
    for (const amps_config& k : c->cfg)
        if (k.stencil == AMPS_STENCIL_REF) a.t_halo = 2;
    a.grid_cap = c->grid_cap;
    a.flags = c->flags;
    a.in = c->buf[from];
    a.out = c->buf[from ^ 1];
    a.aux_in = c->d_aux[from];
    a.aux_out = c->d_aux[from ^ 1];
    return a;
}

int kev_flush(amps_ctx* c) {
    if (c->kev_used == 0) return AMPS_OK;
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < c->kev_used; ++i) {
        float a = 0.0f, b = 0.0f;
        HIPCHK(hipEventElapsedTime(&a, c->kev[4 * i], c->kev[4 * i + 1]));
        if (c->kev_pairs[i] == 2) HIPCHK(hipEventElapsedTime(&b, c->kev[4 * i + 2], c->kev[4 * i + 3]));
        c->kms += (double)a + (double)b;
        c->kms_part[0] += a;
        c->kms_part[1] += b;
    }
    c->klaunches += c->kev_used;
    c->kev_used = 0;
    return AMPS_OK;
}

int launch_step_kernel(amps_ctx* c, const StepArgs& a, const hipEvent_t* ev = nullptr) {
    const bool adaptive = has_low(c);
    const int grid = a.grid_cap > 0 && a.grid_cap < a.nwork ? a.grid_cap : a.nwork;
    if (c->model == AMPS_SNOWFLAKE) return launch_sf_step(a, c->tile, adaptive, c->stream, ev);
    return c->model == AMPS_GRAINGROWTH ? launch_gg_step(a, c->tile, adaptive, grid, c->stream, ev)
                                        : launch_sn_step(a, c->tile, adaptive, grid, c->stream, ev);
}

// Enqueue one explicit-Euler step reading buffer `from`.  With kernel timing
// on, each sweep kernel carries a pair of events bound to its own dispatch
// (start and end of its execution).
int enqueue_step(amps_ctx* c, int from, int step_add, bool capturing = false) {
    StepArgs a = step_args(c, from);
    a.step_add = step_add;
    if (!c->ktiming || capturing) return launch_step_kernel(c, a);
    if (c->kev_used * 4 + 4 > (int)c->kev.size()) RCCHK(kev_flush(c));
    c->kev_pairs[c->kev_used] = two_sweeps(c) ? 2 : 1;
    RCCHK(launch_step_kernel(c, a, &c->kev[4 * c->kev_used]));
    c->kev_used++;
    return AMPS_OK;
}

// Model auxiliary state that must follow the fields (grain growth: the
// per-cell sum of squares, recomputed from current storage).
int refresh_aux(amps_ctx* c, int b, bool changed_only = false) {
    if (c->model != AMPS_GRAINGROWTH) return AMPS_OK;
    StepArgs a = step_args(c, b);
    return launch_gg_aux(a, c->tile, changed_only && a.type ? c->d_type_old : nullptr, c->stream);
}

// Grain growth: forget every "field is exactly 0" bit (fields or types were
// set from the host).
int clear_zmasks(amps_ctx* c) {
    for (int b = 0; b < 2; ++b)
        if (c->buf[b].zmask) HIPCHK(hipMemsetAsync(c->buf[b].zmask, 0, (size_t)c->total_tiles * sizeof(unsigned long long), c->stream));
    return AMPS_OK;
}

// Rebuild the FP64 / low tile work-lists from the current types (snowflake).
// after_monitor: only the tiles the monitor pass converted (type_old !=
// type) lose their same_flags; otherwise (fields or types set from the host)
// every tile does.
int refresh_worklists(amps_ctx* c, bool after_monitor = false) {
    if (c->model != AMPS_SNOWFLAKE || !has_low(c)) return AMPS_OK;
    return launch_worklist(c->d_type, after_monitor ? c->d_type_old : nullptr, c->total_tiles, c->d_wl,
                           c->d_wl + c->total_tiles, c->d_wl_count, c->d_same_flags, c->stream);
}

// One monitor pass on buffer `b`: monitor values, classification, conversion.
int enqueue_monitor(amps_ctx* c, int b) {
    StepArgs a = step_args(c, b);
    int ngroups_eval = 1, field0 = c->mfield, stride = 0, broadcast = 0;
    if (c->model == AMPS_GRAINGROWTH) {
        if (c->mfield < 0) { ngroups_eval = c->q; field0 = 0; stride = 1; }
        else broadcast = 1;
    }
    // the monitored field's own precision decides how it is read
    if (c->gm.fg[field0] < 0) a.type = nullptr;
    // the monitor kernel also clears the classification's per-pass FP64 count
    RCCHK(launch_monitor(a, c->monitor, c->tile, ngroups_eval, field0, stride, c->ngroups, c->d_md, c->d_last,
                         c->d_stats + 1, c->stream));
    RCCHK(launch_classify(c->d_desc, c->d_tile_inst, c->total_tiles, c->ngroups, broadcast, c->ninst == 1 ? 1 : 0,
                          c->d_md, c->thr_eff, c->dilate,
                          c->d_type_old, c->d_type, c->d_stats, c->stream));
    RCCHK(launch_convert(c->d_desc, c->d_tile_inst, c->buf[b], c->total_cells, c->total_tiles, c->d_type_old,
                         c->d_type, c->gm, CONVERT_CHANGED, c->tile, c->stream));
    if (c->model == AMPS_GRAINGROWTH) {
        RCCHK(refresh_aux(c, b, true));
        if (has_low(c))  // converted field-tiles: the buffers no longer agree
            RCCHK(launch_clear_changed(c->d_type_old, c->d_type, c->gm, c->nfields, c->total_tiles, c->d_same_flags,
                                       c->buf[b].zmask, c->stream));
    }
    RCCHK(refresh_worklists(c, true));
    return AMPS_OK;
}

int flush_adv(amps_ctx* c) {
    if (c->pending_adv) {
        RCCHK(launch_advance_step(c->d_step, c->pending_adv, c->stream));
        c->pending_adv = 0;
    }
    return AMPS_OK;
}

// Steps whose index is a multiple of `every` start with a monitor pass.
bool monitor_due(const amps_ctx* c, long long step) {
    return c->mode == AMPS_ADAPTIVE && step % c->every == 0;
}

int enqueue_one(amps_ctx* c) {
    if (monitor_due(c, c->steps_done)) {
        RCCHK(enqueue_monitor(c, c->cur));
        c->monitor_passes++;
    }
    RCCHK(enqueue_step(c, c->cur, c->pending_adv));
    c->pending_adv++;
    c->cur ^= 1;
    c->steps_done++;
    if (c->taping && c->tape_every > 0 && (c->steps_done - c->tape_step0) % c->tape_every == 0) {
        RCCHK(flush_adv(c));
        RCCHK(tape_push(c));
    }
    return AMPS_OK;
}

// Capture `period` steps (monitor pass included) starting from buffer parity p.
int build_graph(amps_ctx* c, int p) {
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int b = p;
    int rc = AMPS_OK;
    for (int k = 0; k < c->graph_period && rc == AMPS_OK; ++k) {
        if (c->mode == AMPS_ADAPTIVE && k % c->every == 0) rc = enqueue_monitor(c, b);
        if (rc == AMPS_OK) rc = enqueue_step(c, b, k, true);
        b ^= 1;
    }
    if (rc == AMPS_OK) rc = launch_advance_step(c->d_step, c->graph_period, c->stream);
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc != AMPS_OK) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return fail(AMPS_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&c->gexec[p], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(AMPS_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    return AMPS_OK;
}

}  // namespace

// ============================================================ tape helpers
int tape_push(amps_ctx* c) {
    const size_t bytes = (size_t)c->nfields * c->total_cells * sizeof(double);
    double* p = nullptr;
    hipError_t e = hipMalloc((void**)&p, bytes);
    if (e != hipSuccess) return fail(AMPS_E_NOMEM, "checkpoint allocation failed");
    e = hipMemcpyAsync(p, c->buf[c->cur].f64, bytes, hipMemcpyDeviceToDevice, c->stream);
    if (e != hipSuccess) { (void)hipFree(p); return fail(AMPS_E_HIP, "checkpoint copy failed"); }
    c->tape.push_back(p);
    c->tape_steps.push_back(c->steps_done);
    return AMPS_OK;
}

namespace {
void tape_clear(amps_ctx* c) {
    for (double* p : c->tape) (void)hipFree(p);
    c->tape.clear();
    c->tape_steps.clear();
    c->taping = false;
}
}  // namespace

namespace amps {
int validate_config(const amps_config& c) { return validate(c); }
void config_desc(InstDesc& d, const amps_config& c) { fill_desc(d, c); }
int fail_with(int code, const char* msg) { return fail(code, msg); }
}  // namespace amps

// ================================================================ C ABI
extern "C" {

const char* amps_status_string(int s) {
    switch (s) {
        case AMPS_OK: return "ok";
        case AMPS_E_ARG: return "invalid argument";
        case AMPS_E_NOMEM: return "out of memory";
        case AMPS_E_HIP: return "HIP runtime error";
        case AMPS_E_PARSE: return "parameter file parse error";
        case AMPS_E_NONFINITE: return "non-finite value in a field";
        case AMPS_E_STATE: return "call not valid in this state";
        case AMPS_E_UNSUPPORTED: return "unsupported configuration";
        case AMPS_E_IO: return "I/O error";
        default: return "unknown status";
    }
}

const char* amps_last_error(void) { return g_last_error.c_str(); }

int amps_default_config(int model, amps_config* cfg) {
    if (!cfg) return fail(AMPS_E_ARG, "cfg is NULL");
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->model = model;
    cfg->mode = AMPS_PURE_FP64;
    cfg->monitor = AMPS_MON_TEMPORAL;
    cfg->nx = cfg->ny = 512;
    cfg->q = 2;
    cfg->tile = 16;
    cfg->monitor_every = 10;
    cfg->monitor_field = 0;
    cfg->stencil = AMPS_STENCIL_REF;
    cfg->bc = AMPS_BC_PERIODIC;
    cfg->init = AMPS_INIT_REF;
    cfg->threshold = 0.99;
    cfg->seed = 1;
    double* p = cfg->params;
    if (model == AMPS_SNOWFLAKE) {  // snowflake/paras/para1.h:4-15, total.h:47
        p[AMPS_SF_DX] = 0.03; p[AMPS_SF_DY] = 0.03; p[AMPS_SF_DT] = 1.0e-4; p[AMPS_SF_TAU] = 0.0003;
        p[AMPS_SF_EPSB] = 0.01; p[AMPS_SF_KAPPA] = 1.8; p[AMPS_SF_DELTA] = 0.02; p[AMPS_SF_ANISO] = 4.0;
        p[AMPS_SF_ALPHA] = 0.9; p[AMPS_SF_GAMMA] = 10.0; p[AMPS_SF_TEQ] = 1.0; p[AMPS_SF_THETA0] = 0.0;
        p[AMPS_SF_PI] = 3.1415926; p[AMPS_SF_NOISE] = 0.0; p[AMPS_SF_R2] = 10.0;
        p[AMPS_SF_CX] = 512 / 2 + 8; p[AMPS_SF_CY] = 512 / 2 + 8;
    } else if (model == AMPS_GRAINGROWTH) {  // graingrowth/paras/para1.h:4-6
        p[AMPS_GG_DX] = 0.5; p[AMPS_GG_DY] = 0.5; p[AMPS_GG_DT] = 0.005; p[AMPS_GG_L] = 5.0;
        p[AMPS_GG_KAPPA] = 0.1; p[AMPS_GG_ALPHA] = 1.0; p[AMPS_GG_BETA] = 1.0; p[AMPS_GG_GAMMA] = 1.0;
        p[AMPS_GG_R] = 50.0; p[AMPS_GG_CX] = 256.0; p[AMPS_GG_CY] = 256.0; p[AMPS_GG_AMP] = 0.0;
        cfg->monitor_field = 1;  // the reference monitors eta2 (graingrowth/end2end_mix.cu:83)
    } else if (model == AMPS_SINTERING) {  // sintering/paras/para1.h:1-9, function.h:72
        p[AMPS_SN_DX] = 0.5; p[AMPS_SN_DY] = 0.5; p[AMPS_SN_DT] = 1.0e-4; p[AMPS_SN_A] = 16.0;
        p[AMPS_SN_B] = 1.0; p[AMPS_SN_KC] = 5.0; p[AMPS_SN_KE] = 2.0; p[AMPS_SN_L] = 5.0;
        p[AMPS_SN_DVOL] = 0.040; p[AMPS_SN_DVAP] = 0.002; p[AMPS_SN_DSURF] = 32.0; p[AMPS_SN_DGB] = 1.6;
        p[AMPS_SN_R1] = 50.0; p[AMPS_SN_R2] = 25.0; p[AMPS_SN_X1] = 264.0; p[AMPS_SN_Y1] = 240.0;
        p[AMPS_SN_X2] = 264.0; p[AMPS_SN_Y2] = 315.0; p[AMPS_SN_WIDTH] = 0.0; p[AMPS_SN_NOISE] = 0.0;
        cfg->q = 2;
    } else {
        return fail(AMPS_E_ARG, "unknown model");
    }
    return AMPS_OK;
}

int amps_create_batch(const amps_config* cfgs, int n, amps_ctx** out) {
    if (!cfgs || !out || n < 1 || n > AMPS_MAX_INST) return fail(AMPS_E_ARG, "need 1..256 configs and an output pointer");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        RCCHK(validate(cfgs[i]));
        const amps_config& a = cfgs[i];
        const amps_config& b = cfgs[0];
        if (a.model != b.model || a.mode != b.mode || a.tile != b.tile || a.monitor != b.monitor ||
            a.monitor_every != b.monitor_every || a.monitor_field != b.monitor_field || a.threshold != b.threshold ||
            a.wg_per_cu != b.wg_per_cu || a.flags != b.flags || a.dilate != b.dilate ||
            (a.model == AMPS_GRAINGROWTH && a.q != b.q))
            return fail(AMPS_E_ARG, "batched instances must share model, mode, tile, monitor settings and q");
    }
    amps_ctx* c = new amps_ctx();
    c->cfg.assign(cfgs, cfgs + n);
    c->ninst = n;
    c->model = cfgs[0].model;
    c->mode = cfgs[0].mode;
    c->monitor = cfgs[0].monitor;
    c->tile = cfgs[0].tile;
    c->every = cfgs[0].monitor_every > 0 ? cfgs[0].monitor_every : 10;
    c->mfield = cfgs[0].monitor_field;
    c->flags = cfgs[0].flags;
    c->dilate = cfgs[0].dilate;
    c->q = c->model == AMPS_GRAINGROWTH ? cfgs[0].q : (c->model == AMPS_SINTERING ? 2 : 0);
    c->nfields = nfields_of(cfgs[0]);
    c->ngroups = c->model == AMPS_GRAINGROWTH ? c->q : 1;
    c->thr_eff = c->monitor == AMPS_MON_GRADIENT ? cfgs[0].threshold * cfgs[0].threshold : cfgs[0].threshold;
    c->gm.nfields = c->nfields;
    c->gm.ngroups = c->ngroups;
    for (int f = 0; f < AMPS_MAX_FIELDS; ++f) c->gm.fg[f] = -1;
    for (int f = 0; f < c->nfields; ++f) {
        if (c->model == AMPS_SNOWFLAKE) c->gm.fg[f] = 0;
        else if (c->model == AMPS_GRAINGROWTH) c->gm.fg[f] = (signed char)f;
        else c->gm.fg[f] = f == 0 ? -1 : 0;  // sintering: c stays FP64
    }
    if (c->mode == AMPS_ADAPTIVE && c->model == AMPS_GRAINGROWTH && c->mfield < 0) { /* per-field types */ }

    c->desc.resize(n);
    long long cells = 0;
    int tiles = 0;
    for (int i = 0; i < n; ++i) {
        InstDesc& d = c->desc[i];
        std::memset(&d, 0, sizeof(d));
        d.nx = cfgs[i].nx;
        d.ny = cfgs[i].ny;
        d.tx = d.nx / c->tile;
        d.ty = d.ny / c->tile;
        d.tile_base = tiles;
        d.bc = cfgs[i].bc;
        d.stencil = cfgs[i].stencil;
        d.cell_base = cells;
        d.seed = cfgs[i].seed;
        derive(d, cfgs[i]);
        cells += (long long)d.nx * d.ny;
        tiles += d.tx * d.ty;
        c->max_inst_cells = std::max(c->max_inst_cells, (long long)d.nx * d.ny);
        for (int t = 0; t < d.tx * d.ty; ++t) c->tile_inst.push_back(i);
    }
    c->total_cells = cells;
    c->total_tiles = tiles;

    int rc = AMPS_OK;
    auto A = [&](void** p, size_t bytes) { if (rc == AMPS_OK) rc = dmalloc(c, p, bytes); };
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; return fail(AMPS_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)); }
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        amps_destroy(c);
        return fail(AMPS_E_HIP, "hipEventCreate failed");
    }
    A((void**)&c->d_desc, sizeof(InstDesc) * n);
    A((void**)&c->d_tile_inst, sizeof(int) * tiles);
    const size_t nf = (size_t)c->nfields;
    for (int b = 0; b < 2; ++b) {
        A((void**)&c->buf[b].f64, nf * cells * sizeof(double));
        if (has_low(c)) {
            A((void**)&c->buf[b].f16, nf * cells * sizeof(__half));
            A((void**)&c->buf[b].off, nf * tiles * sizeof(double));
            A((void**)&c->buf[b].scl, nf * tiles * sizeof(float));
            if (c->model == AMPS_GRAINGROWTH) A((void**)&c->buf[b].zmask, tiles * sizeof(unsigned long long));
        }
        if (c->model == AMPS_GRAINGROWTH) A((void**)&c->d_aux[b], cells * sizeof(double));
    }
    const size_t gt = (size_t)c->ngroups * tiles;
    A((void**)&c->d_type, gt * sizeof(int));
    A((void**)&c->d_type_old, gt * sizeof(int));
    A((void**)&c->d_md, gt * sizeof(double));
    if (c->mode == AMPS_ADAPTIVE && c->monitor == AMPS_MON_TEMPORAL) A((void**)&c->d_last, gt * 16 * sizeof(double));
    A((void**)&c->d_stats, 8 * sizeof(unsigned long long));
    A((void**)&c->d_step, sizeof(long long));
    A((void**)&c->d_scratch, (size_t)c->max_inst_cells * sizeof(double));
    A((void**)&c->d_wl, 2 * (size_t)tiles * sizeof(int));
    A((void**)&c->d_wl_count, 2 * sizeof(int));
    c->same_flags_n = (size_t)(c->model == AMPS_GRAINGROWTH ? c->q : 1) * tiles;
    A((void**)&c->d_same_flags, c->same_flags_n * sizeof(int));
    c->full_cap = (tiles + 7) / 8;
    A((void**)&c->d_full_seg, std::max(8 * (size_t)c->full_cap, (size_t)NSK_SEGS * nsk_cap(tiles)) * sizeof(int));
    A((void**)&c->d_full_cnt, FULL_CNT_INTS * sizeof(int));
    if (rc != AMPS_OK) { std::string m = g_last_error; amps_destroy(c); return fail(rc, m); }
    e = hipMemcpy(c->d_desc, c->desc.data(), sizeof(InstDesc) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->d_tile_inst, c->tile_inst.data(), sizeof(int) * tiles, hipMemcpyHostToDevice);
    if (e != hipSuccess) { amps_destroy(c); return fail(AMPS_E_HIP, std::string("upload: ") + hipGetErrorString(e)); }
    const char* ng = std::getenv("AMPS_NO_GRAPH");
    c->use_graph = !(ng && ng[0] == '1');
    if (cfgs[0].wg_per_cu > 0) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        c->grid_cap = cfgs[0].wg_per_cu * cus;
    }
    c->graph_period = c->mode == AMPS_ADAPTIVE ? (c->every % 2 ? 2 * c->every : c->every) : 10;
    *out = c;
    return AMPS_OK;
}

int amps_create(const amps_config* cfg, amps_ctx** out) { return amps_create_batch(cfg, 1, out); }

void amps_destroy(amps_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    tape_clear(c);
    for (hipEvent_t e : c->kev) (void)hipEventDestroy(e);
    for (int p = 0; p < 2; ++p)
        if (c->gexec[p]) (void)hipGraphExecDestroy(c->gexec[p]);
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int amps_init(amps_ctx* c) {
    if (!c) return fail(AMPS_E_ARG, "ctx is NULL");
    HIPCHK(hipStreamSynchronize(c->stream));
    tape_clear(c);
    c->cur = 0;
    for (int i = 0; i < c->ninst; ++i)
        RCCHK(launch_init(c->d_desc, i, (long long)c->desc[i].nx * c->desc[i].ny, c->model, c->cfg[i].init, c->q,
                          c->buf[0], c->total_cells, c->stream));
    std::vector<int> types((size_t)c->ngroups * c->total_tiles, c->mode == AMPS_PURE_HALF2 ? 1 : 2);
    HIPCHK(hipMemcpyAsync(c->d_type, types.data(), types.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_type_old, types.data(), types.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    long long zero = 0;
    HIPCHK(hipMemcpyAsync(c->d_step, &zero, sizeof(zero), hipMemcpyHostToDevice, c->stream));
    static const unsigned long long z8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    HIPCHK(hipMemcpyAsync(c->d_stats, z8, sizeof(z8), hipMemcpyHostToDevice, c->stream));
    // step 0 appends to parity 0
    if (c->d_full_cnt) HIPCHK(hipMemsetAsync(c->d_full_cnt, 0, FULL_CNT_INTS * sizeof(int), c->stream));
    HIPCHK(hipMemsetAsync(c->d_same_flags, 0, c->same_flags_n * sizeof(int), c->stream));
    RCCHK(clear_zmasks(c));
    if (c->mode == AMPS_PURE_HALF2)
        RCCHK(launch_convert(c->d_desc, c->d_tile_inst, c->buf[0], c->total_cells, c->total_tiles, nullptr, c->d_type,
                             c->gm, CONVERT_ENCODE_LOW, c->tile, c->stream));
    if (c->mode == AMPS_ADAPTIVE && c->monitor == AMPS_MON_TEMPORAL) {
        StepArgs a = step_args(c, 0);
        int ngroups_eval = 1, field0 = c->mfield, stride = 0;
        if (c->model == AMPS_GRAINGROWTH && c->mfield < 0) { ngroups_eval = c->q; field0 = 0; stride = 1; }
        RCCHK(launch_monitor_last_init(a, c->tile, ngroups_eval, field0, stride, c->d_last, c->stream));
    }
    RCCHK(refresh_aux(c, 0));
    RCCHK(refresh_worklists(c));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->steps_done = 0;
    c->monitor_passes = 0;
    c->pending_adv = 0;
    c->timing_open = false;
    c->initialised = true;
    return AMPS_OK;
}

int amps_step(amps_ctx* c, int nsteps) {
    if (!c) return fail(AMPS_E_ARG, "ctx is NULL");
    if (!c->initialised) return fail(AMPS_E_STATE, "amps_init has not been called");
    if (nsteps < 0) return fail(AMPS_E_ARG, "nsteps < 0");
    if (nsteps == 0) return AMPS_OK;
    if (!c->timing_open) {
        HIPCHK(hipEventRecord(c->ev0, c->stream));
        c->timing_open = true;
    }
    int left = nsteps;
    const bool graphs = c->use_graph && !c->taping && !c->ktiming;
    while (left > 0) {
        const long long s = c->steps_done;
        if (graphs && left >= c->graph_period && s % c->graph_period == 0) {
            RCCHK(flush_adv(c));
            if (!c->gexec[c->cur]) RCCHK(build_graph(c, c->cur));
            HIPCHK(hipGraphLaunch(c->gexec[c->cur], c->stream));
            // graph_period is even: the buffer parity is unchanged
            c->steps_done += c->graph_period;
            if (c->mode == AMPS_ADAPTIVE) c->monitor_passes += c->graph_period / c->every;
            left -= c->graph_period;
        } else {
            RCCHK(enqueue_one(c));
            left--;
        }
    }
    RCCHK(flush_adv(c));
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    return AMPS_OK;
}

int amps_sync(amps_ctx* c, double* elapsed_ms) {
    if (!c) return fail(AMPS_E_ARG, "ctx is NULL");
    HIPCHK(hipStreamSynchronize(c->stream));
    if (elapsed_ms) *elapsed_ms = 0.0;
    if (c->timing_open) {
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        if (elapsed_ms) *elapsed_ms = ms;
        c->timing_open = false;
    }
    return AMPS_OK;
}

int amps_monitor_now(amps_ctx* c) {
    if (!c || !c->initialised) return fail(AMPS_E_STATE, "context not initialised");
    if (c->mode != AMPS_ADAPTIVE) return fail(AMPS_E_STATE, "monitor passes need ADAPTIVE mode");
    RCCHK(enqueue_monitor(c, c->cur));
    c->monitor_passes++;
    HIPCHK(hipStreamSynchronize(c->stream));
    return AMPS_OK;
}

int amps_get_field(amps_ctx* c, int inst, int field, double* dst, size_t n) {
    if (!c || !dst) return fail(AMPS_E_ARG, "NULL argument");
    if (inst < 0 || inst >= c->ninst || field < 0 || field >= c->nfields) return fail(AMPS_E_ARG, "instance or field out of range");
    const InstDesc& d = c->desc[inst];
    if (n != (size_t)d.nx * d.ny) return fail(AMPS_E_ARG, "n must equal nx*ny of the instance");
    const int group = has_low(c) ? c->gm.fg[field] : -1;
    RCCHK(launch_gather(c->d_desc, c->buf[c->cur], c->total_cells, c->total_tiles, has_low(c) ? c->d_type : nullptr,
                        group, inst, d.tx * d.ty, field, c->d_scratch, c->tile, c->stream));
    HIPCHK(hipMemcpyAsync(dst, c->d_scratch, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(dst[i])) return fail(AMPS_E_NONFINITE, "field " + std::to_string(field) + " holds non-finite values");
    return AMPS_OK;
}

int amps_set_field(amps_ctx* c, int inst, int field, const double* src, size_t n) {
    if (!c || !src) return fail(AMPS_E_ARG, "NULL argument");
    if (inst < 0 || inst >= c->ninst || field < 0 || field >= c->nfields) return fail(AMPS_E_ARG, "instance or field out of range");
    const InstDesc& d = c->desc[inst];
    if (n != (size_t)d.nx * d.ny) return fail(AMPS_E_ARG, "n must equal nx*ny of the instance");
    tape_clear(c);  // a recorded trajectory no longer leads to the new state
    double* dstp = c->buf[c->cur].f64 + (size_t)field * c->total_cells + d.cell_base;
    HIPCHK(hipMemcpyAsync(dstp, src, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (has_low(c) && c->gm.fg[field] >= 0) {
        GroupMap one = c->gm;
        for (int f = 0; f < c->nfields; ++f) if (f != field) one.fg[f] = -2;
        RCCHK(launch_convert(c->d_desc, c->d_tile_inst, c->buf[c->cur], c->total_cells, c->total_tiles, nullptr,
                             c->d_type, one, CONVERT_ENCODE_LOW, c->tile, c->stream));
    }
    RCCHK(refresh_aux(c, c->cur));
    RCCHK(refresh_worklists(c));  // the other buffer no longer holds the same fields
    HIPCHK(hipMemsetAsync(c->d_same_flags, 0, c->same_flags_n * sizeof(int), c->stream));
    RCCHK(clear_zmasks(c));
    HIPCHK(hipStreamSynchronize(c->stream));
    return AMPS_OK;
}

int amps_get_tile_types(amps_ctx* c, int inst, int32_t* dst, size_t n) {
    if (!c || !dst) return fail(AMPS_E_ARG, "NULL argument");
    if (inst < 0 || inst >= c->ninst) return fail(AMPS_E_ARG, "instance out of range");
    const InstDesc& d = c->desc[inst];
    const size_t nt = (size_t)d.tx * d.ty;
    if (n != nt * c->ngroups) return fail(AMPS_E_ARG, "n must equal tiles * type_groups");
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int g = 0; g < c->ngroups; ++g)
        HIPCHK(hipMemcpy(dst + g * nt, c->d_type + (size_t)g * c->total_tiles + d.tile_base, nt * sizeof(int),
                         hipMemcpyDeviceToHost));
    return AMPS_OK;
}

int amps_set_tile_types(amps_ctx* c, int inst, const int32_t* src, size_t n) {
    if (!c || !src) return fail(AMPS_E_ARG, "NULL argument");
    if (inst < 0 || inst >= c->ninst) return fail(AMPS_E_ARG, "instance out of range");
    if (!has_low(c)) return fail(AMPS_E_STATE, "PURE_FP64 contexts have no tile map");
    const InstDesc& d = c->desc[inst];
    const size_t nt = (size_t)d.tx * d.ty;
    if (n != nt * c->ngroups) return fail(AMPS_E_ARG, "n must equal tiles * type_groups");
    for (size_t i = 0; i < n; ++i)
        if (src[i] != 1 && src[i] != 2) return fail(AMPS_E_ARG, "tile types must be 1 or 2");
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(c->d_type_old, c->d_type, (size_t)c->ngroups * c->total_tiles * sizeof(int), hipMemcpyDeviceToDevice));
    for (int g = 0; g < c->ngroups; ++g)
        HIPCHK(hipMemcpy(c->d_type + (size_t)g * c->total_tiles + d.tile_base, src + g * nt, nt * sizeof(int),
                         hipMemcpyHostToDevice));
    RCCHK(launch_convert(c->d_desc, c->d_tile_inst, c->buf[c->cur], c->total_cells, c->total_tiles, c->d_type_old,
                         c->d_type, c->gm, CONVERT_CHANGED, c->tile, c->stream));
    RCCHK(refresh_aux(c, c->cur));
    RCCHK(refresh_worklists(c));
    HIPCHK(hipMemsetAsync(c->d_same_flags, 0, c->same_flags_n * sizeof(int), c->stream));
    RCCHK(clear_zmasks(c));
    HIPCHK(hipStreamSynchronize(c->stream));
    return AMPS_OK;
}

int amps_get_monitor_values(amps_ctx* c, int inst, double* dst, size_t n) {
    if (!c || !dst) return fail(AMPS_E_ARG, "NULL argument");
    if (inst < 0 || inst >= c->ninst) return fail(AMPS_E_ARG, "instance out of range");
    const InstDesc& d = c->desc[inst];
    const size_t nt = (size_t)d.tx * d.ty;
    if (n != nt * c->ngroups) return fail(AMPS_E_ARG, "n must equal tiles * type_groups");
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int g = 0; g < c->ngroups; ++g)
        HIPCHK(hipMemcpy(dst + g * nt, c->d_md + (size_t)g * c->total_tiles + d.tile_base, nt * sizeof(double),
                         hipMemcpyDeviceToHost));
    return AMPS_OK;
}

int amps_get_raw_field(amps_ctx* c, int inst, int field, uint16_t* words, size_t n, double* off, float* scl,
                       size_t ntiles) {
    if (!c || !words || !off || !scl) return fail(AMPS_E_ARG, "NULL argument");
    if (inst < 0 || inst >= c->ninst || field < 0 || field >= c->nfields) return fail(AMPS_E_ARG, "instance or field out of range");
    if (!has_low(c) || c->gm.fg[field] < 0) return fail(AMPS_E_STATE, "the field has no low-precision storage");
    const InstDesc& d = c->desc[inst];
    if (n != (size_t)d.nx * d.ny || ntiles != (size_t)d.tx * d.ty) return fail(AMPS_E_ARG, "n or ntiles does not match the instance");
    HIPCHK(hipStreamSynchronize(c->stream));
    const Buf& b = c->buf[c->cur];
    HIPCHK(hipMemcpy(words, b.f16 + (size_t)field * c->total_cells + d.cell_base, n * sizeof(uint16_t), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(off, b.off + (size_t)field * c->total_tiles + d.tile_base, ntiles * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(scl, b.scl + (size_t)field * c->total_tiles + d.tile_base, ntiles * sizeof(float), hipMemcpyDeviceToHost));
    return AMPS_OK;
}

int amps_get_info(amps_ctx* c, amps_info* info) {
    if (!c || !info) return fail(AMPS_E_ARG, "NULL argument");
    HIPCHK(hipStreamSynchronize(c->stream));
    unsigned long long st[8];
    HIPCHK(hipMemcpy(st, c->d_stats, sizeof(st), hipMemcpyDeviceToHost));
    std::memset(info, 0, sizeof(*info));
    info->steps_done = c->steps_done;
    info->fp64_tile_updates = (int64_t)st[2];
    info->low_full_tile_updates = (int64_t)st[3];
    info->low_listed_tile_updates = (int64_t)st[4];
    info->monitor_passes = c->monitor_passes;
    int maxt = 0;
    for (auto& d : c->desc) maxt = std::max(maxt, d.tx * d.ty);
    info->ntiles = maxt;
    info->type_groups = c->ngroups;
    info->nfields = c->nfields;
    info->ninstances = c->ninst;
    const double denom = (double)c->ngroups * c->total_tiles;
    if (c->mode == AMPS_ADAPTIVE && c->monitor_passes > 0) {
        info->hi_frac_last = st[1] / denom;
        info->hi_frac_mean = st[0] / (denom * c->monitor_passes);
    } else {
        std::vector<int> t((size_t)c->ngroups * c->total_tiles);
        if (has_low(c)) HIPCHK(hipMemcpy(t.data(), c->d_type, t.size() * sizeof(int), hipMemcpyDeviceToHost));
        size_t hi = 0;
        for (int v : t) hi += (v == 2) || !has_low(c);
        info->hi_frac_last = info->hi_frac_mean = has_low(c) ? hi / denom : 1.0;
    }
    info->cells = (double)c->total_cells;
    return AMPS_OK;
}

}  // extern "C"

// ======================================================= kernel timing
extern "C" int amps_set_kernel_timing(amps_ctx* c, int enable) {
    if (!c) return fail(AMPS_E_ARG, "ctx is NULL");
    RCCHK(kev_flush(c));
    if (enable && c->kev.empty()) {
        c->kev.resize(4 * 512);
        c->kev_pairs.assign(512, 1);
        for (auto& e : c->kev) HIPCHK(hipEventCreate(&e));
    }
    c->ktiming = enable != 0;
    c->klaunches = 0;
    c->kms = 0.0;
    c->kms_part[0] = c->kms_part[1] = 0.0;
    return AMPS_OK;
}

extern "C" int amps_get_kernel_time(amps_ctx* c, int64_t* launches, double* total_ms) {
    return amps_get_sweep_time(c, AMPS_SWEEP_ALL, launches, total_ms);
}

extern "C" int amps_get_sweep_time(amps_ctx* c, int which, int64_t* launches, double* total_ms) {
    if (!c || !launches || !total_ms) return fail(AMPS_E_ARG, "NULL argument");
    if (which < AMPS_SWEEP_ALL || which > AMPS_SWEEP_LOW) return fail(AMPS_E_ARG, "which out of range");
    RCCHK(kev_flush(c));
    *launches = c->klaunches;
    // the snowflake's adaptive step times the low (bulk-hold) sweep first
    const bool low_first = low_sweep_first(c);
    const int part = (which == AMPS_SWEEP_FP64) == low_first ? 1 : 0;
    *total_ms = which == AMPS_SWEEP_ALL ? c->kms : c->kms_part[part];
    return AMPS_OK;
}

// ================================================================ adjoint
extern "C" int amps_tape_begin(amps_ctx* c, int ckpt_every) {
    if (!c || !c->initialised) return fail(AMPS_E_STATE, "context not initialised");
    if (ckpt_every < 1) return fail(AMPS_E_ARG, "ckpt_every must be >= 1");
    if (c->model != AMPS_SNOWFLAKE && c->model != AMPS_SINTERING)
        return fail(AMPS_E_UNSUPPORTED, "the adjoint covers the snowflake and sintering models");
    if (c->mode != AMPS_PURE_FP64) return fail(AMPS_E_UNSUPPORTED, "the adjoint needs a PURE_FP64 context");
    if (c->flags & AMPS_FLAG_CH_CONSERVATIVE)
        return fail(AMPS_E_UNSUPPORTED, "the adjoint covers the reference's Cahn-Hilliard form only");
    tape_clear(c);
    RCCHK(flush_adv(c));
    c->tape_every = ckpt_every;
    c->taping = true;
    c->tape_step0 = c->steps_done;
    return tape_push(c);
}

namespace {
// Gradient slot of a parameter id for the model, or -1.
int grad_slot(int model, int id) {
    if (model == AMPS_SNOWFLAKE) {
        switch (id) {
            case AMPS_GRAD_SF_DELTA: return 0;
            case AMPS_GRAD_SF_KAPPA: return 1;
            case AMPS_GRAD_SF_EPSB: return 2;
            case AMPS_GRAD_SF_TAU: return 3;
        }
    } else if (model == AMPS_SINTERING) {
        switch (id) {
            case AMPS_GRAD_SN_DSURF: return 0;
            case AMPS_GRAD_SN_DGB: return 1;
            case AMPS_GRAD_SN_KC: return 2;
            case AMPS_GRAD_SN_KE: return 3;
            case AMPS_GRAD_SN_L: return 4;
            case AMPS_GRAD_SN_DVOL: return 5;
        }
    }
    return -1;
}

// Device allocations of one adjoint call, released on every exit path.
struct AdjMem {
    std::vector<void*> p;
    ~AdjMem() {
        for (void* q : p) (void)hipFree(q);
    }
    int get(void** out, size_t bytes) {
        *out = nullptr;
        if (hipMalloc(out, bytes) != hipSuccess) return fail(AMPS_E_NOMEM, "adjoint work memory");
        p.push_back(*out);
        return AMPS_OK;
    }
};
}  // namespace

extern "C" int amps_adjoint(amps_ctx* c, const double* dJ_du, const int32_t* param_ids, int np, double* dJ_dp,
                            double* dJ_du0) {
    if (!c || !c->initialised) return fail(AMPS_E_STATE, "context not initialised");
    if (!c->taping || c->tape.empty()) return fail(AMPS_E_STATE, "amps_tape_begin has not been called");
    if (!dJ_du) return fail(AMPS_E_ARG, "dJ_du is NULL");
    if (np < 0 || (np > 0 && (!param_ids || !dJ_dp))) return fail(AMPS_E_ARG, "param_ids/dJ_dp missing");
    std::vector<int> slot(np);
    for (int k = 0; k < np; ++k) {
        slot[k] = grad_slot(c->model, param_ids[k]);
        if (slot[k] < 0) return fail(AMPS_E_ARG, "parameter id not differentiable for this model");
    }
    RCCHK(flush_adv(c));
    const int NG = 6;
    const int F = c->nfields;
    const size_t TC = (size_t)c->total_cells, state = (size_t)F * TC;
    const int k_every = c->tape_every;
    AdjMem mem;
    double* lam[2];
    double* w[6];
    double* gacc;
    long long* dstep;
    std::vector<double*> seg(k_every + 1);
    RCCHK(mem.get((void**)&lam[0], state * 8));
    RCCHK(mem.get((void**)&lam[1], state * 8));
    for (int i = 0; i < 6; ++i) RCCHK(mem.get((void**)&w[i], TC * 8));
    RCCHK(mem.get((void**)&gacc, (size_t)NG * c->total_tiles * 8));
    RCCHK(mem.get((void**)&dstep, 8));
    for (int j = 1; j <= k_every; ++j) RCCHK(mem.get((void**)&seg[j], state * 8));
    // cotangent of the final state: host [inst][field][ny][nx] -> device [field][cells]
    size_t hoff = 0;
    for (int i = 0; i < c->ninst; ++i) {
        const InstDesc& d = c->desc[i];
        const size_t n = (size_t)d.nx * d.ny;
        for (int f = 0; f < F; ++f, hoff += n)
            HIPCHK(hipMemcpyAsync(lam[0] + f * TC + d.cell_base, dJ_du + hoff, n * 8, hipMemcpyHostToDevice,
                                  c->stream));
    }
    HIPCHK(hipMemsetAsync(gacc, 0, (size_t)NG * c->total_tiles * 8, c->stream));
    int cur = 0;
    for (int s = (int)c->tape.size() - 1; s >= 0; --s) {
        const long long S0 = c->tape_steps[s];
        const long long S1 = s + 1 < (int)c->tape.size() ? c->tape_steps[s + 1] : c->steps_done;
        const int m = (int)(S1 - S0);
        if (m <= 0) continue;
        // recompute the segment's states from its checkpoint
        seg[0] = c->tape[s];
        RCCHK(launch_fill_u64((unsigned long long*)dstep, 1, (unsigned long long)S0, c->stream));
        for (int j = 0; j < m; ++j) {
            StepArgs a = step_args(c, 0);
            a.in.f64 = seg[j];
            a.out.f64 = seg[j + 1];
            a.step_ctr = dstep;
            a.step_add = j;
            RCCHK(launch_step_kernel(c, a));
        }
        for (int j = m - 1; j >= 0; --j) {
            AdjArgs a{};
            a.inst = c->d_desc;
            a.tile_inst = c->d_tile_inst;
            a.total_tiles = c->total_tiles;
            a.total_cells = c->total_cells;
            a.u = seg[j];
            a.un = seg[j + 1];
            a.lam_n = lam[cur];
            a.lam = lam[cur ^ 1];
            a.w0 = w[0]; a.w1 = w[1]; a.w2 = w[2]; a.w3 = w[3]; a.w4 = w[4]; a.w5 = w[5];
            a.gacc = gacc;
            a.step_ctr = dstep;
            a.step_add = j;
            RCCHK(launch_adjoint_step(a, c->model, c->tile, c->stream));
            cur ^= 1;
        }
    }
    std::vector<double> g((size_t)NG * c->total_tiles);
    HIPCHK(hipMemcpyAsync(g.data(), gacc, g.size() * 8, hipMemcpyDeviceToHost, c->stream));
    if (dJ_du0) {
        hoff = 0;
        for (int i = 0; i < c->ninst; ++i) {
            const InstDesc& d = c->desc[i];
            const size_t n = (size_t)d.nx * d.ny;
            for (int f = 0; f < F; ++f, hoff += n)
                HIPCHK(hipMemcpyAsync(dJ_du0 + hoff, lam[cur] + f * TC + d.cell_base, n * 8, hipMemcpyDeviceToHost,
                                      c->stream));
        }
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < c->ninst; ++i) {
        const InstDesc& d = c->desc[i];
        const int nt = d.tx * d.ty;
        for (int k = 0; k < np; ++k) {
            double sum = 0.0;
            for (int t = 0; t < nt; ++t) sum += g[(size_t)slot[k] * c->total_tiles + d.tile_base + t];
            dJ_dp[(size_t)i * np + k] = sum;
        }
    }
    return AMPS_OK;
}

// ============================================================ autotuner
namespace {
std::mutex g_tune_mu;
std::map<std::string, amps_tuning> g_tune_cache;

// Every field of every config except the two the tuner varies (tile,
// wg_per_cu), byte for byte, plus the step count: a tiling is reused only
// for the same run.
template <typename T>
void key_add(std::string& k, const T& v) {
    k.append(reinterpret_cast<const char*>(&v), sizeof(v));
}

std::string tune_key(const amps_config* c, int n, int steps) {
    std::string k;
    key_add(k, steps);
    key_add(k, n);
    for (int i = 0; i < n; ++i) {
        const amps_config& x = c[i];
        for (int32_t v : {x.model, x.mode, x.monitor, x.nx, x.ny, x.q, x.monitor_every, x.monitor_field, x.stencil,
                          x.bc, x.init, x.flags, x.dilate})
            key_add(k, v);
        key_add(k, x.threshold);
        key_add(k, x.seed);
        for (double p : x.params) key_add(k, p);
    }
    return k;
}
}  // namespace

extern "C" int amps_autotune(const amps_config* cfgs, int n, int steps, amps_tuning* best) {
    if (!cfgs || !best || n < 1 || n > AMPS_MAX_INST || steps < 1) return fail(AMPS_E_ARG, "bad autotune arguments");
    const std::string key = tune_key(cfgs, n, steps);
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            *best = it->second;
            return AMPS_OK;
        }
    }
    static const int tiles[2] = {16, 32};
    static const int dens[5] = {0, 1, 2, 3, 4};
    amps_tuning b{0, 0, 0.0};
    std::vector<amps_config> v(cfgs, cfgs + n);
    const int warm = steps < 20 ? steps : 20;
    for (int tile : tiles) {
        bool fits = true;
        for (int i = 0; i < n; ++i) fits = fits && cfgs[i].nx % tile == 0 && cfgs[i].ny % tile == 0;
        if (!fits) continue;
        for (int d : dens) {
            for (auto& c : v) { c.tile = tile; c.wg_per_cu = d; }
            amps_ctx* c = nullptr;
            if (amps_create_batch(v.data(), n, &c) != AMPS_OK) continue;
            double ms = 0.0;
            int rc = amps_init(c);
            if (rc == AMPS_OK) rc = amps_step(c, warm);
            if (rc == AMPS_OK) rc = amps_sync(c, nullptr);
            if (rc == AMPS_OK) rc = amps_step(c, steps);
            if (rc == AMPS_OK) rc = amps_sync(c, &ms);
            amps_destroy(c);
            if (rc != AMPS_OK) continue;
            const double us = 1000.0 * ms / steps;
            if (b.tile == 0 || us < b.us_per_step) b = amps_tuning{tile, d, us};
        }
    }
    if (b.tile == 0) return fail(AMPS_E_UNSUPPORTED, "no candidate tiling could run");
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        g_tune_cache[key] = b;
    }
    *best = b;
    return AMPS_OK;
}

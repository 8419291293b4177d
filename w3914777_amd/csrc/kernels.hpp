// kernels.hpp — host-side launch entry points of the gfx950 kernels.
#pragma once

#include "amps_internal.hpp"

namespace amps {

// Which type group each field belongs to (-1: always FP64).
struct GroupMap {
    int nfields;
    int ngroups;
    signed char fg[AMPS_MAX_FIELDS];
};

enum { CONVERT_CHANGED = 0, CONVERT_ENCODE_LOW = 1, CONVERT_DECODE_LOW = 2 };

// Config validation, the descriptor of one instance (tile base and cell
// base 0) and error recording, shared with host code outside api.cpp.
int validate_config(const amps_config& c);
void config_desc(InstDesc& d, const amps_config& c);
int fail_with(int code, const char* msg);

int launch_init(const InstDesc* inst_dev, int inst, long long ncells, int model, int init, int q, const Buf& b,
                long long total_cells, hipStream_t s);
int launch_convert(const InstDesc* inst, const int* tile_inst, const Buf& b, long long total_cells, int total_tiles,
                   const int* type_old, const int* type, const GroupMap& gm, int mode, int tile, hipStream_t s);
// reset (may be null) is set to 0 by the monitor kernel (the counter the
// classification that follows adds to).
int launch_monitor(const StepArgs& a, int kind, int tile, int ngroups_eval, int field0, int field_stride,
                   int type_groups, double* md, double* last, unsigned long long* reset, hipStream_t s);
int launch_monitor_last_init(const StepArgs& a, int tile, int ngroups_eval, int field0, int field_stride,
                             double* last, hipStream_t s);
// single: one instance (descriptor 0 for every tile).
int launch_classify(const InstDesc* inst, const int* tile_inst, int total_tiles, int ngroups, int broadcast,
                    int single, const double* md, double thr, int dilate, int* type_old, int* type,
                    unsigned long long* stats, hipStream_t s);
int launch_gather(const InstDesc* inst, const Buf& b, long long total_cells, int total_tiles, const int* type,
                  int group, int instid, int ntiles_inst, int field, double* dst, int tile, hipStream_t s);
int launch_nonfinite(const double* v, long long n, unsigned long long* cnt, hipStream_t s);
int launch_fill_u64(unsigned long long* p, int n, unsigned long long v, hipStream_t s);
int launch_advance_step(long long* step_ctr, int n, hipStream_t s);

// ev (may be null): timing events bound to the step's sweep kernels: the
// start and end of the first (snowflake adaptive: the bulk-hold sweep) and,
// for the snowflake's adaptive step only, of the second (the mixed sweep).
int launch_sf_step(const StepArgs& a, int tile, bool adaptive, hipStream_t s, const hipEvent_t* ev = nullptr);
// same_flags[f][tile] = 0 and bit f of zmask[tile] cleared for every field f whose
// type group changed type on the tile.
int launch_clear_changed(const int* type_old, const int* type, const GroupMap& gm, int nfields, int total_tiles,
                         int* same_flags, unsigned long long* zmask, hipStream_t s);
int launch_worklist(const int* type, const int* type_old, int total_tiles, int* wl_hi, int* wl_lo, int* counts,
                    int* same_flags, hipStream_t s);
int launch_gg_step(const StepArgs& a, int tile, bool adaptive, int grid, hipStream_t s, const hipEvent_t* ev = nullptr);
// type_old (may be null): refresh only tiles where some field's type differs from a.type.
int launch_gg_aux(const StepArgs& a, int tile, const int* type_old, hipStream_t s);
int launch_sn_step(const StepArgs& a, int tile, bool adaptive, int grid, hipStream_t s, const hipEvent_t* ev = nullptr);

// Reverse-time adjoint of one FP64 step (snowflake, sintering).  Arrays are
// [field][total_cells] like the FP64 state; work arrays hold total_cells.
struct AdjArgs {
    const InstDesc* inst;
    const int* tile_inst;
    int total_tiles;
    long long total_cells;
    const double* u;       // state before the step
    const double* un;      // state after the step
    const double* lam_n;   // dJ/d(state after)
    double* lam;           // dJ/d(state before), written
    double* w0;            // work arrays, one value per cell
    double* w1;
    double* w2;
    double* w3;
    double* w4;
    double* w5;
    double* gacc;          // [gradient slot][total_tiles] accumulators
    const long long* step_ctr;  // step index of the forward step = *step_ctr + step_add
    int step_add;
};
int launch_adjoint_step(const AdjArgs& a, int model, int tile, hipStream_t s);

}  // namespace amps

// noalloc.hip — the FP64 snowflake step on caller-owned device buffers with
// no device allocation: amps_sf_step_fp64_noalloc.
//
// The step kernels read their instance descriptor and step counter through
// device pointers.  amps_step keeps them in buffers a context allocates;
// here they live in `scratch`, a file-scope device array of fixed size that
// every GPU the library runs on holds until the process ends.  A one-thread
// setup kernel writes the descriptor and the step index into it (a kernel,
// not a memset or copy, so the call can be captured into a HIP graph), then
// the FP64 sweep runs from d_in to d_out exactly as in a PURE_FP64 context.
// Every call on a GPU shares that GPU's scratch: calls that overlap on one
// GPU (different streams, or host threads) race.
#include <mutex>

#include "kernels.hpp"

namespace amps {

// Layout of the scratch array.
struct NoAllocSlot {
    InstDesc desc;      // the grid's descriptor
    long long step;     // step index of the step (keys the noise)
    int tile_inst0;     // tile -> instance map of a one-instance grid
};

constexpr size_t NOALLOC_SCRATCH_BYTES = 4096;
static_assert(sizeof(NoAllocSlot) <= NOALLOC_SCRATCH_BYTES, "scratch too small for the descriptor");

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

__global__ void noalloc_setup_kernel(InstDesc d, long long step, NoAllocSlot* slot) {
    if (threadIdx.x == 0) {
        slot->desc = d;
        slot->step = step;
        slot->tile_inst0 = 0;
    }
}

// Device address of `scratch` on the current device, looked up once per device.
static int scratch_address(NoAllocSlot** out) {
    constexpr int MAXDEV = 64;
    static std::mutex mu;
    static void* addr[MAXDEV] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAXDEV)
        return fail_with(AMPS_E_HIP, "no current HIP device");
    std::lock_guard<std::mutex> g(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fail_with(AMPS_E_HIP, "hipGetSymbolAddress(scratch) failed");
        addr[dev] = p;
    }
    *out = static_cast<NoAllocSlot*>(addr[dev]);
    return AMPS_OK;
}

}  // namespace amps

using namespace amps;

extern "C" int amps_sf_step_fp64_noalloc(const amps_config* cfg, const double* d_in, double* d_out, int64_t step,
                                         void* stream) {
    if (!cfg || !d_in || !d_out) return fail_with(AMPS_E_ARG, "NULL argument");
    if (cfg->model != AMPS_SNOWFLAKE) return fail_with(AMPS_E_UNSUPPORTED, "the no-allocation step covers the snowflake model");
    if (cfg->mode != AMPS_PURE_FP64) return fail_with(AMPS_E_UNSUPPORTED, "the no-allocation step is the PURE_FP64 step");
    const int rc = validate_config(*cfg);
    if (rc != AMPS_OK) return rc;
    if (step < 0) return fail_with(AMPS_E_ARG, "step must be >= 0");
    const size_t n = 2 * (size_t)cfg->nx * cfg->ny;
    const char* a0 = reinterpret_cast<const char*>(d_in);
    const char* b0 = reinterpret_cast<const char*>(d_out);
    if (a0 < b0 + n * sizeof(double) && b0 < a0 + n * sizeof(double))
        return fail_with(AMPS_E_ARG, "d_in and d_out overlap");
    NoAllocSlot* slot = nullptr;
    const int sr = scratch_address(&slot);
    if (sr != AMPS_OK) return sr;
    InstDesc d;
    config_desc(d, *cfg);
    if (sizeof(NoAllocSlot) > NOALLOC_SCRATCH_BYTES) return fail_with(AMPS_E_ARG, "descriptor exceeds the scratch");
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(noalloc_setup_kernel, dim3(1), dim3(64), 0, s, d, (long long)step, slot);
    if (hipGetLastError() != hipSuccess) return fail_with(AMPS_E_HIP, "setup kernel launch failed");
    StepArgs a{};
    a.inst = &slot->desc;
    a.tile_inst = &slot->tile_inst0;
    a.total_cells = (long long)cfg->nx * cfg->ny;
    a.total_tiles = d.tx * d.ty;
    a.nwork = a.total_tiles;
    a.nfields = 2;
    a.step_ctr = &slot->step;
    a.step_add = 0;
    a.ninst = 1;
    a.t_halo = cfg->stencil == AMPS_STENCIL_REF ? 2 : 1;
    if (cfg->wg_per_cu > 0) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        a.grid_cap = cfg->wg_per_cu * cus;
    }
    a.in.f64 = const_cast<double*>(d_in);
    a.out.f64 = d_out;
    const int lr = launch_sf_step(a, cfg->tile, false, s);
    return lr == AMPS_OK ? AMPS_OK : fail_with(lr, "FP64 sweep launch failed");
}

// valu_rates.hip — issue cost of the VALU instructions the snowflake sweeps
// are built from, on the device it runs on.
//
// Each kernel runs ITER iterations of 8 independent dependency chains per
// thread (so issue, not latency, bounds it) over 8 workgroups of 256 threads
// per CU.  The host side reports, per kernel, SIMD cycles at the device's
// nominal clock divided by the wave-instructions issued: the cost of one
// wave64 instruction of that kind.  The instructions per iteration of each
// loop body are stated beside its kernel (counted in this file's gfx950
// assembly; the compiler unrolls the chains and keeps nothing else in the
// loops).  Every thread writes one float of `out` (blocks x 256 floats).
#include <hip/hip_runtime.h>

namespace {

typedef _Float16 h2 __attribute__((ext_vector_type(2)));
constexpr int ITER = 2048;

// 8 v_fma_f32 per iteration
__global__ __launch_bounds__(256) void k_fma_f32(const float* in, float* out) {
    float a[8];
    const float b = in[0], c = in[1];
    for (int j = 0; j < 8; ++j) a[j] = in[2 + j] + threadIdx.x;
    for (int i = 0; i < ITER; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] = __builtin_fmaf(a[j], b, c);
    float s = 0;
    for (int j = 0; j < 8; ++j) s += a[j];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

// 8 v_fma_f64 per iteration
__global__ __launch_bounds__(256) void k_fma_f64(const float* in, float* out) {
    double a[8];
    const double b = in[0], c = in[1];
    for (int j = 0; j < 8; ++j) a[j] = in[2 + j] + threadIdx.x;
    for (int i = 0; i < ITER; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] = __builtin_fma(a[j], b, c);
    double s = 0;
    for (int j = 0; j < 8; ++j) s += a[j];
    out[blockIdx.x * 256 + threadIdx.x] = (float)s;
}

// 8 v_pk_fma_f16 per iteration (two fp16 values each)
__global__ __launch_bounds__(256) void k_pk_fma_f16(const float* in, float* out) {
    h2 a[8];
    const h2 b = {(_Float16)in[0], (_Float16)in[0]}, c = {(_Float16)in[1], (_Float16)in[1]};
    for (int j = 0; j < 8; ++j) a[j] = h2{(_Float16)(in[2 + j] + threadIdx.x), (_Float16)in[2 + j]};
    for (int i = 0; i < ITER; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] = __builtin_elementwise_fma(a[j], b, c);
    float s = 0;
    for (int j = 0; j < 8; ++j) s += (float)a[j].x + (float)a[j].y;
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

// 8 v_rcp_f32 per iteration
__global__ __launch_bounds__(256) void k_rcp_f32(const float* in, float* out) {
    float a[8];
    for (int j = 0; j < 8; ++j) a[j] = in[2 + j] + threadIdx.x;
    for (int i = 0; i < ITER; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] = __builtin_amdgcn_rcpf(a[j]);
    float s = 0;
    for (int j = 0; j < 8; ++j) s += a[j];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

// 8 v_rcp_f16 + 1 v_pack_b32_f16 per iteration
__global__ __launch_bounds__(256) void k_rcp_f16(const float* in, float* out) {
    _Float16 a[8];
    for (int j = 0; j < 8; ++j) a[j] = (_Float16)(in[2 + j] + threadIdx.x);
    for (int i = 0; i < ITER; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] = __builtin_amdgcn_rcph(a[j]);
    float s = 0;
    for (int j = 0; j < 8; ++j) s += (float)a[j];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

// 8 v_add_f32 + 8 v_cvt_f16_f32 + 8 v_cvt_f32_f16 per iteration
__global__ __launch_bounds__(256) void k_add_cvt(const float* in, float* out) {
    float a[8];
    const float c = in[1];
    for (int j = 0; j < 8; ++j) a[j] = in[2 + j] + threadIdx.x;
    for (int i = 0; i < ITER; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] = (float)(_Float16)(a[j] + c);
    float s = 0;
    for (int j = 0; j < 8; ++j) s += a[j];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

// 4 v_pk_fma_f16 + 8 v_fma_mix_f32 per iteration
__global__ __launch_bounds__(256) void k_fma_mix(const float* in, float* out) {
    float a[8];
    const float b = in[0];
    h2 hv[4];
    const h2 hb = {(_Float16)in[0], (_Float16)in[0]}, hc = {(_Float16)in[1], (_Float16)in[1]};
    for (int j = 0; j < 4; ++j) hv[j] = h2{(_Float16)in[2 + j], (_Float16)in[3 + j]};
    for (int j = 0; j < 8; ++j) a[j] = in[2 + j] + threadIdx.x;
    for (int i = 0; i < ITER; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            hv[j] = __builtin_elementwise_fma(hv[j], hb, hc);
            a[2 * j] = __builtin_fmaf(b, (float)hv[j].x, a[2 * j]);
            a[2 * j + 1] = __builtin_fmaf(b, (float)hv[j].y, a[2 * j + 1]);
        }
    float s = 0;
    for (int j = 0; j < 8; ++j) s += a[j];
    for (int j = 0; j < 4; ++j) s += (float)hv[j].x;
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

typedef void (*Kern)(const float*, float*);

}  // namespace

// Runs kernel `which` (0..6 in the order above) `reps` times after one
// warm-up launch; writes the elapsed milliseconds, the CU count and the
// nominal clock (kHz).  Returns 0, or a HIP error code.
extern "C" int valu_rate_run(int which, int reps, float* ms, int* cus, int* clock_khz) {
    static const Kern ks[] = {k_fma_f32, k_fma_f64, k_pk_fma_f16, k_rcp_f32, k_rcp_f16, k_add_cvt, k_fma_mix};
    if (which < 0 || which >= 7 || reps < 1) return -1;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(clock_khz, hipDeviceAttributeClockRate, dev);
    if (e != hipSuccess) return (int)e;
    const int blocks = *cus * 8;
    const float h_in[10] = {0.999f, 0.001f, 1, 2, 3, 4, 5, 6, 7, 8};
    float *din = nullptr, *dout = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    e = hipMalloc(&din, sizeof h_in);
    if (e == hipSuccess) e = hipMalloc(&dout, sizeof(float) * blocks * 256);
    if (e == hipSuccess) e = hipMemcpy(din, h_in, sizeof h_in, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(ks[which], dim3(blocks), dim3(256), 0, 0, din, dout);
        e = hipEventRecord(e0);
        for (int r = 0; r < reps && e == hipSuccess; ++r) {
            hipLaunchKernelGGL(ks[which], dim3(blocks), dim3(256), 0, 0, din, dout);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(e1);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (din) (void)hipFree(din);
    if (dout) (void)hipFree(dout);
    return (int)e;
}

// Loop iterations per launch and threads per workgroup, for the host's arithmetic.
extern "C" int valu_rate_iters() { return ITER; }

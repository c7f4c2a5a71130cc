// probe.hip -- measured FP64 and HBM peaks of the current device, the denominators
// of the roofline fractions beside the datasheet's (78.6 TFLOP/s FP64, 8 TB/s HBM).
//
//   FP64: eight independent FMA chains per thread, eight 256-thread workgroups per CU
//         (latency hidden), 2 flop per FMA.
//   HBM:  a grid-stride read of a 1 GiB buffer in 16-byte loads, and a copy of it.
#include <hip/hip_runtime.h>

#include "../../include/aderdg.h"
#include "capi_common.hpp"

namespace {

constexpr int kChains = 8;
constexpr int kFmaIters = 20000;

__global__ __launch_bounds__(256) void k_probe_fma(double *out, double a, double c)
{
    double x[kChains];
#pragma unroll
    for (int k = 0; k < kChains; k++) x[k] = threadIdx.x * 1e-3 + k;
    for (int it = 0; it < kFmaIters; it++) {
#pragma unroll
        for (int k = 0; k < kChains; k++) x[k] = __builtin_fma(x[k], a, c);
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kChains; k++) s += x[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_probe_read(const double2 *__restrict__ a, size_t n2, double *out)
{
    double s = 0.0;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n2; k += (size_t)gridDim.x * blockDim.x) {
        const double2 v = a[k];
        s += v.x + v.y;
    }
    if (s == 1.0e300) out[0] = s;  // never true for the zero-filled buffer: keeps the loads
}

__global__ __launch_bounds__(256) void k_probe_copy(const double2 *__restrict__ a, double2 *__restrict__ b, size_t n2)
{
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n2; k += (size_t)gridDim.x * blockDim.x)
        b[k] = a[k];
}

}  // namespace

extern "C" int aderdg_probe_peaks(int device, double *fp64_tflops, double *hbm_read_gbs, double *hbm_copy_gbs)
{
    using namespace aderdg;
    int cus = 0;
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    const int grid = 8 * (cus > 0 ? cus : 1), block = 256;
    const size_t bytes = (size_t)1 << 30, n2 = bytes / 16;
    double *a = nullptr, *b = nullptr, *out = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int r = 0;
    auto done = [&](int code) {
        if (st) (void)hipStreamSynchronize(st);
        if (a) (void)hipFree(a);
        if (b) (void)hipFree(b);
        if (out) (void)hipFree(out);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (st) (void)hipStreamDestroy(st);
        return code;
    };
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess || hipMalloc((void **)&a, bytes) != hipSuccess ||
        hipMalloc((void **)&b, bytes) != hipSuccess ||
        hipMalloc((void **)&out, sizeof(double) * (size_t)grid * block) != hipSuccess ||
        hipMemsetAsync(a, 0, bytes, st) != hipSuccess)
        return done(fail(ADERDG_EHIP, "aderdg_probe_peaks: setup failed"));
    // best of three launches of each probe (the first warms clocks and caches)
    auto best_ms = [&](auto launch) {
        float best = -1.0f;
        for (int rep = 0; rep < 3; rep++) {
            float ms = 0.0f;
            if (hipEventRecord(e0, st) != hipSuccess) return -1.0f;
            launch();
            if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                return -1.0f;
            if (best < 0.0f || ms < best) best = ms;
        }
        return best;
    };
    const float t_fma = best_ms([&] { hipLaunchKernelGGL(k_probe_fma, dim3(grid), dim3(block), 0, st, out, 0.999999, 1e-9); });
    const float t_read = best_ms([&] {
        hipLaunchKernelGGL(k_probe_read, dim3(grid * 2), dim3(block), 0, st, (const double2 *)a, n2, out);
    });
    const float t_copy = best_ms([&] {
        hipLaunchKernelGGL(k_probe_copy, dim3(grid * 2), dim3(block), 0, st, (const double2 *)a, (double2 *)b, n2);
    });
    if (t_fma <= 0.0f || t_read <= 0.0f || t_copy <= 0.0f || hipGetLastError() != hipSuccess)
        return done(fail(ADERDG_EHIP, "aderdg_probe_peaks: probe launch failed"));
    if (fp64_tflops) *fp64_tflops = 2.0 * kChains * kFmaIters * (double)grid * block / (t_fma * 1e-3) / 1e12;
    if (hbm_read_gbs) *hbm_read_gbs = (double)bytes / (t_read * 1e-3) / 1e9;
    if (hbm_copy_gbs) *hbm_copy_gbs = 2.0 * (double)bytes / (t_copy * 1e-3) / 1e9;
    return done(r);
}

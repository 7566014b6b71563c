// pimc_probe.hip -- a plain streaming copy of known size, used only to scale
// the HBM byte counters (FETCH_SIZE / WRITE_SIZE) that rocprofv3 reports for
// the hot kernels.  Not part of the reference's interface.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/pimc_gpu.h"

namespace {
__global__ __launch_bounds__(256) void k_copy_probe(const double4* __restrict__ src, double4* __restrict__ dst,
                                                   size_t n) {
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}
}  // namespace

extern "C" int pimc_copy_probe(size_t bytes, int iters, double* avg_ms) {
    if (bytes < 32 || iters < 1) return PIMC_E_ARG;
    const size_t n = bytes / sizeof(double4);
    double4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * sizeof(double4)) != hipSuccess) return PIMC_E_NOMEM;
    if (hipMalloc(&b, n * sizeof(double4)) != hipSuccess) {
        (void)hipFree(a);
        return PIMC_E_NOMEM;
    }
    hipStream_t st;
    hipEvent_t e0, e1;
    int rc = PIMC_OK;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) rc = PIMC_E_HIP;
    if (rc == PIMC_OK) {
        (void)hipMemsetAsync(a, 0, n * sizeof(double4), st);
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        const int blocks = 256 * 16;
        hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, st, a, b, n);
        (void)hipEventRecord(e0, st);
        for (int k = 0; k < iters; ++k) hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, st, a, b, n);
        (void)hipEventRecord(e1, st);
        if (hipEventSynchronize(e1) != hipSuccess) rc = PIMC_E_HIP;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (avg_ms) *avg_ms = ms / iters;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(st);
    }
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

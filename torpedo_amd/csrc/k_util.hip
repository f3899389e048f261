// Measurement helper: a plain streaming copy, used to calibrate HBM counters and to state the
// achievable bandwidth beside the kernels' roofline fractions.
#include "tpd_ctx.hpp"

namespace tpd {

// Four independent 16-byte loads in flight per thread per iteration.
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const float4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
        dst[i] = a;
        dst[i + stride] = b;
        dst[i + 2 * stride] = c;
        dst[i + 3 * stride] = d;
    }
    for (; i < n; i += stride) dst[i] = src[i];
}

}  // namespace tpd

using namespace tpd;

extern "C" int tpd_measure_copy(tpd_ctx* c, uint64_t bytes, uint32_t iters, float* ms_per_copy) {
    if (!c || !ms_per_copy || bytes < 16 || iters == 0) return fail(TPD_ERR_INVALID, "tpd_measure_copy: bad argument");
    hipSetDevice(c->device);
    const size_t n = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&a, n * 16);
    if (e == hipSuccess) e = hipMalloc(&b, n * 16);
    if (e == hipSuccess) e = hipMemsetAsync(a, 1, n * 16, c->stream);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        const uint32_t grid = 256 * 8 * 4;
        hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, c->stream, a, b, n);  // warm-up
        hipEventRecord(e0, c->stream);
        for (uint32_t i = 0; i < iters; ++i) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, c->stream, a, b, n);
        hipEventRecord(e1, c->stream);
        e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        *ms_per_copy = ms / (float)iters;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return e == hipSuccess ? TPD_OK : hip_fail(e, "tpd_measure_copy");
}

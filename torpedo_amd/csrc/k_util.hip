// Measurement helpers: plain streaming kernels (copy, nontemporal copy, read-only, write-only),
// hipMemcpyAsync and random gathers, used to calibrate HBM counters and to state the achievable
// bandwidth and random-read rate beside the kernels' roofline fractions.
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

using f4v = float __attribute__((ext_vector_type(4)));

// The same copy with nontemporal loads and stores (no L2 allocation for the streamed data).
__global__ __launch_bounds__(256) void k_copy_nt(const f4v* __restrict__ src, f4v* __restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const f4v a = __builtin_nontemporal_load(src + i), b = __builtin_nontemporal_load(src + i + stride);
        const f4v c = __builtin_nontemporal_load(src + i + 2 * stride), d = __builtin_nontemporal_load(src + i + 3 * stride);
        __builtin_nontemporal_store(a, dst + i);
        __builtin_nontemporal_store(b, dst + i + stride);
        __builtin_nontemporal_store(c, dst + i + 2 * stride);
        __builtin_nontemporal_store(d, dst + i + 3 * stride);
    }
    for (; i < n; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

// Read-only stream: eight 16-byte loads in flight per thread, summed; the sum is stored only if it
// equals a value it never takes, so the loads stay and nothing is written.
__global__ __launch_bounds__(256) void k_read(const float4* __restrict__ src, float* __restrict__ sink, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    float acc = 0.0f;
    for (; i + 7 * stride < n; i += 8 * stride) {
        float4 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = src[i + u * stride];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += (v[u].x + v[u].y) + (v[u].z + v[u].w);
    }
    for (; i < n; i += stride) acc += src[i].x;
    if (acc == -1.2345e30f) sink[threadIdx.x] = acc;
}

// Write-only stream: 16-byte stores.
__global__ __launch_bounds__(256) void k_fill(float4* __restrict__ dst, size_t n, float v) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const float4 f = make_float4(v, v, v, v);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = f;
}

// Random reads: out[i] = src[h(i) mod n_src] for kGatherReads indices h(i) of a 64-bit mix, four
// independent reads in flight per thread (the pattern of a gather by a sorted index).
constexpr uint32_t kGatherReads = 1u << 25;
__device__ __forceinline__ uint32_t gather_index(uint32_t i, uint32_t n_src) {
    uint64_t z = (uint64_t)i * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (uint32_t)((z ^ (z >> 31)) % n_src);
}
template <typename T>
__global__ __launch_bounds__(256) void k_gather(const T* __restrict__ src, uint32_t n_src, T* __restrict__ out) {
    const uint32_t i0 = (blockIdx.x * 256 + threadIdx.x) * 4;  // the grid covers kGatherReads exactly
    const T a = src[gather_index(i0, n_src)], b = src[gather_index(i0 + 1, n_src)];
    const T c = src[gather_index(i0 + 2, n_src)], d = src[gather_index(i0 + 3, n_src)];
    out[i0] = a;
    out[i0 + 1] = b;
    out[i0 + 2] = c;
    out[i0 + 3] = d;
}

// The memory pattern of one stable radix pass without its ranking: block b reads its keys_per_block
// keys and values coalesced, already in digit order inside the block (key & (bins - 1) is the
// digit), and writes each at run_start[b][d] + (j - local_start[b][d]), so that consecutive lanes
// store consecutive places of one digit run, as a pass writes its LDS-reordered keys: first all
// keys, then all values.  Each thread loads its keys_per_block / blockDim.x (at most
// kRunScatterItems) keys and values before it stores any, as a pass holds them in registers.  The
// tables (bins words each per block) come from random digit histograms; they are read into LDS
// once per block.
constexpr uint32_t kRunScatterItems = 32;
template <typename K>
__global__ __launch_bounds__(1024) void k_run_scatter(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                      K* __restrict__ kout, uint32_t* __restrict__ vout,
                                                      const uint32_t* __restrict__ run_start,
                                                      const uint32_t* __restrict__ local_start,
                                                      uint32_t keys_per_block, uint32_t bins) {
    __shared__ uint32_t s_run[512], s_local[512];
    const size_t b = blockIdx.x;
    const uint32_t nt = blockDim.x;
    for (uint32_t d = threadIdx.x; d < bins; d += nt) {
        s_run[d] = run_start[b * bins + d];
        s_local[d] = local_start[b * bins + d];
    }
    const size_t base = b * keys_per_block;
    const uint32_t items = keys_per_block / nt;
    K k[kRunScatterItems];
    uint32_t v[kRunScatterItems];
#pragma unroll
    for (uint32_t i = 0; i < kRunScatterItems; ++i)
        if (i < items) k[i] = kin[base + threadIdx.x + i * nt];
#pragma unroll
    for (uint32_t i = 0; i < kRunScatterItems; ++i)
        if (i < items) v[i] = vin[base + threadIdx.x + i * nt];
    __syncthreads();
    uint32_t dst[kRunScatterItems];
#pragma unroll
    for (uint32_t i = 0; i < kRunScatterItems; ++i)
        if (i < items) {
            const uint32_t j = threadIdx.x + i * nt;
            const uint32_t d = (uint32_t)k[i] & (bins - 1);
            dst[i] = s_run[d] + (j - s_local[d]);
            kout[dst[i]] = k[i];
        }
    __syncthreads();
#pragma unroll
    for (uint32_t i = 0; i < kRunScatterItems; ++i)
        if (i < items) vout[dst[i]] = v[i];
}

}  // namespace tpd

using namespace tpd;

extern "C" int tpd_measure_run_scatter(tpd_ctx* c, uint32_t key_bytes, uint32_t n, uint32_t bins,
                                       uint32_t keys_per_block, uint32_t blocks_per_cu, uint32_t iters, float* ms) {
    if (!c || !ms || iters == 0 || (key_bytes != 4 && key_bytes != 8) || (bins != 256 && bins != 512) ||
        keys_per_block == 0 || keys_per_block % 512 || keys_per_block > 512 * kRunScatterItems || n == 0 ||
        n % keys_per_block)
        return fail(TPD_ERR_INVALID, "tpd_measure_run_scatter: bad argument");
    hipSetDevice(c->device);
    const uint32_t nblocks = n / keys_per_block;
    // 512 threads per block, 1,024 from 12,288 keys per block (the radix passes' shapes)
    const uint32_t threads = keys_per_block >= 12288 && keys_per_block % 1024 == 0 ? 1024 : 512;
    // Residency limit: dynamic LDS so that blocks_per_cu blocks fit in a CU's LDS and one more
    // does not (0: no limit beyond the registers).
    int lds_cu = 0, lds_block = 0;
    if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, c->device) != hipSuccess ||
        hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) != hipSuccess)
        return fail(TPD_ERR_HIP, "tpd_measure_run_scatter: LDS size query");
    const size_t lds_static = 2 * 512 * sizeof(uint32_t);
    size_t lds_dyn = 0;
    if (blocks_per_cu) {
        lds_dyn = ((size_t)lds_cu / (blocks_per_cu + 1) + 1024 - lds_static + 255) / 256 * 256;
        if (lds_dyn + lds_static > (size_t)lds_block || (lds_dyn + lds_static) * blocks_per_cu > (size_t)lds_cu)
            return fail(TPD_ERR_INVALID, "tpd_measure_run_scatter: blocks_per_cu outside what the LDS can express");
    }
    try {
        // Random digits per key (a 64-bit mix of its index), counted per block; the block's keys are
        // laid out in digit order, so only the histogram matters.
        const size_t tn = (size_t)nblocks * bins;
        std::vector<uint32_t> counts(tn, 0), run(tn), local(tn);
        std::vector<uint32_t> keys32(key_bytes == 4 ? n : 0);
        std::vector<uint64_t> keys64(key_bytes == 8 ? n : 0);
        for (uint32_t blk = 0; blk < nblocks; ++blk) {
            uint32_t* cnt = &counts[(size_t)blk * bins];
            for (uint32_t j = 0; j < keys_per_block; ++j) {
                uint64_t z = ((uint64_t)blk * keys_per_block + j + 1) * 0x9E3779B97F4A7C15ull;
                z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
                z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
                ++cnt[(uint32_t)(z ^ (z >> 31)) & (bins - 1)];
            }
            uint32_t at = 0;
            for (uint32_t d = 0; d < bins; ++d) {
                local[(size_t)blk * bins + d] = at;
                for (uint32_t r = 0; r < cnt[d]; ++r, ++at) {
                    const size_t i = (size_t)blk * keys_per_block + at;
                    if (key_bytes == 4) keys32[i] = d | ((uint32_t)i << 9);
                    else keys64[i] = d | ((uint64_t)i << 9);
                }
            }
        }
        // Global positions: digit-major, then block order (the stable order of a radix pass).
        uint64_t at = 0;
        for (uint32_t d = 0; d < bins; ++d)
            for (uint32_t blk = 0; blk < nblocks; ++blk) {
                run[(size_t)blk * bins + d] = (uint32_t)at;
                at += counts[(size_t)blk * bins + d];
            }
        if (at != n) return fail(TPD_ERR_INVALID, "tpd_measure_run_scatter: histograms do not cover the keys");
        void *kin = nullptr, *kout = nullptr;
        uint32_t *vin = nullptr, *vout = nullptr, *d_run = nullptr, *d_local = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        const size_t kb = (size_t)n * key_bytes, tb = (size_t)nblocks * bins * 4;
        hipError_t e = hipMalloc(&kin, kb);
        if (e == hipSuccess) e = hipMalloc(&kout, kb);
        if (e == hipSuccess) e = hipMalloc(&vin, (size_t)n * 4);
        if (e == hipSuccess) e = hipMalloc(&vout, (size_t)n * 4);
        if (e == hipSuccess) e = hipMalloc(&d_run, tb);
        if (e == hipSuccess) e = hipMalloc(&d_local, tb);
        if (e == hipSuccess)
            e = hipMemcpy(kin, key_bytes == 4 ? (const void*)keys32.data() : (const void*)keys64.data(), kb,
                          hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_run, run.data(), tb, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_local, local.data(), tb, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemsetAsync(vin, 0, (size_t)n * 4, c->stream);
        if (e == hipSuccess) e = hipEventCreate(&e0);
        if (e == hipSuccess) e = hipEventCreate(&e1);
        auto one = [&]() -> hipError_t {
            if (key_bytes == 4)
                hipLaunchKernelGGL(k_run_scatter<uint32_t>, dim3(nblocks), dim3(threads), lds_dyn, c->stream,
                                   static_cast<const uint32_t*>(kin), vin, static_cast<uint32_t*>(kout), vout, d_run,
                                   d_local, keys_per_block, bins);
            else
                hipLaunchKernelGGL(k_run_scatter<uint64_t>, dim3(nblocks), dim3(threads), lds_dyn, c->stream,
                                   static_cast<const uint64_t*>(kin), vin, static_cast<uint64_t*>(kout), vout, d_run,
                                   d_local, keys_per_block, bins);
            return hipGetLastError();
        };
        if (e == hipSuccess) e = one();  // warm-up
        if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
        for (uint32_t i = 0; i < iters && e == hipSuccess; ++i) e = one();
        if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float t = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        *ms = t / (float)iters;
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        for (void* p : {kin, kout, (void*)vin, (void*)vout, (void*)d_run, (void*)d_local})
            if (p) (void)hipFree(p);
        return e == hipSuccess ? TPD_OK : hip_fail(e, "tpd_measure_run_scatter");
    } catch (const std::bad_alloc&) {
        return fail(TPD_ERR_CAPACITY, "tpd_measure_run_scatter: host memory for the key layout");
    }
}

// kind: 0 copy kernel, 1 hipMemcpyAsync device to device, 2 nontemporal copy kernel, 3 read-only
// kernel, 4 write-only kernel, 5 / 6 random 8-byte / 4-byte reads (2^25 of them from an array of
// `bytes`).  blocks: grid of the streaming kernels (0: 8192).  *ms: mean time per pass.
extern "C" int tpd_measure_bandwidth(tpd_ctx* c, uint32_t kind, uint64_t bytes, uint32_t iters, uint32_t blocks,
                                     float* ms) {
    if (!c || !ms || bytes < 16 || iters == 0 || kind > 6)
        return fail(TPD_ERR_INVALID, "tpd_measure_bandwidth: bad argument");
    if (kind >= 5 && bytes / 16 * (kind == 5 ? 2u : 4u) > 0xFFFFFFFFull)
        return fail(TPD_ERR_INVALID, "tpd_measure_bandwidth: gather array above 2^32 elements");
    hipSetDevice(c->device);
    const size_t n = bytes / 16;
    const uint32_t grid = blocks ? blocks : 256 * 8 * 4;
    float4 *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&a, n * 16);
    if (e == hipSuccess && kind <= 2) e = hipMalloc(&b, n * 16);
    if (e == hipSuccess && kind == 3) e = hipMalloc(&b, 256 * sizeof(float));
    if (e == hipSuccess && kind >= 5) e = hipMalloc(&b, (size_t)kGatherReads * 8);
    if (e == hipSuccess) e = hipMemsetAsync(a, 1, n * 16, c->stream);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    auto one = [&]() -> hipError_t {
        switch (kind) {
            case 0: hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, c->stream, a, b, n); break;
            case 1: return hipMemcpyAsync(b, a, n * 16, hipMemcpyDeviceToDevice, c->stream);
            case 2:
                hipLaunchKernelGGL(k_copy_nt, dim3(grid), dim3(256), 0, c->stream, reinterpret_cast<const f4v*>(a),
                                   reinterpret_cast<f4v*>(b), n);
                break;
            case 3: hipLaunchKernelGGL(k_read, dim3(grid), dim3(256), 0, c->stream, a, reinterpret_cast<float*>(b), n); break;
            case 4: hipLaunchKernelGGL(k_fill, dim3(grid), dim3(256), 0, c->stream, a, n, 1.0f); break;
            case 5:
                hipLaunchKernelGGL(k_gather<uint64_t>, dim3(kGatherReads / 1024), dim3(256), 0, c->stream,
                                   reinterpret_cast<const uint64_t*>(a), (uint32_t)(n * 2), reinterpret_cast<uint64_t*>(b));
                break;
            default:
                hipLaunchKernelGGL(k_gather<uint32_t>, dim3(kGatherReads / 1024), dim3(256), 0, c->stream,
                                   reinterpret_cast<const uint32_t*>(a), (uint32_t)(n * 4), reinterpret_cast<uint32_t*>(b));
                break;
        }
        return hipGetLastError();
    };
    if (e == hipSuccess) e = one();  // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
    for (uint32_t i = 0; i < iters && e == hipSuccess; ++i) e = one();
    if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    *ms = t / (float)iters;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return e == hipSuccess ? TPD_OK : hip_fail(e, "tpd_measure_bandwidth");
}

extern "C" int tpd_measure_copy(tpd_ctx* c, uint64_t bytes, uint32_t iters, float* ms_per_copy) {
    return tpd_measure_bandwidth(c, 0, bytes, iters, 0, ms_per_copy);
}

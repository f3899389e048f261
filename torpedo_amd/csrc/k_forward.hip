// Forward path kernels: entity matrices, fused preprocess + look-back scan + duplicate,
// tile ranges, and the front-to-back blend.
//
// Reference: torpedo/volumetric/assets/gaussian/{project,prefix,keygen,range,blend}.slang and the
// dispatch sequence of GaussianEngine::recordSplat/recordBlend (GaussianEngine.cpp:777-863).
#include "tpd_blend.hpp"
#include "tpd_kernels.hpp"

namespace tpd {

// ------------------------------------------------------------------ entity matrices

__global__ void k_entity_mats(const float* __restrict__ models, uint32_t n_ent, CameraParams cam,
                              EntityMats* __restrict__ out) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_ent) return;
    entity_mats(models ? models + 16 * (size_t)e : nullptr, cam, out[e]);
}

void launch_entity_mats(const float* models, uint32_t n_ent, const CameraParams& cam, EntityMats* out,
                        hipStream_t s) {
    const uint32_t ne = n_ent ? n_ent : 1u;
    hipLaunchKernelGGL(k_entity_mats, dim3((ne + 63) / 64), dim3(64), 0, s, models, ne, cam, out);
}

// ------------------------------------------------------------------ preprocess + scan + duplicate

// Gaussians per preprocess partition: one per thread, 1,024 threads (a quarter of the look-backs of
// 256; 24 KB of LDS).
constexpr uint32_t kDupBlock = 1024;

__global__ __launch_bounds__(kDupBlock) void k_preprocess_dup(const ForwardArgs a) {
    __shared__ uint32_t s_part;
    __shared__ uint32_t s_wave_sum[kDupBlock / kWave], s_wave_big[kDupBlock / kWave];
    __shared__ uint32_t s_big_incl[kDupBlock], s_start[kDupBlock];
    __shared__ int32_t s_x0[kDupBlock], s_y0[kDupBlock], s_w[kDupBlock];
    __shared__ uint32_t s_depth[kDupBlock];
    __shared__ uint64_t s_base;

    const uint32_t tid = threadIdx.x;
    if (tid == 0) s_part = atomicAdd(&a.counters[kCtrScanTicket], 1u);
    // Tile ranges are cleared here (every block strides over them) so that the ranges kernel,
    // which runs after the sort, only writes the non-empty tiles.
    for (uint32_t t = blockIdx.x * kDupBlock + tid; t < a.ntiles; t += gridDim.x * kDupBlock)
        a.ranges[t] = make_uint2(0u, 0u);
    __syncthreads();
    const uint32_t part = __builtin_amdgcn_readfirstlane(s_part);  // block-uniform: keep it scalar
    if (part >= gridDim.x) return;  // tickets are reset every frame; never true in a consistent state
    const uint32_t i = part * kDupBlock + tid;

    Projected p;
    p.tiles = 0;
    p.radius = 0.0f;
    if (i < a.n) {
        const uint32_t e = a.ent_idx ? a.ent_idx[i] : 0u;
        // The record into registers at once, 16-byte loads when the records are 16-byte aligned (read
        // field by field as the projection needs them, each load waited for in turn, the kernel took
        // a quarter longer, at 63 instead of 78 VGPRs).
        const float* g = a.gauss + (size_t)i * 60;
        float rec[60];
        if ((reinterpret_cast<uintptr_t>(a.gauss) & 15u) == 0u) {  // uniform
            const float4* r4 = reinterpret_cast<const float4*>(g);
#pragma unroll
            for (int q = 0; q < 15; ++q) {
                const float4 t = r4[q];
                rec[4 * q + 0] = t.x; rec[4 * q + 1] = t.y; rec[4 * q + 2] = t.z; rec[4 * q + 3] = t.w;
            }
        } else {
#pragma unroll
            for (int q = 0; q < 60; ++q) rec[q] = g[q];
        }
        const bool vis = project_gaussian(rec, a.ent[e].vm, a.ent[e].pvm, a.cam.view,
                                          (0.5f * (float)a.cam.width) * a.cam.focal_ndc[0],
                                          (0.5f * (float)a.cam.height) * a.cam.focal_ndc[1], a.cam.width,
                                          a.cam.height, a.gx, a.gy, a.sh_degree, p);
        store_splat(a.splats + i, vis, p);
        a.tiles[i] = p.tiles;
    }

    // Block-wide inclusive scans of the tile counts and of the counts above kOwnTiles.
    const uint32_t wave = tid / kWave, lane = __lane_id();
    // Depths that will be emitted but lie outside the sort's depth window: the host then sorts the
    // frame on the wider key.
    if (ballot(p.tiles != 0 && depth_outside_window(p.vz)) != 0 && lane == 0)
        atomicOr(&a.counters[kCtrDepthOut], 1u);
    const uint32_t t = p.tiles, big = t > kOwnTiles ? t : 0u;
    const uint32_t incl_w = wave_incl_scan(t), big_w = wave_incl_scan(big);
    if (lane == kWave - 1) {
        s_wave_sum[wave] = incl_w;
        s_wave_big[wave] = big_w;
    }
    __syncthreads();
    uint32_t wave_off = 0, block_total = 0, big_off = 0, big_total = 0;
#pragma unroll
    for (uint32_t w = 0; w < kDupBlock / kWave; ++w) {
        const uint32_t v = s_wave_sum[w], vb = s_wave_big[w];
        wave_off += (w < wave) ? v : 0u;
        block_total += v;
        big_off += (w < wave) ? vb : 0u;
        big_total += vb;
    }
    const uint32_t excl_in_block = wave_off + incl_w - t;
    s_big_incl[tid] = big_off + big_w;
    if (big) {
        s_start[tid] = excl_in_block;
        s_x0[tid] = p.x0;
        s_y0[tid] = p.y0;
        s_w[tid] = p.x1 - p.x0;
        s_depth[tid] = as_uint(p.vz);
    }

    // Decoupled look-back (atomic ticket order guarantees the predecessors are running).
    if (wave == 0) {
        if (lane == 0) store_relaxed(&a.scan_desc[part], (part == 0 ? kDescP : kDescA) | (uint64_t)block_total);
        uint64_t excl = 0;
        if (part > 0) excl = scan_lookback(a.scan_desc, part);
        if (lane == 0) {
            if (part > 0) store_relaxed(&a.scan_desc[part], kDescP | (excl + block_total));
            s_base = excl;
            const uint32_t nparts = (a.n + kDupBlock - 1) / kDupBlock;
            if (part == nparts - 1) {
                const uint64_t total = excl + block_total;
                a.counters[kCtrPairs] = total > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)total;
                a.counters[kCtrOverflow] = total > a.cap ? 1u : 0u;
            }
        }
    }
    __syncthreads();
    const uint64_t base = s_base;
    if (i < a.n) a.offsets[i] = (uint32_t)(base + excl_in_block);
    if (t && !big)
        emit_own_pairs(t, p.x0, p.y0, p.x1, as_uint(p.vz), base + excl_in_block, a.gx, 0u, a.cap, a.keys, a.vals,
                       a.val_base + i);
    emit_large_pairs<kDupBlock>(s_big_incl, s_start, s_x0, s_y0, s_w, s_depth, big_total, base, a.gx, 0u, a.cap, a.keys, a.vals,
                     a.val_base + part * kDupBlock);
}

void launch_preprocess_dup(const ForwardArgs& a, hipStream_t s) {
    const uint32_t nb = (a.n + kDupBlock - 1) / kDupBlock;
    if (nb == 0) return;
    hipLaunchKernelGGL(k_preprocess_dup, dim3(nb), dim3(kDupBlock), 0, s, a);
}

// ------------------------------------------------------------------ ranges

// Each thread takes runs of kRangeRun consecutive sorted keys, read with 16-byte loads, and
// compares each key's tile with the one before it (the key before the run is one more load).
constexpr uint32_t kRangeRun = 8;

template <typename KT>
__device__ __forceinline__ uint32_t range_tile(KT k, uint32_t tshift) {
    return (uint32_t)(k >> tshift);
}

template <typename KT>
__device__ __forceinline__ void ranges_of_run(const KT* __restrict__ keys, uint32_t i0, uint32_t P, uint32_t tshift,
                                              uint32_t* __restrict__ r) {
    uint32_t t[kRangeRun];
    const uint32_t n = min(kRangeRun, P - i0);
    if (n == kRangeRun) {
        const uint4* v = reinterpret_cast<const uint4*>(keys + i0);  // i0 is a multiple of 8 keys
        if constexpr (sizeof(KT) == 4) {
            const uint4 x = v[0], y = v[1];
            t[0] = x.x >> tshift; t[1] = x.y >> tshift; t[2] = x.z >> tshift; t[3] = x.w >> tshift;
            t[4] = y.x >> tshift; t[5] = y.y >> tshift; t[6] = y.z >> tshift; t[7] = y.w >> tshift;
        } else {  // 64-bit keys: the tile is the high word
#pragma unroll
            for (uint32_t q = 0; q < 4; ++q) {
                const uint4 x = v[q];
                t[2 * q] = x.y;
                t[2 * q + 1] = x.w;
            }
        }
    } else {
#pragma unroll
        for (uint32_t j = 0; j < kRangeRun; ++j) t[j] = j < n ? range_tile(keys[i0 + j], tshift) : 0u;
    }
    uint32_t prev = i0 > 0 ? range_tile(keys[i0 - 1], tshift) : 0u;
#pragma unroll
    for (uint32_t j = 0; j < kRangeRun; ++j) {
        if (j >= n) break;
        const uint32_t idx = i0 + j, cur = t[j];
        if (idx == 0) {
            r[2 * cur] = 0;
        } else if (cur != prev) {
            r[2 * prev + 1] = idx;
            r[2 * cur] = idx;
        }
        if (idx == P - 1) r[2 * cur + 1] = P;
        prev = cur;
    }
}

// range.slang:16-35.  Empty tiles keep the (0, 0) written by the preprocess kernel (the
// reference clears with fillBuffer, GaussianEngine.cpp:843-848).  The kernel also re-zeroes the
// scan descriptors and the scan ticket for the next frame.  A frame sorted on the depth-window key
// although one of its depths lies outside the window (kCtrDepthOut) has some pairs out of depth
// order inside their tiles (never in the wrong tile: the key's tile bits are kept); it is marked
// overflowed, so that every later kernel skips it and the host renders it again on the wider key.
__global__ void k_ranges(const void* keys, uint32_t* __restrict__ counters, uint32_t cap,
                         uint2* __restrict__ ranges, uint64_t* scan_desc, uint32_t n_scan_desc, uint32_t* ticket,
                         uint32_t key_drop, uint32_t lo_bits, uint32_t count_slot) {
    const uint32_t P = min(counters[count_slot], cap);
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nthreads = gridDim.x * blockDim.x;
    for (uint32_t d = gid; d < n_scan_desc; d += nthreads) scan_desc[d] = 0;
    if (gid == 0) {
        *ticket = 0;
        counters[kCtrScanTicket2] = 0;
        if (lo_bits == kDepthWindowBits && counters[kCtrDepthOut]) counters[kCtrOverflow] = 1u;
    }
    uint32_t* r = reinterpret_cast<uint32_t*>(ranges);
    const uint32_t runs = (P + kRangeRun - 1) / kRangeRun;
    if (key_drop == kKeysWide) {
        for (uint32_t run = gid; run < runs; run += nthreads)
            ranges_of_run(static_cast<const uint64_t*>(keys), run * kRangeRun, P, 32u, r);
    } else {
        for (uint32_t run = gid; run < runs; run += nthreads)
            ranges_of_run(static_cast<const uint32_t*>(keys), run * kRangeRun, P, lo_bits - key_drop, r);
    }
}

void launch_ranges(const void* keys, uint32_t* counters, uint32_t P_hint, uint32_t ntiles, uint2* ranges,
                   uint64_t* scan_desc, uint32_t n_scan_desc, uint32_t* ticket, hipStream_t s, uint32_t cap,
                   uint32_t key_drop, uint32_t lo_bits, uint32_t count_slot) {
    (void)ntiles;
    uint32_t work = (P_hint + kRangeRun - 1) / kRangeRun;
    work = work > n_scan_desc ? work : n_scan_desc;
    uint32_t nb = (work + 255) / 256;
    nb = nb < 1 ? 1 : (nb > 8192 ? 8192 : nb);
    hipLaunchKernelGGL(k_ranges, dim3(nb), dim3(256), 0, s, keys, counters, cap, ranges, scan_desc, n_scan_desc, ticket,
                       key_drop, lo_bits, count_slot);
}

__global__ void k_full_keys(const uint32_t* __restrict__ keys32, uint32_t tshift, const uint32_t* __restrict__ vals,
                            const Splat* __restrict__ splats, uint32_t n, uint64_t* __restrict__ out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = ((uint64_t)(keys32[i] >> tshift) << 32) | as_uint(splats[vals[i]].xyzr.z);
}

void launch_full_keys(const uint32_t* keys32, uint32_t drop, uint32_t lo_bits, const uint32_t* vals,
                      const Splat* splats, uint32_t n, uint64_t* out, hipStream_t s) {
    if (n == 0) return;
    uint32_t nb = (n + 255) / 256;
    nb = nb > 8192 ? 8192 : nb;
    hipLaunchKernelGGL(k_full_keys, dim3(nb), dim3(256), 0, s, keys32, lo_bits - drop, vals, splats, n, out);
}

// ------------------------------------------------------------------ tile order

// Tiles are ordered by class (bit length of the pair count, 0 for an empty tile), heaviest class
// first.  Two launches of one 1024-thread block per chunk of 1024 tiles: the first counts each
// chunk's tiles per class into hist[class * chunks + chunk]; the second gives every chunk its
// start in each class (the tiles of all heavier classes, then those of the class in earlier
// chunks) and scatters its tiles.  Within a chunk and class the order is the order in which waves
// claim slots.  Every wave combines its lanes of one class into one LDS atomic.  Inside each
// class's range of launch positions the tiles are placed by xcd_unmap, so that each XCD (blocks
// go to the 8 XCDs round-robin) takes one contiguous run of the class's tiles, in index order by
// chunks: neighbouring tiles, which share splats, then mostly meet in one L2.

__device__ __forceinline__ uint32_t tile_class(const uint2* __restrict__ ranges, uint32_t t) {
    const uint2 r = ranges[t];
    const uint32_t n = r.y > r.x ? r.y - r.x : 0u;
    return n ? 32u - (uint32_t)__builtin_clz(n) : 0u;  // 0 .. 32
}

// Adds, for each lane with `valid`, one to slot[cls] and returns the lane's slot value before
// the addition (lanes of one class in lane order).
__device__ __forceinline__ uint32_t claim_class_slot(uint32_t* slot, uint32_t cls, bool valid) {
    const uint32_t lane = __lane_id();
    const uint64_t lt = lanemask_lt();
    uint64_t todo = ballot(valid);
    uint32_t mine = 0;
    while (todo) {
        const uint32_t leader = (uint32_t)__builtin_ctzll(todo);
        const uint32_t c = __builtin_amdgcn_readlane(cls, leader);
        const uint64_t peers = ballot(valid && cls == c) & todo;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&slot[c], (uint32_t)__popcll(peers));
        base = __builtin_amdgcn_readlane(base, leader);
        if ((peers >> lane) & 1ull) mine = base + (uint32_t)__popcll(peers & lt);
        todo &= ~peers;
    }
    return mine;
}

__global__ __launch_bounds__(kOrderChunk) void k_tile_class_count(const uint2* __restrict__ ranges, uint32_t ntiles,
                                                                  const uint32_t* counters,
                                                                  uint32_t* __restrict__ hist) {
    __shared__ uint32_t cnt[kTileClasses];
    const uint32_t tid = threadIdx.x, chunks = gridDim.x;
    if (counters && counters[kCtrOverflow]) return;
    if (tid < kTileClasses) cnt[tid] = 0;
    __syncthreads();
    const uint32_t t = blockIdx.x * kOrderChunk + tid;
    (void)claim_class_slot(cnt, t < ntiles ? tile_class(ranges, t) : 0u, t < ntiles);
    __syncthreads();
    if (tid < kTileClasses) hist[tid * chunks + blockIdx.x] = cnt[tid];
}

__global__ __launch_bounds__(kOrderChunk) void k_tile_order(const uint2* __restrict__ ranges, uint32_t ntiles,
                                                            const uint32_t* counters, const uint32_t* __restrict__ hist,
                                                            uint32_t* __restrict__ order) {
    __shared__ uint32_t slot[kTileClasses], total[kTileClasses], cstart[kTileClasses];
    const uint32_t tid = threadIdx.x, lane = __lane_id(), wave = tid / kWave, chunks = gridDim.x;
    if (counters && counters[kCtrOverflow]) return;
    for (uint32_t c = wave; c < kTileClasses; c += kOrderChunk / kWave) {
        uint64_t all = 0, before = 0;
        for (uint32_t k = lane; k < chunks; k += kWave) {
            const uint32_t v = hist[c * chunks + k];
            all += v;
            before += k < blockIdx.x ? v : 0u;
        }
        all = wave_sum_u64(all);
        before = wave_sum_u64(before);
        if (lane == 0) {
            total[c] = (uint32_t)all;
            slot[c] = (uint32_t)before;
        }
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t pos = 0;
        for (int c = (int)kTileClasses - 1; c >= 0; --c) {
            cstart[c] = pos;
            slot[c] += pos;
            pos += total[c];
        }
    }
    __syncthreads();
    const uint32_t t = blockIdx.x * kOrderChunk + tid;
    const uint32_t cls = t < ntiles ? tile_class(ranges, t) : 0u;
    const uint32_t pos = claim_class_slot(slot, cls, t < ntiles);
    if (t < ntiles) order[cstart[cls] + xcd_unmap(pos - cstart[cls], total[cls])] = t;
}

void launch_tile_order(const uint2* ranges, uint32_t ntiles, const uint32_t* counters, uint32_t* order, hipStream_t s) {
    if (ntiles == 0) return;
    const uint32_t chunks = (ntiles + kOrderChunk - 1) / kOrderChunk;
    uint32_t* hist = order + ntiles;
    hipLaunchKernelGGL(k_tile_class_count, dim3(chunks), dim3(kOrderChunk), 0, s, ranges, ntiles, counters, hist);
    hipLaunchKernelGGL(k_tile_order, dim3(chunks), dim3(kOrderChunk), 0, s, ranges, ntiles, counters,
                       (const uint32_t*)hist, order);
}

// ------------------------------------------------------------------ blend

// blend.slang:22-105.  One 16x16 tile per 256-thread block (blend_tile); block indices are
// spread so that neighbouring tiles share an XCD's L2.
template <bool COUNT>
__global__ __launch_bounds__(kTilePixels) void k_blend(const BlendArgs a) {
    __shared__ BlendLds lds;
    const uint32_t tile = a.tile_order ? a.tile_order[blockIdx.x] : xcd_remap(blockIdx.x, gridDim.x);
    const TileView tv{tile % a.gx, tile / a.gx, a.width, a.height, a.gx, 0, 0};
    blend_tile<COUNT>(lds, tv, a.ranges[tile], a.vals, a.splats, a.image, a.rgba8, a.eval_count, a.final_T,
                      a.final_last);
}

void launch_blend(const BlendArgs& a, hipStream_t s) {
    if (a.ntiles == 0) return;
    if (a.eval_count)
        hipLaunchKernelGGL(k_blend<true>, dim3(a.ntiles), dim3(kTilePixels), 0, s, a);
    else
        hipLaunchKernelGGL(k_blend<false>, dim3(a.ntiles), dim3(kTilePixels), 0, s, a);
}

}  // namespace tpd

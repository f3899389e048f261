// Batched rendering: many views of many scenes, one launch per stage.  The reference renders one
// camera per GaussianEngine::rasterFrame (GaussianEngine.h:26) and has no batched form; each view
// of a batch is rendered exactly as the plain frame renders it (same per-Gaussian projection,
// same per-view key order, same per-tile blend), only the index spaces are concatenated (the
// layout is described with BatchArgs).
//
//   k_batch_mats            entity matrices of every (view, entity)
//   k_batch_count           per (preprocess block, view) pair counts; decoupled look-back over
//                           the blocks turns them into pair positions; total P and the overflow
//                           flag; clears the tile ranges
//   k_batch_preprocess      each block reads its 256 Gaussian records once and, for every view of
//                           its scene, projects them, writes the splats and emits the pairs
//   (sort, ranges)          the plain path's kernels over the global keys
//   k_batch_blend           one 16x16 tile of some view per block (blend_tile)
//   k_batch_blend_bwd       blend_bwd_tile per tile
//   k_batch_preprocess_bwd  per Gaussian, the chain rule summed over the views of its scene, into
//                           the scene's gradient buffers
#include "tpd_blend.hpp"
#include "tpd_grad.hpp"
#include "tpd_kernels.hpp"

namespace tpd {

// Also clears the frame's depth-window flag, which the count pass that follows may set.
__global__ void k_batch_mats(const BatchSceneDev* __restrict__ scenes, const BatchViewDev* __restrict__ views,
                             EntityMats* __restrict__ mats, uint32_t* __restrict__ counters) {
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) counters[kCtrDepthOut] = 0u;
    const BatchViewDev& vw = views[blockIdx.y];
    const BatchSceneDev& sc = scenes[vw.scene];
    const uint32_t ne = sc.n_ent ? sc.n_ent : 1u;
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    entity_mats(sc.models ? sc.models + 16 * (size_t)e : nullptr, vw.cam, mats[vw.mat_base + e]);
}

void launch_batch_mats(const BatchSceneDev* scenes, const BatchViewDev* views, uint32_t nviews, uint32_t max_ent,
                       EntityMats* mats, uint32_t* counters, hipStream_t s) {
    if (nviews == 0) return;
    const uint32_t ne = max_ent ? max_ent : 1u;
    hipLaunchKernelGGL(k_batch_mats, dim3((ne + 63) / 64, nviews), dim3(64), 0, s, scenes, views, mats, counters);
}

__device__ __forceinline__ bool project_in_view(const float* rec, const EntityMats& m, const BatchViewDev& vw,
                                                uint32_t sh_degree, Projected& p) {
    return project_gaussian(rec, m.vm, m.pvm, vw.cam.view, (0.5f * (float)vw.cam.width) * vw.cam.focal_ndc[0],
                            (0.5f * (float)vw.cam.height) * vw.cam.focal_ndc[1], vw.cam.width, vw.cam.height, vw.gx,
                            vw.gy, sh_degree, p);
}

// ------------------------------------------------------------------ count pass

// The read-only tables come as separate __restrict__ parameters so that the compiler may fetch
// them with scalar loads (uniform addresses, no aliasing stores).
__global__ __launch_bounds__(kPreBlock) void k_batch_count(const BatchArgs a, const BatchSceneDev* __restrict__ scenes,
                                                           const BatchViewDev* __restrict__ views,
                                                           const uint32_t* __restrict__ scene_views,
                                                           const uint32_t* __restrict__ block_scene,
                                                           const EntityMats* __restrict__ mats) {
    constexpr uint32_t kWaves = kPreBlock / kWave;
    __shared__ uint32_t s_part;
    __shared__ uint32_t s_w[kWaves][kWave];  // wave sums of the current chunk of 64 views

    const uint32_t tid = threadIdx.x, wave = tid / kWave, lane = __lane_id();
    if (tid == 0) s_part = atomicAdd(&a.counters[kCtrScanTicket], 1u);
    for (uint32_t t = blockIdx.x * kPreBlock + tid; t < a.ntiles; t += gridDim.x * kPreBlock)
        a.ranges[t] = make_uint2(0u, 0u);
    if (a.nblocks == 0 && blockIdx.x == 0 && tid == 0) {  // no Gaussians: the one block reports P = 0
        a.counters[kCtrPairs] = 0;
        a.counters[kCtrOverflow] = 0;
    }
    __syncthreads();
    // The ticket is block-uniform: read it into a scalar register so that the scene and view
    // tables below are fetched with scalar loads.
    const uint32_t b = __builtin_amdgcn_readfirstlane(s_part);
    if (b >= a.nblocks) return;  // tickets are reset every frame; true only without Gaussians
    const BatchSceneDev& sc = scenes[block_scene[b]];
    const uint32_t lb = b - sc.first_block, nviews = sc.nviews;
    const uint32_t g = lb * kPreBlock + tid;
    const bool valid = g < sc.n;
    const uint32_t e = (valid && sc.ent_idx) ? sc.ent_idx[g] : 0u;
    const uint32_t entry0 = sc.entry_base + lb * nviews;
    // Only position, opacity, quaternion and scale decide the tile count; the colour part of the
    // projection is unused here.
    float rec[60];
#pragma unroll
    for (int q = 0; q < 60; ++q) rec[q] = 0.0f;
    if (valid) {
        const float4* r4 = reinterpret_cast<const float4*>(sc.gauss + (size_t)g * 60);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const float4 t = r4[q];
            rec[4 * q + 0] = t.x; rec[4 * q + 1] = t.y; rec[4 * q + 2] = t.z; rec[4 * q + 3] = t.w;
        }
    }

    // Per view: the block's pair count, stored in bases[] until the look-back gives positions.
    uint32_t my_total = 0;  // threads 0..63: sum of the counts they stored
    bool depth_out = false;  // an emitted depth outside the sort's depth window
    for (uint32_t k0 = 0; k0 < nviews; k0 += kWave) {
        const uint32_t kn = min(kWave, nviews - k0);
        for (uint32_t kk = 0; kk < kn; ++kk) {
            const BatchViewDev& vw = views[scene_views[sc.view_begin + k0 + kk]];
            uint32_t t = 0;
            if (valid) {
                Projected p;
                project_in_view(rec, mats[vw.mat_base + e], vw, sc.sh_degree, p);
                t = p.tiles;
                depth_out |= t != 0 && depth_outside_window(p.vz);
            }
            const uint32_t ws = wave_incl_scan(t);
            if (lane == kWave - 1) s_w[wave][kk] = ws;
        }
        __syncthreads();
        if (tid < kn) {
            uint32_t c = 0;
#pragma unroll
            for (uint32_t w = 0; w < kWaves; ++w) c += s_w[w][tid];
            a.bases[entry0 + k0 + tid] = c;
            my_total += c;
        }
        __syncthreads();
    }

    if (ballot(depth_out) != 0 && lane == 0) atomicOr(&a.counters[kCtrDepthOut], 1u);

    if (wave == 0) {
        const uint64_t total = wave_sum_u64(my_total);
        if (lane == 0) store_relaxed(&a.scan_desc[b], (b == 0 ? kDescP : kDescA) | total);
        uint64_t excl = 0;
        if (b > 0) excl = scan_lookback(a.scan_desc, b);
        if (lane == 0) {
            if (b > 0) store_relaxed(&a.scan_desc[b], kDescP | (excl + total));
            if (b == a.nblocks - 1) {
                const uint64_t P = excl + total;
                a.counters[kCtrPairs] = P > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)P;
                a.counters[kCtrOverflow] = P > a.cap ? 1u : 0u;
            }
        }
        // counts -> first pair position of each (block, view), views in the scene's order
        uint64_t carry = excl;
        for (uint32_t k0 = 0; k0 < nviews; k0 += kWave) {
            const uint32_t k = k0 + lane;
            const uint32_t c = k < nviews ? a.bases[entry0 + k] : 0u;  // stored by this same thread
            const uint32_t incl = wave_incl_scan(c);
            if (k < nviews) {
                const uint64_t pos = carry + incl - c;
                a.bases[entry0 + k] = pos > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)pos;
            }
            carry += __shfl(incl, kWave - 1, kWave);
        }
    }
}

void launch_batch_count(const BatchArgs& a, hipStream_t s) {
    const uint32_t nb = a.nblocks ? a.nblocks : 1u;  // one block clears the ranges of an empty batch
    hipLaunchKernelGGL(k_batch_count, dim3(nb), dim3(kPreBlock), 0, s, a, a.scenes, a.views, a.scene_views,
                       a.block_scene, a.mats);
}

// ------------------------------------------------------------------ preprocess + duplicate

__global__ __launch_bounds__(kPreBlock) void k_batch_preprocess(const BatchArgs a,
                                                                const BatchSceneDev* __restrict__ scenes,
                                                                const BatchViewDev* __restrict__ views,
                                                                const uint32_t* __restrict__ scene_views,
                                                                const uint32_t* __restrict__ block_scene,
                                                                const EntityMats* __restrict__ mats) {
    // Views are software-pipelined so that one barrier per view suffices: the segment between two
    // barriers emits the large splats' pairs of view k-1, derives the pair positions of view k and
    // writes its own pairs, and projects view k+1.  The LDS arrays are double-buffered by view
    // parity: a segment reads view k-1's emission arrays and view k's wave sums while it writes
    // view k's emission arrays and view k+1's wave sums.
    __shared__ uint32_t s_wave_sum[2][kPreBlock / kWave], s_wave_big[2][kPreBlock / kWave];
    __shared__ uint32_t s_big_incl[2][kPreBlock], s_start[2][kPreBlock];
    __shared__ int32_t s_x0[2][kPreBlock], s_y0[2][kPreBlock], s_w[2][kPreBlock];
    __shared__ uint32_t s_depth[2][kPreBlock];

    const uint32_t tid = threadIdx.x, wave = tid / kWave, lane = __lane_id();
    const uint32_t b = blockIdx.x;
    const BatchSceneDev& sc = scenes[block_scene[b]];
    const uint32_t lb = b - sc.first_block, nviews = sc.nviews;
    const uint32_t g = lb * kPreBlock + tid;
    const bool valid = g < sc.n;
    const uint32_t e = (valid && sc.ent_idx) ? sc.ent_idx[g] : 0u;
    const uint32_t entry0 = sc.entry_base + lb * nviews;
    const uint32_t sh_degree = sc.sh_degree;
    if (nviews == 0) return;

    // The record is read once and kept in registers for every view.
    float rec[60];
    if (valid) {
        const float4* r4 = reinterpret_cast<const float4*>(sc.gauss + (size_t)g * 60);
#pragma unroll
        for (int q = 0; q < 15; ++q) {
            const float4 t = r4[q];
            rec[4 * q + 0] = t.x; rec[4 * q + 1] = t.y; rec[4 * q + 2] = t.z; rec[4 * q + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < 60; ++q) rec[q] = 0.0f;
    }

    // This thread's Gaussian in the view last projected, and its wave-inclusive counts.
    Projected p;
    uint32_t incl_w = 0, big_w = 0;
    // Projects view k, writes the splat record and tile count, and publishes the wave totals of
    // the tile counts and of the counts above kOwnTiles.
    const auto project_view = [&](uint32_t k) {
        const BatchViewDev& vw = views[scene_views[sc.view_begin + k]];
        const uint32_t item = vw.item_base + g;
        p.tiles = 0;
        if (valid) {
            if (project_in_view(rec, mats[vw.mat_base + e], vw, sh_degree, p)) store_splat(a.splats + item, true, p);
            if (a.tiles) a.tiles[item] = p.tiles;
        }
        const uint32_t t = p.tiles, big = t > kOwnTiles ? t : 0u;
        incl_w = wave_incl_scan(t);
        big_w = wave_incl_scan(big);
        if (lane == kWave - 1) {
            s_wave_sum[k & 1u][wave] = incl_w;
            s_wave_big[k & 1u][wave] = big_w;
        }
    };

    project_view(0);
    uint32_t big_total_prev = 0, base_prev = 0;
    for (uint32_t k = 0; k < nviews; ++k) {
        __syncthreads();
        if (k > 0) {
            const uint32_t q = (k - 1) & 1u;
            const BatchViewDev& pv = views[scene_views[sc.view_begin + k - 1]];
            emit_large_pairs(s_big_incl[q], s_start[q], s_x0[q], s_y0[q], s_w[q], s_depth[q], big_total_prev,
                             base_prev, pv.gx, pv.tile_base, a.cap, a.keys, a.vals, pv.item_base + lb * kPreBlock);
        }
        {
            const uint32_t q = k & 1u;
            const BatchViewDev& vw = views[scene_views[sc.view_begin + k]];
            const uint32_t t = p.tiles, big = t > kOwnTiles ? t : 0u;
            uint32_t wave_off = 0, big_off = 0, big_total = 0;
#pragma unroll
            for (uint32_t w = 0; w < kPreBlock / kWave; ++w) {
                wave_off += (w < wave) ? s_wave_sum[q][w] : 0u;
                const uint32_t vb = s_wave_big[q][w];
                big_off += (w < wave) ? vb : 0u;
                big_total += vb;
            }
            const uint32_t excl = wave_off + incl_w - t;
            const uint32_t base = a.bases[entry0 + k];
            s_big_incl[q][tid] = big_off + big_w;
            if (big) {
                s_start[q][tid] = excl;
                s_x0[q][tid] = p.x0;
                s_y0[q][tid] = p.y0;
                s_w[q][tid] = p.x1 - p.x0;
                s_depth[q][tid] = as_uint(p.vz);
            } else if (t) {
                emit_own_pairs(t, p.x0, p.y0, p.x1, as_uint(p.vz), (uint64_t)base + excl, vw.gx, vw.tile_base, a.cap,
                               a.keys, a.vals, vw.item_base + g);
            }
            if (a.offsets && valid) a.offsets[vw.item_base + g] = base + excl;
            big_total_prev = big_total;
            base_prev = base;
        }
        if (k + 1 < nviews) project_view(k + 1);
    }
    __syncthreads();
    {
        const uint32_t q = (nviews - 1) & 1u;
        const BatchViewDev& pv = views[scene_views[sc.view_begin + nviews - 1]];
        emit_large_pairs(s_big_incl[q], s_start[q], s_x0[q], s_y0[q], s_w[q], s_depth[q], big_total_prev, base_prev,
                         pv.gx, pv.tile_base, a.cap, a.keys, a.vals, pv.item_base + lb * kPreBlock);
    }
}

void launch_batch_preprocess(const BatchArgs& a, hipStream_t s) {
    if (a.nblocks == 0) return;
    hipLaunchKernelGGL(k_batch_preprocess, dim3(a.nblocks), dim3(kPreBlock), 0, s, a, a.scenes, a.views,
                       a.scene_views, a.block_scene, a.mats);
}

// ------------------------------------------------------------------ blend, forward and backward

__device__ __forceinline__ TileView batch_tile_view(const BatchViewDev* __restrict__ views,
                                                    const uint32_t* __restrict__ tile_view, uint32_t tile) {
    const BatchViewDev& vw = views[tile_view[tile]];
    const uint32_t local = tile - vw.tile_base;
    return TileView{local % vw.gx, local / vw.gx, vw.cam.width, vw.cam.height, vw.gx, vw.gy, vw.pix_base};
}

__global__ __launch_bounds__(kTilePixels) void k_batch_blend(const BatchBlendArgs a,
                                                             const BatchViewDev* __restrict__ views,
                                                             const uint32_t* __restrict__ tile_view) {
    __shared__ BlendLds lds;
    if (a.counters[kCtrOverflow]) return;  // the frame is rendered again with larger buffers
    const uint32_t tile = a.tile_order ? a.tile_order[blockIdx.x] : xcd_remap(blockIdx.x, gridDim.x);
    const TileView tv = batch_tile_view(views, tile_view, tile);
    blend_tile<false>(lds, tv, a.ranges[tile], a.vals, a.splats, a.image, a.rgba8, nullptr, a.final_T, a.final_last);
}

void launch_batch_blend(const BatchBlendArgs& a, hipStream_t s) {
    if (a.ntiles == 0) return;
    hipLaunchKernelGGL(k_batch_blend, dim3(a.ntiles), dim3(kTilePixels), 0, s, a, a.views, a.tile_view);
}

__global__ __launch_bounds__(kTilePixels) void k_batch_blend_bwd(const BatchBlendArgs a,
                                                                 const BatchViewDev* __restrict__ views,
                                                                 const uint32_t* __restrict__ tile_view) {
    __shared__ BlendBwdLds lds;
    if (a.counters[kCtrOverflow]) return;
    const uint32_t tile = a.tile_order ? a.tile_order[blockIdx.x] : xcd_remap(blockIdx.x, gridDim.x);
    const TileView tv = batch_tile_view(views, tile_view, tile);
    blend_bwd_tile(lds, tv, a.ranges[tile], a.vals, a.splats, a.dL_dimg, a.offsets, a.partials, a.final_T,
                   a.final_last);
}

void launch_batch_blend_bwd(const BatchBlendArgs& a, hipStream_t s) {
    if (a.ntiles == 0) return;
    hipLaunchKernelGGL(k_batch_blend_bwd, dim3(a.ntiles), dim3(kTilePixels), 0, s, a, a.views, a.tile_view);
}

// ------------------------------------------------------------------ preprocess backward

__global__ __launch_bounds__(kPreBlock) void k_batch_preprocess_bwd(const BatchPreBwdArgs a,
                                                                    const BatchSceneDev* __restrict__ scenes,
                                                                    const BatchViewDev* __restrict__ views,
                                                                    const uint32_t* __restrict__ scene_views,
                                                                    const uint32_t* __restrict__ block_scene,
                                                                    const EntityMats* __restrict__ mats,
                                                                    const BatchGradsDev* __restrict__ grads) {
    if (a.counters[kCtrOverflow]) return;  // gradients are left untouched; the step is run again
    const uint32_t b = blockIdx.x;
    const uint32_t si = block_scene[b];
    const BatchSceneDev& sc = scenes[si];
    const uint32_t g = (b - sc.first_block) * kPreBlock + threadIdx.x;
    if (g >= sc.n) return;
    const float* rec = sc.gauss + (size_t)g * 60;
    const uint32_t e = sc.ent_idx ? sc.ent_idx[g] : 0u;
    float gm[3] = {0.f, 0.f, 0.f}, gls[3] = {0.f, 0.f, 0.f}, gq[4] = {0.f, 0.f, 0.f, 0.f}, gop = 0.f;
    float gsh[48];
#pragma unroll
    for (int k = 0; k < 48; ++k) gsh[k] = 0.f;
    for (uint32_t k = 0; k < sc.nviews; ++k) {
        const BatchViewDev& vw = views[scene_views[sc.view_begin + k]];
        const uint32_t item = vw.item_base + g;
        const uint32_t tiles = a.tiles[item];
        if (tiles == 0) continue;
        float gr[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) gr[q] = 0.0f;
        const float* src = a.partials + (size_t)a.offsets[item] * 9;
        for (uint32_t t = 0; t < tiles; ++t)
#pragma unroll
            for (int q = 0; q < 9; ++q) gr[q] += src[t * 9 + q];
        const EntityMats& m = mats[vw.mat_base + e];
        accumulate_param_grads(rec, m.vm, m.pvm, vw.cam, sc.sh_degree, gr, gm, gls, gq, gop, gsh);
    }
    const BatchGradsDev& gd = grads[si];
    store_param_grads(g, a.accumulate != 0, gm, gls, gq, gop, gsh, gd.means, gd.log_scales, gd.quats,
                      gd.opacity_logits, gd.sh);
}

void launch_batch_preprocess_bwd(const BatchPreBwdArgs& a, hipStream_t s) {
    if (a.nblocks == 0) return;
    hipLaunchKernelGGL(k_batch_preprocess_bwd, dim3(a.nblocks), dim3(kPreBlock), 0, s, a, a.scenes, a.views,
                       a.scene_views, a.block_scene, a.mats, a.grads);
}

}  // namespace tpd

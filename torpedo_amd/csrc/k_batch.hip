// Batched rendering: many views of many scenes, one launch per stage.  The reference renders one
// camera per GaussianEngine::rasterFrame (GaussianEngine.h:26) and has no batched form; each view
// of a batch is rendered exactly as the plain frame renders it (same per-Gaussian projection,
// same per-view key order, same per-tile blend), only the index spaces are concatenated (the
// layout is described with BatchArgs).
//
//   k_batch_mats            entity matrices of every (view, entity)
//   k_batch_items           each block reads its 256 Gaussian records once and, for every view of
//                           its scene, projects them and writes per item the splat, the tile
//                           rectangle and a depth key; clears the tile ranges
//   (item sort)             the items by depth key (32-bit radix passes)
//   k_batch_emit            the pairs in depth order: key = global tile, value = item, positions
//                           by a decoupled look-back scan over the sorted items; total P and the
//                           overflow flag
//   (pair sort, ranges)     the pairs by global tile alone, then the plain path's ranges
//   k_batch_blend           one 16x16 tile of some view per block (blend_tile)
//   k_batch_blend_bwd       blend_bwd_tile per tile
//   k_batch_preprocess_bwd  per Gaussian, the chain rule summed over the views of its scene, into
//                           the scene's gradient buffers
#include "tpd_blend.hpp"
#include "tpd_grad.hpp"
#include "tpd_kernels.hpp"

#include <type_traits>

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

__device__ __forceinline__ bool project_in_view(const float* rec, const Cov3& cov, const EntityMats& m,
                                                const BatchViewDev& vw, uint32_t sh_degree, Projected& p) {
    return project_gaussian_cov(rec, [&cov] { return cov; }, m.vm, m.pvm, vw.cam.view,
                                (0.5f * (float)vw.cam.width) * vw.cam.focal_ndc[0],
                                (0.5f * (float)vw.cam.height) * vw.cam.focal_ndc[1], vw.cam.width, vw.cam.height,
                                vw.gx, vw.gy, sh_degree, p);
}

// ------------------------------------------------------------------ items

// The item kernel's digit histograms of the two sorts that follow it.
struct ItemCountLds {
    uint32_t item[kMaxCountedPasses][kMaxRadix];      // item key (depth key) digits, per pass
    uint32_t pair[kMaxCountedPairPasses][kMaxRadix];  // pair key (global tile) digits, per pass
    uint32_t pair_diff[kMaxRadix];  // pass 0: difference array (modulo 2^32) of the digit counts
    uint32_t pair_all;              // pass 0: added to every digit
    uint32_t wave_sum[kPreBlock / kWave];
};

// Counts the pair keys of a tile rectangle: tiles tile0 + r * gx + c, r < h, c < w (w, h >= 1).
// Row r's w tiles are consecutive, so their pass-0 digits are the circular run of w digits from
// d0 = first tile mod R: +1 at d0 and -1 at (d0 + w) mod R in the difference array, whose prefix
// sums then fall one short on every digit for each time the run passes digit R - 1, i.e.
// (d0 + w) >> bits times; that number goes into all0 (the thread's share of pair_all).  A higher
// pass adds the whole rectangle to one digit when its first and last tile share that digit's
// prefix, else each row's segments.
__device__ __forceinline__ void count_pair_digits(ItemCountLds& L, uint32_t& all0, uint32_t tile0, uint32_t gx,
                                                  uint32_t w, uint32_t h, uint32_t bits, uint32_t passes) {
    const uint32_t dmask = (1u << bits) - 1u;
    const uint32_t last = tile0 + (h - 1u) * gx + (w - 1u);
    for (uint32_t q = 1; q < passes; ++q) {
        const uint32_t sh = q * bits;
        if ((tile0 >> sh) == (last >> sh)) {
            atomicAdd(&L.pair[q][(tile0 >> sh) & dmask], w * h);
            continue;
        }
        for (uint32_t r = 0; r < h; ++r) {
            uint32_t lo = tile0 + r * gx;
            const uint32_t hi = lo + (w - 1u);
            for (uint32_t seg = lo >> sh; seg <= (hi >> sh); ++seg) {
                const uint32_t end = min(hi, ((seg + 1u) << sh) - 1u);
                atomicAdd(&L.pair[q][seg & dmask], end - lo + 1u);
                lo = end + 1u;
            }
        }
    }
    uint32_t t0 = tile0;
    for (uint32_t r = 0; r < h; ++r, t0 += gx) {
        const uint32_t d0 = t0 & dmask;
        atomicAdd(&L.pair_diff[d0], 1u);
        atomicSub(&L.pair_diff[(d0 + w) & dmask], 1u);
        all0 += (d0 + w) >> bits;
    }
}

// One Gaussian's items, for every view of its scene (the body of k_batch_items); with L, the
// digits of each item key (and, with a.pair_counts, of each visible item's pair keys) are counted
// in the block's LDS histograms.  The record's 3D covariance is computed once for all views.
// ENT: the scene has entities (a per-Gaussian entity index); without, every Gaussian of the block
// uses the view's first entity matrices, a block-uniform address (scalar loads).
template <bool ENT>
__device__ __forceinline__ void items_of_gaussian(const BatchArgs& a, const BatchSceneDev& sc,
                                                  const BatchViewDev* __restrict__ views,
                                                  const uint32_t* __restrict__ scene_views,
                                                  const EntityMats* __restrict__ mats, uint32_t g,
                                                  ItemCountLds* L, uint32_t& all0) {
    const uint32_t e = ENT ? sc.ent_idx[g] : 0u;
    const uint32_t culled_key = a.lo_bits >= 32 ? 0xFFFFFFFFu : (1u << a.lo_bits) - 1u;

    // The record is read once and kept in registers for every view.
    float rec[60];
    const float4* r4 = reinterpret_cast<const float4*>(sc.gauss + (size_t)g * 60);
#pragma unroll
    for (int q = 0; q < 15; ++q) {
        const float4 t = r4[q];
        rec[4 * q + 0] = t.x; rec[4 * q + 1] = t.y; rec[4 * q + 2] = t.z; rec[4 * q + 3] = t.w;
    }
    bool depth_out = false;  // a visible depth outside the sort's depth window
    const OpacityTerms ot = opacity_terms(rec[3]);  // the same for every view
    const Cov3 cov = covariance3d(rec);
    for (uint32_t k = 0; k < sc.nviews; ++k) {
        const BatchViewDev& vw = views[scene_views[sc.view_begin + k]];
        const uint32_t item = vw.item_base + g;
        Projected p;
        const bool vis = project_in_view(rec, cov, mats[vw.mat_base + (ENT ? e : 0u)], vw, sc.sh_degree, p);
        uint64_t key = culled_key;
        if (vis) {
            store_splat(a.splats + item, true, p, ot);
            const uint32_t tile0 = vw.tile_base + (uint32_t)p.y0 * vw.gx + (uint32_t)p.x0;
            const uint32_t w = (uint32_t)(p.x1 - p.x0), h = (uint32_t)(p.y1 - p.y0);
            const uint32_t r32 = inline_rect(tile0, vw.gx, w, h);
            if (r32 == kRectLookup) a.rects[item] = pack_emit_rect(tile0, vw.gx, w, h);
            key = ((uint64_t)r32 << 32) | ((as_uint(p.vz) - a.lo_base) & culled_key);
            depth_out |= depth_outside_window(p.vz);
            if (L && a.pair_counts) count_pair_digits(*L, all0, tile0, vw.gx, w, h, a.pair_bits, a.pair_passes);
        }
        a.item_keys[item] = key;
        if (a.item_vals) a.item_vals[item] = item;
        if (a.tiles) a.tiles[item] = p.tiles;
        if (L && vis) {  // the item sort drops the culled items in its first pass
            const uint32_t k32 = (uint32_t)key;  // the depth key: the bits the item sort orders by
            const uint32_t dmask = (1u << a.hist_bits) - 1u;
            for (uint32_t q = 0; q < a.hist_passes; ++q)
                atomicAdd(&L->item[q][(k32 >> (q * a.hist_bits)) & dmask], 1u);
        }
    }
    if (depth_out) atomicOr(&a.counters[kCtrDepthOut], 1u);
}

// Per (Gaussian, view) item: project, write the splat record of a visible item, its emission
// record (and tile count for the backward) and its depth key and index for the item sort.
// Each block reads its 256 Gaussian records once, for all views of their scene.  With
// a.hist_counts it also counts the item keys' digits for the item sort in LDS and adds its counts
// into the sort's digit counts (the sort's own histogram pass, a second read of every key, is then
// left out), and clears the first pass's look-back descriptors and the tickets; with
// a.pair_counts likewise the digits of the pair keys the emission will write (every tile of each
// visible item's rectangle), for the pair sort.  The block also clears its share of the tile
// ranges, and block 0 the pair sort's count, which the emission sets.
__global__ __launch_bounds__(kPreBlock) void k_batch_items(const BatchArgs a, const BatchSceneDev* __restrict__ scenes,
                                                           const BatchViewDev* __restrict__ views,
                                                           const uint32_t* __restrict__ scene_views,
                                                           const uint32_t* __restrict__ block_scene,
                                                           const EntityMats* __restrict__ mats) {
    __shared__ ItemCountLds L;
    const uint32_t tid = threadIdx.x, lane = __lane_id(), wave = tid / kWave;
    for (uint32_t t = blockIdx.x * kPreBlock + tid; t < a.ntiles; t += gridDim.x * kPreBlock)
        a.ranges[t] = make_uint2(0u, 0u);
    if (blockIdx.x == 0 && tid == 0) a.counters[kCtrSortPairs] = 0u;
    const uint32_t b = blockIdx.x;
    if (b >= a.nblocks) return;  // the one block of a batch without Gaussians only clears
    const bool counting = a.hist_counts != nullptr;  // block-uniform
    const bool pairs = counting && a.pair_counts != nullptr;
    if (counting) {
        for (uint32_t x = blockIdx.x * kPreBlock + tid; x < a.hist_desc0_words; x += gridDim.x * kPreBlock)
            a.hist_desc0[x] = 0u;
        if (b == 0 && tid < kMaxPasses) a.hist_tickets[tid] = 0u;
        for (uint32_t x = tid; x < kMaxCountedPasses * kMaxRadix; x += kPreBlock) (&L.item[0][0])[x] = 0u;
        if (pairs) {
            for (uint32_t x = tid; x < kMaxCountedPairPasses * kMaxRadix; x += kPreBlock) (&L.pair[0][0])[x] = 0u;
            for (uint32_t x = tid; x < kMaxRadix; x += kPreBlock) L.pair_diff[x] = 0u;
            if (tid == 0) L.pair_all = 0u;
        }
        __syncthreads();
    }
    const BatchSceneDev& sc = scenes[block_scene[b]];
    const uint32_t g = (b - sc.first_block) * kPreBlock + tid;
    uint32_t all0 = 0;
    if (g < sc.n) {
        if (sc.ent_idx)
            items_of_gaussian<true>(a, sc, views, scene_views, mats, g, counting ? &L : nullptr, all0);
        else
            items_of_gaussian<false>(a, sc, views, scene_views, mats, g, counting ? &L : nullptr, all0);
    }
    if (!counting) return;
    if (pairs) {
        const uint32_t s_all = wave_incl_scan(all0);
        if (lane == kWave - 1 && s_all) atomicAdd(&L.pair_all, s_all);
    }
    __syncthreads();
    const uint32_t R = 1u << a.hist_bits;
    uint32_t* counts = a.hist_counts + (b % kCountCopies) * (kMaxPasses * kMaxRadix);
    for (uint32_t x = tid; x < a.hist_passes * R; x += kPreBlock) {
        const uint32_t v = L.item[x / R][x % R];
        if (v) atomicAdd(&counts[x], v);  // x = pass * R + digit, the layout of the sort's counts
    }
    if (!pairs) return;
    // pass 0 of the pair keys: pair_all plus the inclusive prefix sum of the difference array
    // (modulo 2^32, where its negative entries live); each thread takes R / 256 consecutive
    // digits (R = 256 or 512)
    const uint32_t PR = 1u << a.pair_bits, per = PR / kPreBlock;
    uint32_t dv[2] = {0u, 0u}, local = 0;
    for (uint32_t i = 0; i < per; ++i) {
        dv[i] = L.pair_diff[tid * per + i];
        local += dv[i];
    }
    const uint32_t incl = wave_incl_scan(local);
    if (lane == kWave - 1) L.wave_sum[wave] = incl;
    __syncthreads();
    uint32_t run = incl - local + L.pair_all;
    for (uint32_t w = 0; w < wave; ++w) run += L.wave_sum[w];
    for (uint32_t i = 0; i < per; ++i) {
        run += dv[i];
        L.pair[0][tid * per + i] = run;
    }
    __syncthreads();
    uint32_t* pcounts = a.pair_counts + (b % kCountCopies) * (kMaxPasses * kMaxRadix);
    for (uint32_t x = tid; x < a.pair_passes * PR; x += kPreBlock) {
        const uint32_t v = L.pair[x / PR][x % PR];
        if (v) atomicAdd(&pcounts[x], v);
    }
}

void launch_batch_items(const BatchArgs& a, hipStream_t s) {
    const uint32_t nb = a.nblocks ? a.nblocks : 1u;  // one block clears the ranges of an empty batch
    hipLaunchKernelGGL(k_batch_items, dim3(nb), dim3(kPreBlock), 0, s, a, a.scenes, a.views, a.scene_views,
                       a.block_scene, a.mats);
}

// ------------------------------------------------------------------ pair emission in depth order

// Blocks take partitions of kEmitItems depth-sorted items (kEmitPerThread consecutive ones per
// thread, so that a thread has that many independent record loads in flight) from an atomic
// ticket; item i's pairs go to base + the exclusive scan of the tile counts over the sorted items
// (decoupled look-back over the partitions, as the plain preprocess), key = global tile, value =
// item.  Items with at most kOwnTiles tiles write their own pairs; the larger ones are spread over
// the block (pair j of them belongs to the item whose inclusive count of large tiles first exceeds
// j).  The last partition stores the pair total and the overflow flag.
constexpr uint32_t kEmitItems = kEmitBlockItems;
constexpr uint32_t kEmitThreads = 512;
constexpr uint32_t kEmitPerThread = kEmitItems / kEmitThreads;
static_assert(kEmitPerThread % 4 == 0, "items are loaded four at a time");

// Pairs of the block's small rectangles staged per chunk before they are written out: the large
// items' tables (start, rectangle; their values are re-read from the sorted items) are done with
// by then and lend their space.
// 2 x 4,096 - 512 pairs: with the large items' tables (48 KB) in the same space, the kernel's LDS
// stays under 80 KB, two blocks of 512 threads per CU.
constexpr uint32_t kEmitStage = 2 * kEmitItems - kEmitThreads;
constexpr uint32_t kEmitMaxChunks = 4;
constexpr uint32_t kNoPair = 0xFFFFFFFFu;  // an empty staging slot (no item has this index)
union EmitLds {
    struct {
        uint32_t start[kEmitItems];
        uint64_t rect[kEmitItems];
    } big;
    struct {
        uint32_t key[kEmitStage], val[kEmitStage];
    } stage;
};

__global__ __launch_bounds__(kEmitThreads) void k_batch_emit(const BatchArgs a) {
    __shared__ uint32_t s_part;
    __shared__ uint32_t s_wave_sum[kEmitThreads / kWave], s_wave_big[kEmitThreads / kWave];
    __shared__ uint32_t s_big_incl[kEmitItems];
    __shared__ EmitLds L;
    __shared__ uint64_t s_base;
    uint32_t* const s_start = L.big.start;
    uint64_t* const s_rect = L.big.rect;

    const uint32_t tid = threadIdx.x, wave = tid / kWave, lane = __lane_id();
    // the pair sort's counts come from the item kernel: its first pass's descriptors (which the
    // item sort's last pass may have used) and the tickets are cleared here
    for (uint32_t x = blockIdx.x * kEmitThreads + tid; x < a.pair_desc0_words; x += gridDim.x * kEmitThreads)
        a.hist_desc0[x] = 0u;
    if (a.pair_desc0_words && blockIdx.x == 0 && tid < kMaxPasses) a.hist_tickets[tid] = 0u;
    // The sorted items are the visible ones (the item sort dropped the culled).  The grid is sized
    // for all items; the blocks past the visible items' partitions leave before taking a ticket
    // (the tickets, one atomic each on one address, pace the blocks' starts), so that exactly the
    // first nparts blocks take tickets 0 .. nparts - 1.
    const uint32_t nitems =
        __builtin_amdgcn_readfirstlane(a.kept_items ? min(*a.kept_items, a.nitems) : a.nitems);
    const uint32_t nparts = (nitems + kEmitItems - 1) / kEmitItems;
    if (blockIdx.x >= nparts) {
        if (blockIdx.x == 0 && tid == 0) {  // no visible item: the totals the last partition would store
            a.counters[kCtrPairs] = 0u;
            a.counters[kCtrOverflow] = 0u;
            a.counters[kCtrSortPairs] = 0u;
        }
        return;
    }
    if (tid == 0) s_part = atomicAdd(&a.counters[kCtrScanTicket], 1u);
    __syncthreads();
    const uint32_t part = __builtin_amdgcn_readfirstlane(s_part);
    if (part >= nparts) return;  // never: the first nparts blocks share nparts tickets
    const uint32_t i0 = part * kEmitItems + tid * kEmitPerThread;

    uint32_t item[kEmitPerThread];
    uint64_t rect[kEmitPerThread];
    if (i0 + kEmitPerThread <= nitems) {
#pragma unroll
        for (uint32_t j = 0; j < kEmitPerThread; j += 4) {  // i0 is a multiple of 4
            const uint4 v = *reinterpret_cast<const uint4*>(a.sorted_items + i0 + j);
            item[j] = v.x; item[j + 1] = v.y; item[j + 2] = v.z; item[j + 3] = v.w;
            const ulonglong2* k2 = reinterpret_cast<const ulonglong2*>(a.sorted_keys64 + i0 + j);
            const ulonglong2 k01 = k2[0], k23 = k2[1];
            rect[j] = k01.x; rect[j + 1] = k01.y; rect[j + 2] = k23.x; rect[j + 3] = k23.y;
        }
    } else {
#pragma unroll
        for (uint32_t j = 0; j < kEmitPerThread; ++j) {
            item[j] = i0 + j < nitems ? a.sorted_items[i0 + j] : 0u;
            rect[j] = i0 + j < nitems ? a.sorted_keys64[i0 + j] : 0ull;
        }
    }
#pragma unroll
    for (uint32_t j = 0; j < kEmitPerThread; ++j) rect[j] = rect_of((uint32_t)(rect[j] >> 32), a.rects, item[j]);
    uint32_t t[kEmitPerThread], mine = 0, mine_big = 0;
#pragma unroll
    for (uint32_t j = 0; j < kEmitPerThread; ++j) {
        t[j] = emit_rect_w(rect[j]) * emit_rect_h(rect[j]);
        mine += t[j];
        mine_big += t[j] > kOwnTiles ? t[j] : 0u;
    }
    const uint32_t incl_w = wave_incl_scan(mine), big_w = wave_incl_scan(mine_big);
    if (lane == kWave - 1) {
        s_wave_sum[wave] = incl_w;
        s_wave_big[wave] = big_w;
    }
    __syncthreads();
    uint32_t wave_off = 0, block_total = 0, big_off = 0, big_total = 0;
#pragma unroll
    for (uint32_t w = 0; w < kEmitThreads / kWave; ++w) {
        const uint32_t v = s_wave_sum[w], vb = s_wave_big[w];
        wave_off += (w < wave) ? v : 0u;
        block_total += v;
        big_off += (w < wave) ? vb : 0u;
        big_total += vb;
    }
    // this thread's items: their block-local starts, and their inclusive counts of large tiles
    uint32_t start = wave_off + incl_w - mine, big_incl = big_off + big_w - mine_big;
    uint32_t first[kEmitPerThread];
#pragma unroll
    for (uint32_t j = 0; j < kEmitPerThread; ++j) {
        const uint32_t o = tid * kEmitPerThread + j;
        first[j] = start;
        if (t[j] > kOwnTiles) {
            big_incl += t[j];
            s_start[o] = start;
            s_rect[o] = rect[j];
        }
        s_big_incl[o] = big_incl;
        start += t[j];
    }
    if (wave == 0) {
        if (lane == 0) store_relaxed(&a.scan_desc[part], (part == 0 ? kDescP : kDescA) | (uint64_t)block_total);
        uint64_t excl = 0;
        if (part > 0) excl = scan_lookback(a.scan_desc, part);
        if (lane == 0) {
            if (part > 0) store_relaxed(&a.scan_desc[part], kDescP | (excl + block_total));
            s_base = excl;
            if (part == nparts - 1) {
                const uint64_t total = excl + block_total;
                a.counters[kCtrPairs] = total > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)total;
                a.counters[kCtrOverflow] = total > a.cap ? 1u : 0u;
                a.counters[kCtrSortPairs] = total > a.cap ? 0u : (uint32_t)total;
            }
        }
    }
    __syncthreads();
    const uint64_t base = s_base;
    // (a large item's value is read back from the sorted items, which this block has just read)
    emit_large_rects(s_big_incl, s_start, s_rect, a.sorted_items + part * kEmitItems, kEmitItems, big_total, base,
                     a.cap, a.keys, a.vals);
    // The small rectangles' pairs (at most kOwnTiles per item, written by the item's thread) go
    // through LDS, chunk by chunk of the block's output range, so that the stores to HBM are
    // consecutive across lanes; slots of the large items' pairs stay empty and are skipped.  A
    // block whose large items make its range longer than kEmitMaxChunks chunks writes them directly.
    if (block_total > kEmitMaxChunks * kEmitStage) {
#pragma unroll
        for (uint32_t j = 0; j < kEmitPerThread; ++j)
            if (t[j] && t[j] <= kOwnTiles) emit_own_rect(t[j], rect[j], base + first[j], a.cap, a.keys, a.vals, item[j]);
        return;
    }
    for (uint32_t c0 = 0; c0 < block_total; c0 += kEmitStage) {
        const uint32_t n = min(block_total - c0, kEmitStage);
        __syncthreads();  // the large items' tables, or the previous chunk, are done with
        for (uint32_t p = tid; p < n; p += kEmitThreads) L.stage.val[p] = kNoPair;
        __syncthreads();
#pragma unroll
        for (uint32_t j = 0; j < kEmitPerThread; ++j)
            if (t[j] && t[j] <= kOwnTiles) stage_own_rect(t[j], rect[j], first[j], c0, n, L.stage.key, L.stage.val, item[j]);
        __syncthreads();
        for (uint32_t p = tid; p < n; p += kEmitThreads) {
            const uint32_t v = L.stage.val[p];
            const uint64_t g = base + c0 + p;
            if (v != kNoPair && g < a.cap) {
                a.keys[g] = L.stage.key[p];
                a.vals[g] = v;
            }
        }
    }
}

// For frames kept for a backward: the exclusive scan of the items' tile counts in item order,
// where the backward keeps each item's per-pair partial sums (consecutive Gaussians of a view then
// have adjacent partials, which the parameter-gradient pass reads item by item).  kEmitPerThread
// items per thread, decoupled look-back over the second half of the scan descriptors.
__global__ __launch_bounds__(kEmitThreads) void k_batch_item_offsets(const BatchArgs a) {
    __shared__ uint32_t s_part;
    __shared__ uint32_t s_wave_sum[kEmitThreads / kWave];
    __shared__ uint64_t s_base;
    const uint32_t tid = threadIdx.x, wave = tid / kWave, lane = __lane_id();
    if (tid == 0) s_part = atomicAdd(&a.counters[kCtrScanTicket2], 1u);
    __syncthreads();
    const uint32_t part = __builtin_amdgcn_readfirstlane(s_part);
    const uint32_t nparts = (a.nitems + kEmitItems - 1) / kEmitItems;
    if (part >= nparts) return;
    uint64_t* desc = a.scan_desc + nparts;
    const uint32_t i0 = part * kEmitItems + tid * kEmitPerThread;
    uint32_t t[kEmitPerThread], mine = 0;
    if (i0 + kEmitPerThread <= a.nitems) {
#pragma unroll
        for (uint32_t j = 0; j < kEmitPerThread; j += 4) {  // i0 is a multiple of 4
            const uint4 v = *reinterpret_cast<const uint4*>(a.tiles + i0 + j);
            t[j] = v.x; t[j + 1] = v.y; t[j + 2] = v.z; t[j + 3] = v.w;
        }
    } else {
#pragma unroll
        for (uint32_t j = 0; j < kEmitPerThread; ++j) t[j] = i0 + j < a.nitems ? a.tiles[i0 + j] : 0u;
    }
#pragma unroll
    for (uint32_t j = 0; j < kEmitPerThread; ++j) mine += t[j];
    const uint32_t incl_w = wave_incl_scan(mine);
    if (lane == kWave - 1) s_wave_sum[wave] = incl_w;
    __syncthreads();
    uint32_t wave_off = 0, block_total = 0;
#pragma unroll
    for (uint32_t w = 0; w < kEmitThreads / kWave; ++w) {
        const uint32_t v = s_wave_sum[w];
        wave_off += (w < wave) ? v : 0u;
        block_total += v;
    }
    if (wave == 0) {
        if (lane == 0) store_relaxed(&desc[part], (part == 0 ? kDescP : kDescA) | (uint64_t)block_total);
        uint64_t excl = 0;
        if (part > 0) excl = scan_lookback(desc, part);
        if (lane == 0) {
            if (part > 0) store_relaxed(&desc[part], kDescP | (excl + block_total));
            s_base = excl;
        }
    }
    __syncthreads();
    uint32_t off = (uint32_t)(s_base + wave_off + incl_w - mine);
    uint32_t o[kEmitPerThread];
#pragma unroll
    for (uint32_t j = 0; j < kEmitPerThread; ++j) {
        o[j] = off;
        off += t[j];
    }
    if (i0 + kEmitPerThread <= a.nitems) {
#pragma unroll
        for (uint32_t j = 0; j < kEmitPerThread; j += 4)
            *reinterpret_cast<uint4*>(a.offsets + i0 + j) = make_uint4(o[j], o[j + 1], o[j + 2], o[j + 3]);
    } else {
#pragma unroll
        for (uint32_t j = 0; j < kEmitPerThread; ++j)
            if (i0 + j < a.nitems) a.offsets[i0 + j] = o[j];
    }
}

void launch_batch_item_offsets(const BatchArgs& a, hipStream_t s) {
    const uint32_t nparts = (a.nitems + kEmitItems - 1) / kEmitItems;
    if (nparts == 0 || !a.offsets || !a.tiles) return;
    hipLaunchKernelGGL(k_batch_item_offsets, dim3(nparts), dim3(kEmitThreads), 0, s, a);
}

void launch_batch_emit(const BatchArgs& a, hipStream_t s) {
    const uint32_t nparts = (a.nitems + kEmitItems - 1) / kEmitItems;
    if (nparts == 0) return;
    hipLaunchKernelGGL(k_batch_emit, dim3(nparts), dim3(kEmitThreads), 0, s, a);
}

// ------------------------------------------------------------------ blend, forward and backward

__device__ __forceinline__ TileView batch_tile_view(const BatchViewDev* __restrict__ views,
                                                    const uint32_t* __restrict__ tile_view, uint32_t tile) {
    const BatchViewDev& vw = views[tile_view[tile]];
    const uint32_t local = tile - vw.tile_base;
    return TileView{local % vw.gx, local / vw.gx, vw.cam.width, vw.cam.height, vw.gx, vw.gy, vw.pix_base};
}

// COUNT: also adds the evaluated pixel-splat pairs to a.eval_count (a separate launch for the
// FLOP count, so that the frame's own blend is timed without the counting).
template <bool COUNT>
__global__ __launch_bounds__(kTilePixels) void k_batch_blend(const BatchBlendArgs a,
                                                             const BatchViewDev* __restrict__ views,
                                                             const uint32_t* __restrict__ tile_view) {
    __shared__ BlendLds lds;
    if (a.counters[kCtrOverflow]) return;  // the frame is rendered again with larger buffers
    const uint32_t tile = a.tile_order ? a.tile_order[blockIdx.x] : xcd_remap(blockIdx.x, gridDim.x);
    const TileView tv = batch_tile_view(views, tile_view, tile);
    blend_tile<COUNT>(lds, tv, a.ranges[tile], a.vals, a.splats, a.image, a.rgba8, a.eval_count, a.final_T,
                      a.final_last);
}

void launch_batch_blend(const BatchBlendArgs& a, hipStream_t s) {
    if (a.ntiles == 0) return;
    if (a.eval_count)
        hipLaunchKernelGGL(k_batch_blend<true>, dim3(a.ntiles), dim3(kTilePixels), 0, s, a, a.views, a.tile_view);
    else
        hipLaunchKernelGGL(k_batch_blend<false>, dim3(a.ntiles), dim3(kTilePixels), 0, s, a, a.views, a.tile_view);
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
    // The block's records (consecutive Gaussians, 240 B each) are read from HBM once, as consecutive
    // 16-byte loads of the whole block (one thread's record at a time would make every load touch
    // 64 cache lines), and kept in LDS for the views of their scene: read again per view, the
    // records of the resident blocks (about 6 MB per XCD) outgrow the XCD's L2.  A stride of 61
    // floats keeps the lanes' reads of one field on distinct banks.  The same LDS later holds the
    // block's SH gradients (48 floats per Gaussian), stored to HBM the same way.
    constexpr uint32_t kRecStride = 61;
    __shared__ alignas(16) float s_rec[kPreBlock * kRecStride];
    if (a.counters[kCtrOverflow]) return;  // gradients are left untouched; the step is run again
    const uint32_t b = blockIdx.x;
    const uint32_t si = block_scene[b];
    const BatchSceneDev& sc = scenes[si];
    const uint32_t g0 = (b - sc.first_block) * kPreBlock;
    const uint32_t g = g0 + threadIdx.x;
    const uint32_t nvalid = min(kPreBlock, sc.n - g0);  // >= 1: the block exists for these Gaussians
    const bool valid = threadIdx.x < nvalid;           // every thread stays for the barriers
    {
        const float4* src = reinterpret_cast<const float4*>(sc.gauss + (size_t)g0 * 60);
        for (uint32_t x = threadIdx.x; x < nvalid * 15u; x += kPreBlock) {
            const float4 v = src[x];
            float* d = s_rec + (x / 15u) * kRecStride + 4u * (x % 15u);
            d[0] = v.x; d[1] = v.y; d[2] = v.z; d[3] = v.w;
        }
    }
    __syncthreads();
    const float* rec = s_rec + threadIdx.x * kRecStride;
    const uint32_t e = (valid && sc.ent_idx) ? sc.ent_idx[g] : 0u;
    // The same FP32 operations as a plain backward per view accumulated into the gradient buffers:
    // every log-scale, quaternion, opacity and SH entry receives one term per view, added to the
    // running sum directly; a mean receives three terms per view, so they are summed on their own
    // first and then added.
    float gm[3] = {0.f, 0.f, 0.f}, gls[3] = {0.f, 0.f, 0.f}, gq[4] = {0.f, 0.f, 0.f, 0.f}, gop = 0.f;
    float gsh[48];
#pragma unroll
    for (int k = 0; k < 48; ++k) gsh[k] = 0.f;
    // ent: the scene has entities; without, every Gaussian uses the view's first entity matrices,
    // a block-uniform address (scalar loads)
    const auto views_of = [&](auto ent) {
        for (uint32_t k = 0; k < sc.nviews; ++k) {
            const BatchViewDev& vw = views[scene_views[sc.view_begin + k]];
            const uint32_t item = vw.item_base + g;
            const uint32_t tiles = a.tiles[item];
            if (tiles == 0) continue;
            float gr[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) gr[q] = 0.0f;
            sum_pair_partials(a.partials + (size_t)a.offsets[item] * 9, tiles, gr);
            const EntityMats& m = mats[vw.mat_base + (decltype(ent)::value ? e : 0u)];
            float vm[3] = {0.f, 0.f, 0.f};
            accumulate_param_grads(rec, m.vm, m.pvm, vw.cam, sc.sh_degree, gr, vm, gls, gq, gop, gsh);
#pragma unroll
            for (int c = 0; c < 3; ++c) gm[c] += vm[c];
        }
    };
    if (valid) {
        if (sc.ent_idx)
            views_of(std::true_type{});
        else
            views_of(std::false_type{});
    }
    const BatchGradsDev& gd = grads[si];
    if (valid) store_param_grads_but_sh(g, a.accumulate != 0, gm, gls, gq, gop, gd.means, gd.log_scales, gd.quats,
                                        gd.opacity_logits);
    __syncthreads();  // the records are done with: the LDS takes the SH gradients
    if (valid) {
        float4* d = reinterpret_cast<float4*>(s_rec + threadIdx.x * 48u);
#pragma unroll
        for (int k = 0; k < 12; ++k) d[k] = make_float4(gsh[4 * k], gsh[4 * k + 1], gsh[4 * k + 2], gsh[4 * k + 3]);
    }
    __syncthreads();
    store_sh_grads_block(reinterpret_cast<const float4*>(s_rec), nvalid, a.accumulate != 0,
                         reinterpret_cast<float4*>(gd.sh + (size_t)g0 * 48));
}

void launch_batch_preprocess_bwd(const BatchPreBwdArgs& a, hipStream_t s) {
    if (a.nblocks == 0) return;
    hipLaunchKernelGGL(k_batch_preprocess_bwd, dim3(a.nblocks), dim3(kPreBlock), 0, s, a, a.scenes, a.views,
                       a.scene_views, a.block_scene, a.mats, a.grads);
}

}  // namespace tpd

// Batched rendering: scenes owned by the context, a view table, and the frame recorded once into
// a HIP graph and replayed.  The reference renders one camera of one scene per rasterFrame
// (GaussianEngine.h:26, one Scene per compile, GaussianEngine.cpp:359-399); a batch renders every
// view as the plain frame would, in one launch per stage.
//
// A frame's pair count is known only on the device, so the graph sizes the sort and the pair
// buffers by a capacity; the emission raises an overflow flag when the capacity is too small,
// every later kernel then skips its work, and the validation step (tpd_batch_synchronize, or any
// read) grows the buffers, re-captures and renders again.
#include "tpd_ctx.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace tpd {

struct HostScene {
    float* d_gauss = nullptr;
    bool own = true;
    uint32_t* d_ent_idx = nullptr;
    float* d_models = nullptr;
    uint32_t n = 0, sh_degree = 3, n_ent = 1;
    std::vector<float> h_models;  // n_ent x 16 (empty: identity)
};

struct BatchState {
    std::vector<HostScene> scenes;
    std::vector<tpd_view> views;
    bool layout_dirty = true;  // scenes, view sizes or view scenes changed
    bool tables_dirty = true;  // view/scene tables need an upload (cameras, transforms)

    // layout
    std::vector<BatchSceneDev> h_scenes;
    std::vector<BatchViewDev> h_views;
    std::vector<uint32_t> h_scene_views, h_block_scene, h_tile_view;
    uint32_t nblocks = 0, nitems = 0, ntiles = 0, nmats = 0, max_ent = 1;
    uint64_t npix = 0;
    uint32_t item_shape = 0;  // radix block shape of the item sort (ws.shape: that of the pair sort)

    // device tables
    BatchSceneDev* d_scenes = nullptr; size_t cap_scenes = 0;
    BatchViewDev* d_views = nullptr; size_t cap_views = 0;
    uint32_t* d_scene_views = nullptr; size_t cap_scene_views = 0;
    uint32_t* d_block_scene = nullptr; size_t cap_block_scene = 0;
    uint32_t* d_tile_view = nullptr; size_t cap_tile_view = 0;
    EntityMats* d_mats = nullptr; size_t cap_mats = 0;
    uint64_t* d_scan_desc = nullptr; size_t cap_scan_desc = 0;  // the emission's look-back, per partition
    uint32_t* d_counters = nullptr;
    uint32_t* h_counters = nullptr;  // pinned, written by the frame's last copy

    // per item / tile / pixel
    Splat* d_splats = nullptr; size_t cap_splats = 0;
    uint32_t* d_tiles = nullptr; size_t cap_tiles = 0;
    uint32_t* d_offsets = nullptr; size_t cap_offsets = 0;
    uint64_t* d_rects = nullptr; size_t cap_rects = 0;  // per item emission record
    uint64_t* d_ikeys = nullptr; size_t cap_ikeys = 0;  // item sort: depth keys (+ emission records) and
    uint32_t* d_ivals = nullptr; size_t cap_ivals = 0;  // items, ping-pong
    uint64_t* d_ikeys_alt = nullptr; size_t cap_ikeys_alt = 0;
    uint32_t* d_ivals_alt = nullptr; size_t cap_ivals_alt = 0;
    uint2* d_ranges = nullptr; size_t cap_ranges = 0;
    uint32_t* d_tile_order = nullptr; size_t cap_tile_order = 0;
    float* d_image = nullptr; size_t cap_image = 0;
    // per pixel final transmittance and last blended splat, written by frames kept for a backward
    float* d_final_T = nullptr; size_t cap_final_T = 0;
    int32_t* d_final_last = nullptr; size_t cap_final_last = 0;
    uint8_t* d_rgba8 = nullptr; size_t cap_rgba8 = 0;

    // pairs: global tile keys and item values, ping-pong
    uint32_t* d_keys = nullptr; size_t cap_keys = 0;
    uint32_t* d_vals = nullptr; size_t cap_vals = 0;
    uint32_t* d_keys_alt = nullptr; size_t cap_keys_alt = 0;
    uint32_t* d_vals_alt = nullptr; size_t cap_vals_alt = 0;
    uint64_t* d_full_keys = nullptr; size_t cap_full_keys = 0;  // 64-bit keys rebuilt for readers
    uint64_t* d_eval_count = nullptr;  // tpd_batch_get_blend_evaluations
    uint32_t cap = 0;  // pair capacity used by the recorded frame
    SortWorkspace ws;  // shared by the item sort and the pair sort (they run one after the other)
    const uint32_t* sorted_items = nullptr;
    const uint32_t* sorted_keys = nullptr;
    const uint32_t* sorted_vals = nullptr;
    // Depth key of the recorded frame's item sort: the depth window unless a depth of this layout
    // fell outside it (then the positive-float form until the layout changes).
    bool depth_wide = false;
    uint32_t lo_bits() const { return depth_wide ? kDepthWideBits : kDepthWindowBits; }
    uint32_t lo_base() const { return depth_wide ? 0u : kDepthBase; }

    // backward
    float* d_partials = nullptr; size_t cap_partials = 0;
    BatchGradsDev* d_grads = nullptr; size_t cap_grads = 0;
    std::vector<BatchGradsDev> h_grads;

    // graphs: [0] forward, [1] forward keeping the backward's per-item state, [2] forward + backward
    bool use_graph = true;
    hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
    const float* step_dL = nullptr;
    uint32_t step_accumulate = 0;

    // frame state
    bool pending = false;       // launched, not validated
    int pending_mode = 0;
    const float* pending_dL = nullptr;
    uint32_t pending_accumulate = 0;
    hipStream_t pending_stream = nullptr;
    bool rendered = false;      // last frame validated
    bool kept_bwd_state = false;
    uint32_t last_pairs = 0;
    bool items_dropped = false;  // the recorded frame's item sort drops the culled items
    uint32_t last_kept = 0;      // the last frame's items kept by its item sort (the visible ones)
    hipEvent_t ev[11] = {};  // eager stage events [0..6], frame bracket [7], [8], after each sort's histogram [9], [10]
    bool timed = false;
    uint32_t item_passes = 0, pair_passes = 0;  // radix passes of the recorded frame's two sorts
};

static void destroy_graphs(BatchState* b) {
    for (auto& x : b->exec)
        if (x) {
            (void)hipGraphExecDestroy(x);
            x = nullptr;
        }
}

void batch_release(tpd_ctx* c) {
    BatchState* b = c->batch;
    if (!b) return;
    destroy_graphs(b);
    for (auto& s : b->scenes) {
        if (s.own) release(s.d_gauss);
        release(s.d_ent_idx);
        release(s.d_models);
    }
    release(b->d_scenes); release(b->d_views); release(b->d_scene_views); release(b->d_block_scene);
    release(b->d_tile_view); release(b->d_mats); release(b->d_scan_desc);
    release(b->d_counters);
    if (b->h_counters) (void)hipHostFree(b->h_counters);
    release(b->d_splats); release(b->d_tiles); release(b->d_offsets); release(b->d_rects);
    release(b->d_ikeys); release(b->d_ivals); release(b->d_ikeys_alt); release(b->d_ivals_alt);
    release(b->d_full_keys); release(b->d_eval_count);
    release(b->d_ranges); release(b->d_tile_order); release(b->d_image); release(b->d_rgba8);
    release(b->d_final_T); release(b->d_final_last);
    release(b->d_keys); release(b->d_vals); release(b->d_keys_alt); release(b->d_vals_alt);
    release(b->ws.desc); release(b->ws.partials); release(b->ws.offsets); release(b->ws.tickets);
    release(b->d_partials); release(b->d_grads);
    for (auto& ev : b->ev)
        if (ev) (void)hipEventDestroy(ev);
    delete b;
    c->batch = nullptr;
}

static BatchState* state(tpd_ctx* c) {
    if (!c->batch) {
        c->batch = new BatchState();
        c->batch->ws.shape = c->ws.shape;  // the context's radix block shape until tuned separately
        c->batch->item_shape = c->ws.shape;
    }
    return c->batch;
}


static int validate(tpd_ctx* c, BatchState* b);

// Finishes a launched frame before host-side table or buffer changes, or before another frame is
// launched over it: waits for it and, when it overflowed its pair buffers, grows them and renders
// it again, so that an asynchronous step's images and gradients are complete even when the caller
// never synchronised it.
static int finish_pending(tpd_ctx* c, BatchState* b) {
    if (!b->pending) return TPD_OK;
    return validate(c, b);
}

void batch_set_sort_shape(tpd_ctx* c, uint32_t shape) {
    if (!c->batch) return;
    (void)finish_pending(c, c->batch);
    c->batch->ws.shape = shape;
    c->batch->item_shape = shape;
    destroy_graphs(c->batch);
}

// ------------------------------------------------------------------ layout

static int build_layout(tpd_ctx* c, BatchState* b) {
    const uint32_t ns = (uint32_t)b->scenes.size(), nv = (uint32_t)b->views.size();
    std::vector<std::vector<uint32_t>> per_scene(ns);
    for (uint32_t v = 0; v < nv; ++v) per_scene[b->views[v].scene].push_back(v);
    b->h_scenes.assign(ns, BatchSceneDev{});
    b->h_scene_views.clear();
    b->h_block_scene.clear();
    uint64_t blocks = 0;
    b->max_ent = 1;
    for (uint32_t s = 0; s < ns; ++s) {
        const HostScene& hs = b->scenes[s];
        BatchSceneDev& d = b->h_scenes[s];
        d.gauss = hs.d_gauss;
        d.ent_idx = hs.d_ent_idx;
        d.models = hs.d_models;
        d.n = hs.n;
        d.sh_degree = hs.sh_degree;
        d.n_ent = hs.n_ent;
        d.first_block = (uint32_t)blocks;
        d.view_begin = (uint32_t)b->h_scene_views.size();
        d.nviews = (uint32_t)per_scene[s].size();
        const uint32_t nb = (hs.n + kPreBlock - 1) / kPreBlock;
        for (uint32_t i = 0; i < nb; ++i) b->h_block_scene.push_back(s);
        b->h_scene_views.insert(b->h_scene_views.end(), per_scene[s].begin(), per_scene[s].end());
        blocks += nb;
        b->max_ent = std::max(b->max_ent, std::max<uint32_t>(hs.n_ent, 1u));
    }
    b->h_views.assign(nv, BatchViewDev{});
    b->h_tile_view.clear();
    uint64_t tiles = 0, items = 0, mats = 0, pix = 0;
    for (uint32_t v = 0; v < nv; ++v) {
        const tpd_view& u = b->views[v];
        BatchViewDev& d = b->h_views[v];
        std::memcpy(d.cam.view, u.view, sizeof(d.cam.view));
        std::memcpy(d.cam.proj_view, u.proj_view, sizeof(d.cam.proj_view));
        d.cam.focal_ndc[0] = u.focal_ndc[0];
        d.cam.focal_ndc[1] = u.focal_ndc[1];
        d.cam.width = u.width;
        d.cam.height = u.height;
        d.scene = u.scene;
        d.gx = (u.width + kTile - 1) / kTile;
        d.gy = (u.height + kTile - 1) / kTile;
        d.tile_base = (uint32_t)tiles;
        d.item_base = (uint32_t)items;
        d.mat_base = (uint32_t)mats;
        d.pix_base = pix;
        const uint32_t nt = d.gx * d.gy;
        b->h_tile_view.insert(b->h_tile_view.end(), nt, v);
        tiles += nt;
        items += b->scenes[u.scene].n;
        mats += std::max<uint32_t>(b->scenes[u.scene].n_ent, 1u);
        pix += (uint64_t)u.width * u.height;
    }
    if (blocks > 0xFFFFFFFFull || items > (uint64_t)kMaxSortKeys || tiles >= (1ull << 31) || mats > 0xFFFFFFFFull)
        return fail(TPD_ERR_CAPACITY, "tpd_batch: batch too large (items beyond 2^30 - 1, tiles or blocks beyond 32 bits)");
    b->nblocks = (uint32_t)blocks;
    b->nitems = (uint32_t)items;
    b->ntiles = (uint32_t)tiles;
    b->nmats = (uint32_t)mats;
    b->npix = pix;

    hipError_t e;
    const size_t it = std::max<size_t>(b->nitems, 1);
    if ((e = grow(b->d_scenes, b->cap_scenes, std::max<size_t>(ns, 1))) != hipSuccess ||
        (e = grow(b->d_views, b->cap_views, std::max<size_t>(nv, 1))) != hipSuccess ||
        (e = grow(b->d_scene_views, b->cap_scene_views, std::max<size_t>(nv, 1))) != hipSuccess ||
        (e = grow(b->d_block_scene, b->cap_block_scene, std::max<size_t>(blocks, 1))) != hipSuccess ||
        (e = grow(b->d_tile_view, b->cap_tile_view, std::max<size_t>(tiles, 1))) != hipSuccess ||
        (e = grow(b->d_mats, b->cap_mats, std::max<size_t>(mats, 1))) != hipSuccess ||
        (e = grow(b->d_splats, b->cap_splats, it)) != hipSuccess || (e = grow(b->d_tiles, b->cap_tiles, it)) != hipSuccess ||
        (e = grow(b->d_offsets, b->cap_offsets, it)) != hipSuccess ||
        (e = grow(b->d_rects, b->cap_rects, it)) != hipSuccess ||
        (e = grow(b->d_ikeys, b->cap_ikeys, it)) != hipSuccess || (e = grow(b->d_ivals, b->cap_ivals, it)) != hipSuccess ||
        (e = grow(b->d_ikeys_alt, b->cap_ikeys_alt, it)) != hipSuccess ||
        (e = grow(b->d_ivals_alt, b->cap_ivals_alt, it)) != hipSuccess ||
        (e = grow(b->d_ranges, b->cap_ranges, std::max<size_t>(tiles, 1))) != hipSuccess ||
        (e = grow(b->d_tile_order, b->cap_tile_order, std::max<size_t>(tile_order_words(tiles), 1))) != hipSuccess ||
        (e = grow(b->d_image, b->cap_image, std::max<size_t>(3 * pix, 1))) != hipSuccess ||
        (e = grow(b->d_final_T, b->cap_final_T, std::max<size_t>(pix, 1))) != hipSuccess ||
        (e = grow(b->d_final_last, b->cap_final_last, std::max<size_t>(pix, 1))) != hipSuccess ||
        (e = grow(b->d_rgba8, b->cap_rgba8, std::max<size_t>(4 * pix, 1))) != hipSuccess)
        return hip_fail(e, "tpd_batch: alloc");
    // look-back descriptors: the emission's, then the item offsets' (frames kept for a backward)
    const size_t nparts = 2 * std::max<size_t>((items + kEmitBlockItems - 1) / kEmitBlockItems, 1);
    if (!b->d_scan_desc || b->cap_scan_desc < nparts) {
        if ((e = grow(b->d_scan_desc, b->cap_scan_desc, nparts)) != hipSuccess) return hip_fail(e, "tpd_batch: alloc");
    }
    // look-back descriptors and counters start at zero; afterwards every frame leaves them so
    if ((e = hipMemset(b->d_scan_desc, 0, b->cap_scan_desc * sizeof(uint64_t))) != hipSuccess)
        return hip_fail(e, "tpd_batch: clear");
    if (!b->d_counters) {
        if ((e = hipMalloc(&b->d_counters, sizeof(uint32_t) * kCtrCount)) != hipSuccess)
            return hip_fail(e, "tpd_batch: alloc");
        if ((e = hipHostMalloc(reinterpret_cast<void**>(&b->h_counters), sizeof(uint32_t) * kCtrCount,
                               hipHostMallocDefault)) != hipSuccess)
            return hip_fail(e, "tpd_batch: alloc");
    }
    if ((e = hipMemset(b->d_counters, 0, sizeof(uint32_t) * kCtrCount)) != hipSuccess)
        return hip_fail(e, "tpd_batch: clear");
    if (nv) {
        e = hipMemcpy(b->d_scene_views, b->h_scene_views.data(), nv * sizeof(uint32_t), hipMemcpyHostToDevice);
        if (e == hipSuccess && tiles)
            e = hipMemcpy(b->d_tile_view, b->h_tile_view.data(), tiles * sizeof(uint32_t), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && blocks)
        e = hipMemcpy(b->d_block_scene, b->h_block_scene.data(), blocks * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "tpd_batch: upload layout");
    destroy_graphs(b);
    b->cap = 0;  // the next frame sizes its pair buffers first
    b->depth_wide = false;  // and sorts on the depth window until a depth falls outside it
    b->rendered = false;
    b->layout_dirty = false;
    b->tables_dirty = true;
    (void)c;
    return TPD_OK;
}

static int upload_tables(BatchState* b) {
    hipError_t e = hipSuccess;
    if (!b->h_scenes.empty())
        e = hipMemcpy(b->d_scenes, b->h_scenes.data(), b->h_scenes.size() * sizeof(BatchSceneDev), hipMemcpyHostToDevice);
    if (e == hipSuccess && !b->h_views.empty())
        e = hipMemcpy(b->d_views, b->h_views.data(), b->h_views.size() * sizeof(BatchViewDev), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "tpd_batch: upload tables");
    b->tables_dirty = false;
    return TPD_OK;
}

// The sort workspace of both sorts: look-back descriptors for the larger of the pair capacity
// and the item count.
static int ensure_sort_ws(BatchState* b) {
    hipError_t e;
    if (!b->ws.partials) {
        if ((e = hipMalloc(&b->ws.partials, sizeof(uint32_t) * kCountCopies * kMaxPasses * kMaxRadix)) != hipSuccess ||
            (e = hipMemset(b->ws.partials, 0, sizeof(uint32_t) * kCountCopies * kMaxPasses * kMaxRadix)) != hipSuccess ||
            (e = hipMalloc(&b->ws.offsets, sizeof(uint32_t) * kMaxPasses * kMaxRadix)) != hipSuccess ||
            (e = hipMalloc(&b->ws.tickets, sizeof(uint32_t) * kMaxPasses)) != hipSuccess)
            return hip_fail(e, "tpd_batch: alloc sort workspace");
    }
    const uint32_t parts = sort_parts(std::max(b->cap, b->nitems));
    if (b->ws.parts_cap < parts) {
        release(b->ws.desc);
        if ((e = hipMalloc(&b->ws.desc, sizeof(uint32_t) * 2 * (size_t)parts * kMaxRadix)) != hipSuccess)
            return hip_fail(e, "tpd_batch: alloc sort workspace");
        b->ws.parts_cap = parts;
        destroy_graphs(b);
    }
    return TPD_OK;
}

static int ensure_pairs(BatchState* b, size_t pairs) {
    pairs = std::max<size_t>(pairs, 1);
    if (b->cap_keys >= pairs && b->cap_vals >= pairs && b->cap_keys_alt >= pairs && b->cap_vals_alt >= pairs) {
        // The buffers fit already (a new layout after a larger one): only the frame's capacity,
        // which a new layout sets to zero, is restored.
        const uint32_t cap = (uint32_t)std::min(
            {b->cap_keys, b->cap_vals, b->cap_keys_alt, b->cap_vals_alt, (size_t)kMaxSortKeys});
        if (b->cap != cap) {
            destroy_graphs(b);
            b->cap = cap;
        }
        return ensure_sort_ws(b);
    }
    const double h = 1.15;
    hipError_t e;
    if ((e = grow(b->d_keys, b->cap_keys, pairs, h)) != hipSuccess ||
        (e = grow(b->d_vals, b->cap_vals, pairs, h)) != hipSuccess ||
        (e = grow(b->d_keys_alt, b->cap_keys_alt, pairs, h)) != hipSuccess ||
        (e = grow(b->d_vals_alt, b->cap_vals_alt, pairs, h)) != hipSuccess)
        return hip_fail(e, "tpd_batch: alloc pairs");
    destroy_graphs(b);
    b->cap = (uint32_t)std::min({b->cap_keys, b->cap_vals, b->cap_keys_alt, b->cap_vals_alt, (size_t)kMaxSortKeys});
    return ensure_sort_ws(b);
}

static int ensure_partials(BatchState* b, size_t pairs) {
    const float* before = b->d_partials;
    hipError_t e = grow(b->d_partials, b->cap_partials, 9 * std::max<size_t>(pairs, 1));
    if (b->d_partials != before) destroy_graphs(b);  // recorded frames hold the old address
    return e == hipSuccess ? TPD_OK : hip_fail(e, "tpd_batch: alloc backward buffers");
}

// ------------------------------------------------------------------ the recorded frame

static uint32_t tile_key_bits(const BatchState* b) { return bit_length(b->ntiles > 1 ? b->ntiles - 1 : 1); }

static BatchArgs forward_args(BatchState* b, bool keep_bwd) {
    BatchArgs a{};
    a.scenes = b->d_scenes;
    a.views = b->d_views;
    a.scene_views = b->d_scene_views;
    a.block_scene = b->d_block_scene;
    a.nblocks = b->nblocks;
    a.nviews = (uint32_t)b->views.size();
    a.nitems = b->nitems;
    a.mats = b->d_mats;
    a.scan_desc = b->d_scan_desc;
    a.counters = b->d_counters;
    a.cap = b->cap;
    a.splats = b->d_splats;
    a.tiles = keep_bwd ? b->d_tiles : nullptr;
    a.offsets = keep_bwd ? b->d_offsets : nullptr;
    a.rects = b->d_rects;
    a.item_keys = b->d_ikeys;
    a.item_vals = b->d_ivals;
    a.lo_bits = b->lo_bits();
    a.lo_base = b->lo_base();
    // the item sort's digit counts come from the item kernel when its passes fit the LDS histograms
    const uint32_t ipasses = sort_shape_passes(b->item_shape, a.lo_bits);
    if (b->nitems && ipasses <= kMaxCountedPasses) {
        const uint32_t ibits = sort_shape_bits(b->item_shape);
        const uint32_t iparts = (b->nitems + sort_shape_keys(b->item_shape) - 1) / sort_shape_keys(b->item_shape);
        a.hist_counts = b->ws.partials;
        a.hist_desc0 = b->ws.desc;
        a.hist_tickets = b->ws.tickets;
        a.hist_bits = ibits;
        a.hist_passes = ipasses;
        a.hist_desc0_words = iparts << ibits;
        // and the pair sort's from the items' tile rectangles, when its passes fit too
        const uint32_t pshape = b->ws.shape, pbits = sort_shape_bits(pshape);
        const uint32_t ppasses = sort_shape_passes(pshape, tile_key_bits(b));
        if (b->cap && ppasses >= 1 && ppasses <= kMaxCountedPairPasses) {
            const uint32_t pparts = (b->cap + sort_shape_keys(pshape) - 1) / sort_shape_keys(pshape);
            a.pair_counts = b->ws.partials + kPairCountRow;
            a.pair_bits = pbits;
            a.pair_passes = ppasses;
            a.pair_desc0_words = pparts << pbits;
        }
    }
    // with its counts from the item kernel the item sort drops the culled items and its first pass
    // takes each item's index as its value: the item kernel writes no values
    if (a.hist_counts) a.item_vals = nullptr;
    a.keys = b->d_keys;
    a.vals = b->d_vals;
    a.ranges = b->d_ranges;
    a.ntiles = b->ntiles;
    return a;
}


static BatchBlendArgs blend_args(BatchState* b) {
    BatchBlendArgs a{};
    a.views = b->d_views;
    a.tile_view = b->d_tile_view;
    a.tile_order = b->d_tile_order;
    a.ranges = b->d_ranges;
    a.vals = b->sorted_vals;
    a.splats = b->d_splats;
    a.ntiles = b->ntiles;
    a.counters = b->d_counters;
    a.image = b->d_image;
    a.rgba8 = b->d_rgba8;
    a.offsets = b->d_offsets;
    a.partials = b->d_partials;
    a.final_T = b->d_final_T;
    a.final_last = b->d_final_last;
    return a;
}

// Issues the forward frame on s (no host synchronisation, no allocation, no memset: capturable).
// With profiling on, stage events ev[0..6] are recorded (eager frames only): items, item sort,
// emission, pair sort, ranges + tile order, blend.
static void record_forward(BatchState* b, bool keep_bwd, hipStream_t s, bool events) {
    if (events) hipEventRecord(b->ev[0], s);
    launch_batch_mats(b->d_scenes, b->d_views, (uint32_t)b->views.size(), b->max_ent, b->d_mats, b->d_counters, s);
    BatchArgs a = forward_args(b, keep_bwd);
    launch_batch_items(a, s);
    launch_batch_item_offsets(a, s);  // frames kept for a backward only (a.offsets set)
    if (events) hipEventRecord(b->ev[1], s);
    b->item_passes = b->nitems ? sort_shape_passes(b->item_shape, b->lo_bits()) : 0;
    bool items_alt = false;
    if (b->nitems) {
        // 64-bit keys ordered by their low lo_bits bits (the depth key); the high word (the item's
        // emission record) rides along.  The item sort has a block shape of its own.
        const uint32_t pair_shape = b->ws.shape;
        b->ws.shape = b->item_shape;
        // with the counts from the item kernel, which counts the visible items only, the first
        // pass drops the culled items (zero high word) and the rest of the frame runs on the
        // visible ones (their number in kCtrItemsKept)
        const bool counted = a.hist_counts != nullptr;
        b->items_dropped = counted;
        items_alt = launch_sort_pairs(b->ws, b->d_ikeys, b->d_ivals, b->d_ikeys_alt, b->d_ivals_alt, nullptr,
                                      b->nitems, b->lo_bits(), s, events ? b->ev[9] : nullptr, 32u, nullptr, 0u,
                                      counted, counted ? &b->d_counters[kCtrItemsKept] : nullptr);
        a.kept_items = counted ? &b->d_counters[kCtrItemsKept] : nullptr;
        b->ws.shape = pair_shape;
    } else if (events)
        hipEventRecord(b->ev[9], s);
    b->sorted_items = items_alt ? b->d_ivals_alt : b->d_ivals;
    if (events) hipEventRecord(b->ev[2], s);
    a.sorted_items = b->sorted_items;
    a.sorted_keys64 = items_alt ? b->d_ikeys_alt : b->d_ikeys;
    launch_batch_emit(a, s);
    if (events) hipEventRecord(b->ev[3], s);
    const uint32_t key_bits = tile_key_bits(b);
    b->pair_passes = sort_shape_passes(b->ws.shape, key_bits);
    // kCtrSortPairs is P, or 0 after an overflow: the counted digits then cover pairs that were
    // never written, and the sort and the ranges must not run on them
    const bool in_alt = launch_sort_pairs32(b->ws, b->d_keys, b->d_vals, b->d_keys_alt, b->d_vals_alt,
                                            &b->d_counters[kCtrSortPairs], b->cap, key_bits, s,
                                            events ? b->ev[10] : nullptr, a.pair_counts != nullptr);
    b->sorted_keys = in_alt ? b->d_keys_alt : b->d_keys;
    b->sorted_vals = in_alt ? b->d_vals_alt : b->d_vals;
    if (events) hipEventRecord(b->ev[4], s);
    // the keys are global tiles (a shift of 0 past key_drop = lo_bits); lo_bits also tells the
    // ranges kernel whether the items were sorted on the depth window
    launch_ranges(b->sorted_keys, b->d_counters, b->cap, b->ntiles, b->d_ranges, b->d_scan_desc,
                  2 * ((b->nitems + kEmitBlockItems - 1) / kEmitBlockItems), &b->d_counters[kCtrScanTicket], s, b->cap,
                  b->lo_bits(),
                  b->lo_bits(), kCtrSortPairs);
    launch_tile_order(b->d_ranges, b->ntiles, b->d_counters, b->d_tile_order, s);
    if (events) hipEventRecord(b->ev[5], s);
    BatchBlendArgs ba = blend_args(b);
    if (!keep_bwd) {  // a forward-only frame stores no per-pixel state for the backward
        ba.final_T = nullptr;
        ba.final_last = nullptr;
    }
    launch_batch_blend(ba, s);
    if (events) hipEventRecord(b->ev[6], s);
    (void)hipMemcpyAsync(b->h_counters, b->d_counters, sizeof(uint32_t) * kCtrCount, hipMemcpyDeviceToHost, s);
}

static void record_backward(BatchState* b, const float* dL, uint32_t accumulate, hipStream_t s) {
    BatchBlendArgs a = blend_args(b);
    a.image = nullptr;
    a.rgba8 = nullptr;
    a.dL_dimg = dL;
    launch_batch_blend_bwd(a, s);
    BatchPreBwdArgs p{};
    p.scenes = b->d_scenes;
    p.views = b->d_views;
    p.scene_views = b->d_scene_views;
    p.block_scene = b->d_block_scene;
    p.nblocks = b->nblocks;
    p.mats = b->d_mats;
    p.tiles = b->d_tiles;
    p.offsets = b->d_offsets;
    p.partials = b->d_partials;
    p.grads = b->d_grads;
    p.counters = b->d_counters;
    p.accumulate = accumulate;
    launch_batch_preprocess_bwd(p, s);
}

// First frame of a layout: pair buffers for twice the items (or 2^20 pairs) to start with; a frame
// with more pairs overflows them and is rendered again in grown buffers (validate).
static int size_pairs(tpd_ctx* c, BatchState* b) {
    (void)c;
    return ensure_pairs(b, std::max<size_t>(2 * (size_t)b->nitems, 1u << 20));
}

static int prepare(tpd_ctx* c, BatchState* b) {
    int rc;
    if (b->scenes.empty() || b->views.empty()) return fail(TPD_ERR_STATE, "tpd_batch: add scenes and set views first");
    if ((rc = finish_pending(c, b)) != TPD_OK) return rc;
    if (b->layout_dirty && (rc = build_layout(c, b)) != TPD_OK) return rc;
    if (b->tables_dirty && (rc = upload_tables(b)) != TPD_OK) return rc;
    if (b->cap == 0 && (rc = size_pairs(c, b)) != TPD_OK) return rc;
    return TPD_OK;
}

// mode 0: forward, 1: forward keeping the backward's state, 2: forward + backward
static int launch_frame(tpd_ctx* c, BatchState* b, int mode, const float* dL, uint32_t accumulate, hipStream_t s) {
    hipError_t e;
    if (mode == 2 && (b->exec[2] && (b->step_dL != dL || b->step_accumulate != accumulate))) {
        (void)hipGraphExecDestroy(b->exec[2]);
        b->exec[2] = nullptr;
    }
    const bool events = c->profile && !b->use_graph;
    if (c->profile)
        for (auto& ev : b->ev)
            if (!ev && (e = hipEventCreate(&ev)) != hipSuccess) return hip_fail(e, "hipEventCreate");
    if (c->profile) hipEventRecord(b->ev[7], s);
    if (b->use_graph) {
        if (!b->exec[mode]) {
            hipGraph_t g = nullptr;
            if ((e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
                return hip_fail(e, "tpd_batch: begin capture");
            record_forward(b, mode != 0, c->stream, false);
            if (mode == 2) record_backward(b, dL, accumulate, c->stream);
            if ((e = hipStreamEndCapture(c->stream, &g)) != hipSuccess) return hip_fail(e, "tpd_batch: end capture");
            e = hipGraphInstantiate(&b->exec[mode], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                b->exec[mode] = nullptr;
                return hip_fail(e, "tpd_batch: instantiate graph");
            }
            if (mode == 2) {
                b->step_dL = dL;
                b->step_accumulate = accumulate;
            }
        }
        if ((e = hipGraphLaunch(b->exec[mode], s)) != hipSuccess) return hip_fail(e, "tpd_batch: graph launch");
    } else {
        record_forward(b, mode != 0, s, events);
        if (mode == 2) record_backward(b, dL, accumulate, s);
    }
    if (c->profile) {
        hipEventRecord(b->ev[8], s);
        b->timed = true;
    }
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "tpd_batch: launch");
    b->pending = true;
    b->pending_mode = mode;
    b->pending_dL = dL;
    b->pending_accumulate = accumulate;
    b->pending_stream = s;
    b->rendered = false;
    return TPD_OK;
}

// Waits for the launched frame; on a capacity overflow grows the pair buffers and renders again.
static int validate(tpd_ctx* c, BatchState* b) {
    if (!b->pending) return b->rendered ? TPD_OK : fail(TPD_ERR_STATE, "tpd_batch: render first");
    for (int attempt = 0;; ++attempt) {
        hipError_t e = hipStreamSynchronize(b->pending_stream);
        if (e != hipSuccess) return hip_fail(e, "tpd_batch: frame");
        b->pending = false;
        const uint32_t P = b->h_counters[kCtrPairs];
        if (P > kMaxSortKeys) return fail(TPD_ERR_CAPACITY, "tpd_batch: more than 2^30 - 1 tile/Gaussian pairs");
        if (!b->h_counters[kCtrOverflow]) {
            b->last_pairs = P;
            b->last_kept = b->items_dropped ? b->h_counters[kCtrItemsKept] : b->nitems;
            b->rendered = true;
            b->kept_bwd_state = b->pending_mode != 0;
            return TPD_OK;
        }
        if (attempt > 2) return fail(TPD_ERR_STATE, "tpd_batch: pair buffers did not converge");
        // An overflowed frame either had more pairs than the buffers hold or (sorted on the depth
        // window) a depth outside the window; either way nothing after the sort ran.
        const bool depth_out = !b->depth_wide && b->h_counters[kCtrDepthOut] != 0;
        if (depth_out) {
            b->depth_wide = true;
            destroy_graphs(b);
        }
        int rc = TPD_OK;
        if (!depth_out || P > b->cap) {
            rc = ensure_pairs(b, (size_t)P + P / 8);
            if (rc == TPD_OK && b->pending_mode == 2) rc = ensure_partials(b, b->cap);
        }
        if (rc != TPD_OK) return rc;
        if ((rc = launch_frame(c, b, b->pending_mode, b->pending_dL, b->pending_accumulate, b->pending_stream)) != TPD_OK)
            return rc;
    }
}

static int upload_grads(BatchState* b, const tpd_scene_grads* grads) {
    const size_t ns = b->scenes.size();
    std::vector<BatchGradsDev> g(ns);
    for (size_t s = 0; s < ns; ++s) {
        const tpd_scene_grads& u = grads[s];
        if (b->scenes[s].n && (!u.means || !u.log_scales || !u.quats || !u.opacity_logits || !u.sh))
            return fail(TPD_ERR_INVALID, "tpd_batch: null gradient buffer");
        g[s] = BatchGradsDev{u.means, u.log_scales, u.quats, u.opacity_logits, u.sh};
    }
    const BatchGradsDev* before = b->d_grads;
    hipError_t e = grow(b->d_grads, b->cap_grads, std::max<size_t>(ns, 1));
    if (b->d_grads != before) {
        destroy_graphs(b);
        b->h_grads.clear();
    }
    if (e == hipSuccess && (b->h_grads.size() != ns ||
                            std::memcmp(b->h_grads.data(), g.data(), ns * sizeof(BatchGradsDev)) != 0)) {
        e = hipMemcpy(b->d_grads, g.data(), ns * sizeof(BatchGradsDev), hipMemcpyHostToDevice);
        if (e == hipSuccess) b->h_grads = g;
    }
    return e == hipSuccess ? TPD_OK : hip_fail(e, "tpd_batch: gradient table");
}

static int add_scene(tpd_ctx* c, HostScene&& hs, uint32_t* scene_id) {
    BatchState* b = state(c);
    int rc = finish_pending(c, b);
    if (rc != TPD_OK) return rc;
    b->scenes.push_back(std::move(hs));
    if (scene_id) *scene_id = (uint32_t)(b->scenes.size() - 1);
    b->layout_dirty = true;
    return TPD_OK;
}

// The RGBA8 image of one view of the last batched frame, for the presentation calls: device
// pointer (valid until the next frame) and size.  The frame has finished when this returns.
int batch_view_rgba8(tpd_ctx* c, uint32_t view, const uint8_t** d_rgba, uint32_t* width, uint32_t* height) {
    BatchState* b = state(c);
    int rc = validate(c, b);
    if (rc != TPD_OK) return rc;
    if (view >= b->h_views.size()) return fail(TPD_ERR_INVALID, "batch presentation: no such view");
    const BatchViewDev& v = b->h_views[view];
    *d_rgba = b->d_rgba8 + 4 * v.pix_base;
    *width = v.cam.width;
    *height = v.cam.height;
    return TPD_OK;
}

}  // namespace tpd

using namespace tpd;

extern "C" {

int tpd_scene_add(tpd_ctx* c, const void* gaussians, uint32_t n, const uint32_t* entity_idx, const float* models,
                  uint32_t n_ent, uint32_t sh_degree, uint32_t* scene_id) {
    if (!c) return fail(TPD_ERR_INVALID, "null context");
    if (sh_degree > 3) return fail(TPD_ERR_INVALID, "tpd_scene_add: sh_degree must be 0..3");
    if (n && !gaussians) return fail(TPD_ERR_INVALID, "tpd_scene_add: null data");
    if (models && n_ent == 0) return fail(TPD_ERR_INVALID, "tpd_scene_add: models without entities");
    if (entity_idx && !models) return fail(TPD_ERR_INVALID, "tpd_scene_add: entity indices need models");
    const uint32_t ne = models ? n_ent : 1u;
    if (entity_idx)
        for (uint32_t i = 0; i < n; ++i)
            if (entity_idx[i] >= ne) return fail(TPD_ERR_INVALID, "tpd_scene_add: entity index out of range");
    hipSetDevice(c->device);
    HostScene hs;
    hs.n = n;
    hs.sh_degree = sh_degree;
    hs.n_ent = ne;
    hipError_t e = hipSuccess;
    if (n) {
        e = hipMalloc(&hs.d_gauss, sizeof(float) * 60 * (size_t)n);
        if (e == hipSuccess) e = hipMemcpy(hs.d_gauss, gaussians, sizeof(float) * 60 * (size_t)n, hipMemcpyHostToDevice);
        if (e == hipSuccess && entity_idx) {
            e = hipMalloc(&hs.d_ent_idx, sizeof(uint32_t) * (size_t)n);
            if (e == hipSuccess)
                e = hipMemcpy(hs.d_ent_idx, entity_idx, sizeof(uint32_t) * (size_t)n, hipMemcpyHostToDevice);
        }
    }
    if (e == hipSuccess && models) {
        hs.h_models.assign(models, models + 16 * (size_t)ne);
        e = hipMalloc(&hs.d_models, sizeof(float) * 16 * (size_t)ne);
        if (e == hipSuccess) e = hipMemcpy(hs.d_models, models, sizeof(float) * 16 * (size_t)ne, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        release(hs.d_gauss);
        release(hs.d_ent_idx);
        release(hs.d_models);
        return hip_fail(e, "tpd_scene_add");
    }
    return add_scene(c, std::move(hs), scene_id);
}

int tpd_scene_add_device(tpd_ctx* c, const float* d_gaussians, uint32_t n, uint32_t sh_degree, uint32_t* scene_id) {
    if (!c) return fail(TPD_ERR_INVALID, "null context");
    if (sh_degree > 3) return fail(TPD_ERR_INVALID, "tpd_scene_add_device: sh_degree must be 0..3");
    if (n && !d_gaussians) return fail(TPD_ERR_INVALID, "tpd_scene_add_device: null data");
    if (reinterpret_cast<uintptr_t>(d_gaussians) % 16 != 0)
        return fail(TPD_ERR_INVALID, "tpd_scene_add_device: records must be 16-byte aligned");
    HostScene hs;
    hs.d_gauss = const_cast<float*>(d_gaussians);
    hs.own = false;
    hs.n = n;
    hs.sh_degree = sh_degree;
    return add_scene(c, std::move(hs), scene_id);
}

int tpd_scene_set_transform(tpd_ctx* c, uint32_t scene, uint32_t entity, const float model[16]) {
    if (!c || !model) return fail(TPD_ERR_INVALID, "tpd_scene_set_transform: bad argument");
    BatchState* b = state(c);
    if (scene >= b->scenes.size()) return fail(TPD_ERR_INVALID, "tpd_scene_set_transform: no such scene");
    HostScene& hs = b->scenes[scene];
    if (!hs.d_models || entity >= hs.n_ent)
        return fail(TPD_ERR_INVALID, "tpd_scene_set_transform: no such entity (scene added without models)");
    hipSetDevice(c->device);
    int rc = finish_pending(c, b);
    if (rc != TPD_OK) return rc;
    std::memcpy(&hs.h_models[16 * (size_t)entity], model, 16 * sizeof(float));
    hipError_t e = hipMemcpy(hs.d_models + 16 * (size_t)entity, model, 16 * sizeof(float), hipMemcpyHostToDevice);
    return e == hipSuccess ? TPD_OK : hip_fail(e, "tpd_scene_set_transform");
}

int tpd_scene_clear(tpd_ctx* c) {
    if (!c) return fail(TPD_ERR_INVALID, "null context");
    hipSetDevice(c->device);
    batch_release(c);
    return TPD_OK;
}

int tpd_batch_set_views(tpd_ctx* c, const tpd_view* views, uint32_t nviews) {
    if (!c || (nviews && !views)) return fail(TPD_ERR_INVALID, "tpd_batch_set_views: bad argument");
    BatchState* b = state(c);
    for (uint32_t v = 0; v < nviews; ++v) {
        if (views[v].scene >= b->scenes.size()) return fail(TPD_ERR_INVALID, "tpd_batch_set_views: no such scene");
        if (views[v].width == 0 || views[v].height == 0 || views[v].width > 16384 || views[v].height > 16384)
            return fail(TPD_ERR_INVALID, "tpd_batch_set_views: view size must be 1..16384");
    }
    hipSetDevice(c->device);
    int rc = finish_pending(c, b);
    if (rc != TPD_OK) return rc;
    bool same_layout = !b->layout_dirty && b->views.size() == nviews;
    for (uint32_t v = 0; same_layout && v < nviews; ++v)
        same_layout = b->views[v].scene == views[v].scene && b->views[v].width == views[v].width &&
                      b->views[v].height == views[v].height;
    b->views.assign(views, views + nviews);
    if (same_layout) {
        // cameras only: the recorded frame reads the view table from device memory
        for (uint32_t v = 0; v < nviews; ++v) {
            BatchViewDev& d = b->h_views[v];
            std::memcpy(d.cam.view, views[v].view, sizeof(d.cam.view));
            std::memcpy(d.cam.proj_view, views[v].proj_view, sizeof(d.cam.proj_view));
            d.cam.focal_ndc[0] = views[v].focal_ndc[0];
            d.cam.focal_ndc[1] = views[v].focal_ndc[1];
        }
        b->tables_dirty = true;
    } else {
        b->layout_dirty = true;
    }
    b->rendered = false;
    return TPD_OK;
}

int tpd_batch_set_graph(tpd_ctx* c, int enabled) {
    if (!c) return fail(TPD_ERR_INVALID, "null context");
    BatchState* b = state(c);
    int rc = finish_pending(c, b);
    if (rc != TPD_OK) return rc;
    b->use_graph = enabled != 0;
    return TPD_OK;
}

int tpd_batch_render(tpd_ctx* c, int for_backward, void* stream) {
    if (!c) return fail(TPD_ERR_INVALID, "null context");
    hipSetDevice(c->device);
    BatchState* b = state(c);
    int rc = prepare(c, b);
    if (rc != TPD_OK) return rc;
    hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : c->stream;
    return launch_frame(c, b, for_backward ? 1 : 0, nullptr, 0, s);
}

int tpd_batch_synchronize(tpd_ctx* c) {
    if (!c) return fail(TPD_ERR_INVALID, "null context");
    hipSetDevice(c->device);
    return validate(c, state(c));
}

int tpd_batch_backward(tpd_ctx* c, const float* d_dL_dimages, const tpd_scene_grads* grads, int accumulate,
                       void* stream) {
    if (!c || !d_dL_dimages || !grads) return fail(TPD_ERR_INVALID, "tpd_batch_backward: bad argument");
    hipSetDevice(c->device);
    BatchState* b = state(c);
    int rc = validate(c, b);
    if (rc != TPD_OK) return rc;
    if (!b->kept_bwd_state) return fail(TPD_ERR_STATE, "tpd_batch_backward: render with for_backward = 1 first");
    if ((rc = upload_grads(b, grads)) != TPD_OK) return rc;
    if ((rc = ensure_partials(b, b->last_pairs)) != TPD_OK) return rc;
    hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : c->stream;
    record_backward(b, d_dL_dimages, accumulate ? 1u : 0u, s);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TPD_OK : hip_fail(e, "tpd_batch_backward");
}

int tpd_batch_step(tpd_ctx* c, const float* d_dL_dimages, const tpd_scene_grads* grads, int accumulate, void* stream) {
    if (!c || !d_dL_dimages || !grads) return fail(TPD_ERR_INVALID, "tpd_batch_step: bad argument");
    hipSetDevice(c->device);
    BatchState* b = state(c);
    int rc = prepare(c, b);
    if (rc != TPD_OK) return rc;
    if ((rc = upload_grads(b, grads)) != TPD_OK) return rc;
    if ((rc = ensure_partials(b, b->cap)) != TPD_OK) return rc;
    hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : c->stream;
    return launch_frame(c, b, 2, d_dL_dimages, accumulate ? 1u : 0u, s);
}

int tpd_batch_get_pair_count(tpd_ctx* c, uint32_t* P) {
    if (!c || !P) return fail(TPD_ERR_INVALID, "tpd_batch_get_pair_count: bad argument");
    hipSetDevice(c->device);
    BatchState* b = state(c);
    int rc = validate(c, b);
    if (rc != TPD_OK) return rc;
    *P = b->last_pairs;
    return TPD_OK;
}

int tpd_batch_view_layout(tpd_ctx* c, uint32_t view, uint32_t* tile_base, uint32_t* item_base, uint64_t* pix_base) {
    if (!c) return fail(TPD_ERR_INVALID, "null context");
    BatchState* b = state(c);
    if (b->layout_dirty) return fail(TPD_ERR_STATE, "tpd_batch_view_layout: render first");
    if (view >= b->h_views.size()) return fail(TPD_ERR_INVALID, "tpd_batch_view_layout: no such view");
    if (tile_base) *tile_base = b->h_views[view].tile_base;
    if (item_base) *item_base = b->h_views[view].item_base;
    if (pix_base) *pix_base = b->h_views[view].pix_base;
    return TPD_OK;
}

int tpd_batch_read_image_f32(tpd_ctx* c, uint32_t view, float* host) {
    if (!c || !host) return fail(TPD_ERR_INVALID, "tpd_batch_read_image_f32: bad argument");
    hipSetDevice(c->device);
    BatchState* b = state(c);
    int rc = validate(c, b);
    if (rc != TPD_OK) return rc;
    if (view >= b->h_views.size()) return fail(TPD_ERR_INVALID, "tpd_batch_read_image_f32: no such view");
    const BatchViewDev& v = b->h_views[view];
    hipError_t e = hipMemcpy(host, b->d_image + 3 * v.pix_base, sizeof(float) * 3 * (size_t)v.cam.width * v.cam.height,
                             hipMemcpyDeviceToHost);
    return e == hipSuccess ? TPD_OK : hip_fail(e, "tpd_batch_read_image_f32");
}

int tpd_batch_read_image_rgba8(tpd_ctx* c, uint32_t view, uint8_t* host) {
    if (!c || !host) return fail(TPD_ERR_INVALID, "tpd_batch_read_image_rgba8: bad argument");
    hipSetDevice(c->device);
    BatchState* b = state(c);
    int rc = validate(c, b);
    if (rc != TPD_OK) return rc;
    if (view >= b->h_views.size()) return fail(TPD_ERR_INVALID, "tpd_batch_read_image_rgba8: no such view");
    const BatchViewDev& v = b->h_views[view];
    hipError_t e = hipMemcpy(host, b->d_rgba8 + 4 * v.pix_base, 4 * (size_t)v.cam.width * v.cam.height,
                             hipMemcpyDeviceToHost);
    return e == hipSuccess ? TPD_OK : hip_fail(e, "tpd_batch_read_image_rgba8");
}

int tpd_batch_device_images(tpd_ctx* c, float** d_images, uint64_t* total_pixels) {
    if (!c || !d_images) return fail(TPD_ERR_INVALID, "tpd_batch_device_images: bad argument");
    hipSetDevice(c->device);
    BatchState* b = state(c);
    int rc = validate(c, b);
    if (rc != TPD_OK) return rc;
    *d_images = b->d_image;
    if (total_pixels) *total_pixels = b->npix;
    return TPD_OK;
}

int tpd_batch_read_pairs(tpd_ctx* c, uint64_t* keys, uint32_t* vals, uint32_t capacity, uint32_t* P) {
    if (!c || !P) return fail(TPD_ERR_INVALID, "tpd_batch_read_pairs: bad argument");
    hipSetDevice(c->device);
    BatchState* b = state(c);
    int rc = validate(c, b);
    if (rc != TPD_OK) return rc;
    *P = b->last_pairs;
    if (!keys && !vals) return TPD_OK;
    if (capacity < b->last_pairs) return fail(TPD_ERR_INVALID, "tpd_batch_read_pairs: capacity too small");
    hipError_t e = hipSuccess;
    if (keys && b->last_pairs) {
        // the sorted keys are global tiles: the 64-bit keys (tile << 32) | depth bits are rebuilt
        // from the tile and the depth of each value's splat
        e = grow(b->d_full_keys, b->cap_full_keys, b->last_pairs);
        if (e == hipSuccess) {
            launch_full_keys(b->sorted_keys, 0u, 0u, b->sorted_vals, b->d_splats, b->last_pairs, b->d_full_keys,
                             c->stream);
            e = hipStreamSynchronize(c->stream);
        }
        if (e == hipSuccess) e = hipMemcpy(keys, b->d_full_keys, sizeof(uint64_t) * b->last_pairs, hipMemcpyDeviceToHost);
    }
    if (e == hipSuccess && vals && b->last_pairs)
        e = hipMemcpy(vals, b->sorted_vals, sizeof(uint32_t) * b->last_pairs, hipMemcpyDeviceToHost);
    return e == hipSuccess ? TPD_OK : hip_fail(e, "tpd_batch_read_pairs");
}

int tpd_batch_get_sort_shapes(tpd_ctx* c, uint32_t* item_shape, uint32_t* pair_shape) {
    if (!c || !item_shape || !pair_shape) return fail(TPD_ERR_INVALID, "tpd_batch_get_sort_shapes: bad argument");
    const BatchState* b = state(c);
    *item_shape = b->item_shape;
    *pair_shape = b->ws.shape;
    return TPD_OK;
}

// Blends the last frame again with the counting kernel (same tiles, splats and order; it writes the
// same image, and not the backward's per-pixel state) and returns the pixel-splat pairs whose power
// the blend evaluated: the FLOP count of the frame's blend.
int tpd_batch_get_blend_evaluations(tpd_ctx* c, uint64_t* evaluations) {
    if (!c || !evaluations) return fail(TPD_ERR_INVALID, "tpd_batch_get_blend_evaluations: bad argument");
    hipSetDevice(c->device);
    BatchState* b = state(c);
    int rc = validate(c, b);
    if (rc != TPD_OK) return rc;
    if (!b->rendered) return fail(TPD_ERR_STATE, "tpd_batch_get_blend_evaluations: no batch frame rendered");
    hipError_t e = hipSuccess;
    if (!b->d_eval_count) e = hipMalloc(&b->d_eval_count, sizeof(uint64_t));
    if (e == hipSuccess) e = hipMemsetAsync(b->d_eval_count, 0, sizeof(uint64_t), c->stream);
    if (e != hipSuccess) return hip_fail(e, "tpd_batch_get_blend_evaluations: counter");
    BatchBlendArgs ba = blend_args(b);
    ba.final_T = nullptr;
    ba.final_last = nullptr;
    ba.eval_count = b->d_eval_count;
    launch_batch_blend(ba, c->stream);
    if ((e = hipGetLastError()) == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(evaluations, b->d_eval_count, sizeof(uint64_t), hipMemcpyDeviceToHost);
    return e == hipSuccess ? TPD_OK : hip_fail(e, "tpd_batch_get_blend_evaluations");
}

// ms[0] = whole frame (first launch to last); for eager profiled forward frames ms[1..6] = items
// (with entity matrices), item sort, pair emission, pair sort, ranges + tile order, blend (else
// -1).  *passes = radix passes per frame (item sort + pair sort).
int tpd_batch_get_times(tpd_ctx* c, float* ms, uint32_t* passes) {
    if (!c || !ms) return fail(TPD_ERR_INVALID, "tpd_batch_get_times: bad argument");
    hipSetDevice(c->device);
    BatchState* b = state(c);
    if (!c->profile || !b->timed) return fail(TPD_ERR_STATE, "tpd_batch_get_times: no profiled batch frame");
    hipError_t e = hipEventSynchronize(b->ev[8]);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms[0], b->ev[7], b->ev[8]);
    for (int i = 1; i <= 6; ++i) ms[i] = -1.0f;
    if (e == hipSuccess && !b->use_graph && b->pending_mode == 0)
        for (int i = 1; i <= 6 && e == hipSuccess; ++i) e = hipEventElapsedTime(&ms[i], b->ev[i - 1], b->ev[i]);
    if (passes) *passes = b->item_passes + b->pair_passes;
    return e == hipSuccess ? TPD_OK : hip_fail(e, "tpd_batch_get_times");
}

// The last eager profiled forward frame's two sorts: ms = {item histogram, item radix passes, pair
// histogram, pair radix passes}; counts = {items, item passes, pairs, pair passes}.
int tpd_batch_get_sort_times(tpd_ctx* c, float* ms, uint32_t* counts) {
    if (!c || !ms) return fail(TPD_ERR_INVALID, "tpd_batch_get_sort_times: bad argument");
    hipSetDevice(c->device);
    BatchState* b = state(c);
    if (!c->profile || !b->timed || b->use_graph || b->pending_mode != 0)
        return fail(TPD_ERR_STATE, "tpd_batch_get_sort_times: no eager profiled forward frame");
    hipError_t e = hipEventSynchronize(b->ev[8]);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms[0], b->ev[1], b->ev[9]);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms[1], b->ev[9], b->ev[2]);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms[2], b->ev[3], b->ev[10]);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms[3], b->ev[10], b->ev[4]);
    if (counts) {
        counts[0] = b->nitems;
        counts[1] = b->item_passes;
        counts[2] = b->last_pairs;
        counts[3] = b->pair_passes;
    }
    return e == hipSuccess ? TPD_OK : hip_fail(e, "tpd_batch_get_sort_times");
}

int tpd_batch_get_item_counts(tpd_ctx* c, uint32_t* items, uint32_t* kept) {
    if (!c || !items || !kept) return fail(TPD_ERR_INVALID, "tpd_batch_get_item_counts: bad argument");
    hipSetDevice(c->device);
    BatchState* b = state(c);
    int rc = validate(c, b);
    if (rc != TPD_OK) return rc;
    *items = b->nitems;
    *kept = b->last_kept;
    return TPD_OK;
}

// Times every radix-pass block shape on the batch's own items and pairs (both sorts of full
// eager frames) and keeps the fastest for each sort; sort_ms[shape] = both sorts' time with that
// shape, *best = the pair sort's shape (tpd_batch_get_sort_shapes gives both).
int tpd_batch_autotune(tpd_ctx* c, float* sort_ms, uint32_t capacity, uint32_t* n_shapes, uint32_t* best) {
    if (!c) return fail(TPD_ERR_INVALID, "null context");
    hipSetDevice(c->device);
    BatchState* b = state(c);
    int rc = prepare(c, b);
    if (rc != TPD_OK) return rc;
    const uint32_t ns = sort_num_shapes();
    if (n_shapes) *n_shapes = ns;
    hipStream_t s = c->stream;
    const bool keep_graph = b->use_graph, keep_profile = c->profile;
    b->use_graph = false;
    c->profile = true;
    std::vector<float> med(ns, 0.0f), med_items(ns, 0.0f), med_pairs(ns, 0.0f);
    for (uint32_t sh = 0; sh < ns && rc == TPD_OK; ++sh) {
        b->ws.shape = sh;
        b->item_shape = sh;
        std::vector<float> t, ti, tp;
        for (int rep = 0; rep < 4 && rc == TPD_OK; ++rep) {
            if ((rc = launch_frame(c, b, 0, nullptr, 0, s)) != TPD_OK) break;
            if ((rc = validate(c, b)) != TPD_OK) break;
            float ms_items = 0.0f, ms_pairs = 0.0f;  // both sorts
            if (hipEventElapsedTime(&ms_items, b->ev[1], b->ev[2]) != hipSuccess ||
                hipEventElapsedTime(&ms_pairs, b->ev[3], b->ev[4]) != hipSuccess) {
                rc = fail(TPD_ERR_HIP, "tpd_batch_autotune: timing");
                break;
            }
            if (rep > 0) {
                t.push_back(ms_items + ms_pairs);
                ti.push_back(ms_items);
                tp.push_back(ms_pairs);
            }
        }
        if (!t.empty()) {
            std::sort(t.begin(), t.end());
            std::sort(ti.begin(), ti.end());
            std::sort(tp.begin(), tp.end());
            med[sh] = t[t.size() / 2];
            med_items[sh] = ti[ti.size() / 2];
            med_pairs[sh] = tp[tp.size() / 2];
        }
    }
    b->use_graph = keep_graph;
    c->profile = keep_profile;
    if (rc != TPD_OK) return rc;
    uint32_t bi = 0, bp = 0;
    for (uint32_t sh = 1; sh < ns; ++sh) {
        if (med_items[sh] < med_items[bi]) bi = sh;
        if (med_pairs[sh] < med_pairs[bp]) bp = sh;
    }
    b->item_shape = bi;
    b->ws.shape = bp;
    destroy_graphs(b);
    if (best) *best = bp;
    if (sort_ms)
        for (uint32_t sh = 0; sh < ns && sh < capacity; ++sh) sort_ms[sh] = med[sh];
    return TPD_OK;
}

}  // extern "C"

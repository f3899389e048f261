// The plain forward frame (GaussianEngine::rasterFrame, GaussianEngine.cpp:621-712).
#include "tpd_ctx.hpp"

#include <algorithm>

namespace tpd {

static int clear_frame_state(tpd_ctx* c, hipStream_t s) {
    const size_t nd = (std::max<uint32_t>(c->n, 1u) + kPreBlock - 1) / kPreBlock;
    hipError_t e = hipMemsetAsync(c->d_scan_desc, 0, nd * sizeof(uint64_t), s);
    if (e == hipSuccess) e = hipMemsetAsync(c->d_counters, 0, sizeof(uint32_t) * kCtrCount, s);
    return e == hipSuccess ? TPD_OK : hip_fail(e, "clear frame state");
}

int render_forward(tpd_ctx* c, hipStream_t s) {
    const uint32_t W = c->cam.width, H = c->cam.height;
    const uint32_t gx = (W + kTile - 1) / kTile, gy = (H + kTile - 1) / kTile;
    const uint32_t ntiles = gx * gy;
    hipError_t e;
    int rc;
    if ((rc = ensure_scene_buffers(c)) != TPD_OK) return rc;
    if ((rc = ensure_image_buffers(c)) != TPD_OK) return rc;

    if (c->n == 0) {
        // Empty scene: black image, empty ranges (GaussianEngine.cpp:362-365 renders nothing).
        e = hipMemsetAsync(c->d_image, 0, sizeof(float) * 3 * (size_t)W * H, s);
        if (e == hipSuccess) e = hipMemsetAsync(c->d_rgba8, 0, 4 * (size_t)W * H, s);
        if (e == hipSuccess) e = hipMemsetAsync(c->d_ranges, 0, sizeof(uint2) * ntiles, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_fail(e, "render empty scene");
        // RGBA8 alpha is 1 for every pixel (blend.slang:104 stores float4(C, 1)).
        std::vector<uint8_t> black(4 * (size_t)W * H, 0);
        for (size_t i = 3; i < black.size(); i += 4) black[i] = 255;
        e = hipMemcpy(c->d_rgba8, black.data(), black.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, "render empty scene");
        c->last_pairs = 0;
        c->sorted_keys = c->d_keys;
        c->sorted_vals = c->d_vals;
        c->key_drop = kKeysWide;
        return TPD_OK;
    }

    if (c->profile) hipEventRecord(c->ev[0], s);
    launch_entity_mats(c->d_models, c->n_ent, c->cam, c->d_ent_mats, s);
    if (c->cap_keys == 0 && (rc = ensure_pair_buffers(c, std::max<size_t>(4 * (size_t)c->n, 1u << 20))) != TPD_OK)
        return rc;

    uint32_t P = 0;
    for (int attempt = 0;; ++attempt) {
        ForwardArgs a{};
        a.gauss = c->d_gauss;
        a.ent_idx = c->d_ent_idx;
        a.ent = c->d_ent_mats;
        a.cam = c->cam;
        a.n = c->n;
        a.sh_degree = c->sh_degree;
        a.gx = gx;
        a.gy = gy;
        a.cap = (uint32_t)std::min<size_t>(c->cap_keys, kMaxSortKeys);
        a.splats = c->d_splats;
        a.tiles = c->d_tiles;
        a.offsets = c->d_offsets;
        a.scan_desc = c->d_scan_desc;
        a.counters = c->d_counters;
        a.keys = c->d_keys;
        a.vals = c->d_vals;
        a.ranges = c->d_ranges;
        a.ntiles = ntiles;
        launch_preprocess_dup(a, s);
        if (c->profile) hipEventRecord(c->ev[1], s);
        // The reference reads tilesRendered back here (GaussianEngine.cpp:659-674) to size its
        // buffers; this read also detects a capacity overflow, which is replayed after growing.
        e = hipMemcpyAsync(c->h_counters, c->d_counters, sizeof(uint32_t) * kCtrCount, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_fail(e, "preprocess");
        P = c->h_counters[kCtrPairs];
        if (P > kMaxSortKeys) {
            clear_frame_state(c, s);
            return fail(TPD_ERR_CAPACITY, "tpd_render: more than 2^30 - 1 tile/Gaussian pairs");
        }
        if (!c->h_counters[kCtrOverflow]) break;
        if ((rc = clear_frame_state(c, s)) != TPD_OK) return rc;
        if ((rc = ensure_pair_buffers(c, P)) != TPD_OK) return rc;
        if (attempt > 2) return fail(TPD_ERR_STATE, "tpd_render: pair buffers did not converge");
    }
    c->last_pairs = P;
    if (c->keep_unsorted && P) {
        if ((rc = ensure_pair_buffers(c, P)) != TPD_OK) return rc;
        e = hipMemcpyAsync(c->d_keys_dbg, c->d_keys, sizeof(uint64_t) * P, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(c->d_vals_dbg, c->d_vals, sizeof(uint32_t) * P, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return hip_fail(e, "copy unsorted pairs");
    }

    bool in_alt = false;
    // Emitted depths are positive floats (view z > 0), so bit 31 of every key's low word is clear
    // and the sort may run on the compacted (tile << 31) | depth key: same order, one bit less;
    // when the preprocess saw no depth outside [2^-8, 2^8), on (tile << 27) | (depth - 2^-8 bits).
    const bool depth_out = c->h_counters[kCtrDepthOut] != 0;
    c->lo_bits = depth_out ? kDepthWideBits : kDepthWindowBits;
    c->lo_base = depth_out ? 0u : kDepthBase;
    if (depth_out && (e = hipMemsetAsync(&c->d_counters[kCtrDepthOut], 0, sizeof(uint32_t), s)) != hipSuccess)
        return hip_fail(e, "clear depth flag");
    const uint32_t key_bits = bit_length(ntiles > 1 ? ntiles - 1 : 1) + c->lo_bits;
    c->last_passes = P > 0 ? sort_shape_passes(c->ws.shape, key_bits) : 0;
    if (c->profile) hipEventRecord(c->ev[2], s);
    if (P > 0)
        in_alt = launch_sort_pairs(c->ws, c->d_keys, c->d_vals, c->d_keys_alt, c->d_vals_alt, nullptr, P, key_bits, s,
                                   c->profile ? c->ev[3] : nullptr, c->lo_bits, &c->key_drop, c->lo_base);
    else
        c->key_drop = kKeysWide;
    if (P == 0 && c->profile) hipEventRecord(c->ev[3], s);
    if (c->profile) hipEventRecord(c->ev[4], s);
    c->sorted_keys = in_alt ? c->d_keys_alt : c->d_keys;
    c->sorted_vals = in_alt ? c->d_vals_alt : c->d_vals;
    const uint32_t nd = (c->n + kPreBlock - 1) / kPreBlock;
    launch_ranges(c->sorted_keys, c->d_counters, P, ntiles, c->d_ranges, c->d_scan_desc, nd,
                  &c->d_counters[kCtrScanTicket], s, 0xFFFFFFFFu, c->key_drop, c->lo_bits);
    launch_tile_order(c->d_ranges, ntiles, nullptr, c->d_tile_order, s);
    if (c->profile) hipEventRecord(c->ev[5], s);
    BlendArgs b{};
    b.ranges = c->d_ranges;
    b.tile_order = c->d_tile_order;
    b.vals = c->sorted_vals;
    b.splats = c->d_splats;
    b.width = W;
    b.height = H;
    b.gx = gx;
    b.ntiles = ntiles;
    b.image = c->d_image;
    b.rgba8 = c->d_rgba8;
    b.eval_count = nullptr;  // counted on request, by blending the frame once more
    b.final_T = c->d_final_T;  // kept for a backward of this frame
    b.final_last = c->d_final_last;
    launch_blend(b, s);
    c->last_blend = b;
    c->has_blend = true;
    if (c->profile) hipEventRecord(c->ev[6], s);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "forward launch");
    return TPD_OK;
}

}  // namespace tpd

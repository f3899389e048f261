// Autotuner: picks launch shapes by timing them on the context's own last frame.  Only choices
// that leave every output bit-identical are tuned (radix-pass block shape here); the reference
// has no counterpart (its shapes are compile-time constants, GaussianEngine.h:121-124).
#include "tpd_ctx.hpp"

#include <algorithm>
#include <vector>

namespace tpd {

static int rerun_duplicate(tpd_ctx* c, hipStream_t s) {
    const uint32_t W = c->cam.width, H = c->cam.height;
    const uint32_t gx = (W + kTile - 1) / kTile, gy = (H + kTile - 1) / kTile;
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
    a.ntiles = gx * gy;
    launch_preprocess_dup(a, s);
    const size_t nd = (c->n + kPreBlock - 1) / kPreBlock;
    hipError_t e = hipMemsetAsync(c->d_scan_desc, 0, nd * sizeof(uint64_t), s);
    if (e == hipSuccess) e = hipMemsetAsync(&c->d_counters[kCtrScanTicket], 0, sizeof(uint32_t), s);
    if (e == hipSuccess) e = hipMemsetAsync(&c->d_counters[kCtrDepthOut], 0, sizeof(uint32_t), s);
    return e == hipSuccess ? TPD_OK : hip_fail(e, "autotune");
}

}  // namespace tpd

using namespace tpd;

extern "C" {

int tpd_set_sort_shape(tpd_ctx* c, uint32_t shape) {
    if (!c) return fail(TPD_ERR_INVALID, "null context");
    if (shape >= sort_num_shapes()) return fail(TPD_ERR_INVALID, "tpd_set_sort_shape: no such shape");
    c->ws.shape = shape;
    batch_set_sort_shape(c, shape);
    return TPD_OK;
}

int tpd_autotune(tpd_ctx* c, float* sort_ms, uint32_t capacity, uint32_t* n_shapes, uint32_t* best) {
    if (!c) return fail(TPD_ERR_INVALID, "null context");
    hipSetDevice(c->device);
    const uint32_t ns = sort_num_shapes();
    if (n_shapes) *n_shapes = ns;
    if (!c->rendered || c->n == 0 || c->last_pairs == 0)
        return fail(TPD_ERR_STATE, "tpd_autotune: render a non-empty frame first");
    hipStream_t s = c->stream;
    const uint32_t W = c->cam.width, H = c->cam.height;
    const uint32_t ntiles = ((W + kTile - 1) / kTile) * ((H + kTile - 1) / kTile);
    const uint32_t key_bits = bit_length(ntiles > 1 ? ntiles - 1 : 1) + c->lo_bits;  // the last frame's key
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return fail(TPD_ERR_HIP, "tpd_autotune: hipEventCreate");
    const uint32_t keep = c->ws.shape;
    std::vector<float> med(ns, 0.0f);
    int rc = TPD_OK;
    for (uint32_t sh = 0; sh < ns && rc == TPD_OK; ++sh) {
        c->ws.shape = sh;
        std::vector<float> t;
        for (int rep = 0; rep < 4 && rc == TPD_OK; ++rep) {
            if ((rc = rerun_duplicate(c, s)) != TPD_OK) break;
            hipEventRecord(e0, s);
            uint32_t drop = kKeysWide;  // with the narrowed passes, as render_forward sorts
            launch_sort_pairs(c->ws, c->d_keys, c->d_vals, c->d_keys_alt, c->d_vals_alt, nullptr, c->last_pairs,
                              key_bits, s, nullptr, c->lo_bits, &drop, c->lo_base);
            hipEventRecord(e1, s);
            if (hipEventSynchronize(e1) != hipSuccess) {
                rc = fail(TPD_ERR_HIP, "tpd_autotune: sort failed");
                break;
            }
            float ms = 0.0f;
            hipEventElapsedTime(&ms, e0, e1);
            if (rep > 0) t.push_back(ms);  // first run warms up
        }
        if (!t.empty()) {
            std::sort(t.begin(), t.end());
            med[sh] = t[t.size() / 2];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != TPD_OK) {
        c->ws.shape = keep;
        return rc;
    }
    uint32_t b = 0;
    for (uint32_t sh = 1; sh < ns; ++sh)
        if (med[sh] < med[b]) b = sh;
    c->ws.shape = b;
    if (best) *best = b;
    if (sort_ms)
        for (uint32_t sh = 0; sh < ns && sh < capacity; ++sh) sort_ms[sh] = med[sh];
    // Leave a consistent frame behind (keys/ranges/image of the current camera).
    return render_forward(c, s);
}

}  // extern "C"

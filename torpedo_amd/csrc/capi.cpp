// C ABI of libtpd_hip.so: context lifetime, scene upload, camera, the forward frame, readbacks,
// the stand-alone sort and the host camera math.  Declarations and reference citations are in
// include/tpd_hip.h.
#include "tpd_ctx.hpp"

#include <torpedo/math.hpp>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>

using namespace tpd;

namespace tpd {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(TPD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

uint32_t bit_length(uint32_t v) {
    uint32_t b = 0;
    while (v) {
        ++b;
        v >>= 1;
    }
    return b;
}

int ensure_scene_buffers(tpd_ctx* c) {
    const size_t n = std::max<uint32_t>(c->n, 1u);
    hipError_t e;
    if ((e = grow(c->d_splats, c->cap_splats, n)) != hipSuccess) return hip_fail(e, "alloc splats");
    if ((e = grow(c->d_tiles, c->cap_tiles, n)) != hipSuccess) return hip_fail(e, "alloc tiles");
    if ((e = grow(c->d_offsets, c->cap_offsets, n)) != hipSuccess) return hip_fail(e, "alloc offsets");
    const size_t nd = (n + kPreBlock - 1) / kPreBlock;
    if (!c->d_scan_desc || c->cap_scan_desc < nd) {
        if ((e = grow(c->d_scan_desc, c->cap_scan_desc, nd)) != hipSuccess) return hip_fail(e, "alloc scan desc");
        if ((e = hipMemsetAsync(c->d_scan_desc, 0, c->cap_scan_desc * sizeof(uint64_t), c->stream)) != hipSuccess)
            return hip_fail(e, "clear scan desc");
    }
    const size_t ne = std::max<uint32_t>(c->n_ent, 1u);
    if ((e = grow(c->d_ent_mats, c->cap_ent_mats, ne)) != hipSuccess) return hip_fail(e, "alloc entity mats");
    return TPD_OK;
}

int ensure_image_buffers(tpd_ctx* c) {
    const uint32_t gx = (c->cam.width + kTile - 1) / kTile, gy = (c->cam.height + kTile - 1) / kTile;
    const size_t ntiles = (size_t)gx * gy;
    const size_t npix = (size_t)c->cam.width * c->cam.height;
    hipError_t e;
    if ((e = grow(c->d_ranges, c->cap_ranges, std::max<size_t>(ntiles, 1))) != hipSuccess ||
        (e = grow(c->d_tile_order, c->cap_tile_order, std::max<size_t>(tile_order_words(ntiles), 1))) != hipSuccess)
        return hip_fail(e, "alloc ranges");
    if ((e = grow(c->d_image, c->cap_image, std::max<size_t>(3 * npix, 1))) != hipSuccess)
        return hip_fail(e, "alloc image");
    if ((e = grow(c->d_final_T, c->cap_final_T, std::max<size_t>(npix, 1))) != hipSuccess ||
        (e = grow(c->d_final_last, c->cap_final_last, std::max<size_t>(npix, 1))) != hipSuccess)
        return hip_fail(e, "alloc pixel state");
    if ((e = grow(c->d_rgba8, c->cap_rgba8, std::max<size_t>(4 * npix, 1))) != hipSuccess)
        return hip_fail(e, "alloc rgba8");
    return TPD_OK;
}

// Pair buffers and the sort workspace for `pairs` pairs (grown with 25% headroom).
int ensure_pair_buffers(tpd_ctx* c, size_t pairs) {
    pairs = std::max<size_t>(pairs, 1);
    if (c->cap_keys >= pairs && c->cap_vals >= pairs && c->cap_keys_alt >= pairs && c->cap_vals_alt >= pairs &&
        (!c->keep_unsorted || c->cap_keys_dbg >= pairs))
        return TPD_OK;
    const double h = 1.25;
    hipError_t e;
    if ((e = grow(c->d_keys, c->cap_keys, pairs, h)) != hipSuccess) return hip_fail(e, "alloc keys");
    if ((e = grow(c->d_vals, c->cap_vals, pairs, h)) != hipSuccess) return hip_fail(e, "alloc vals");
    if ((e = grow(c->d_keys_alt, c->cap_keys_alt, pairs, h)) != hipSuccess) return hip_fail(e, "alloc keys");
    if ((e = grow(c->d_vals_alt, c->cap_vals_alt, pairs, h)) != hipSuccess) return hip_fail(e, "alloc vals");
    if (c->keep_unsorted) {
        if ((e = grow(c->d_keys_dbg, c->cap_keys_dbg, pairs, h)) != hipSuccess) return hip_fail(e, "alloc keys");
        if ((e = grow(c->d_vals_dbg, c->cap_vals_dbg, pairs, h)) != hipSuccess) return hip_fail(e, "alloc vals");
    }
    return ensure_sort_workspace(c, c->cap_keys);
}

int ensure_sort_workspace(tpd_ctx* c, size_t pairs) {
    const uint32_t parts = sort_parts((uint32_t)std::min<size_t>(pairs, kMaxSortKeys));
    hipError_t e;
    if (!c->ws.partials) {
        if ((e = hipMalloc(&c->ws.partials, sizeof(uint32_t) * kCountCopies * kMaxPasses * kMaxRadix)) != hipSuccess ||
            (e = hipMemset(c->ws.partials, 0, sizeof(uint32_t) * kCountCopies * kMaxPasses * kMaxRadix)) != hipSuccess)
            return hip_fail(e, "alloc sort counts");
        if ((e = hipMalloc(&c->ws.offsets, sizeof(uint32_t) * kMaxPasses * kMaxRadix)) != hipSuccess)
            return hip_fail(e, "alloc sort offsets");
        if ((e = hipMalloc(&c->ws.tickets, sizeof(uint32_t) * kMaxPasses)) != hipSuccess)
            return hip_fail(e, "alloc sort tickets");
    }
    if (c->ws.parts_cap < parts) {
        release(c->ws.desc);
        const uint32_t want = parts + parts / 4 + 1;
        if ((e = hipMalloc(&c->ws.desc, sizeof(uint32_t) * 2 * (size_t)want * kMaxRadix)) != hipSuccess)
            return hip_fail(e, "alloc sort descriptors");
        c->ws.parts_cap = want;
    }
    return TPD_OK;
}

}  // namespace tpd

static bool check_ctx(tpd_ctx* c) {
    if (!c) {
        g_last_error = "null context";
        return false;
    }
    hipSetDevice(c->device);
    return true;
}

static hipStream_t pick_stream(tpd_ctx* c, void* stream) {
    return stream ? reinterpret_cast<hipStream_t>(stream) : c->stream;
}

extern "C" {

const char* tpd_last_error(void) { return g_last_error.c_str(); }

int tpd_create(tpd_ctx** out, const tpd_config* cfg) {
    if (!out) return fail(TPD_ERR_INVALID, "tpd_create: null out");
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) return fail(TPD_ERR_NO_DEVICE, "tpd_create: no HIP device");
    const int dev = cfg ? cfg->device : 0;
    if (dev < 0 || dev >= ndev) return fail(TPD_ERR_INVALID, "tpd_create: bad device ordinal");
    if ((e = hipSetDevice(dev)) != hipSuccess) return hip_fail(e, "hipSetDevice");
    int rc = noalloc_setup(dev);
    if (rc != TPD_OK) return rc;
    auto* c = new tpd_ctx();
    c->device = dev;
    c->keep_unsorted = cfg ? cfg->keep_unsorted != 0 : false;
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) {
        delete c;
        return hip_fail(e, "hipStreamCreate");
    }
    if ((e = hipHostMalloc(reinterpret_cast<void**>(&c->h_counters), sizeof(uint32_t) * kCtrCount,
                           hipHostMallocDefault)) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return hip_fail(e, "hipHostMalloc");
    }
    if ((e = hipMalloc(&c->d_counters, sizeof(uint32_t) * kCtrCount)) != hipSuccess ||
        (e = hipMemsetAsync(c->d_counters, 0, sizeof(uint32_t) * kCtrCount, c->stream)) != hipSuccess) {
        tpd_destroy(c);
        return hip_fail(e, "alloc counters");
    }
    *out = c;
    return TPD_OK;
}

int tpd_destroy(tpd_ctx* c) {
    if (!c) return TPD_OK;
    hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    batch_release(c);
    backward_release(c);
    if (c->own_gauss) release(c->d_gauss);
    release(c->d_ent_idx);
    release(c->d_models);
    release(c->d_ent_mats);
    release(c->d_splats);
    release(c->d_tiles);
    release(c->d_offsets);
    release(c->d_scan_desc);
    release(c->d_counters);
    release(c->d_keys);
    release(c->d_vals);
    release(c->d_keys_alt);
    release(c->d_vals_alt);
    release(c->d_keys_dbg);
    release(c->d_vals_dbg);
    release(c->ws.desc);
    release(c->ws.partials);
    release(c->ws.offsets);
    release(c->ws.tickets);
    release(c->d_ranges);
    release(c->d_tile_order);
    release(c->d_image);
    release(c->d_rgba8);
    release(c->d_final_T);
    release(c->d_final_last);
    release(c->d_eval_count);
    if (c->h_counters) (void)hipHostFree(c->h_counters);
    for (auto& ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return TPD_OK;
}

int tpd_upload(tpd_ctx* c, const float* gaussians, uint32_t n, const uint32_t* entity_idx, const float* models,
               uint32_t n_entities, uint32_t sh_degree) {
    if (!check_ctx(c)) return TPD_ERR_INVALID;
    if (n > 0 && !gaussians) return fail(TPD_ERR_INVALID, "tpd_upload: null gaussians");
    if (entity_idx && (!models || n_entities == 0))
        return fail(TPD_ERR_INVALID, "tpd_upload: entity indices without models");
    if (entity_idx) {
        for (uint32_t i = 0; i < n; ++i)
            if (entity_idx[i] >= n_entities) return fail(TPD_ERR_INVALID, "tpd_upload: entity index out of range");
    }
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(e, "sync");
    if (c->own_gauss) release(c->d_gauss);
    c->d_gauss = nullptr;
    c->own_gauss = true;
    c->n = n;
    c->sh_degree = std::min<uint32_t>(sh_degree, 3u);
    if (n) {
        if ((e = hipMalloc(&c->d_gauss, sizeof(float) * 60 * (size_t)n)) != hipSuccess) return hip_fail(e, "alloc");
        if ((e = hipMemcpy(c->d_gauss, gaussians, sizeof(float) * 60 * (size_t)n, hipMemcpyHostToDevice)) !=
            hipSuccess)
            return hip_fail(e, "upload gaussians");
    }
    release(c->d_ent_idx);
    release(c->d_models);
    c->n_ent = models ? n_entities : 1u;
    c->h_models.assign(16 * (size_t)c->n_ent, 0.0f);
    for (uint32_t k = 0; k < c->n_ent; ++k)
        for (int i = 0; i < 16; ++i) c->h_models[16 * k + i] = models ? models[16 * k + i] : ((i % 5) == 0 ? 1.f : 0.f);
    if ((e = hipMalloc(&c->d_models, sizeof(float) * 16 * (size_t)c->n_ent)) != hipSuccess) return hip_fail(e, "alloc");
    if ((e = hipMemcpy(c->d_models, c->h_models.data(), sizeof(float) * 16 * (size_t)c->n_ent,
                       hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, "upload models");
    if (entity_idx && n) {
        if ((e = hipMalloc(&c->d_ent_idx, sizeof(uint32_t) * (size_t)n)) != hipSuccess) return hip_fail(e, "alloc");
        if ((e = hipMemcpy(c->d_ent_idx, entity_idx, sizeof(uint32_t) * (size_t)n, hipMemcpyHostToDevice)) !=
            hipSuccess)
            return hip_fail(e, "upload entity indices");
    }
    c->have_scene = true;
    c->rendered = false;
    return ensure_scene_buffers(c);
}

int tpd_bind_device_scene(tpd_ctx* c, const float* d_gaussians, uint32_t n, uint32_t sh_degree) {
    if (!check_ctx(c)) return TPD_ERR_INVALID;
    if (n > 0 && !d_gaussians) return fail(TPD_ERR_INVALID, "tpd_bind_device_scene: null pointer");
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(e, "sync");
    if (c->own_gauss) release(c->d_gauss);
    c->own_gauss = false;
    c->d_gauss = const_cast<float*>(d_gaussians);
    c->n = n;
    c->sh_degree = std::min<uint32_t>(sh_degree, 3u);
    // A bound scene is one entity with the identity model: whatever entity indices and transforms an
    // earlier upload left (sized for that scene's point count) are dropped.
    release(c->d_ent_idx);
    release(c->d_models);
    c->n_ent = 1;
    c->h_models = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    if ((e = hipMalloc(&c->d_models, sizeof(float) * 16)) != hipSuccess) return hip_fail(e, "alloc");
    if ((e = hipMemcpy(c->d_models, c->h_models.data(), sizeof(float) * 16, hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, "upload models");
    c->have_scene = true;
    c->rendered = false;
    return ensure_scene_buffers(c);
}

int tpd_set_transform(tpd_ctx* c, uint32_t entity, const float model[16]) {
    if (!check_ctx(c)) return TPD_ERR_INVALID;
    if (!model) return fail(TPD_ERR_INVALID, "tpd_set_transform: null matrix");
    // TransformHost::transform ignores unknown entities (TransformHost.cpp:4-6); so does this.
    if (entity >= c->n_ent || !c->d_models) return TPD_OK;
    std::memcpy(&c->h_models[16 * (size_t)entity], model, 16 * sizeof(float));
    hipError_t e = hipMemcpyAsync(c->d_models + 16 * (size_t)entity, &c->h_models[16 * (size_t)entity],
                                  16 * sizeof(float), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(e, "upload transform");
    return TPD_OK;
}

int tpd_set_camera(tpd_ctx* c, const float view[16], const float proj_view[16], const float focal_ndc[2],
                   uint32_t width, uint32_t height) {
    if (!check_ctx(c)) return TPD_ERR_INVALID;
    if (!view || !proj_view || !focal_ndc) return fail(TPD_ERR_INVALID, "tpd_set_camera: null pointer");
    if (width == 0 || height == 0 || width > 32768 || height > 32768)
        return fail(TPD_ERR_INVALID, "tpd_set_camera: image size out of range");
    std::memcpy(c->cam.view, view, sizeof(c->cam.view));
    std::memcpy(c->cam.proj_view, proj_view, sizeof(c->cam.proj_view));
    c->cam.focal_ndc[0] = focal_ndc[0];
    c->cam.focal_ndc[1] = focal_ndc[1];
    c->cam.width = width;
    c->cam.height = height;
    c->have_camera = true;
    c->rendered = false;
    return ensure_image_buffers(c);
}

int tpd_render(tpd_ctx* c, void* stream) {
    if (!check_ctx(c)) return TPD_ERR_INVALID;
    if (!c->have_scene) return fail(TPD_ERR_NO_SCENE, "tpd_render: no scene uploaded");
    if (!c->have_camera) return fail(TPD_ERR_NO_CAMERA, "tpd_render: no camera set");
    hipStream_t s = pick_stream(c, stream);
    int rc = render_forward(c, s);
    if (rc == TPD_OK) c->rendered = true;
    return rc;
}

int tpd_get_pair_count(tpd_ctx* c, uint32_t* pairs) {
    if (!check_ctx(c) || !pairs) return fail(TPD_ERR_INVALID, "tpd_get_pair_count: bad argument");
    if (!c->rendered) return fail(TPD_ERR_STATE, "tpd_get_pair_count: nothing rendered");
    *pairs = c->last_pairs;
    return TPD_OK;
}

static int copy_back(tpd_ctx* c, void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return TPD_OK;
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "readback");
    return TPD_OK;
}

int tpd_read_image_f32(tpd_ctx* c, float* rgb) {
    if (!check_ctx(c) || !rgb) return fail(TPD_ERR_INVALID, "tpd_read_image_f32: bad argument");
    if (!c->rendered) return fail(TPD_ERR_STATE, "tpd_read_image_f32: nothing rendered");
    return copy_back(c, rgb, c->d_image, sizeof(float) * 3 * (size_t)c->cam.width * c->cam.height);
}

int tpd_read_image_rgba8(tpd_ctx* c, uint8_t* rgba) {
    if (!check_ctx(c) || !rgba) return fail(TPD_ERR_INVALID, "tpd_read_image_rgba8: bad argument");
    if (!c->rendered) return fail(TPD_ERR_STATE, "tpd_read_image_rgba8: nothing rendered");
    return copy_back(c, rgba, c->d_rgba8, 4 * (size_t)c->cam.width * c->cam.height);
}

int tpd_read_splats(tpd_ctx* c, float* out) {
    if (!check_ctx(c) || !out) return fail(TPD_ERR_INVALID, "tpd_read_splats: bad argument");
    if (!c->rendered) return fail(TPD_ERR_STATE, "tpd_read_splats: nothing rendered");
    const size_t n = c->n;
    if (n == 0) return TPD_OK;
    std::vector<tpd::Splat> sp(n);
    std::vector<uint32_t> tiles(n);
    int rc;
    if ((rc = copy_back(c, sp.data(), c->d_splats, n * sizeof(tpd::Splat))) != TPD_OK) return rc;
    if ((rc = copy_back(c, tiles.data(), c->d_tiles, n * sizeof(uint32_t))) != TPD_OK) return rc;
    for (size_t i = 0; i < n; ++i) {
        float* o = out + 12 * i;
        o[0] = sp[i].rgb.x; o[1] = sp[i].rgb.y; o[2] = sp[i].rgb.z;
        std::memcpy(&o[3], &tiles[i], 4);
        o[4] = sp[i].xyzr.x; o[5] = sp[i].xyzr.y; o[6] = sp[i].xyzr.z; o[7] = sp[i].xyzr.w;
        o[8] = sp[i].conop.x; o[9] = sp[i].conop.y; o[10] = sp[i].conop.z; o[11] = sp[i].conop.w;
    }
    return TPD_OK;
}

int tpd_read_tiles(tpd_ctx* c, uint32_t* counts, uint32_t* offsets) {
    if (!check_ctx(c) || !counts) return fail(TPD_ERR_INVALID, "tpd_read_tiles: bad argument");
    if (!c->rendered) return fail(TPD_ERR_STATE, "tpd_read_tiles: nothing rendered");
    int rc = copy_back(c, counts, c->d_tiles, sizeof(uint32_t) * (size_t)c->n);
    if (rc == TPD_OK && offsets) rc = copy_back(c, offsets, c->d_offsets, sizeof(uint32_t) * (size_t)c->n);
    return rc;
}

int tpd_read_pairs(tpd_ctx* c, uint64_t* keys, uint32_t* vals, uint32_t capacity, int unsorted) {
    if (!check_ctx(c) || !keys || !vals) return fail(TPD_ERR_INVALID, "tpd_read_pairs: bad argument");
    if (!c->rendered) return fail(TPD_ERR_STATE, "tpd_read_pairs: nothing rendered");
    if (capacity < c->last_pairs) return fail(TPD_ERR_INVALID, "tpd_read_pairs: capacity below pair count");
    if (unsorted && !c->keep_unsorted) return fail(TPD_ERR_STATE, "tpd_read_pairs: keep_unsorted was not set");
    const uint64_t* k = unsorted ? c->d_keys_dbg : c->sorted_keys;
    const uint32_t* v = unsorted ? c->d_vals_dbg : c->sorted_vals;
    if (!unsorted && c->key_drop != kKeysWide && c->last_pairs) {
        // the sort left 32-bit keys: rebuild the 64-bit ones in the other (free) key buffer
        uint64_t* full = c->sorted_keys == c->d_keys ? c->d_keys_alt : c->d_keys;
        launch_full_keys(reinterpret_cast<const uint32_t*>(c->sorted_keys), c->key_drop, c->lo_bits, c->sorted_vals,
                         c->d_splats, c->last_pairs, full, c->stream);
        k = full;
    }
    int rc = copy_back(c, keys, k, sizeof(uint64_t) * (size_t)c->last_pairs);
    if (rc == TPD_OK) rc = copy_back(c, vals, v, sizeof(uint32_t) * (size_t)c->last_pairs);
    return rc;
}

int tpd_read_ranges(tpd_ctx* c, uint32_t* ranges) {
    if (!check_ctx(c) || !ranges) return fail(TPD_ERR_INVALID, "tpd_read_ranges: bad argument");
    if (!c->rendered) return fail(TPD_ERR_STATE, "tpd_read_ranges: nothing rendered");
    const size_t ntiles = (size_t)((c->cam.width + 15) / 16) * ((c->cam.height + 15) / 16);
    return copy_back(c, ranges, c->d_ranges, sizeof(uint2) * ntiles);
}

int tpd_read_tile_order(tpd_ctx* c, uint32_t* order) {
    if (!check_ctx(c) || !order) return fail(TPD_ERR_INVALID, "tpd_read_tile_order: bad argument");
    if (!c->rendered) return fail(TPD_ERR_STATE, "tpd_read_tile_order: nothing rendered");
    if (c->n == 0) return fail(TPD_ERR_STATE, "tpd_read_tile_order: an empty scene blends no tiles");
    const size_t ntiles = (size_t)((c->cam.width + 15) / 16) * ((c->cam.height + 15) / 16);
    return copy_back(c, order, c->d_tile_order, sizeof(uint32_t) * ntiles);
}

int tpd_synchronize(tpd_ctx* c) {
    if (!check_ctx(c)) return TPD_ERR_INVALID;
    hipError_t e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? TPD_OK : hip_fail(e, "hipStreamSynchronize");
}

int tpd_set_profiling(tpd_ctx* c, int enable) {
    if (!check_ctx(c)) return TPD_ERR_INVALID;
    if (enable && !c->ev[0]) {
        for (auto& ev : c->ev) {
            hipError_t e = hipEventCreate(&ev);
            if (e != hipSuccess) return hip_fail(e, "hipEventCreate");
        }
    }
    c->profile = enable != 0;
    return TPD_OK;
}

int tpd_get_stage_times(tpd_ctx* c, float* ms, uint32_t* passes) {
    if (!check_ctx(c) || !ms) return fail(TPD_ERR_INVALID, "tpd_get_stage_times: bad argument");
    if (!c->profile || !c->rendered || c->n == 0) return fail(TPD_ERR_STATE, "tpd_get_stage_times: no profiled frame");
    hipError_t e = hipEventSynchronize(c->ev[6]);
    if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize");
    const int from[6] = {0, 2, 3, 4, 5, 0};
    const int to[6] = {1, 3, 4, 5, 6, 6};
    for (int k = 0; k < 6; ++k) {
        e = hipEventElapsedTime(&ms[k], c->ev[from[k]], c->ev[to[k]]);
        if (e != hipSuccess) return hip_fail(e, "hipEventElapsedTime");
    }
    if (passes) *passes = c->last_passes;
    return TPD_OK;
}

// The last rendered frame is blended once more by the counting kernel: the same image, transmittance
// and last contributors are written again, and the frame's own blend runs uncounted.
int tpd_get_blend_evaluations(tpd_ctx* c, uint64_t* evaluations) {
    if (!check_ctx(c) || !evaluations) return fail(TPD_ERR_INVALID, "tpd_get_blend_evaluations: bad argument");
    if (!c->rendered || !c->has_blend) return fail(TPD_ERR_STATE, "tpd_get_blend_evaluations: nothing rendered");
    hipSetDevice(c->device);
    hipError_t e = hipSuccess;
    if (!c->d_eval_count && (e = hipMalloc(&c->d_eval_count, sizeof(uint64_t))) != hipSuccess)
        return hip_fail(e, "alloc evaluation counter");
    if ((e = hipMemsetAsync(c->d_eval_count, 0, sizeof(uint64_t), c->stream)) != hipSuccess)
        return hip_fail(e, "clear evaluation counter");
    BlendArgs b = c->last_blend;
    b.eval_count = c->d_eval_count;
    launch_blend(b, c->stream);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "counting blend");
    return copy_back(c, evaluations, c->d_eval_count, sizeof(uint64_t));
}

int tpd_alloc(tpd_ctx* c, uint64_t bytes, void** d_ptr) {
    if (!check_ctx(c) || !d_ptr) return fail(TPD_ERR_INVALID, "tpd_alloc: bad argument");
    *d_ptr = nullptr;
    hipError_t e = hipMalloc(d_ptr, bytes ? bytes : 1);
    return e == hipSuccess ? TPD_OK : hip_fail(e, "tpd_alloc");
}

int tpd_free(tpd_ctx* c, void* d_ptr) {
    if (!check_ctx(c)) return TPD_ERR_INVALID;
    if (!d_ptr) return TPD_OK;
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipFree(d_ptr);
    return e == hipSuccess ? TPD_OK : hip_fail(e, "tpd_free");
}

int tpd_memcpy(tpd_ctx* c, void* dst, const void* src, uint64_t bytes, int kind) {
    if (!check_ctx(c) || (bytes && (!dst || !src)) || kind < 0 || kind > 2)
        return fail(TPD_ERR_INVALID, "tpd_memcpy: bad argument");
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess && bytes) e = hipMemcpy(dst, src, bytes, k);
    return e == hipSuccess ? TPD_OK : hip_fail(e, "tpd_memcpy");
}

int tpd_copy_image(tpd_ctx* c, float* d_dst, void* stream) {
    if (!check_ctx(c) || !d_dst) return fail(TPD_ERR_INVALID, "tpd_copy_image: bad argument");
    if (!c->rendered) return fail(TPD_ERR_STATE, "tpd_copy_image: nothing rendered");
    hipError_t e = hipMemcpyAsync(d_dst, c->d_image, sizeof(float) * 3 * (size_t)c->cam.width * c->cam.height,
                                  hipMemcpyDeviceToDevice, pick_stream(c, stream));
    return e == hipSuccess ? TPD_OK : hip_fail(e, "tpd_copy_image");
}

int tpd_device_image(tpd_ctx* c, float** d_rgb) {
    if (!check_ctx(c) || !d_rgb) return fail(TPD_ERR_INVALID, "tpd_device_image: bad argument");
    *d_rgb = c->d_image;
    return TPD_OK;
}

int tpd_sort_pairs(tpd_ctx* c, uint64_t* d_keys, uint32_t* d_vals, uint64_t* d_keys_alt, uint32_t* d_vals_alt,
                   uint32_t n, uint32_t key_bits, void* stream, int* in_alt) {
    if (!check_ctx(c)) return TPD_ERR_INVALID;
    if (!in_alt) return fail(TPD_ERR_INVALID, "tpd_sort_pairs: null in_alt");
    if (key_bits == 0 || key_bits > 64) return fail(TPD_ERR_INVALID, "tpd_sort_pairs: key_bits must be 1..64");
    if (n > kMaxSortKeys) return fail(TPD_ERR_CAPACITY, "tpd_sort_pairs: more than 2^30 - 1 pairs");
    *in_alt = 0;
    if (n == 0) return TPD_OK;
    if (!d_keys || !d_vals || !d_keys_alt || !d_vals_alt) return fail(TPD_ERR_INVALID, "tpd_sort_pairs: null pointer");
    int rc = ensure_sort_workspace(c, n);
    if (rc != TPD_OK) return rc;
    hipStream_t s = pick_stream(c, stream);
    *in_alt = launch_sort_pairs(c->ws, d_keys, d_vals, d_keys_alt, d_vals_alt, nullptr, n, key_bits, s) ? 1 : 0;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "sort launch");
    return TPD_OK;
}

// ----- host camera math

void tpd_look_at(const float eye[3], const float center[3], const float up[3], float view[16]) {
    const vec3 e{eye[0], eye[1], eye[2]}, ce{center[0], center[1], center[2]}, u{up[0], up[1], up[2]};
    const vec3 z = math::normalize(ce - e);
    const vec3 x = math::normalize(math::cross(z, u));
    const vec3 y = math::normalize(math::cross(z, x));
    const float r[16] = {x.x, x.y, x.z, -math::dot(x, e), y.x, y.y, y.z, -math::dot(y, e),
                         z.x, z.y, z.z, -math::dot(z, e), 0.f, 0.f, 0.f, 1.f};
    std::memcpy(view, r, sizeof(r));
}

void tpd_perspective(uint32_t width, uint32_t height, float fy, float znear, float zfar, float proj[16]) {
    const float aspect = static_cast<float>(width) / static_cast<float>(height);
    const float fx = fy / aspect;
    const float za = znear / (znear - zfar);
    const float zb = znear * zfar / (zfar - znear);
    const float r[16] = {fx, 0.f, 0.f, 0.f, 0.f, fy, 0.f, 0.f, 0.f, 0.f, za, zb, 0.f, 0.f, 1.f, 0.f};
    std::memcpy(proj, r, sizeof(r));
}

float tpd_fy_from_fov(float degrees) {
    const float fov = degrees * 3.14159265358979323846f / 180.f;
    return 1.f / std::tan(fov * 0.5f);
}

void tpd_mat4_mul(const float a[16], const float b[16], float out[16]) {
    const mat4 r = math::mul(mat4{a}, mat4{b});
    std::memcpy(out, r.data_ptr(), 16 * sizeof(float));
}

void tpd_camera_upload(const float view[16], const float proj[16], float proj_view[16], float focal_ndc[2]) {
    tpd_mat4_mul(proj, view, proj_view);
    focal_ndc[0] = proj[0];
    focal_ndc[1] = proj[5];
}

}  // extern "C"

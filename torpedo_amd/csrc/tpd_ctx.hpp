// The context behind the opaque tpd_ctx handle, and the host-side helpers shared by the C-ABI
// translation units.
#pragma once

#include "tpd_kernels.hpp"
#include "../../include/tpd_hip.h"

#include <string>
#include <vector>

namespace tpd {
struct BackwardState;
struct BatchState;
}  // namespace tpd

struct tpd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool keep_unsorted = false;

    // scene (GaussianEngine::compile)
    float* d_gauss = nullptr;  // N x 60
    bool own_gauss = true;
    uint32_t n = 0, sh_degree = 3;
    uint32_t* d_ent_idx = nullptr;
    float* d_models = nullptr;
    uint32_t n_ent = 1;
    std::vector<float> h_models;
    bool have_scene = false;

    // camera
    tpd::CameraParams cam{};
    bool have_camera = false;

    // per-Gaussian frame buffers
    tpd::EntityMats* d_ent_mats = nullptr; size_t cap_ent_mats = 0;
    tpd::Splat* d_splats = nullptr; size_t cap_splats = 0;
    uint32_t* d_tiles = nullptr; size_t cap_tiles = 0;
    uint32_t* d_offsets = nullptr; size_t cap_offsets = 0;
    uint64_t* d_scan_desc = nullptr; size_t cap_scan_desc = 0;
    uint32_t* d_counters = nullptr;
    uint32_t* h_counters = nullptr;  // pinned

    // pairs + sort
    uint64_t* d_keys = nullptr; size_t cap_keys = 0;
    uint32_t* d_vals = nullptr; size_t cap_vals = 0;
    uint64_t* d_keys_alt = nullptr; size_t cap_keys_alt = 0;
    uint32_t* d_vals_alt = nullptr; size_t cap_vals_alt = 0;
    uint64_t* d_keys_dbg = nullptr; size_t cap_keys_dbg = 0;
    uint32_t* d_vals_dbg = nullptr; size_t cap_vals_dbg = 0;
    tpd::SortWorkspace ws;
    const uint64_t* sorted_keys = nullptr;   // 64-bit keys, or 32-bit ones when key_drop != kKeysWide
    uint32_t key_drop = tpd::kKeysWide;
    // The last frame's sort key: (tile << lo_bits) | (depth bits - lo_base), the depth window
    // (27, kDepthBase) or the positive-float form (31, 0).
    uint32_t lo_bits = tpd::kDepthWideBits, lo_base = 0;
    const uint32_t* sorted_vals = nullptr;

    // image
    uint2* d_ranges = nullptr; size_t cap_ranges = 0;
    uint32_t* d_tile_order = nullptr; size_t cap_tile_order = 0;
    float* d_image = nullptr; size_t cap_image = 0;
    uint8_t* d_rgba8 = nullptr; size_t cap_rgba8 = 0;
    // per pixel final transmittance and last blended splat of the frame, for the backward
    float* d_final_T = nullptr; size_t cap_final_T = 0;
    int32_t* d_final_last = nullptr; size_t cap_final_last = 0;

    uint32_t last_pairs = 0;
    uint32_t last_passes = 0;
    bool rendered = false;

    // stage timing with HIP events on the render stream (tpd_set_profiling)
    bool profile = false;
    hipEvent_t ev[8] = {};
    uint64_t* d_eval_count = nullptr;  // blend evaluations, counted on request (tpd_get_blend_evaluations)
    tpd::BlendArgs last_blend{};       // the last rendered frame's blend, for that count
    bool has_blend = false;

    // backward and batched state (backward.cpp, batch.cpp)
    tpd::BackwardState* bwd = nullptr;
    tpd::BatchState* batch = nullptr;
};

namespace tpd {
extern thread_local std::string g_last_error;
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
uint32_t bit_length(uint32_t v);
int ensure_scene_buffers(tpd_ctx* c);
int ensure_image_buffers(tpd_ctx* c);
int ensure_pair_buffers(tpd_ctx* c, size_t pairs);
int ensure_sort_workspace(tpd_ctx* c, size_t pairs);
int render_forward(tpd_ctx* c, hipStream_t s);
void backward_release(tpd_ctx* c);
void batch_release(tpd_ctx* c);
void batch_set_sort_shape(tpd_ctx* c, uint32_t shape);
int batch_view_rgba8(tpd_ctx* c, uint32_t view, const uint8_t** d_rgba, uint32_t* width, uint32_t* height);
int noalloc_setup(int device);  // device address of the allocation-free sort's scratch (noalloc.hip)

// Grows a device allocation to hold at least `count` elements (contents are not preserved).
template <typename T>
inline hipError_t grow(T*& p, size_t& cap, size_t count, double headroom = 1.0) {
    if (p && cap >= count) return hipSuccess;
    if (p) {
        hipError_t e = hipFree(p);
        if (e != hipSuccess) return e;
        p = nullptr;
    }
    size_t want = (size_t)((double)count * headroom);
    if (want < count) want = count;
    if (want < 1) want = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), want * sizeof(T));
    if (e != hipSuccess) {
        cap = 0;
        p = nullptr;
        return e;
    }
    cap = want;
    return hipSuccess;
}

template <typename T>
inline void release(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}
}  // namespace tpd

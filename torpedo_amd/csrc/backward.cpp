// tpd_backward: gradients of the last rendered frame.  The reference has no backward
// (torpedo/differential/include/torpedo/differential/GaussianRasterizer.h:1-8 is an empty stub).
#include "tpd_ctx.hpp"

#include <algorithm>

namespace tpd {

struct BackwardState {
    float* g2d = nullptr;       // N x 9 per-Gaussian screen-space gradients
    size_t cap = 0;
    float* partials = nullptr;  // P x 9 per (tile, splat) pair
    size_t cap_partials = 0;
    hipEvent_t ev[3] = {};
    bool timed = false;
};

void backward_release(tpd_ctx* c) {
    if (!c->bwd) return;
    if (c->bwd->g2d) (void)hipFree(c->bwd->g2d);
    if (c->bwd->partials) (void)hipFree(c->bwd->partials);
    for (auto& ev : c->bwd->ev)
        if (ev) (void)hipEventDestroy(ev);
    delete c->bwd;
    c->bwd = nullptr;
}

}  // namespace tpd

using namespace tpd;

extern "C" int tpd_backward(tpd_ctx* c, const float* d_dL_dimage, float* d_grad_means, float* d_grad_log_scales,
                            float* d_grad_quats, float* d_grad_opacity_logits, float* d_grad_sh, int accumulate,
                            void* stream) {
    if (!c) return fail(TPD_ERR_INVALID, "null context");
    hipSetDevice(c->device);
    if (!d_dL_dimage || !d_grad_means || !d_grad_log_scales || !d_grad_quats || !d_grad_opacity_logits || !d_grad_sh)
        return fail(TPD_ERR_INVALID, "tpd_backward: null pointer");
    if (!c->rendered) return fail(TPD_ERR_STATE, "tpd_backward: render the frame first");
    hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : c->stream;
    if (!c->bwd) c->bwd = new BackwardState();
    BackwardState* b = c->bwd;
    const size_t need = 9 * (size_t)std::max<uint32_t>(c->n, 1u);
    hipError_t e;
    if (b->cap < need) {
        if (b->g2d) (void)hipFree(b->g2d);
        b->g2d = nullptr;
        if ((e = hipMalloc(&b->g2d, need * sizeof(float))) != hipSuccess) return hip_fail(e, "alloc backward buffers");
        b->cap = need;
    }
    const size_t need_p = 9 * (size_t)std::max<uint32_t>(c->last_pairs, 1u);
    if (b->cap_partials < need_p) {
        if (b->partials) (void)hipFree(b->partials);
        b->partials = nullptr;
        const size_t want = need_p + need_p / 4;
        if ((e = hipMalloc(&b->partials, want * sizeof(float))) != hipSuccess) return hip_fail(e, "alloc backward buffers");
        b->cap_partials = want;
    }
    if (c->n == 0) return TPD_OK;
    if (c->profile) {
        for (auto& ev : b->ev)
            if (!ev && (e = hipEventCreate(&ev)) != hipSuccess) return hip_fail(e, "hipEventCreate");
        hipEventRecord(b->ev[0], s);
    }
    const uint32_t W = c->cam.width, H = c->cam.height;
    const uint32_t gx = (W + kTile - 1) / kTile, gy = (H + kTile - 1) / kTile;
    if (c->last_pairs > 0) {
        BackwardArgs a{};
        a.ranges = c->d_ranges;
        a.tile_order = c->d_tile_order;
        a.vals = c->sorted_vals;
        a.splats = c->d_splats;
        a.width = W;
        a.height = H;
        a.gx = gx;
        a.gy = gy;
        a.ntiles = gx * gy;
        a.dL_dimg = d_dL_dimage;
        a.offsets = c->d_offsets;
        a.partials = b->partials;
        a.final_T = c->d_final_T;
        a.final_last = c->d_final_last;
        launch_blend_bwd(a, s);
    }
    if (c->profile) hipEventRecord(b->ev[1], s);
    PreBackwardArgs p{};
    p.gauss = c->d_gauss;
    p.ent_idx = c->d_ent_idx;
    p.ent = c->d_ent_mats;
    p.cam = c->cam;
    p.n = c->n;
    p.sh_degree = c->sh_degree;
    p.tiles = c->d_tiles;
    p.offsets = c->d_offsets;
    p.partials = b->partials;
    p.g2d = b->g2d;
    p.grad_means = d_grad_means;
    p.grad_log_scales = d_grad_log_scales;
    p.grad_quats = d_grad_quats;
    p.grad_opacity_logits = d_grad_opacity_logits;
    p.grad_sh = d_grad_sh;
    p.accumulate = accumulate ? 1u : 0u;
    launch_preprocess_bwd(p, s);
    if (c->profile) {
        hipEventRecord(b->ev[2], s);
        b->timed = true;
    }
    e = hipGetLastError();
    return e == hipSuccess ? TPD_OK : hip_fail(e, "backward launch");
}

// ms[0] = splat-gradient pass (blend backward, incl. clearing), ms[1] = parameter-gradient pass
// (preprocess backward), ms[2] = forward start -> backward end of the last profiled frame.
extern "C" int tpd_get_backward_times(tpd_ctx* c, float* ms) {
    if (!c || !ms) return fail(TPD_ERR_INVALID, "tpd_get_backward_times: bad argument");
    hipSetDevice(c->device);
    if (!c->profile || !c->bwd || !c->bwd->timed) return fail(TPD_ERR_STATE, "tpd_get_backward_times: no profiled backward");
    BackwardState* b = c->bwd;
    hipError_t e = hipEventSynchronize(b->ev[2]);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms[0], b->ev[0], b->ev[1]);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms[1], b->ev[1], b->ev[2]);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms[2], c->ev[0], b->ev[2]);
    return e == hipSuccess ? TPD_OK : hip_fail(e, "tpd_get_backward_times");
}

extern "C" int tpd_read_splat_grads(tpd_ctx* c, float* g2d) {
    if (!c || !g2d) return fail(TPD_ERR_INVALID, "tpd_read_splat_grads: bad argument");
    hipSetDevice(c->device);
    if (!c->bwd || !c->bwd->g2d) return fail(TPD_ERR_STATE, "tpd_read_splat_grads: no backward run");
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(g2d, c->bwd->g2d, 9 * (size_t)c->n * sizeof(float), hipMemcpyDeviceToHost);
    return e == hipSuccess ? TPD_OK : hip_fail(e, "tpd_read_splat_grads");
}

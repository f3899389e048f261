// Backward pass of the forward frame.  The reference has no backward (torpedo/differential is
// an empty stub, GaussianRasterizer.h:1-8); this differentiates the forward of project.slang and
// blend.slang as restated in k_forward.hip:
//   k_blend_bwd       per tile: starts from each pixel's final transmittance and last
//                     contributor (stored by the forward blend) and walks the tile's splats back
//                     to front recovering T_i = T_{i+1} / (1 - a_i), producing per splat
//                     dL/d(image x, y), dL/d(conic a, b, c), dL/d(rgb), dL/d(opacity); row sums
//                     are combined in LDS and stored once per (tile, splat) pair at the pair's
//                     pre-sort position (blend_bwd_tile).
//   k_preprocess_bwd  per Gaussian: chain rule through the conic inverse, the EWA projection
//                     T Sigma T^T, the Jacobian, R(q) S, the projection of the mean and the SH
//                     colour (band-3 typo included), to means, log-scales, quaternions, opacity
//                     logits and SH coefficients.
// Non-differentiable decisions (culling, tile rectangles, alpha < 1/255 skips, early
// termination, the 0.99 clamp, max(colour, 0)) pass zero gradient, as the forward's branches do.
#include "tpd_blend.hpp"
#include "tpd_grad.hpp"
#include "tpd_kernels.hpp"

namespace tpd {

// ------------------------------------------------------------------ blend backward

__global__ __launch_bounds__(kTilePixels) void k_blend_bwd(const BackwardArgs a) {
    __shared__ BlendBwdLds lds;
    const uint32_t tile = a.tile_order ? a.tile_order[blockIdx.x] : xcd_remap(blockIdx.x, gridDim.x);
    const TileView tv{tile % a.gx, tile / a.gx, a.width, a.height, a.gx, a.gy, 0};
    blend_bwd_tile(lds, tv, a.ranges[tile], a.vals, a.splats, a.dL_dimg, a.offsets, a.partials, a.final_T,
                   a.final_last);
}

void launch_blend_bwd(const BackwardArgs& a, hipStream_t s) {
    if (a.ntiles == 0) return;
    hipLaunchKernelGGL(k_blend_bwd, dim3(a.ntiles), dim3(kTilePixels), 0, s, a);
}

// ------------------------------------------------------------------ preprocess backward

// Per Gaussian: sums its per-tile partials (contiguous, in tile order: deterministic) and applies
// the chain rule of accumulate_param_grads.
__global__ __launch_bounds__(256) void k_preprocess_bwd(const PreBackwardArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    float gm[3] = {0.f, 0.f, 0.f}, gls[3] = {0.f, 0.f, 0.f}, gq[4] = {0.f, 0.f, 0.f, 0.f}, gop = 0.f;
    float gsh[48];
#pragma unroll
    for (int k = 0; k < 48; ++k) gsh[k] = 0.f;
    const uint32_t tiles = a.tiles[i];
    float gr[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) gr[q] = 0.0f;
    if (tiles > 0) {
        sum_pair_partials(a.partials + (size_t)a.offsets[i] * 9, tiles, gr);
        const uint32_t e = a.ent_idx ? a.ent_idx[i] : 0u;
        accumulate_param_grads(a.gauss + (size_t)i * 60, a.ent[e].vm, a.ent[e].pvm, a.cam, a.sh_degree, gr, gm, gls,
                               gq, gop, gsh);
    }
    if (a.g2d)
#pragma unroll
        for (int q = 0; q < 9; ++q) a.g2d[(size_t)i * 9 + q] = gr[q];
    store_param_grads(i, a.accumulate != 0, gm, gls, gq, gop, gsh, a.grad_means, a.grad_log_scales, a.grad_quats,
                      a.grad_opacity_logits, a.grad_sh);
}

void launch_preprocess_bwd(const PreBackwardArgs& a, hipStream_t s) {
    if (a.n == 0) return;
    hipLaunchKernelGGL(k_preprocess_bwd, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
}

}  // namespace tpd

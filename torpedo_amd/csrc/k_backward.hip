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
// the chain rule of accumulate_param_grads.  The block's 256 records (240 B each, consecutive) are
// staged in LDS by consecutive loads of the whole block (one thread's record at a time would make
// every load touch 64 cache lines): 16-byte loads when the records are 16-byte aligned, else 4-byte
// ones.  A stride of 61 floats keeps the lanes' reads of one field on distinct banks.  The same LDS
// then holds the block's SH gradients (48 floats per Gaussian), stored to HBM the same way when
// the gradient buffer is 16-byte aligned (else by each thread).
constexpr uint32_t kPreBwdBlock = 256;
constexpr uint32_t kRecStride = 61;
__global__ __launch_bounds__(kPreBwdBlock) void k_preprocess_bwd(const PreBackwardArgs a) {
    __shared__ alignas(16) float s_rec[kPreBwdBlock * kRecStride];
    const uint32_t i0 = blockIdx.x * kPreBwdBlock, i = i0 + threadIdx.x;
    const uint32_t nvalid = min(kPreBwdBlock, a.n - i0);  // >= 1: the grid covers n
    const bool valid = threadIdx.x < nvalid;             // every thread stays for the barriers
    const float* src = a.gauss + (size_t)i0 * 60;
    if ((reinterpret_cast<uintptr_t>(src) & 15u) == 0u) {  // block-uniform
        const float4* src4 = reinterpret_cast<const float4*>(src);
        for (uint32_t x = threadIdx.x; x < nvalid * 15u; x += kPreBwdBlock) {
            const float4 v = src4[x];
            float* d = s_rec + (x / 15u) * kRecStride + 4u * (x % 15u);
            d[0] = v.x; d[1] = v.y; d[2] = v.z; d[3] = v.w;
        }
    } else {
        for (uint32_t x = threadIdx.x; x < nvalid * 60u; x += kPreBwdBlock)
            s_rec[(x / 60u) * kRecStride + x % 60u] = src[x];
    }
    __syncthreads();
    float gm[3] = {0.f, 0.f, 0.f}, gls[3] = {0.f, 0.f, 0.f}, gq[4] = {0.f, 0.f, 0.f, 0.f}, gop = 0.f;
    float gsh[48];
#pragma unroll
    for (int k = 0; k < 48; ++k) gsh[k] = 0.f;
    float gr[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) gr[q] = 0.0f;
    const uint32_t tiles = valid ? a.tiles[i] : 0u;
    if (tiles > 0) {
        sum_pair_partials(a.partials + (size_t)a.offsets[i] * 9, tiles, gr);
        const uint32_t e = a.ent_idx ? a.ent_idx[i] : 0u;
        accumulate_param_grads(s_rec + threadIdx.x * kRecStride, a.ent[e].vm, a.ent[e].pvm, a.cam, a.sh_degree, gr,
                               gm, gls, gq, gop, gsh);
    }
    if (valid && a.g2d)
#pragma unroll
        for (int q = 0; q < 9; ++q) a.g2d[(size_t)i * 9 + q] = gr[q];
    if (valid)
        store_param_grads_but_sh(i, a.accumulate != 0, gm, gls, gq, gop, a.grad_means, a.grad_log_scales,
                                 a.grad_quats, a.grad_opacity_logits);
    float* const dsh = a.grad_sh + (size_t)i0 * 48;
    if ((reinterpret_cast<uintptr_t>(dsh) & 15u) != 0u) {  // block-uniform: each thread its own
        if (valid) {
            float* o = dsh + (size_t)threadIdx.x * 48;
#pragma unroll
            for (int k = 0; k < 48; ++k) o[k] = a.accumulate ? o[k] + gsh[k] : gsh[k];
        }
        return;
    }
    __syncthreads();  // the records are done with: the LDS takes the SH gradients
    if (valid) {
        float4* d = reinterpret_cast<float4*>(s_rec + threadIdx.x * 48u);
#pragma unroll
        for (int k = 0; k < 12; ++k) d[k] = make_float4(gsh[4 * k], gsh[4 * k + 1], gsh[4 * k + 2], gsh[4 * k + 3]);
    }
    __syncthreads();
    store_sh_grads_block(reinterpret_cast<const float4*>(s_rec), nvalid, a.accumulate != 0,
                         reinterpret_cast<float4*>(dsh));
}

void launch_preprocess_bwd(const PreBackwardArgs& a, hipStream_t s) {
    if (a.n == 0) return;
    hipLaunchKernelGGL(k_preprocess_bwd, dim3((a.n + kPreBwdBlock - 1) / kPreBwdBlock), dim3(kPreBwdBlock), 0, s, a);
}

}  // namespace tpd

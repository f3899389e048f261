// Chain rule from one view's screen-space splat gradients to the Gaussian's parameters, shared by
// the plain and the batched preprocess-backward kernels.  Differentiates the preprocess of
// project.slang:33-90 as restated in project_gaussian (tpd_device.hpp): the conic inverse, the
// EWA projection T Sigma T^T with T = J W (no clamp of x/z, y/z), the covariance (R(q) S)(R(q) S)^T
// with the unnormalised quaternion (volume.slang:28-35), the projection of the mean, and the SH
// colour (band-3 "+ feature 12" typo included, common.slang:69) through max(colour, 0).
#pragma once

#include "tpd_device.hpp"

namespace tpd {

// gr[q] += the tiles' partial sums src[t * 9 + q], t = 0 .. tiles - 1, in that order (the sums are
// the same floats as a loop over t).  Four tiles' 36 loads are issued together: one at a time, each
// pair waited for its loads, and a lane walked as many pairs as the largest footprint of its wave.
__device__ __forceinline__ void sum_pair_partials(const float* __restrict__ src, uint32_t tiles, float gr[9]) {
    uint32_t t = 0;
    for (; t + 4u <= tiles; t += 4u) {
        float x[36];
#pragma unroll
        for (int k = 0; k < 36; ++k) x[k] = src[t * 9u + (uint32_t)k];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int q = 0; q < 9; ++q) gr[q] += x[u * 9 + q];
    }
    for (; t < tiles; ++t)
#pragma unroll
        for (int q = 0; q < 9; ++q) gr[q] += src[t * 9u + (uint32_t)q];
}

// Adds this view's contribution to gm (means), gls (log-scales), gq (quaternion), gop (opacity
// logit) and gsh (48 SH coefficients).  g is the Gaussian record (60 floats), VM / PVM the entity's
// view-model and projection-view-model matrices, gr the summed dL/d(image x, y, conic a, b, c,
// r, g, b, opacity) of the splat in this view.
__device__ __forceinline__ void accumulate_param_grads(const float* g, const float* VM, const float* PVM,
                                                       const CameraParams& cam, uint32_t sh_degree, const float gr[9],
                                                       float gm[3], float gls[3], float gq[4], float& gop,
                                                       float gsh[48]) {
    const float* V = cam.view;
    const float W = (float)cam.width, H = (float)cam.height;
    const float fx = (0.5f * W) * cam.focal_ndc[0], fy = (0.5f * H) * cam.focal_ndc[1];
    const float gpx = gr[0], gpy = gr[1], gA = gr[2], gB = gr[3], gC = gr[4];
    const float gcol[3] = {gr[5], gr[6], gr[7]};
    const float gopa = gr[8];

    const float mx = g[0], my = g[1], mz = g[2];
    // forward recompute
    const float vx = row_dot_p1(VM, 0, mx, my, mz), vy = row_dot_p1(VM, 1, mx, my, mz),
                vz = row_dot_p1(VM, 2, mx, my, mz);
    const float cx = row_dot_p1(PVM, 0, mx, my, mz), cy = row_dot_p1(PVM, 1, mx, my, mz),
                cw = row_dot_p1(PVM, 3, mx, my, mz);
    const float inv_w = 1.0f / cw;
    const float mod = g[11];
    const float s[3] = {g[8] * mod, g[9] * mod, g[10] * mod};
    const float qx = g[4], qy = g[5], qz = g[6], qw = g[7];
    const float R[9] = {
        1.0f - 2.0f * (qy * qy + qz * qz), 2.0f * (qx * qy - qw * qz), 2.0f * (qx * qz + qw * qy),
        2.0f * (qx * qy + qw * qz), 1.0f - 2.0f * (qx * qx + qz * qz), 2.0f * (qy * qz - qw * qx),
        2.0f * (qx * qz - qw * qy), 2.0f * (qy * qz + qw * qx), 1.0f - 2.0f * (qx * qx + qy * qy)};
    float M[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) M[r * 3 + c] = R[r * 3 + c] * s[c];
    float Cv[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            Cv[r * 3 + c] = M[r * 3 + 0] * M[c * 3 + 0] + M[r * 3 + 1] * M[c * 3 + 1] + M[r * 3 + 2] * M[c * 3 + 2];
    const float fxz = fx / vz, fyz = fy / vz;
    const float j02 = -fxz * (vx / vz), j12 = -fyz * (vy / vz);
    float Tm[6];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        Tm[0 * 3 + c] = fxz * V[0 * 4 + c] + j02 * V[2 * 4 + c];
        Tm[1 * 3 + c] = fyz * V[1 * 4 + c] + j12 * V[2 * 4 + c];
    }
    // Sigma2 = T C T^T
    float TC[6];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            TC[r * 3 + c] = Tm[r * 3 + 0] * Cv[0 * 3 + c] + Tm[r * 3 + 1] * Cv[1 * 3 + c] + Tm[r * 3 + 2] * Cv[2 * 3 + c];
    const float s00 = TC[0] * Tm[0] + TC[1] * Tm[1] + TC[2] * Tm[2];
    const float s01 = TC[0] * Tm[3] + TC[1] * Tm[4] + TC[2] * Tm[5];
    const float s11 = TC[3] * Tm[3] + TC[4] * Tm[4] + TC[5] * Tm[5];
    const float ca = s00 + 0.3f, cb = s01, cc = s11 + 0.3f;
    const float det = ca * cc - cb * cb;
    const float ia = cc / det, ib = -cb / det, ic = ca / det;  // conic

    // conic -> 2D covariance: dSigma = -A dA A with A = [[ia, ib], [ib, ic]], dA = [[gA, gB/2], [gB/2, gC]]
    const float hB = 0.5f * gB;
    const float t00 = ia * gA + ib * hB, t01 = ia * hB + ib * gC;
    const float t10 = ib * gA + ic * hB, t11 = ib * hB + ic * gC;
    const float G00 = -(t00 * ia + t01 * ib);
    const float G01 = -(t00 * ib + t01 * ic);
    const float G11 = -(t10 * ib + t11 * ic);
    // dL/dC = T^T G T (3x3), dL/dT = 2 G T C (2x3)
    float GT[6];  // G T
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        GT[0 * 3 + c] = G00 * Tm[c] + G01 * Tm[3 + c];
        GT[1 * 3 + c] = G01 * Tm[c] + G11 * Tm[3 + c];
    }
    float gCv[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) gCv[r * 3 + c] = Tm[r] * GT[c] + Tm[3 + r] * GT[3 + c];
    float gT[6];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            gT[r * 3 + c] = 2.0f * (GT[r * 3 + 0] * Cv[0 * 3 + c] + GT[r * 3 + 1] * Cv[1 * 3 + c] + GT[r * 3 + 2] * Cv[2 * 3 + c]);
    // T = J W3: dL/dJ = dL/dT W3^T (entries J00, J02, J11, J12)
    const float gJ00 = gT[0] * V[0] + gT[1] * V[1] + gT[2] * V[2];
    const float gJ02 = gT[0] * V[8] + gT[1] * V[9] + gT[2] * V[10];
    const float gJ11 = gT[3] * V[4] + gT[4] * V[5] + gT[5] * V[6];
    const float gJ12 = gT[3] * V[8] + gT[4] * V[9] + gT[5] * V[10];
    const float iz = 1.0f / vz, iz2 = iz * iz, iz3 = iz2 * iz;
    const float gvx = -fx * iz2 * gJ02;
    const float gvy = -fy * iz2 * gJ12;
    const float gvz = -fx * iz2 * gJ00 - fy * iz2 * gJ11 + 2.0f * fx * vx * iz3 * gJ02 + 2.0f * fy * vy * iz3 * gJ12;
    // view position -> mean (VM 3x3 transposed)
#pragma unroll
    for (int c = 0; c < 3; ++c) gm[c] += VM[0 * 4 + c] * gvx + VM[1 * 4 + c] * gvy + VM[2 * 4 + c] * gvz;
    // image position -> clip -> mean
    const float gnx = gpx * (0.5f * W), gny = gpy * (0.5f * H);
    const float gcx = gnx * inv_w, gcy = gny * inv_w;
    const float gcw = -(gnx * cx + gny * cy) * inv_w * inv_w;
#pragma unroll
    for (int c = 0; c < 3; ++c) gm[c] += PVM[0 * 4 + c] * gcx + PVM[1 * 4 + c] * gcy + PVM[3 * 4 + c] * gcw;
    // Sigma3 = M M^T -> dL/dM = 2 dL/dC M
    float gM[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            gM[r * 3 + c] = 2.0f * (gCv[r * 3 + 0] * M[0 * 3 + c] + gCv[r * 3 + 1] * M[1 * 3 + c] + gCv[r * 3 + 2] * M[2 * 3 + c]);
    float gR[9];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float gs = gM[0 * 3 + c] * R[0 * 3 + c] + gM[1 * 3 + c] * R[1 * 3 + c] + gM[2 * 3 + c] * R[2 * 3 + c];
        gls[c] += gs * mod * g[8 + c];  // d scale / d log-scale = scale
#pragma unroll
        for (int r = 0; r < 3; ++r) gR[r * 3 + c] = gM[r * 3 + c] * s[c];
    }
    // R(q) (unnormalised quaternion, volume.slang:28-35)
    gq[0] += 2.0f * (qy * (gR[1] + gR[3]) + qz * (gR[2] + gR[6]) - 2.0f * qx * (gR[4] + gR[8]) + qw * (gR[7] - gR[5]));
    gq[1] += 2.0f * (qx * (gR[1] + gR[3]) + qz * (gR[5] + gR[7]) - 2.0f * qy * (gR[0] + gR[8]) + qw * (gR[2] - gR[6]));
    gq[2] += 2.0f * (qx * (gR[2] + gR[6]) + qy * (gR[5] + gR[7]) - 2.0f * qz * (gR[0] + gR[4]) + qw * (gR[3] - gR[1]));
    gq[3] += 2.0f * (qx * (gR[7] - gR[5]) + qy * (gR[2] - gR[6]) + qz * (gR[3] - gR[1]));
    // opacity logit
    const float o = g[3];
    gop += gopa * o * (1.0f - o);

    // colour: SH of dir = normalize(mean - camera position)
    const float camx = -((V[0] * V[3] + V[4] * V[7]) + V[8] * V[11]);
    const float camy = -((V[1] * V[3] + V[5] * V[7]) + V[9] * V[11]);
    const float camz = -((V[2] * V[3] + V[6] * V[7]) + V[10] * V[11]);
    const float dx = mx - camx, dy = my - camy, dz = mz - camz;
    const float len = __builtin_sqrtf((dx * dx + dy * dy) + dz * dz);
    const float il = 1.0f / len;
    const float x = dx * il, y = dy * il, z = dz * il;
    float col[3];
    eval_sh(g + 12, x, y, z, sh_degree, col);
    const float* sh = g + 12;
    float gdx = 0.f, gdy = 0.f, gdz = 0.f;
    const float C0 = 0.28209479177387814f, C1 = 0.4886025119029199f;
    const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, zx = z * x;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float gc = col[c] > 0.0f ? gcol[c] : 0.0f;  // max(colour, 0)
        if (gc == 0.0f) continue;
        const float* f = sh + 3 + c * 15;
        float* gf = gsh + 3 + c * 15;
        gsh[c] += C0 * gc;
        if (sh_degree > 0) {
            gf[0] += -C1 * y * gc;
            gf[1] += C1 * z * gc;
            gf[2] += -C1 * x * gc;
            gdx += -C1 * f[2] * gc;
            gdy += -C1 * f[0] * gc;
            gdz += C1 * f[1] * gc;
            if (sh_degree > 1) {
                const float c20 = 1.0925484305920792f, c21 = -1.0925484305920792f, c22 = 0.31539156525252005f,
                            c23 = -1.0925484305920792f, c24 = 0.5462742152960396f;
                gf[3] += c20 * xy * gc;
                gf[4] += c21 * yz * gc;
                gf[5] += c22 * (2.0f * zz - xx - yy) * gc;
                gf[6] += c23 * zx * gc;
                gf[7] += c24 * (xx - yy) * gc;
                gdx += gc * (c20 * y * f[3] - 2.0f * c22 * x * f[5] + c23 * z * f[6] + 2.0f * c24 * x * f[7]);
                gdy += gc * (c20 * x * f[3] + c21 * z * f[4] - 2.0f * c22 * y * f[5] - 2.0f * c24 * y * f[7]);
                gdz += gc * (c21 * y * f[4] + 4.0f * c22 * z * f[5] + c23 * x * f[6]);
                if (sh_degree > 2) {
                    const float c30 = -0.5900435899266435f, c31 = 2.890611442640554f, c32 = -0.4570457994644658f,
                                c33 = 0.3731763325901154f, c34 = -0.4570457994644658f, c35 = 1.445305721320277f,
                                c36 = -0.5900435899266435f;
                    gf[8] += c30 * y * (3.0f * xx - yy) * gc;
                    gf[9] += c31 * xy * z * gc;
                    gf[10] += c32 * y * (4.0f * zz - xx - yy) * gc;
                    gf[11] += gc;  // "+ feature 12" (common.slang:69)
                    gf[12] += c34 * x * (4.0f * zz - xx - yy) * gc;
                    gf[13] += c35 * z * (xx - yy) * gc;
                    gf[14] += c36 * x * (xx - 3.0f * yy) * gc;
                    gdx += gc * (6.0f * c30 * xy * f[8] + c31 * yz * f[9] - 2.0f * c32 * xy * f[10] - 6.0f * c33 * zx +
                                 c34 * (4.0f * zz - 3.0f * xx - yy) * f[12] + 2.0f * c35 * zx * f[13] +
                                 c36 * (3.0f * xx - 3.0f * yy) * f[14]);
                    gdy += gc * (c30 * (3.0f * xx - 3.0f * yy) * f[8] + c31 * zx * f[9] +
                                 c32 * (4.0f * zz - xx - 3.0f * yy) * f[10] - 6.0f * c33 * yz - 2.0f * c34 * xy * f[12] -
                                 2.0f * c35 * yz * f[13] - 6.0f * c36 * xy * f[14]);
                    gdz += gc * (c31 * xy * f[9] + 8.0f * c32 * yz * f[10] + c33 * (6.0f * zz - 3.0f * xx - 3.0f * yy) +
                                 8.0f * c34 * zx * f[12] + c35 * (xx - yy) * f[13]);
                }
            }
        }
    }
    // dir = d / |d|
    const float dot = x * gdx + y * gdy + z * gdz;
    gm[0] += (gdx - x * dot) * il;
    gm[1] += (gdy - y * dot) * il;
    gm[2] += (gdz - z * dot) * il;
}

// Writes (or adds, when accumulate) one Gaussian's parameter gradients but the SH ones, which the
// parameter-gradient passes store for a whole block at once (store_sh_grads_block).
__device__ __forceinline__ void store_param_grads_but_sh(uint32_t i, bool accumulate, const float gm[3],
                                                         const float gls[3], const float gq[4], float gop,
                                                         float* grad_means, float* grad_log_scales, float* grad_quats,
                                                         float* grad_opacity_logits) {
    float* om = grad_means + (size_t)i * 3;
    float* os = grad_log_scales + (size_t)i * 3;
    float* oq = grad_quats + (size_t)i * 4;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        om[c] = accumulate ? om[c] + gm[c] : gm[c];
        os[c] = accumulate ? os[c] + gls[c] : gls[c];
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) oq[c] = accumulate ? oq[c] + gq[c] : gq[c];
    grad_opacity_logits[i] = accumulate ? grad_opacity_logits[i] + gop : gop;
}

// The SH gradients of n consecutive Gaussians (48 floats each, src in LDS, Gaussian-major) to dst,
// as consecutive 16-byte stores by the whole block (accumulate: dst + src).
__device__ __forceinline__ void store_sh_grads_block(const float4* src, uint32_t n, bool accumulate, float4* dst) {
    for (uint32_t x = threadIdx.x; x < n * 12u; x += blockDim.x) {
        float4 v = src[x];
        if (accumulate) {
            const float4 o = dst[x];
            v = make_float4(o.x + v.x, o.y + v.y, o.z + v.z, o.w + v.w);
        }
        dst[x] = v;
    }
}

}  // namespace tpd

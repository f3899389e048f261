// Backward pass of the forward frame.  The reference has no backward (torpedo/differential is
// an empty stub, GaussianRasterizer.h:1-8); this differentiates the forward of project.slang and
// blend.slang as restated in k_forward.hip:
//   k_blend_bwd       per tile: re-runs the front-to-back blend to find each pixel's final
//                     transmittance and last contributor (nothing is stored by the forward), then
//                     walks the tile's splats back to front recovering T_i = T_{i+1} / (1 - a_i),
//                     producing per splat dL/d(image x, y), dL/d(conic a, b, c), dL/d(rgb),
//                     dL/d(opacity); wave sums are combined in LDS and added to the splat's
//                     global accumulator once per tile.
//   k_preprocess_bwd  per Gaussian: chain rule through the conic inverse, the EWA projection
//                     T Sigma T^T, the Jacobian, R(q) S, the projection of the mean and the SH
//                     colour (band-3 typo included), to means, log-scales, quaternions, opacity
//                     logits and SH coefficients.
// Non-differentiable decisions (culling, tile rectangles, alpha < 1/255 skips, early
// termination, the 0.99 clamp, max(colour, 0)) pass zero gradient, as the forward's branches do.
#include "tpd_kernels.hpp"

namespace tpd {

// ------------------------------------------------------------------ blend backward

__global__ __launch_bounds__(kTilePixels) void k_blend_bwd(const BackwardArgs a) {
    constexpr uint32_t kWaves = kTilePixels / kWave;
    constexpr uint32_t kQuad = 8;
    constexpr uint32_t kG = 9;  // gradient components per splat
    __shared__ float4 s_a[kTilePixels];  // x, y, conic a, conic b
    __shared__ float4 s_b[kTilePixels];  // conic c, opacity, skip threshold, -
    __shared__ float4 s_c[kTilePixels];  // colour
    __shared__ uint32_t s_g[kTilePixels];  // Gaussian index
    __shared__ uint8_t s_list[kWaves][kTilePixels];
    __shared__ uint32_t s_cnt[kWaves][kWaves];
    __shared__ float s_acc[kTilePixels][kG];
    __shared__ int32_t s_last;

    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t tx = tile % a.gx, ty = tile / a.gx;
    const uint32_t tid = threadIdx.x, wave = tid / kWave, lane = __lane_id();
    const uint32_t pxl = tx * kTile + (wave & 1u) * kQuad + (lane % kQuad);
    const uint32_t pyl = ty * kTile + (wave >> 1) * kQuad + (lane / kQuad);
    const bool inside = pxl < a.width && pyl < a.height;
    const float fx = (float)pxl, fy = (float)pyl;
    const uint2 range = a.ranges[tile];
    const uint64_t lt = lanemask_lt();
    if (tid == 0) s_last = -1;
    for (uint32_t x = tid; x < kTilePixels * kG; x += kTilePixels) (&s_acc[0][0])[x] = 0.0f;

    // Stage chunk [base, base + 256) of the tile list and build the quadrant lists; returns this
    // wave's list length.  Identical to the forward's staging.
    auto stage = [&](uint32_t base) -> uint32_t {
        const uint32_t k = base + tid;
        uint32_t mask = 0;
        if (k < range.y) {
            const uint32_t g = a.vals[k];
            const float4 p = a.xyzr[g];
            const float4 c = a.conop[g];
            s_a[tid] = make_float4(p.x, p.y, c.x, c.y);
            s_b[tid] = make_float4(c.z, c.w, __logf(1.0f / (255.0f * c.w)) - 0.01f, 0.0f);
            s_c[tid] = a.rgb[g];
            s_g[tid] = g;
            float ex = __builtin_inff(), ey = __builtin_inff();
            const float det = c.x * c.z - c.y * c.y;
            const float q0 = (__logf(255.0f * c.w) + 0.05f) * (2.0f / 0.99f);
            if (c.x > 0.0f && c.z > 0.0f && c.y * c.y < 0.98f * (c.x * c.z) && q0 >= 0.0f) {
                ex = 1.01f * __builtin_sqrtf(q0 * c.z / det) + 1e-3f;
                ey = 1.01f * __builtin_sqrtf(q0 * c.x / det) + 1e-3f;
            }
#pragma unroll
            for (uint32_t q = 0; q < kWaves; ++q) {
                const uint32_t x0 = tx * kTile + (q & 1u) * kQuad, y0 = ty * kTile + (q >> 1) * kQuad;
                if (x0 >= a.width || y0 >= a.height) continue;
                const float fx0 = (float)x0, fy0 = (float)y0;
                const float fx1 = (float)(min(x0 + kQuad, a.width) - 1), fy1 = (float)(min(y0 + kQuad, a.height) - 1);
                if (p.x + ex >= fx0 && p.x - ex <= fx1 && p.y + ey >= fy0 && p.y - ey <= fy1 &&
                    splat_touches_rect(p.x, p.y, c.x, c.y, c.z, c.w, fx0, fx1, fy0, fy1))
                    mask |= 1u << q;
            }
        }
        uint64_t bal[kWaves];
#pragma unroll
        for (uint32_t q = 0; q < kWaves; ++q) bal[q] = __ballot((mask >> q) & 1u);
        if (lane < kWaves) {
            uint64_t b = bal[0];
#pragma unroll
            for (uint32_t q = 1; q < kWaves; ++q) b = lane == q ? bal[q] : b;
            s_cnt[wave][lane] = (uint32_t)__popcll(b);
        }
        __syncthreads();
#pragma unroll
        for (uint32_t q = 0; q < kWaves; ++q) {
            if ((mask >> q) & 1u) {
                uint32_t off = (uint32_t)__popcll(bal[q] & lt);
                for (uint32_t w = 0; w < wave; ++w) off += s_cnt[w][q];
                s_list[q][off] = (uint8_t)tid;
            }
        }
        uint32_t cnt = 0;
#pragma unroll
        for (uint32_t w = 0; w < kWaves; ++w) cnt += s_cnt[w][wave];
        __syncthreads();
        return cnt;
    };

    // ---- pass 1: forward re-run (same decisions as k_blend) -> T_final, last contributor
    float T = 1.0f;
    bool done = !inside;
    int32_t last = -1;
    for (uint32_t base = range.x; base < range.y; base += kTilePixels) {
        if (__syncthreads_count(done) == (int)kTilePixels) break;
        const uint32_t cnt = stage(base);
        if (__ballot(!done) == 0) continue;
        for (uint32_t i = 0; i < cnt; ++i) {
            const uint32_t j = s_list[wave][i];
            const float4 A = s_a[j];
            const float4 B = s_b[j];
            const float dx = A.x - fx, dy = A.y - fy;
            const float power = -0.5f * ((A.z * dx) * dx + (B.x * dy) * dy) - (A.w * dx) * dy;
            const bool act = !done && !(power > 0.0f) && !(power < B.z);
            if (__ballot(act) == 0) continue;
            float alpha = B.y * soft_expf(power);
            alpha = alpha < 0.99f ? alpha : 0.99f;
            const bool ok = act && !(alpha < 1.0f / 255.0f);
            const float tn = T * (1.0f - alpha);
            const bool stop = ok && tn < 0.0001f;
            const bool upd = ok && !(tn < 0.0001f);
            T = upd ? tn : T;
            last = upd ? (int32_t)(base + j) : last;
            done = done || stop;
            if (__ballot(stop) != 0 && __ballot(!done) == 0) break;
        }
    }
    if (last >= 0) atomicMax(&s_last, last);
    __syncthreads();
    const int32_t block_last = s_last;
    if (block_last < 0) return;  // uniform: no pixel of this tile blended anything

    // ---- pass 2: back to front
    float g_r = 0.0f, g_g = 0.0f, g_b = 0.0f;
    if (inside) {
        const size_t o = (size_t)pyl * a.width + pxl;
        g_r = a.dL_dimg[o * 3 + 0];
        g_g = a.dL_dimg[o * 3 + 1];
        g_b = a.dL_dimg[o * 3 + 2];
    }
    float acc_r = 0.0f, acc_g = 0.0f, acc_b = 0.0f;  // colour accumulated behind the current splat
    float last_alpha = 0.0f, last_r = 0.0f, last_g = 0.0f, last_b = 0.0f;
    const uint32_t first_chunk = ((uint32_t)block_last - range.x) / kTilePixels;
    for (int32_t chunk = (int32_t)first_chunk; chunk >= 0; --chunk) {
        const uint32_t base = range.x + (uint32_t)chunk * kTilePixels;
        __syncthreads();
        const uint32_t cnt = stage(base);
        for (int32_t i = (int32_t)cnt - 1; i >= 0; --i) {
            const uint32_t j = s_list[wave][i];
            const int32_t kk = (int32_t)(base + j);
            const bool mine = kk <= last;
            if (__ballot(mine) == 0) continue;
            const float4 A = s_a[j];
            const float4 B = s_b[j];
            const float dx = A.x - fx, dy = A.y - fy;
            const float power = -0.5f * ((A.z * dx) * dx + (B.x * dy) * dy) - (A.w * dx) * dy;
            const bool act = mine && !(power > 0.0f) && !(power < B.z);
            if (__ballot(act) == 0) continue;
            const float G = soft_expf(power);
            const float raw = B.y * G;
            const float alpha = raw < 0.99f ? raw : 0.99f;
            const bool upd = act && !(alpha < 1.0f / 255.0f);
            float v[kG];
#pragma unroll
            for (uint32_t q = 0; q < kG; ++q) v[q] = 0.0f;
            if (upd) {
                const float4 C = s_c[j];
                T = T / (1.0f - alpha);  // transmittance in front of this splat
                acc_r = last_alpha * last_r + (1.0f - last_alpha) * acc_r;
                acc_g = last_alpha * last_g + (1.0f - last_alpha) * acc_g;
                acc_b = last_alpha * last_b + (1.0f - last_alpha) * acc_b;
                const float w = alpha * T;
                v[5] = w * g_r;
                v[6] = w * g_g;
                v[7] = w * g_b;
                const float dL_dalpha = T * ((C.x - acc_r) * g_r + (C.y - acc_g) * g_g + (C.z - acc_b) * g_b);
                last_alpha = alpha;
                last_r = C.x; last_g = C.y; last_b = C.z;
                if (raw < 0.99f) {  // unclamped: alpha = o * G
                    v[8] = dL_dalpha * G;
                    const float dL_dpower = dL_dalpha * alpha;
                    v[0] = dL_dpower * (-A.z * dx - A.w * dy);
                    v[1] = dL_dpower * (-B.x * dy - A.w * dx);
                    v[2] = dL_dpower * (-0.5f * dx * dx);
                    v[3] = dL_dpower * (-dx * dy);
                    v[4] = dL_dpower * (-0.5f * dy * dy);
                }
            }
            if (__ballot(upd) != 0) {
#pragma unroll
                for (uint32_t q = 0; q < kG; ++q) {
                    const float sum = wave_sum_to_lane63(v[q]);
                    if (lane == 63 && sum != 0.0f) atomicAdd(&s_acc[j][q], sum);
                }
            }
        }
        __syncthreads();
        // Store the chunk's sums (slot tid) at the pair's pre-sort position: offsets[g] plus the
        // row-major rank of this tile in g's rectangle (keygen.slang:47-53 order).
        if (base + tid < range.y) {
            float vals[kG];
            bool any = false;
#pragma unroll
            for (uint32_t q = 0; q < kG; ++q) {
                vals[q] = s_acc[tid][q];
                any = any || vals[q] != 0.0f;
                s_acc[tid][q] = 0.0f;
            }
            if (any) {
                const uint32_t g = s_g[tid];
                const float4 p = a.xyzr[g];
                int32_t x0, y0, x1, y1;
                bounding_rect(p.x, p.y, p.w, a.gx, a.gy, x0, y0, x1, y1);
                const size_t pos = (size_t)a.offsets[g] + (size_t)((int32_t)ty - y0) * (size_t)(x1 - x0) +
                                   (size_t)((int32_t)tx - x0);
                float* dst = a.partials + pos * kG;
#pragma unroll
                for (uint32_t q = 0; q < kG; ++q) dst[q] = vals[q];
            }
        }
    }
}

void launch_blend_bwd(const BackwardArgs& a, hipStream_t s) {
    if (a.ntiles == 0) return;
    hipLaunchKernelGGL(k_blend_bwd, dim3(a.ntiles), dim3(kTilePixels), 0, s, a);
}

// ------------------------------------------------------------------ preprocess backward

__global__ __launch_bounds__(256) void k_preprocess_bwd(const PreBackwardArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const float* g = a.gauss + (size_t)i * 60;
    float gm[3] = {0.f, 0.f, 0.f}, gls[3] = {0.f, 0.f, 0.f}, gq[4] = {0.f, 0.f, 0.f, 0.f}, gop = 0.f;
    float gsh[48];
#pragma unroll
    for (int k = 0; k < 48; ++k) gsh[k] = 0.f;
    if (a.g2d && a.tiles[i] == 0)
#pragma unroll
        for (int q = 0; q < 9; ++q) a.g2d[(size_t)i * 9 + q] = 0.0f;

    if (a.tiles[i] > 0) {
        const uint32_t e = a.ent_idx ? a.ent_idx[i] : 0u;
        const float* VM = a.ent[e].vm;
        const float* PVM = a.ent[e].pvm;
        const float* V = a.cam.view;
        const float W = (float)a.cam.width, H = (float)a.cam.height;
        const float fx = (0.5f * W) * a.cam.focal_ndc[0], fy = (0.5f * H) * a.cam.focal_ndc[1];
        // sum the per-tile partials of this Gaussian (contiguous, in tile order: deterministic)
        float gr[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) gr[q] = 0.0f;
        const float* src = a.partials + (size_t)a.offsets[i] * 9;
        for (uint32_t t = 0; t < a.tiles[i]; ++t)
#pragma unroll
            for (int q = 0; q < 9; ++q) gr[q] += src[t * 9 + q];
        if (a.g2d)
#pragma unroll
            for (int q = 0; q < 9; ++q) a.g2d[(size_t)i * 9 + q] = gr[q];
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
            gls[c] = gs * mod * g[8 + c];  // d scale / d log-scale = scale
#pragma unroll
            for (int r = 0; r < 3; ++r) gR[r * 3 + c] = gM[r * 3 + c] * s[c];
        }
        // R(q) (unnormalised quaternion, volume.slang:28-35)
        gq[0] = 2.0f * (qy * (gR[1] + gR[3]) + qz * (gR[2] + gR[6]) - 2.0f * qx * (gR[4] + gR[8]) + qw * (gR[7] - gR[5]));
        gq[1] = 2.0f * (qx * (gR[1] + gR[3]) + qz * (gR[5] + gR[7]) - 2.0f * qy * (gR[0] + gR[8]) + qw * (gR[2] - gR[6]));
        gq[2] = 2.0f * (qx * (gR[2] + gR[6]) + qy * (gR[5] + gR[7]) - 2.0f * qz * (gR[0] + gR[4]) + qw * (gR[3] - gR[1]));
        gq[3] = 2.0f * (qx * (gR[7] - gR[5]) + qy * (gR[2] - gR[6]) + qz * (gR[3] - gR[1]));
        // opacity logit
        const float o = g[3];
        gop = gopa * o * (1.0f - o);

        // colour: SH of dir = normalize(mean - camera position)
        const float camx = -((V[0] * V[3] + V[4] * V[7]) + V[8] * V[11]);
        const float camy = -((V[1] * V[3] + V[5] * V[7]) + V[9] * V[11]);
        const float camz = -((V[2] * V[3] + V[6] * V[7]) + V[10] * V[11]);
        const float dx = mx - camx, dy = my - camy, dz = mz - camz;
        const float len = __builtin_sqrtf((dx * dx + dy * dy) + dz * dz);
        const float il = 1.0f / len;
        const float x = dx * il, y = dy * il, z = dz * il;
        float col[3];
        eval_sh(g + 12, x, y, z, a.sh_degree, col);
        const float* sh = g + 12;
        float gdx = 0.f, gdy = 0.f, gdz = 0.f;
        const float C0 = 0.28209479177387814f, C1 = 0.4886025119029199f;
        const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, zx = z * x;
        for (int c = 0; c < 3; ++c) {
            const float gc = col[c] > 0.0f ? gcol[c] : 0.0f;  // max(colour, 0)
            if (gc == 0.0f) continue;
            const float* f = sh + 3 + c * 15;
            float* gf = gsh + 3 + c * 15;
            gsh[c] = C0 * gc;
            if (a.sh_degree > 0) {
                gf[0] = -C1 * y * gc;
                gf[1] = C1 * z * gc;
                gf[2] = -C1 * x * gc;
                gdx += -C1 * f[2] * gc;
                gdy += -C1 * f[0] * gc;
                gdz += C1 * f[1] * gc;
                if (a.sh_degree > 1) {
                    const float c20 = 1.0925484305920792f, c21 = -1.0925484305920792f, c22 = 0.31539156525252005f,
                                c23 = -1.0925484305920792f, c24 = 0.5462742152960396f;
                    gf[3] = c20 * xy * gc;
                    gf[4] = c21 * yz * gc;
                    gf[5] = c22 * (2.0f * zz - xx - yy) * gc;
                    gf[6] = c23 * zx * gc;
                    gf[7] = c24 * (xx - yy) * gc;
                    gdx += gc * (c20 * y * f[3] - 2.0f * c22 * x * f[5] + c23 * z * f[6] + 2.0f * c24 * x * f[7]);
                    gdy += gc * (c20 * x * f[3] + c21 * z * f[4] - 2.0f * c22 * y * f[5] - 2.0f * c24 * y * f[7]);
                    gdz += gc * (c21 * y * f[4] + 4.0f * c22 * z * f[5] + c23 * x * f[6]);
                    if (a.sh_degree > 2) {
                        const float c30 = -0.5900435899266435f, c31 = 2.890611442640554f, c32 = -0.4570457994644658f,
                                    c33 = 0.3731763325901154f, c34 = -0.4570457994644658f, c35 = 1.445305721320277f,
                                    c36 = -0.5900435899266435f;
                        gf[8] = c30 * y * (3.0f * xx - yy) * gc;
                        gf[9] = c31 * xy * z * gc;
                        gf[10] = c32 * y * (4.0f * zz - xx - yy) * gc;
                        gf[11] = gc;  // "+ feature 12" (common.slang:69)
                        gf[12] = c34 * x * (4.0f * zz - xx - yy) * gc;
                        gf[13] = c35 * z * (xx - yy) * gc;
                        gf[14] = c36 * x * (xx - 3.0f * yy) * gc;
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
    float* om = a.grad_means + (size_t)i * 3;
    float* os = a.grad_log_scales + (size_t)i * 3;
    float* oq = a.grad_quats + (size_t)i * 4;
    const float acc = a.accumulate ? 1.0f : 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        om[c] = acc * om[c] + gm[c];
        os[c] = acc * os[c] + gls[c];
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) oq[c] = acc * oq[c] + gq[c];
    a.grad_opacity_logits[i] = acc * a.grad_opacity_logits[i] + gop;
    float* osh = a.grad_sh + (size_t)i * 48;
#pragma unroll
    for (int k = 0; k < 48; ++k) osh[k] = acc * osh[k] + gsh[k];
}

void launch_preprocess_bwd(const PreBackwardArgs& a, hipStream_t s) {
    if (a.n == 0) return;
    hipLaunchKernelGGL(k_preprocess_bwd, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
}

}  // namespace tpd

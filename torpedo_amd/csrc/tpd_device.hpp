// Device-side building blocks shared by the forward, backward and batched kernels.
//
// Every translation unit that includes this header is compiled with -ffp-contract=off and
// IEEE-rounded division and square root, so the float operation sequence written here is
// exactly the one executed.  That sequence follows the reference's Slang source operation by
// operation (files cited per function); the CPU oracle restates the same sequence, which makes
// splats, keys, ranges and pixels bit-identical between the two.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tpd {

constexpr uint32_t kTile = 16;                 // BLOCK_X = BLOCK_Y (splat.slang:8-9)
constexpr uint32_t kTilePixels = kTile * kTile;
constexpr uint32_t kPreBlock = 256;            // Gaussians per preprocess partition (WORKGROUP_SIZE)
constexpr uint32_t kWave = 64;

// Camera uniform block (splat.slang:18-22) plus the image size the reference reads from the target.
struct CameraParams {
    float view[16];      // world -> view, row-major
    float proj_view[16]; // world -> clip, row-major (projection * view, compensated on the host)
    float focal_ndc[2];  // (P[0][0], P[1][1])
    uint32_t width, height;
};

// One projected splat (splat.slang:33-39 holds the same quantities), 64 bytes: the blend and its
// backward fetch it whole, in one aligned memory segment, for each (tile, splat) pair.
struct alignas(64) Splat {
    float4 xyzr;   // image x, image y, view depth, radius
    float4 conop;  // conic a, b, c, opacity
    float4 rgb;    // colour rgb; SubBlockReach::inva
    float4 aux;    // the blend's skip threshold log(1 / (255 opacity)) - 0.01; SubBlockReach ey, v*, a q0
};

// Per-entity products view*model and projView*model, 32 floats per entity.
struct EntityMats {
    float vm[16];
    float pvm[16];
};

__device__ __forceinline__ float as_float(uint32_t u) { return __builtin_bit_cast(float, u); }
__device__ __forceinline__ uint32_t as_uint(float f) { return __builtin_bit_cast(uint32_t, f); }

// exp(x) for x in [-86, 88]: x = n ln2 + r with n = round(x / ln2) taken from the low bits of
// x log2(e) + 1.5 * 2^23 (one fused multiply-add rounds the product to an integer), r by a
// Cody-Waite two-constant split of ln 2 (exact), e^r by a degree-5 polynomial fitted for relative
// error on [-ln2/2, ln2/2] (Horner form with fused multiply-adds), and 2^n added into the result's
// exponent field (the result is a normal number over that range).  Relative error below
// 1.5 x 2^-23 against exp for every float in [-86, 0].  Only correctly rounded operations and an
// integer add, so the oracle's tpo_expf gives the same bits.  Outside [-86, 88] the result is
// garbage (never a fault): callers keep their arguments inside.
__device__ __forceinline__ float soft_expf_inrange(float x) {
    const float t = __builtin_fmaf(x, 1.44269504088896341f, 12582912.0f);
    const float n = t - 12582912.0f;
    float r = __builtin_fmaf(n, -0.693145751953125f, x);
    r = __builtin_fmaf(n, -1.42860682030941723212e-6f, r);
    float p = 0.008286180905997753f;
    p = __builtin_fmaf(p, r, 0.04189886897802353f);
    p = __builtin_fmaf(p, r, 0.16667655110359192f);
    p = __builtin_fmaf(p, r, 0.49999144673347473f);
    p = __builtin_fmaf(p, r, 0.9999997615814209f);
    p = __builtin_fmaf(p, r, 1.0f);
    return as_float(as_uint(p) + (as_uint(t) << 23));
}

// Float -> UNORM8 for the RGBA8 target (blend.slang:104): clamp to [0, 1], scale by 255, round to
// nearest even.
__device__ __forceinline__ uint8_t unorm8(float v) {
    if (!(v > 0.0f)) return 0;
    if (v > 1.0f) v = 1.0f;
    return (uint8_t)__builtin_rintf(v * 255.0f);
}

// Saturating float -> int32 truncation (NaN -> 0), written out so that the behaviour does not
// depend on the conversion instruction.
__device__ __forceinline__ int32_t sat_i32(float v) {
    if (v != v) return 0;
    if (v >= 2147483648.0f) return 2147483647;
    if (v <= -2147483648.0f) return (int32_t)0x80000000u;
    return (int32_t)v;
}

__device__ __forceinline__ int32_t clamp_grid(int32_t v, uint32_t g) {
    v = v < 0 ? 0 : v;
    return v > (int32_t)g ? (int32_t)g : v;
}

// volume.slang:12-17 getBoundingRect.
__device__ __forceinline__ void bounding_rect(float px, float py, float radius, uint32_t gx, uint32_t gy,
                                              int32_t& x0, int32_t& y0, int32_t& x1, int32_t& y1) {
    const float b = 16.0f;
    x0 = clamp_grid(sat_i32((px - radius) / b), gx);
    y0 = clamp_grid(sat_i32((py - radius) / b), gy);
    x1 = clamp_grid(sat_i32((((px + radius) + b) - 1.0f) / b), gx);
    y1 = clamp_grid(sat_i32((((py + radius) + b) - 1.0f) / b), gy);
}

__device__ __forceinline__ float row_dot_p1(const float* m, int i, float x, float y, float z) {
    return ((m[i * 4 + 0] * x + m[i * 4 + 1] * y) + m[i * 4 + 2] * z) + m[i * 4 + 3];
}

// splat/common.slang:35-80 (band-3 "+ feature 12" typo included).  sh points at the 48-float
// coefficient block; dir is the unit view direction.
__device__ __forceinline__ void eval_sh(const float* sh, float x, float y, float z, uint32_t degree, float* out) {
    const float C0 = 0.28209479177387814f;
    const float C1 = 0.4886025119029199f;
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c] = C0 * sh[c];
    if (degree > 0) {
        const float a = C1 * y, b = C1 * z, d = C1 * x;
#pragma unroll
        for (int c = 0; c < 3; ++c)
            out[c] = ((out[c] - a * sh[3 + c * 15 + 0]) + b * sh[3 + c * 15 + 1]) - d * sh[3 + c * 15 + 2];
        if (degree > 1) {
            const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, zx = z * x;
            const float b4 = 1.0925484305920792f * xy, b5 = -1.0925484305920792f * yz,
                        b6 = 0.31539156525252005f * ((2.0f * zz - xx) - yy), b7 = -1.0925484305920792f * zx,
                        b8 = 0.5462742152960396f * (xx - yy);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float* f = sh + 3 + c * 15;
                out[c] = ((((out[c] + b4 * f[3]) + b5 * f[4]) + b6 * f[5]) + b7 * f[6]) + b8 * f[7];
            }
            if (degree > 2) {
                const float b9 = (-0.5900435899266435f * y) * (3.0f * xx - yy);
                const float b10 = (2.890611442640554f * xy) * z;
                const float b11 = (-0.4570457994644658f * y) * ((4.0f * zz - xx) - yy);
                const float s12 = (0.3731763325901154f * z) * ((2.0f * zz - 3.0f * xx) - 3.0f * yy);
                const float b13 = (-0.4570457994644658f * x) * ((4.0f * zz - xx) - yy);
                const float b14 = (1.445305721320277f * z) * (xx - yy);
                const float b15 = (-0.5900435899266435f * x) * (xx - 3.0f * yy);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float* f = sh + 3 + c * 15;
                    out[c] = ((((((((out[c] + b9 * f[8]) + b10 * f[9]) + b11 * f[10]) + s12) + f[11]) + b13 * f[12])
                               + b14 * f[13]) + b15 * f[14]);
                }
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float v = out[c] + 0.5f;
        out[c] = v > 0.0f ? v : 0.0f;
    }
}

struct Projected {
    float px, py, vz, radius;
    float con_a, con_b, con_c, opacity;
    float r, g, b;
    uint32_t tiles;
    int32_t x0, y0, x1, y1;
};

// volume.slang:20-41 computeCovariance: the 3D covariance M M^T, M = R(q) diag(scale * modifier),
// of record g.  Its products commute exactly, so (M M^T)[i][j] and [j][i] are the same float and
// six values hold it.
struct Cov3 {
    float c00, c01, c02, c11, c12, c22;
};
__device__ __forceinline__ Cov3 covariance3d(const float* __restrict__ g) {
    const float sx = g[8] * g[11], sy = g[9] * g[11], sz = g[10] * g[11];
    const float qx = g[4], qy = g[5], qz = g[6], qw = g[7];
    const float R00 = 1.0f - 2.0f * (qy * qy + qz * qz), R01 = 2.0f * (qx * qy - qw * qz),
                R02 = 2.0f * (qx * qz + qw * qy);
    const float R10 = 2.0f * (qx * qy + qw * qz), R11 = 1.0f - 2.0f * (qx * qx + qz * qz),
                R12 = 2.0f * (qy * qz - qw * qx);
    const float R20 = 2.0f * (qx * qz - qw * qy), R21 = 2.0f * (qy * qz + qw * qx),
                R22 = 1.0f - 2.0f * (qx * qx + qy * qy);
    const float M00 = R00 * sx, M01 = R01 * sy, M02 = R02 * sz;
    const float M10 = R10 * sx, M11 = R11 * sy, M12 = R12 * sz;
    const float M20 = R20 * sx, M21 = R21 * sy, M22 = R22 * sz;
    Cov3 c;
    c.c00 = (M00 * M00 + M01 * M01) + M02 * M02;
    c.c01 = (M00 * M10 + M01 * M11) + M02 * M12;
    c.c02 = (M00 * M20 + M01 * M21) + M02 * M22;
    c.c11 = (M10 * M10 + M11 * M11) + M12 * M12;
    c.c12 = (M10 * M20 + M11 * M21) + M12 * M22;
    c.c22 = (M20 * M20 + M21 * M21) + M22 * M22;
    return c;
}

// project.slang:33-90 for the Gaussian record g (60 floats: position, opacity, quaternion xyzw,
// scale xyz + modifier, sh[48]).  Returns false for culled / degenerate / tile-less Gaussians,
// in which case p.tiles = 0 and p.radius = 0.  cov_of() gives the record's 3D covariance (computed
// here, past the frustum test, or once for all the views of a batch).
template <class CovOf>
__device__ __forceinline__ bool project_gaussian_cov(const float* __restrict__ g, CovOf cov_of, const float* vm,
                                                     const float* pvm, const float* V, float focal_x, float focal_y,
                                                     uint32_t W, uint32_t H, uint32_t gx, uint32_t gy,
                                                     uint32_t sh_degree, Projected& p) {
    p.tiles = 0;
    p.radius = 0.0f;
    const float mx = g[0], my = g[1], mz = g[2];
    // splat/common.slang:98-119 passFrustumClipping
    const float vx = row_dot_p1(vm, 0, mx, my, mz), vy = row_dot_p1(vm, 1, mx, my, mz),
                vz = row_dot_p1(vm, 2, mx, my, mz);
    if (vz <= 0.0f) return false;
    const float cx = row_dot_p1(pvm, 0, mx, my, mz), cy = row_dot_p1(pvm, 1, mx, my, mz),
                cz = row_dot_p1(pvm, 2, mx, my, mz), cw = row_dot_p1(pvm, 3, mx, my, mz);
    if (cx < -1.3f * cw || cx > 1.3f * cw) return false;
    if (cy < -1.3f * cw || cy > 1.3f * cw) return false;
    if (cz < 0.0f || cz > cw) return false;
    const float inv_w = 1.0f / cw;
    const float ndc_x = cx * inv_w, ndc_y = cy * inv_w;

    const Cov3 cv = cov_of();
    const float C00 = cv.c00, C01 = cv.c01, C02 = cv.c02;
    const float C10 = cv.c01, C11 = cv.c11, C12 = cv.c12;
    const float C20 = cv.c02, C21 = cv.c12, C22 = cv.c22;

    // volume.slang:44-63 projectCovariance + computeJacobian (no clamp of x/z, y/z)
    const float fxz = focal_x / vz, fyz = focal_y / vz;
    const float tx = vx / vz, ty = vy / vz;
    const float j02 = -fxz * tx, j12 = -fyz * ty;
    const float T00 = fxz * V[0] + j02 * V[8], T01 = fxz * V[1] + j02 * V[9], T02 = fxz * V[2] + j02 * V[10];
    const float T10 = fyz * V[4] + j12 * V[8], T11 = fyz * V[5] + j12 * V[9], T12 = fyz * V[6] + j12 * V[10];
    // (C T^T)[k][s]
    const float A00 = (C00 * T00 + C01 * T01) + C02 * T02, A01 = (C00 * T10 + C01 * T11) + C02 * T12;
    const float A10 = (C10 * T00 + C11 * T01) + C12 * T02, A11 = (C10 * T10 + C11 * T11) + C12 * T12;
    const float A20 = (C20 * T00 + C21 * T01) + C22 * T02, A21 = (C20 * T10 + C21 * T11) + C22 * T12;
    const float c00 = (T00 * A00 + T01 * A10) + T02 * A20;
    const float c10 = (T10 * A00 + T11 * A10) + T12 * A20;
    const float c11 = (T10 * A01 + T11 * A11) + T12 * A21;

    // project.slang:57-70
    const float ca = c00 + 0.3f, cb = c10, cc = c11 + 0.3f;
    const float det = ca * cc - cb * cb;
    if (det == 0.0f) return false;
    const float det_inv = 1.0f / det;
    const float mid = 0.5f * (ca + cc);
    float disc = mid * mid - det;
    disc = disc > 0.1f ? disc : 0.1f;
    const float sq = __builtin_sqrtf(disc);
    const float l1 = mid + sq, l2 = mid - sq;
    const float radius = __builtin_ceilf(3.0f * __builtin_sqrtf(l1 > l2 ? l1 : l2));

    // volume.slang:3-8 ndc2pix, then the tile rectangle
    const float px = ((ndc_x + 1.0f) * (float)W - 1.0f) * 0.5f;
    const float py = ((ndc_y + 1.0f) * (float)H - 1.0f) * 0.5f;
    int32_t x0, y0, x1, y1;
    bounding_rect(px, py, radius, gx, gy, x0, y0, x1, y1);
    const uint32_t tiles = (uint32_t)(x1 - x0) * (uint32_t)(y1 - y0);
    if (tiles == 0) return false;

    // common.slang:88-92 camera position, colour from the untransformed mean (project.slang:75-76)
    const float camx = -((V[0] * V[3] + V[4] * V[7]) + V[8] * V[11]);
    const float camy = -((V[1] * V[3] + V[5] * V[7]) + V[9] * V[11]);
    const float camz = -((V[2] * V[3] + V[6] * V[7]) + V[10] * V[11]);
    const float dx = mx - camx, dy = my - camy, dz = mz - camz;
    const float len = __builtin_sqrtf((dx * dx + dy * dy) + dz * dz);
    const float inv_len = 1.0f / len;
    float col[3];
    eval_sh(g + 12, dx * inv_len, dy * inv_len, dz * inv_len, sh_degree, col);

    p.px = px; p.py = py; p.vz = vz; p.radius = radius;
    p.con_a = cc * det_inv; p.con_b = -cb * det_inv; p.con_c = ca * det_inv; p.opacity = g[3];
    p.r = col[0]; p.g = col[1]; p.b = col[2];
    p.tiles = tiles;
    p.x0 = x0; p.y0 = y0; p.x1 = x1; p.y1 = y1;
    return true;
}
__device__ __forceinline__ bool project_gaussian(const float* __restrict__ g, const float* vm, const float* pvm,
                                                 const float* V, float focal_x, float focal_y, uint32_t W, uint32_t H,
                                                 uint32_t gx, uint32_t gy, uint32_t sh_degree, Projected& p) {
    return project_gaussian_cov(g, [g] { return covariance3d(g); }, vm, pvm, V, focal_x, focal_y, W, H, gx, gy,
                                sh_degree, p);
}

// Which 4x4-pixel sub-blocks of the 16x16 tile with top-left pixel (ox, oy) a splat may reach:
// bit 4 * row + column is set unless no pixel centre of that sub-block can pass the blend's
// alpha >= 1/255 test, so dropping the splat from the other sub-blocks' lists leaves every pixel
// bit-identical.
//
// power = -Q/2 with Q(u, v) = a u^2 + 2 b u v + c v^2, (u, v) = pixel - centre.  A pixel is
// certainly skipped when 0.99 * Q/2 > ln(255 o) + 0.05: with |b| < 0.99 sqrt(ac) the float error
// of the per-pixel power stays below 2e-4 relative, so its power is then below -ln(255 o) - 0.04
// and o * exp(power) < 1/255 (exp is accurate to 2.5 ulp).  The mask covers the region
// Q <= q0 = 1.02 * (ln(255 o) + 0.05) * 2 / 0.99, that bound inflated by 2 %.  For each band of
// four pixel rows the region's x-extent has a closed form: with det = ac - b^2 the chord at height
// v is u = (-b v +- sqrt(a q0 - det v^2)) / a, its right end is largest at the rightmost point of
// the ellipse, v* = -b ex / c (ex = sqrt(q0 c / det)), and its left end smallest at -v*; clamping
// v* (and -v*) into the band gives the band's extent.  A slack of 0.01 px + 2^-20 |coordinate|
// covers the rounding of these few operations.  Opacity below 1/255 reaches nothing; a conic that
// is not safely positive definite (b^2 >= 0.98 ac, NaN) may reach everything.
//
// The square roots, reciprocals and the logarithm are the hardware's approximate instructions
// on the device (errors of a few ulp, against 16 and 12 instructions for the correctly rounded
// square root and division).  Such errors move the band extents by a few parts in 2^23, inside the
// slack; the host build lets tests/cpp/sub_block_mask.cpp run these operations with relative
// errors of 2^-21 in either direction (mask_host_error) and check each variant against brute force.
inline float mask_host_error = 0.0f;  // host variable: relative error given to the operations below
__host__ __device__ inline float mask_sqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x) * (1.0f + mask_host_error);
#endif
}
__host__ __device__ inline float mask_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return (1.0f / x) * (1.0f + mask_host_error);
#endif
}
__host__ __device__ inline float mask_log(float x) {  // natural logarithm
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_logf(x) * 0.693147182f;
#else
    const float l = logf(x);
    return l + (fabsf(l) + 0.5f) * mask_host_error;
#endif
}

// The tile-independent part, computed once per splat by the preprocess and kept in the splat
// record: the ellipse's half-height ey, the height v* of its rightmost point, a q0 and 1/a.
// ey < 0: the splat reaches nothing; ey NaN: it may reach everything.
struct SubBlockReach {
    float ey, vstar, aq0, inva;
};

// q0 of the mask, which depends on the opacity alone.
__host__ __device__ inline float mask_q0(float o) { return 1.02f * (mask_log(255.0f * o) + 0.05f) * (2.0f / 0.99f); }

__host__ __device__ inline SubBlockReach sub_block_reach_q0(float a, float b, float c, float o, float q0) {
    SubBlockReach r{-1.0f, 0.0f, 0.0f, 0.0f};
    if (o < 1.0f / 255.0f) return r;  // alpha <= o for power <= 0
    if (!(a > 0.0f && c > 0.0f && b * b < 0.98f * (a * c) && q0 >= 0.0f)) {
        r.ey = __builtin_nanf("");
        return r;
    }
    const float det = a * c - b * b;
    const float inv_det = mask_rcp(det);
    const float ex = mask_sqrt(q0 * c * inv_det);
    r.ey = mask_sqrt(q0 * a * inv_det);
    r.vstar = -b * ex * mask_rcp(c);
    r.aq0 = a * q0;
    r.inva = mask_rcp(a);
    return r;
}

__host__ __device__ inline SubBlockReach sub_block_reach(float a, float b, float c, float o) {
    return sub_block_reach_q0(a, b, c, o, mask_q0(o));
}

// The per-tile part: the band extents of the splat at (gx, gy) with conic (a, b, c) against the
// tile with top-left pixel (ox, oy).
__host__ __device__ inline uint32_t sub_block_mask(float gx, float gy, float a, float b, float c,
                                                   const SubBlockReach& r, float ox, float oy) {
    if (r.ey < 0.0f) return 0u;
    if (!(r.ey == r.ey)) return 0xFFFFu;
    const float det = a * c - b * b;
    const float ey = r.ey, vstar = r.vstar, aq0 = r.aq0, inva = r.inva;
    const float sx = 0.01f + fabsf(gx) * (1.0f / 1048576.0f), sy = 0.01f + fabsf(gy) * (1.0f / 1048576.0f);
    uint32_t mask = 0;
#pragma unroll
    for (uint32_t band = 0; band < 4; ++band) {
        const float lo = (oy + 4.0f * (float)band) - gy, hi = lo + 3.0f;
        if (hi < -ey - sy || lo > ey + sy) continue;
        const float vlo = fmaxf(lo, -ey), vhi = fminf(hi, ey);
        const float vr = fminf(fmaxf(vstar, vlo), vhi), vl = fminf(fmaxf(-vstar, vlo), vhi);
        const float xr = gx + (-b * vr + mask_sqrt(fmaxf(aq0 - det * vr * vr, 0.0f))) * inva + sx;
        const float xl = gx + (-b * vl - mask_sqrt(fmaxf(aq0 - det * vl * vl, 0.0f))) * inva - sx;
        if (xr < ox || xl > ox + 15.0f) continue;
        const uint32_t c0 = (uint32_t)floorf(fminf(fmaxf((xl - ox) * 0.25f, 0.0f), 3.0f));
        const uint32_t c1 = (uint32_t)floorf(fminf(fmaxf((xr - ox) * 0.25f, 0.0f), 3.0f));
        mask |= (((2u << c1) - 1u) & ~((1u << c0) - 1u)) << (4u * band);
    }
    return mask;
}

__host__ __device__ inline uint32_t splat_sub_block_mask(float gx, float gy, float a, float b, float c, float o,
                                                         float ox, float oy) {
    return sub_block_mask(gx, gy, a, b, c, sub_block_reach(a, b, c, o), ox, oy);
}

// The parts of a splat record that depend on the opacity alone (the same for every view of a
// Gaussian): the blend's skip threshold log(1/(255 o)) - 0.01 and the sub-block mask's q0.
struct OpacityTerms {
    float skip, q0;
};
__device__ __forceinline__ OpacityTerms opacity_terms(float o) {
    return OpacityTerms{__logf(1.0f / (255.0f * o)) - 0.01f, mask_q0(o)};
}

// Writes one splat record: the projected values (and the tile-independent part of the blend's
// sub-block mask), or zeros for a culled Gaussian.  ot = opacity_terms(p.opacity).
__device__ __forceinline__ void store_splat(Splat* dst, bool vis, const Projected& p, const OpacityTerms& ot) {
    Splat s;
    if (vis) {
        const SubBlockReach r = sub_block_reach_q0(p.con_a, p.con_b, p.con_c, p.opacity, ot.q0);
        s.xyzr = make_float4(p.px, p.py, p.vz, p.radius);
        s.conop = make_float4(p.con_a, p.con_b, p.con_c, p.opacity);
        s.rgb = make_float4(p.r, p.g, p.b, r.inva);
        s.aux = make_float4(ot.skip, r.ey, r.vstar, r.aq0);
    } else {
        s.xyzr = s.conop = s.rgb = s.aux = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    *dst = s;
}
__device__ __forceinline__ void store_splat(Splat* dst, bool vis, const Projected& p) {
    store_splat(dst, vis, p, vis ? opacity_terms(p.opacity) : OpacityTerms{0.0f, 0.0f});
}


// ------------------------------------------------------------------ wave64 helpers

// Wave-wide vote: bit l set when lane l's predicate holds (a scalar AND with exec).
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// The inverse: this lane's bit of a wave-uniform mask, as a predicate held in a scalar lane mask
// (no vector instruction).
__device__ __forceinline__ bool lane_on(uint64_t mask) { return __builtin_amdgcn_inverse_ballot_w64(mask); }

// Number of lanes below this one whose bit is set in the wave-uniform mask m.
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const uint32_t l = __lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Inclusive prefix sum across the 64 lanes of a wave.
// DPP moves only (no LDS permutes): row_shr 1, 2, 4, 8 scan each 16-lane row (lanes the shift
// leaves without a source read 0), row_bcast15 adds row 0's total to row 1 and row 2's to row 3,
// row_bcast31 adds the total of rows 0-1 to rows 2-3.  Every lane of the wave must be active.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    int x = (int)v;
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, true);   // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, true);   // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, true);   // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, true);   // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xA, 0xF, false);  // row_bcast:15 into rows 1, 3
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xC, 0xF, false);  // row_bcast:31 into rows 2, 3
    return (uint32_t)x;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// The value of another lane of the wave by a DPP move (CTRL: quad_perm, row rotation, ...; every
// lane must be active).
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}

// Sum over each row of 16 lanes with DPP moves only: quad_perm [1,0,3,2] and [2,3,0,1] -> quad
// sums, row_half_mirror -> 8-lane sums, row_mirror -> every lane holds its row's sum.
__device__ __forceinline__ float row_sum16(float v) {
    int x = __builtin_bit_cast(int, v);
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false));
    x = __builtin_bit_cast(int, v);
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false));
    x = __builtin_bit_cast(int, v);
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, x, 0x141, 0xF, 0xF, false));
    x = __builtin_bit_cast(int, v);
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, x, 0x140, 0xF, 0xF, false));
    return v;
}

// Relaxed device-scope atomic load/store used by the decoupled look-back descriptors.
template <typename T>
__device__ __forceinline__ T load_relaxed(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void store_relaxed(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Logical block index remapped so that each of the 8 XCDs (which receive workgroups round-robin)
// works on one contiguous range of the index space; neighbouring tiles then share an L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb) {
    const uint32_t q = nb / 8u, r = nb % 8u;
    const uint32_t xcd = b % 8u, pos = b / 8u;
    return xcd < r ? xcd * (q + 1u) + pos : r * (q + 1u) + (xcd - r) * q + pos;
}

// The inverse: the block position (0 .. nb - 1) that xcd_remap maps to logical index r.
__host__ __device__ __forceinline__ uint32_t xcd_unmap(uint32_t r, uint32_t nb) {
    const uint32_t q = nb / 8u, m = nb % 8u;
    uint32_t xcd, pos;
    if (r < m * (q + 1u)) {
        xcd = r / (q + 1u);
        pos = r % (q + 1u);
    } else {
        xcd = m + (r - m * (q + 1u)) / q;
        pos = (r - m * (q + 1u)) % q;
    }
    return pos * 8u + xcd;
}

// ------------------------------------------------------------------ shared kernel pieces

// VM = view * model, PVM = projView * model (Slang mul order: each entry summed k = 0..3 left to
// right; project.slang:48/57 via common.slang:102,106).  models null = identity.
__device__ __forceinline__ void entity_mats(const float* __restrict__ model, const CameraParams& cam, EntityMats& out) {
    float M[16];
    if (model) {
        for (int i = 0; i < 16; ++i) M[i] = model[i];
    } else {
        for (int i = 0; i < 16; ++i) M[i] = (i % 5 == 0) ? 1.0f : 0.0f;
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            out.vm[i * 4 + j] = ((cam.view[i * 4 + 0] * M[0 * 4 + j] + cam.view[i * 4 + 1] * M[1 * 4 + j])
                                 + cam.view[i * 4 + 2] * M[2 * 4 + j]) + cam.view[i * 4 + 3] * M[3 * 4 + j];
            out.pvm[i * 4 + j] = ((cam.proj_view[i * 4 + 0] * M[0 * 4 + j] + cam.proj_view[i * 4 + 1] * M[1 * 4 + j])
                                  + cam.proj_view[i * 4 + 2] * M[2 * 4 + j]) + cam.proj_view[i * 4 + 3] * M[3 * 4 + j];
        }
}

// Scan look-back descriptor: bits 63..62 = state (0 none, 1 aggregate, 2 inclusive prefix),
// bits 61..0 = value.  One descriptor per partition.
constexpr uint64_t kDescA = 1ull << 62;
constexpr uint64_t kDescP = 2ull << 62;
constexpr uint64_t kDescMask = (1ull << 62) - 1;

// Exclusive prefix of partition `part` from its predecessors' descriptors; called by one whole
// wave (lane l inspects partition part-1-l, 64 at a time).  Partitions are numbered in the order
// their blocks took an atomic ticket, so every predecessor is running and publishes eventually.
__device__ __forceinline__ uint64_t scan_lookback(uint64_t* desc, uint32_t part) {
    uint64_t excl = 0;
    int64_t j = (int64_t)part - 1;
    const uint32_t l = __lane_id();
    while (true) {
        const int64_t idx = j - (int64_t)l;
        const uint64_t d = idx >= 0 ? load_relaxed(&desc[idx]) : kDescP;
        const uint32_t st = (uint32_t)(d >> 62);
        const uint64_t mP = ballot(st == 2);
        const uint64_t mX = ballot(st == 0);
        const uint32_t first_p = mP ? (uint32_t)__builtin_ctzll(mP) : 64u;
        const uint64_t upto = first_p >= 63 ? ~0ull : ((2ull << first_p) - 1ull);
        if (mX & upto) {
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        const uint64_t mine = (l <= first_p) ? (d & kDescMask) : 0ull;
        excl += wave_sum_u64(mine);
        if (first_p < 64) return excl;
        j -= 64;
    }
}

// Pair emission of one block of 256 Gaussians for one view (keygen.slang:21-54): Gaussian i's
// pairs go to positions base + excl_i + k, k row-major over its tile rectangle, with key
// ((tile_base + tile) << 32) | depth bits and value val_first + i; positions at or past cap are
// dropped (the frame overflowed).  A Gaussian with at most kOwnTiles tiles writes its pairs from
// its own thread (emit_own_pairs, no search); the larger ones are spread over the block
// (emit_large_pairs): pair j of the large ones belongs to the thread whose inclusive count of large
// tiles s_big_incl first exceeds j, which balances footprints of thousands of tiles.
constexpr uint32_t kOwnTiles = 4;

__device__ __forceinline__ void emit_own_pairs(uint32_t tiles, int32_t x0, int32_t y0, int32_t x1, uint32_t depth,
                                               uint64_t pos, uint32_t gx, uint32_t tile_base, uint64_t cap,
                                               uint64_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                               uint32_t val) {
    int32_t x = x0, y = y0;
    for (uint32_t k = 0; k < tiles; ++k) {
        if (pos + k < cap) {
            keys[pos + k] = ((uint64_t)(tile_base + (uint32_t)y * gx + (uint32_t)x) << 32) | depth;
            vals[pos + k] = val;
        }
        if (++x == x1) {
            x = x0;
            ++y;
        }
    }
}

template <uint32_t NB>  // threads of the block, one Gaussian each
__device__ __forceinline__ void emit_large_pairs(const uint32_t* s_big_incl, const uint32_t* s_start,
                                                 const int32_t* s_x0, const int32_t* s_y0, const int32_t* s_w,
                                                 const uint32_t* s_depth, uint32_t big_total, uint64_t base,
                                                 uint32_t gx, uint32_t tile_base, uint64_t cap,
                                                 uint64_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                 uint32_t val_first) {
    for (uint32_t j = threadIdx.x; j < big_total; j += NB) {
        uint32_t lo = 0, hi = NB - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_big_incl[mid] > j) hi = mid; else lo = mid + 1;
        }
        const uint32_t owner = lo;
        const uint32_t k = j - (owner ? s_big_incl[owner - 1] : 0u);
        const uint32_t w = (uint32_t)s_w[owner];
        // k / w through a float reciprocal, corrected by one step either way (exact below 2^22)
        uint32_t ry, rx;
        if (k < (1u << 22)) {
            int32_t q = (int32_t)((float)k * (1.0f / (float)w));
            int32_t r = (int32_t)k - q * (int32_t)w;
            if (r < 0) { --q; r += (int32_t)w; }
            if (r >= (int32_t)w) { ++q; r -= (int32_t)w; }
            ry = (uint32_t)q; rx = (uint32_t)r;
        } else {
            ry = k / w; rx = k - ry * w;
        }
        const uint32_t tile = tile_base + ((uint32_t)s_y0[owner] + ry) * gx + ((uint32_t)s_x0[owner] + rx);
        const uint64_t gpos = base + s_start[owner] + k;
        if (gpos < cap) {
            keys[gpos] = ((uint64_t)tile << 32) | s_depth[owner];
            vals[gpos] = val_first + owner;
        }
    }
}

// Pair emission of the batch, in depth order (k_batch_emit).  Each item carries one 64-bit
// emission record: the global tile of its rectangle's first tile (31 bits: global tiles stay below
// 2^31), its view's tile-grid width and its rectangle's width and height in tiles (11 bits each:
// grids are at most 1024 tiles wide and high); zero for an item without tiles.  Its k-th pair
// (row-major over the rectangle) has key tile0 + (k / w) gx + k % w and value the item; positions
// at or past cap are dropped (the frame overflowed).
__host__ __device__ inline uint64_t pack_emit_rect(uint32_t tile0, uint32_t gx, uint32_t w, uint32_t h) {
    return (uint64_t)tile0 | ((uint64_t)gx << 31) | ((uint64_t)w << 42) | ((uint64_t)h << 53);
}
__host__ __device__ inline uint32_t emit_rect_tile0(uint64_t r) { return (uint32_t)(r & 0x7FFFFFFFull); }
__host__ __device__ inline uint32_t emit_rect_gx(uint64_t r) { return (uint32_t)(r >> 31) & 0x7FFu; }
__host__ __device__ inline uint32_t emit_rect_w(uint64_t r) { return (uint32_t)(r >> 42) & 0x7FFu; }
__host__ __device__ inline uint32_t emit_rect_h(uint64_t r) { return (uint32_t)(r >> 53) & 0x7FFu; }

// An item's emission record as carried through the item sort in the high word of its 64-bit key
// (the sort orders by the low word alone): 0 for a culled item; for a rectangle of at most 4 x 4
// tiles in a view at most 127 tiles wide whose first global tile is below 2^21,
// tile0 | (w - 1) << 21 | (h - 1) << 23 | gx << 25; for any other, kRectLookup (the full record is
// then in the item's slot of the per-item records).
constexpr uint32_t kRectLookup = 1u;
__host__ __device__ inline uint32_t inline_rect(uint32_t tile0, uint32_t gx, uint32_t w, uint32_t h) {
    return (w <= 4u && h <= 4u && gx <= 127u && tile0 < (1u << 21))
               ? tile0 | ((w - 1u) << 21) | ((h - 1u) << 23) | (gx << 25)
               : kRectLookup;
}
__device__ __forceinline__ uint64_t rect_of(uint32_t u, const uint64_t* __restrict__ rects, uint32_t item) {
    if (u == 0u) return 0ull;
    if ((u >> 25) == 0u) return rects[item];
    return pack_emit_rect(u & 0x1FFFFFu, u >> 25, ((u >> 21) & 3u) + 1u, ((u >> 23) & 3u) + 1u);
}

__device__ __forceinline__ void emit_own_rect(uint32_t tiles, uint64_t rect, uint64_t pos, uint64_t cap,
                                              uint32_t* __restrict__ keys, uint32_t* __restrict__ vals, uint32_t val) {
    const uint32_t w = emit_rect_w(rect), gx = emit_rect_gx(rect);
    uint32_t row = emit_rect_tile0(rect), x = 0;
    for (uint32_t k = 0; k < tiles; ++k) {
        if (pos + k < cap) {
            keys[pos + k] = row + x;
            vals[pos + k] = val;
        }
        if (++x == w) {
            x = 0;
            row += gx;
        }
    }
}

// The same pairs into a staging chunk: pair k (block-local position first + k) is placed at slot
// first + k - c0 when that lies in [0, n).
__device__ __forceinline__ void stage_own_rect(uint32_t tiles, uint64_t rect, uint32_t first, uint32_t c0,
                                               uint32_t n, uint32_t* keys, uint32_t* vals, uint32_t val) {
    if (first + tiles <= c0 || first >= c0 + n) return;
    const uint32_t w = emit_rect_w(rect), gx = emit_rect_gx(rect);
    uint32_t row = emit_rect_tile0(rect), x = 0;
    for (uint32_t k = 0; k < tiles; ++k) {
        const uint32_t slot = first + k - c0;  // wraps past 2^32 below c0: then >= n
        if (slot < n) {
            keys[slot] = row + x;
            vals[slot] = val;
        }
        if (++x == w) {
            x = 0;
            row += gx;
        }
    }
}

// The large items of a block: s_big_incl[o] is the inclusive count of large tiles over the block's
// items 0..o (n of them), s_start/s_rect the block-local first pair and the record of a large item
// o, item_vals[o] its value (the block's sorted items).
__device__ __forceinline__ void emit_large_rects(const uint32_t* s_big_incl, const uint32_t* s_start,
                                                 const uint64_t* s_rect, const uint32_t* __restrict__ item_vals,
                                                 uint32_t n,
                                                 uint32_t big_total, uint64_t base, uint64_t cap,
                                                 uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    for (uint32_t j = threadIdx.x; j < big_total; j += blockDim.x) {
        uint32_t lo = 0, hi = n - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_big_incl[mid] > j) hi = mid; else lo = mid + 1;
        }
        const uint32_t owner = lo;
        const uint32_t k = j - (owner ? s_big_incl[owner - 1] : 0u);
        const uint64_t rect = s_rect[owner];
        const uint32_t w = emit_rect_w(rect);
        uint32_t ry, rx;
        if (k < (1u << 22)) {  // k / w through a float reciprocal, corrected by one step
            int32_t q = (int32_t)((float)k * (1.0f / (float)w));
            int32_t r = (int32_t)k - q * (int32_t)w;
            if (r < 0) { --q; r += (int32_t)w; }
            if (r >= (int32_t)w) { ++q; r -= (int32_t)w; }
            ry = (uint32_t)q; rx = (uint32_t)r;
        } else {
            ry = k / w; rx = k - ry * w;
        }
        const uint64_t gpos = base + s_start[owner] + k;
        if (gpos < cap) {
            keys[gpos] = emit_rect_tile0(rect) + ry * emit_rect_gx(rect) + rx;
            vals[gpos] = item_vals[owner];
        }
    }
}

}  // namespace tpd

// Per-tile blend bodies shared by the plain and the batched kernels: the forward front-to-back
// composite (blend.slang:22-105) and its backward.  A kernel maps its block to a tile (TileView)
// and calls blend_tile / blend_bwd_tile with its block's LDS.
#pragma once

#include "tpd_device.hpp"

namespace tpd {

// Where a tile's pixels lie: tile coordinates inside its view, the view's size and tile grid,
// and the view's first pixel in the (concatenated) image buffers.
struct TileView {
    uint32_t tx, ty;
    uint32_t width, height;
    uint32_t gx, gy;
    uint64_t pix_base;
};

constexpr uint32_t kBlendWaves = kTilePixels / kWave;  // 4 waves, one 8x8 quadrant each
constexpr uint32_t kQuad = 8;
constexpr uint32_t kSplatGrads = 9;  // dL/d(image x, y, conic a, b, c, r, g, b, opacity)

struct BlendLds {
    float4 a[kTilePixels];  // x, y, conic a, conic b
    float4 b[kTilePixels];  // conic c, skip threshold, opacity, -
    float4 c[kTilePixels];  // colour
    uint8_t list[kBlendWaves][kTilePixels];
    uint32_t cnt[kBlendWaves][kBlendWaves];  // [staging wave][quadrant]
};

struct BlendBwdLds {
    BlendLds s;
    uint32_t g[kTilePixels];                // splat index of each staged slot
    float acc[kTilePixels][kSplatGrads];    // per staged splat: the tile's gradient sums
    int32_t last;
};

// Pixel of this thread: wave w owns the 8x8 quadrant (w & 1, w >> 1) of the tile.
__device__ __forceinline__ void tile_pixel(const TileView& tv, uint32_t& px, uint32_t& py) {
    const uint32_t wave = threadIdx.x / kWave, lane = __lane_id();
    px = tv.tx * kTile + (wave & 1u) * kQuad + (lane % kQuad);
    py = tv.ty * kTile + (wave >> 1) * kQuad + (lane / kQuad);
}

// Stages splats [base, min(base + 256, end)) of the tile's sorted list in LDS and builds one
// compacted list per quadrant (sorted order kept) of the splats that can reach alpha >= 1/255 at
// some pixel centre of the quadrant (a bounding-box test, then splat_touches_rect).  Returns the
// length of this wave's list.  Ends with a barrier.
template <bool STORE_G>
__device__ __forceinline__ uint32_t blend_stage(BlendLds& s, uint32_t* s_g, const TileView& tv, uint32_t base,
                                                uint32_t end, const uint32_t* __restrict__ vals,
                                                const float4* __restrict__ xyzr, const float4* __restrict__ conop,
                                                const float4* __restrict__ rgb) {
    const uint32_t tid = threadIdx.x, wave = tid / kWave, lane = __lane_id();
    const uint32_t k = base + tid;
    uint32_t mask = 0;
    if (k < end) {
        const uint32_t g = vals[k];
        const float4 p = xyzr[g];
        const float4 c = conop[g];
        s.a[tid] = make_float4(p.x, p.y, c.x, c.y);
        s.b[tid] = make_float4(c.z, __logf(1.0f / (255.0f * c.w)) - 0.01f, c.w, 0.0f);
        s.c[tid] = rgb[g];
        if (STORE_G) s_g[tid] = g;
        // Bounding box of the region Q <= q0 outside which splat_touches_rect always says no
        // (half-extents sqrt(q0 c / det), sqrt(q0 a / det), widened by 1%); only quadrants the
        // box reaches get the exact test.  Degenerate or near-degenerate conics (|b| close to
        // sqrt(ac), where det loses precision) give an unbounded box.
        float ex = __builtin_inff(), ey = __builtin_inff();
        const float det = c.x * c.z - c.y * c.y;
        const float q0 = (__logf(255.0f * c.w) + 0.05f) * (2.0f / 0.99f);
        if (c.x > 0.0f && c.z > 0.0f && c.y * c.y < 0.98f * (c.x * c.z) && q0 >= 0.0f) {
            ex = 1.01f * __builtin_sqrtf(q0 * c.z / det) + 1e-3f;
            ey = 1.01f * __builtin_sqrtf(q0 * c.x / det) + 1e-3f;
        }
#pragma unroll
        for (uint32_t q = 0; q < kBlendWaves; ++q) {
            const uint32_t x0 = tv.tx * kTile + (q & 1u) * kQuad, y0 = tv.ty * kTile + (q >> 1) * kQuad;
            if (x0 >= tv.width || y0 >= tv.height) continue;
            const float fx0 = (float)x0, fy0 = (float)y0;
            const float fx1 = (float)(min(x0 + kQuad, tv.width) - 1), fy1 = (float)(min(y0 + kQuad, tv.height) - 1);
            if (p.x + ex >= fx0 && p.x - ex <= fx1 && p.y + ey >= fy0 && p.y - ey <= fy1 &&
                splat_touches_rect(p.x, p.y, c.x, c.y, c.z, c.w, fx0, fx1, fy0, fy1))
                mask |= 1u << q;
        }
    }
    const uint64_t lt = lanemask_lt();
    uint64_t bal[kBlendWaves];
#pragma unroll
    for (uint32_t q = 0; q < kBlendWaves; ++q) bal[q] = ballot((mask >> q) & 1u);
    if (lane < kBlendWaves) {
        uint64_t b = bal[0];
#pragma unroll
        for (uint32_t q = 1; q < kBlendWaves; ++q) b = lane == q ? bal[q] : b;
        s.cnt[wave][lane] = (uint32_t)__popcll(b);
    }
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < kBlendWaves; ++q) {
        if ((mask >> q) & 1u) {
            uint32_t off = (uint32_t)__popcll(bal[q] & lt);
            for (uint32_t w = 0; w < wave; ++w) off += s.cnt[w][q];
            s.list[q][off] = (uint8_t)tid;
        }
    }
    uint32_t cnt = 0;
#pragma unroll
    for (uint32_t w = 0; w < kBlendWaves; ++w) cnt += s.cnt[w][wave];
    __syncthreads();
    return cnt;
}

// Forward composite of one tile.  Per pixel the operations and their order are the reference's;
// the shortcuts around them change no pixel:
//   - the quadrant lists of blend_stage drop only splats no pixel of the quadrant can take;
//   - skip = log(1/(255 o)) - 0.01: a power below it gives o * exp(power) < 0.99/255, so the
//     alpha < 1/255 test would discard the splat anyway; exp() runs only if some lane needs it;
//   - a wave leaves the round as soon as all its pixels have terminated.
// COUNT adds up the pixel-splat pairs whose power is evaluated (for the FLOP roofline).
template <bool COUNT>
__device__ __forceinline__ void blend_tile(BlendLds& s, const TileView& tv, uint2 range,
                                           const uint32_t* __restrict__ vals, const float4* __restrict__ xyzr,
                                           const float4* __restrict__ conop, const float4* __restrict__ rgb,
                                           float* __restrict__ image, uint8_t* __restrict__ rgba8,
                                           uint64_t* eval_count) {
    const uint32_t tid = threadIdx.x, wave = tid / kWave, lane = __lane_id();
    uint32_t pxl, pyl;
    tile_pixel(tv, pxl, pyl);
    const bool inside = pxl < tv.width && pyl < tv.height;
    const float fx = (float)pxl, fy = (float)pyl;

    // The per-pixel state machine runs on wave-uniform lane masks (scalar instructions) rather
    // than per-lane booleans: live = pixels not yet terminated.
    float T = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
    uint64_t live = ballot(inside);
    uint32_t evals = 0;
    const uint8_t* list = s.list[__builtin_amdgcn_readfirstlane(wave)];
    for (uint32_t base = range.x; base < range.y; base += kTilePixels) {
        if (__syncthreads_count(!lane_on(live)) == (int)kTilePixels) break;
        const uint32_t cnt = __builtin_amdgcn_readfirstlane(
            blend_stage<false>(s, nullptr, tv, base, range.y, vals, xyzr, conop, rgb));
        if (live == 0) continue;
        for (uint32_t i = 0; i < cnt; ++i) {
            const uint32_t j = list[i];
            const float4 A = s.a[j];
            const float2 B = *reinterpret_cast<const float2*>(&s.b[j]);  // conic c, skip threshold
            if (COUNT) evals += lane_on(live) ? 1u : 0u;
            const float dx = A.x - fx, dy = A.y - fy;
            const float power = -0.5f * ((A.z * dx) * dx + (B.x * dy) * dy) - (A.w * dx) * dy;
            // power <= 0 equals !(power > 0) except for NaN, whose exp is 0 and whose alpha
            // therefore fails the 1/255 test: either way the splat is skipped.  Active lanes have
            // power in [skip, 0] with skip >= log(1/255) - 0.01, inside soft_expf's plain range.
            const uint64_t act = ballot(power <= 0.0f) & ballot(!(power < B.y)) & live;
            if (act == 0) continue;
            float alpha = s.b[j].z * soft_expf_inrange(power);
            alpha = alpha < 0.99f ? alpha : 0.99f;
            const float tn = T * (1.0f - alpha);
            const uint64_t ok = act & ballot(!(alpha < 1.0f / 255.0f));
            const uint64_t low = ballot(tn < 0.0001f);
            if (lane_on(ok & ~low)) {
                const float4 C = s.c[j];
                cr = cr + (C.x * alpha) * T;
                cg = cg + (C.y * alpha) * T;
                cb = cb + (C.z * alpha) * T;
                T = tn;
            }
            live &= ~(ok & low);
            if (live == 0) break;
        }
    }
    if (COUNT) {
        const uint64_t sum = wave_sum_u64(evals);
        if (lane == 0) atomicAdd(reinterpret_cast<unsigned long long*>(eval_count), (unsigned long long)sum);
    }
    if (inside) {
        const size_t o = tv.pix_base + (size_t)pyl * tv.width + pxl;
        if (image) {
            image[o * 3 + 0] = cr;
            image[o * 3 + 1] = cg;
            image[o * 3 + 2] = cb;
        }
        if (rgba8) {
            uchar4 q;
            q.x = unorm8(cr); q.y = unorm8(cg); q.z = unorm8(cb); q.w = 255;
            reinterpret_cast<uchar4*>(rgba8)[o] = q;
        }
    }
}

// Stores the gradient sums of the pair (this tile, splat g) at its pre-sort position; zeros when
// out is null.
__device__ __forceinline__ void store_pair_partial(const TileView& tv, uint32_t g, const float4* __restrict__ xyzr,
                                                   const uint32_t* __restrict__ offsets, float* __restrict__ partials,
                                                   const float* out) {
    const float4 p = xyzr[g];
    int32_t x0, y0, x1, y1;
    bounding_rect(p.x, p.y, p.w, tv.gx, tv.gy, x0, y0, x1, y1);
    const size_t pos = (size_t)offsets[g] + (size_t)((int32_t)tv.ty - y0) * (size_t)(x1 - x0) +
                       (size_t)((int32_t)tv.tx - x0);
    float* dst = partials + pos * kSplatGrads;
#pragma unroll
    for (uint32_t q = 0; q < kSplatGrads; ++q) dst[q] = out ? out[q] : 0.0f;
}

// Backward of one tile.  Pass 1 re-runs the forward decisions to find each pixel's final
// transmittance and last contributor (nothing is stored by the forward); pass 2 walks the tile's
// splats back to front recovering T_i = T_{i+1} / (1 - a_i) and the colour accumulated behind
// each splat, producing per splat dL/d(image x, y), dL/d(conic a, b, c), dL/d(rgb), dL/d(opacity).
// Wave sums are combined in LDS; each (tile, splat) pair's sums are stored once, at the pair's
// pre-sort position offsets[g] + row-major rank of the tile in g's rectangle (keygen.slang:47-53
// order), so no global atomics are needed and the per-Gaussian sum is deterministic.  Every pair
// of the tile's range is written (zeros behind the last contributor), so the partials need no
// clearing beforehand.
__device__ __forceinline__ void blend_bwd_tile(BlendBwdLds& L, const TileView& tv, uint2 range,
                                               const uint32_t* __restrict__ vals, const float4* __restrict__ xyzr,
                                               const float4* __restrict__ conop, const float4* __restrict__ rgb,
                                               const float* __restrict__ dL_dimg, const uint32_t* __restrict__ offsets,
                                               float* __restrict__ partials) {
    constexpr uint32_t kG = kSplatGrads;
    const uint32_t tid = threadIdx.x, wave = tid / kWave, lane = __lane_id();
    uint32_t pxl, pyl;
    tile_pixel(tv, pxl, pyl);
    const bool inside = pxl < tv.width && pyl < tv.height;
    const float fx = (float)pxl, fy = (float)pyl;
    BlendLds& s = L.s;
    if (tid == 0) L.last = -1;
    for (uint32_t x = tid; x < kTilePixels * kG; x += kTilePixels) (&L.acc[0][0])[x] = 0.0f;

    // ---- pass 1: forward re-run (same decisions as blend_tile) -> T_final, last contributor
    float T = 1.0f;
    uint64_t live = ballot(inside);
    int32_t last = -1;
    const uint8_t* list = s.list[__builtin_amdgcn_readfirstlane(wave)];
    for (uint32_t base = range.x; base < range.y; base += kTilePixels) {
        if (__syncthreads_count(!lane_on(live)) == (int)kTilePixels) break;
        const uint32_t cnt = __builtin_amdgcn_readfirstlane(
            blend_stage<true>(s, L.g, tv, base, range.y, vals, xyzr, conop, rgb));
        if (live == 0) continue;
        for (uint32_t i = 0; i < cnt; ++i) {
            const uint32_t j = list[i];
            const float4 A = s.a[j];
            const float2 B = *reinterpret_cast<const float2*>(&s.b[j]);
            const float dx = A.x - fx, dy = A.y - fy;
            const float power = -0.5f * ((A.z * dx) * dx + (B.x * dy) * dy) - (A.w * dx) * dy;
            const uint64_t act = ballot(power <= 0.0f) & ballot(!(power < B.y)) & live;
            if (act == 0) continue;
            float alpha = s.b[j].z * soft_expf_inrange(power);
            alpha = alpha < 0.99f ? alpha : 0.99f;
            const float tn = T * (1.0f - alpha);
            const uint64_t ok = act & ballot(!(alpha < 1.0f / 255.0f));
            const uint64_t low = ballot(tn < 0.0001f);
            if (lane_on(ok & ~low)) {
                T = tn;
                last = (int32_t)(base + j);
            }
            live &= ~(ok & low);
            if (live == 0) break;
        }
    }
    if (last >= 0) atomicMax(&L.last, last);
    __syncthreads();
    const int32_t block_last = L.last;
    // Pairs behind the tile's last contributor (all of them when no pixel blended anything) get
    // zero gradients.
    const uint32_t zero_from = block_last < 0 ? range.x
                                              : range.x + (((uint32_t)block_last - range.x) / kTilePixels + 1u) * kTilePixels;
    for (uint32_t k = zero_from + tid; k < range.y; k += kTilePixels)
        store_pair_partial(tv, vals[k], xyzr, offsets, partials, nullptr);
    if (block_last < 0) return;  // uniform

    // ---- pass 2: back to front
    float g_r = 0.0f, g_g = 0.0f, g_b = 0.0f;
    if (inside) {
        const size_t o = tv.pix_base + (size_t)pyl * tv.width + pxl;
        g_r = dL_dimg[o * 3 + 0];
        g_g = dL_dimg[o * 3 + 1];
        g_b = dL_dimg[o * 3 + 2];
    }
    float acc_r = 0.0f, acc_g = 0.0f, acc_b = 0.0f;  // colour accumulated behind the current splat
    float last_alpha = 0.0f, last_r = 0.0f, last_g = 0.0f, last_b = 0.0f;
    const uint32_t first_chunk = ((uint32_t)block_last - range.x) / kTilePixels;
    for (int32_t chunk = (int32_t)first_chunk; chunk >= 0; --chunk) {
        const uint32_t base = range.x + (uint32_t)chunk * kTilePixels;
        __syncthreads();
        const uint32_t cnt = __builtin_amdgcn_readfirstlane(
            blend_stage<true>(s, L.g, tv, base, range.y, vals, xyzr, conop, rgb));
        for (int32_t i = (int32_t)cnt - 1; i >= 0; --i) {
            const uint32_t j = list[i];
            const int32_t kk = (int32_t)(base + j);
            const uint64_t mine = ballot(kk <= last);
            if (mine == 0) continue;
            const float4 A = s.a[j];
            const float4 B = s.b[j];
            const float dx = A.x - fx, dy = A.y - fy;
            const float power = -0.5f * ((A.z * dx) * dx + (B.x * dy) * dy) - (A.w * dx) * dy;
            const uint64_t act = ballot(power <= 0.0f) & ballot(!(power < B.y)) & mine;
            if (act == 0) continue;
            const float G = soft_expf_inrange(power);
            const float raw = B.z * G;
            const float alpha = raw < 0.99f ? raw : 0.99f;
            const uint64_t updm = act & ballot(!(alpha < 1.0f / 255.0f));
            const bool upd = lane_on(updm);
            float v[kG];
#pragma unroll
            for (uint32_t q = 0; q < kG; ++q) v[q] = 0.0f;
            if (upd) {
                const float4 C = s.c[j];
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
            if (updm != 0) {
                // row sums by DPP, then one LDS atomic per row of 16 lanes
#pragma unroll
                for (uint32_t q = 0; q < kG; ++q) {
                    const float sum = row_sum16(v[q]);
                    if ((lane & 15u) == 15u && sum != 0.0f) atomicAdd(&L.acc[j][q], sum);
                }
            }
        }
        __syncthreads();
        if (base + tid < range.y) {
            float out[kG];
#pragma unroll
            for (uint32_t q = 0; q < kG; ++q) {
                out[q] = L.acc[tid][q];
                L.acc[tid][q] = 0.0f;
            }
            store_pair_partial(tv, L.g[tid], xyzr, offsets, partials, out);
        }
    }
}

}  // namespace tpd

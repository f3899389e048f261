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
constexpr uint32_t kSub = 4;                              // 4x4-pixel sub-blocks ...
constexpr uint32_t kSubBlocks = kTilePixels / (kSub * kSub);  // ... 16 per tile, one per 16-lane row
constexpr uint32_t kSlotBytes = 16;                       // one float4 per staged splat and array
constexpr uint32_t kSentinel = kTilePixels * kSlotBytes;  // byte offset of the slot no pixel takes
constexpr uint32_t kSplatGrads = 9;  // dL/d(image x, y, conic a, b, c, r, g, b, opacity)

// A staging round holds up to 256 splats of the tile's sorted list in slots 0..255; slot 256 is a
// sentinel that no pixel can take (skip threshold +inf).  Sub-block lists hold the byte offsets
// (slot * 16) of the splats that can reach the sub-block, in sorted order, padded with the
// sentinel up to the longest list of the wave.
struct BlendLds {
    float4 a[kTilePixels + 1];  // x, y, conic a, conic b
    float4 b[kTilePixels + 1];  // conic c, skip threshold, opacity, -
    float4 c[kTilePixels + 1];  // colour
    uint16_t list[kSubBlocks][kTilePixels];
    uint32_t cnt[kBlendWaves][kSubBlocks];  // [staging wave][sub-block]
    alignas(16) uint32_t done[kBlendWaves];  // per wave: all of its pixels have terminated
};

struct BlendBwdLds {
    BlendLds s;
    double acc[kTilePixels][kSplatGrads];   // per staged splat: the tile's gradient sums (see blend_bwd_tile)
    int32_t last;
};

__device__ __forceinline__ float4 slot4(const float4* arr, uint32_t off) {
    return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(arr) + off);
}
__device__ __forceinline__ float2 slot2(const float4* arr, uint32_t off) {
    return *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(arr) + off);
}

// Sub-blocks are numbered row-major inside the tile: q = 4 * row + column (4x4 pixels each).
__device__ __forceinline__ void sub_block_origin(const TileView& tv, uint32_t q, uint32_t& x0, uint32_t& y0) {
    x0 = tv.tx * kTile + (q & 3u) * kSub;
    y0 = tv.ty * kTile + (q >> 2) * kSub;
}

// Sub-block walked by row g (0..3) of wave w: wave w owns the 8x8 quadrant (w & 1, w >> 1) of the
// tile, row g the 4x4 block (g & 1, g >> 1) inside it.
__device__ __forceinline__ uint32_t sub_block_of(uint32_t w, uint32_t g) {
    return (((w >> 1) * 2u + (g >> 1)) * 4u) + (w & 1u) * 2u + (g & 1u);
}

// Pixel of this thread: the sub-block of its wave and 16-lane row, lanes row-major inside it.
__device__ __forceinline__ void tile_pixel(const TileView& tv, uint32_t& px, uint32_t& py) {
    const uint32_t wave = threadIdx.x / kWave, lane = __lane_id();
    uint32_t x0, y0;
    sub_block_origin(tv, sub_block_of(wave, lane >> 4), x0, y0);
    px = x0 + (lane & 3u);
    py = y0 + ((lane >> 2) & 3u);
}

// Stages splats [base, min(base + 256, end)) of the tile's sorted list in LDS and builds one
// compacted list per 4x4 sub-block (sorted order kept) of the splats that may reach
// alpha >= 1/255 at some pixel centre of the sub-block (sub_block_mask, with the per-splat part
// the preprocess stored in the record).  Returns the number
// of list entries this wave walks: the longest of its four lists (the shorter ones are padded
// with the sentinel).  Ends with a barrier.
// This thread's splat of a staging round: whether the slot holds one, its index and its record.
struct StagedSplat {
    bool valid;
    uint32_t g;
    Splat sp;
};

// Both loads are unconditional (slot past the end: value 0 and record 0, which every non-empty
// buffer has), so that a load issued a round ahead stays in flight in its own registers.
__device__ __forceinline__ uint32_t load_value(const uint32_t* __restrict__ vals, uint32_t k, uint32_t end) {
    return vals[k < end ? k : 0u];
}

__device__ __forceinline__ StagedSplat load_staged(const Splat* __restrict__ splats, uint32_t g, uint32_t k,
                                                   uint32_t end) {
    StagedSplat t;
    t.valid = k < end;
    t.g = g;
    t.sp = splats[t.valid ? g : 0u];
    return t;
}

__device__ __forceinline__ uint32_t blend_stage_loaded(BlendLds& s, const TileView& tv, const StagedSplat& in) {
    const uint32_t tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid / kWave), lane = __lane_id();
    {
        constexpr uint32_t s2 = kSentinel | (kSentinel << 16);
        uint4* lw = reinterpret_cast<uint4*>(&s.list[0][0]);
        static_assert(sizeof(s.list) == 2 * kTilePixels * sizeof(uint4), "list fill");
        lw[tid] = make_uint4(s2, s2, s2, s2);
        lw[tid + kTilePixels] = make_uint4(s2, s2, s2, s2);
        if (tid == 0) {
            s.a[kTilePixels] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            s.b[kTilePixels] = make_float4(0.0f, __builtin_inff(), 0.0f, 0.0f);
            s.c[kTilePixels] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
    uint32_t mask = 0;
    if (in.valid) {
        const Splat& sp = in.sp;
        const float4 p = sp.xyzr;
        const float4 c = sp.conop;
        s.a[tid] = make_float4(p.x, p.y, c.x, c.y);
        s.b[tid] = make_float4(c.z, sp.aux.x, c.w, 0.0f);
        s.c[tid] = sp.rgb;
        const uint32_t reach = sub_block_mask(p.x, p.y, c.x, c.y, c.z,
                                              SubBlockReach{sp.aux.y, sp.aux.z, sp.aux.w, sp.rgb.w},
                                              (float)(tv.tx * kTile), (float)(tv.ty * kTile));
        // sub-blocks with at least one pixel inside the view
        const uint32_t vx = min((tv.width - tv.tx * kTile + kSub - 1) / kSub, 4u);
        const uint32_t vy = min((tv.height - tv.ty * kTile + kSub - 1) / kSub, 4u);
        const uint32_t vcols = (1u << vx) - 1u;
        const uint32_t vrows = (1u << vy) - 1u;
        const uint32_t vspread = (vrows & 1u) | ((vrows & 2u) << 3) | ((vrows & 4u) << 6) | ((vrows & 8u) << 9);
        mask = reach & (vcols * vspread);
    }
    uint64_t bal[kSubBlocks];
#pragma unroll
    for (uint32_t q = 0; q < kSubBlocks; ++q) bal[q] = ballot((mask >> q) & 1u);
    if (lane == 0) {
#pragma unroll
        for (uint32_t q = 0; q < kSubBlocks; ++q) s.cnt[wave][q] = (uint32_t)__popcll(bal[q]);
    }
    __syncthreads();
    // lane q < 16: this wave's start in sub-block q's list, and the list's length
    uint32_t start = 0, total = 0;
    if (lane < kSubBlocks) {
#pragma unroll
        for (uint32_t w = 0; w < kBlendWaves; ++w) {
            const uint32_t n = s.cnt[w][lane];
            start += w < wave ? n : 0u;
            total += n;
        }
    }
    // The lanes with bit q set are exactly bal[q]: it serves as the execution mask, and the
    // lane's rank among them is a masked bit count.
#pragma unroll
    for (uint32_t q = 0; q < kSubBlocks; ++q) {
        if (lane_on(bal[q])) {
            const uint32_t off = (uint32_t)__builtin_amdgcn_readlane(start, q) + lanes_below(bal[q]);
            s.list[q][off] = (uint16_t)(tid * kSlotBytes);
        }
    }
    uint32_t walk = 0;
#pragma unroll
    for (uint32_t g = 0; g < 4; ++g) walk = max(walk, (uint32_t)__builtin_amdgcn_readlane(total, sub_block_of(wave, g)));
    __syncthreads();
    return walk;
}

// Forward composite of one tile.  Per pixel the operations and their order are the reference's;
// the shortcuts around them change no pixel:
//   - the sub-block lists of blend_stage drop only splats no pixel of the sub-block can take;
//   - skip = log(1/(255 o)) - 0.01: a power below it gives o * exp(power) < 0.99/255, so the
//     alpha < 1/255 test would discard the splat anyway; exp() runs only if some lane needs it;
//   - a wave leaves the round as soon as all its pixels have terminated.
// Each 16-lane row of a wave walks its own sub-block's list, so one wave step evaluates up to four
// different splats.  COUNT adds up the pixel-splat pairs whose power is evaluated (for the FLOP
// roofline; sentinel entries are not counted).  With final_T / final_last, each pixel's final
// transmittance and the sorted index of the last splat it blended (-1: none) are stored for the
// backward, which then need not run the composite again.
template <bool COUNT>
__device__ __forceinline__ void blend_tile(BlendLds& s, const TileView& tv, uint2 range,
                                           const uint32_t* __restrict__ vals, const Splat* __restrict__ splats,
                                           float* __restrict__ image, uint8_t* __restrict__ rgba8,
                                           uint64_t* eval_count, float* __restrict__ final_T = nullptr,
                                           int32_t* __restrict__ final_last = nullptr) {
    const uint32_t tid = threadIdx.x, wave = tid / kWave, lane = __lane_id();
    uint32_t pxl, pyl;
    tile_pixel(tv, pxl, pyl);
    const bool inside = pxl < tv.width && pyl < tv.height;
    const float fx = (float)pxl, fy = (float)pyl;

    // The per-pixel state machine runs on wave-uniform lane masks (scalar instructions) rather
    // than per-lane booleans: live = pixels not yet terminated.
    float T = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
    int32_t last = -1;
    uint64_t live = ballot(inside);
    uint32_t evals = 0;
    const uint16_t* list = s.list[sub_block_of(wave, lane >> 4)];
    // The global loads of staging are issued a round ahead: the record of round r + 1 and the
    // value of round r + 2 are in flight while round r blends.
    StagedSplat next = load_staged(splats, load_value(vals, range.x + tid, range.y), range.x + tid, range.y);
    uint32_t g_after = load_value(vals, range.x + kTilePixels + tid, range.y);
    for (uint32_t base = range.x; base < range.y; base += kTilePixels) {
        // The block leaves once every pixel has terminated: one flag per wave and one barrier
        // (a wave rewrites its flag only after this round's staging barriers, so no wave can
        // overwrite a flag another is still reading).
        if (lane == 0) s.done[wave] = live == 0 ? 1u : 0u;
        __syncthreads();
        const uint4 done = *reinterpret_cast<const uint4*>(s.done);
        if ((done.x & done.y & done.z & done.w) != 0u) break;
        const uint32_t walk = blend_stage_loaded(s, tv, next);
        next = load_staged(splats, g_after, base + kTilePixels + tid, range.y);
        g_after = load_value(vals, base + 2 * kTilePixels + tid, range.y);
        if (live == 0) continue;
        // Two list entries per step: both splats' powers are formed before the first one updates
        // the pixel (power does not depend on T), then the two updates run in list order.  An odd
        // walk reads one sentinel entry (every list is padded with it), which no lane takes.
        const auto power_of = [&](const float4& A, const float2& B) {
            const float dx = A.x - fx, dy = A.y - fy;
            return -0.5f * ((A.z * dx) * dx + (B.x * dy) * dy) - (A.w * dx) * dy;
        };
        // power <= 0 equals !(power > 0) except for NaN, whose exp is 0 and whose alpha therefore
        // fails the 1/255 test: either way the splat is skipped.  Active lanes have power in
        // [skip, 0] with skip >= log(1/255) - 0.01, inside soft_expf_inrange's range.
        const auto blend_one = [&](uint32_t off, float power, uint64_t act) {
            float alpha = slot4(s.b, off).z * soft_expf_inrange(power);
            alpha = alpha < 0.99f ? alpha : 0.99f;
            const float tn = T * (1.0f - alpha);
            const uint64_t ok = act & ballot(!(alpha < 1.0f / 255.0f));
            const uint64_t low = ballot(tn < 0.0001f);
            if (lane_on(ok & ~low)) {
                const float4 C = slot4(s.c, off);
                cr = cr + (C.x * alpha) * T;
                cg = cg + (C.y * alpha) * T;
                cb = cb + (C.z * alpha) * T;
                T = tn;
                last = (int32_t)(base + off / kSlotBytes);
            }
            live &= ~(ok & low);
        };
        for (uint32_t i = 0; i < walk; i += 2) {
            const uint32_t o1 = list[i], o2 = list[i + 1];
            const float4 A1 = slot4(s.a, o1), A2 = slot4(s.a, o2);
            const float2 B1 = slot2(s.b, o1), B2 = slot2(s.b, o2);  // conic c, skip threshold
            const float p1 = power_of(A1, B1), p2 = power_of(A2, B2);
            const uint64_t act1 = ballot(p1 <= 0.0f) & ballot(!(p1 < B1.y)) & live;
            const uint64_t can2 = ballot(p2 <= 0.0f) & ballot(!(p2 < B2.y));
            if (COUNT) evals += (lane_on(live) && o1 != kSentinel) ? 1u : 0u;
            if (act1 != 0) blend_one(o1, p1, act1);
            if (COUNT) evals += (lane_on(live) && o2 != kSentinel) ? 1u : 0u;
            const uint64_t act2 = can2 & live;
            if (act2 != 0) blend_one(o2, p2, act2);
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
        if (final_T) {
            final_T[o] = T;
            final_last[o] = last;
        }
    }
}

// Pre-sort position of the pair (this tile, splat g): g's first pair (offset) + the row-major rank of
// the tile in g's rectangle (from its record's position and radius).
__device__ __forceinline__ size_t pair_position(const TileView& tv, const float4& xyzr, uint32_t offset) {
    int32_t x0, y0, x1, y1;
    bounding_rect(xyzr.x, xyzr.y, xyzr.w, tv.gx, tv.gy, x0, y0, x1, y1);
    return (size_t)offset + (size_t)((int32_t)tv.ty - y0) * (size_t)(x1 - x0) + (size_t)((int32_t)tv.tx - x0);
}

// Stores the gradient sums of a pair at its pre-sort position; zeros when out is null.
__device__ __forceinline__ void store_partial_at(size_t pos, float* __restrict__ partials, const float* out) {
    float* dst = partials + pos * kSplatGrads;
#pragma unroll
    for (uint32_t q = 0; q < kSplatGrads; ++q) dst[q] = out ? out[q] : 0.0f;
}

// The same for the pair (this tile, splat g), its record and offset read here.
__device__ __forceinline__ void store_pair_partial(const TileView& tv, uint32_t g, const Splat* __restrict__ splats,
                                                   const uint32_t* __restrict__ offsets, float* __restrict__ partials,
                                                   const float* out) {
    store_partial_at(pair_position(tv, splats[g].xyzr, offsets[g]), partials, out);
}

// Backward of one tile.  The forward frame stored each pixel's final transmittance and last
// contributor (blend_tile's final_T / final_last); the tile's splats are walked back to front
// recovering T_i = T_{i+1} / (1 - a_i) and the colour accumulated behind each splat, producing per splat dL/d(image x, y), dL/d(conic a, b, c), dL/d(rgb), dL/d(opacity).
// Each 16-lane row (one 4x4 sub-block) sums its lanes with DPP row operations in a fixed order;
// the rows' sums for a splat (at most one per sub-block, 16 in all) are added in LDS with float64
// atomics.  Those FP32 row sums add exactly in float64 unless their magnitudes span more than
// 2^25, so the tile's total, rounded once to FP32, does not depend on the order in which the rows
// arrive, and a frame's gradients repeat bit for bit from run to run (the exception would need
// row sums 2^25 apart whose float64 rounding then also moves the FP32 result).  Each (tile,
// splat) pair's sums are stored once, at the pair's pre-sort position offsets[g] + row-major rank
// of the tile in g's rectangle, so no global atomics are needed, and the per-Gaussian sum over
// its tiles runs in tile order.  Every pair of the tile's range is written (zeros behind the last
// contributor), so the partials need no clearing beforehand.
__device__ __forceinline__ void blend_bwd_tile(BlendBwdLds& L, const TileView& tv, uint2 range,
                                               const uint32_t* __restrict__ vals, const Splat* __restrict__ splats,
                                               const float* __restrict__ dL_dimg, const uint32_t* __restrict__ offsets,
                                               float* __restrict__ partials, const float* __restrict__ final_T,
                                               const int32_t* __restrict__ final_last) {
    constexpr uint32_t kG = kSplatGrads;
    const uint32_t tid = threadIdx.x, wave = tid / kWave, lane = __lane_id();
    uint32_t pxl, pyl;
    tile_pixel(tv, pxl, pyl);
    const bool inside = pxl < tv.width && pyl < tv.height;
    const float fx = (float)pxl, fy = (float)pyl;
    BlendLds& s = L.s;
    if (tid == 0) L.last = -1;
    for (uint32_t x = tid; x < kTilePixels * kG; x += kTilePixels) (&L.acc[0][0])[x] = 0.0;
    const uint16_t* list = s.list[sub_block_of(wave, lane >> 4)];

    // The forward frame stored each pixel's final transmittance and last blended splat.
    float T = 1.0f;
    int32_t last = -1;
    const size_t o = tv.pix_base + (size_t)pyl * tv.width + pxl;
    if (inside) {
        T = final_T[o];
        last = final_last[o];
    }
    __syncthreads();  // L.last initialised
    if (last >= 0) atomicMax(&L.last, last);
    __syncthreads();
    const int32_t block_last = L.last;
    // Pairs behind the tile's last contributor (all of them when no pixel blended anything) get
    // zero gradients.
    const uint32_t zero_from = block_last < 0 ? range.x
                                              : range.x + (((uint32_t)block_last - range.x) / kTilePixels + 1u) * kTilePixels;
    for (uint32_t k = zero_from + tid; k < range.y; k += kTilePixels)
        store_pair_partial(tv, vals[k], splats, offsets, partials, nullptr);
    if (block_last < 0) return;  // uniform

    // ---- back to front
    float g_r = 0.0f, g_g = 0.0f, g_b = 0.0f;
    if (inside) {
        g_r = dL_dimg[o * 3 + 0];
        g_g = dL_dimg[o * 3 + 1];
        g_b = dL_dimg[o * 3 + 2];
    }
    float acc_r = 0.0f, acc_g = 0.0f, acc_b = 0.0f;  // colour accumulated behind the current splat
    float last_alpha = 0.0f, last_r = 0.0f, last_g = 0.0f, last_b = 0.0f;
    const uint32_t first_chunk = ((uint32_t)block_last - range.x) / kTilePixels;
    // Loads a chunk ahead, as the forward does: while chunk c is walked, the record and the offset of
    // this thread's splat of chunk c - 1 and the value of chunk c - 2 are in flight.  The offset gives
    // the pair's pre-sort position once the record is there, so the partial sums of the thread's slot
    // are stored after the walk without reading anything more.
    const auto value_of = [&](int32_t c) -> uint32_t {
        return c >= 0 ? load_value(vals, range.x + (uint32_t)c * kTilePixels + tid, range.y) : 0u;
    };
    const auto offset_of = [&](int32_t c, uint32_t g) -> uint32_t {
        return c >= 0 && range.x + (uint32_t)c * kTilePixels + tid < range.y ? offsets[g] : 0u;
    };
    uint32_t g_next = value_of((int32_t)first_chunk);
    StagedSplat next = load_staged(splats, g_next, range.x + first_chunk * kTilePixels + tid, range.y);
    uint32_t off_next = offset_of((int32_t)first_chunk, g_next);
    uint32_t g_after = value_of((int32_t)first_chunk - 1);
    for (int32_t chunk = (int32_t)first_chunk; chunk >= 0; --chunk) {
        const uint32_t base = range.x + (uint32_t)chunk * kTilePixels;
        const StagedSplat cur = next;
        const size_t pos = cur.valid ? pair_position(tv, cur.sp.xyzr, off_next) : 0;
        __syncthreads();
        const uint32_t walk = blend_stage_loaded(s, tv, cur);
        if (chunk > 0) {
            next = load_staged(splats, g_after, base - kTilePixels + tid, range.y);
            off_next = offset_of(chunk - 1, g_after);
            g_after = value_of(chunk - 2);
        }
        // Two list entries per step, back to front: both splats' powers, exponentials and alphas
        // are formed first (none depends on the transmittance), then the two updates run in order.
        // A step past the list's start reads the sentinel entry, which no lane takes.
        const auto entry_update = [&](uint32_t off, const float4& A, const float4& B, float dx, float dy,
                                      float G, float raw, float alpha, uint64_t updm) {
            const uint32_t j = off / kSlotBytes;
            const bool upd = lane_on(updm);
            float v[kG];
#pragma unroll
            for (uint32_t q = 0; q < kG; ++q) v[q] = 0.0f;
            if (upd) {
                const float4 C = slot4(s.c, off);
                // transmittance in front of this splat: 1 - alpha >= 0.01, so the hardware reciprocal
                // (about 1 ulp) stands in for the correctly rounded division (the gradients are
                // checked against float64, not bit for bit; plain and batched frames run this same code)
                T = T * __builtin_amdgcn_rcpf(1.0f - alpha);
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
            // Row sums.  v[0..7] by a reduce-scatter: lanes exchange half of their values with the
            // lane across bit 0, then across bit 1 (quad_perm), keeping the halves whose index bits
            // match their own, and the four quads of the row add up by rotations; lane (b0, b1) of
            // every quad then holds the row sums of values 4 b0 + 2 b1 + {0, 1}, and the first quad
            // adds them into the tile's float64 sums (2 LDS atomics per row instead of 8).  v[8] by a
            // plain row reduction.  The order of every addition is fixed.
            const bool b0 = (lane & 1u) != 0u, b1 = (lane & 2u) != 0u;
            float x[4];
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) {
                const float send = b0 ? v[k] : v[4 + k], keep = b0 ? v[4 + k] : v[k];
                x[k] = keep + dpp_f32<0xB1>(send);  // quad_perm [1, 0, 3, 2]
            }
            float y[2];
#pragma unroll
            for (uint32_t m = 0; m < 2; ++m) {
                const float send = b1 ? x[m] : x[2 + m], keep = b1 ? x[2 + m] : x[m];
                y[m] = keep + dpp_f32<0x4E>(send);  // quad_perm [2, 3, 0, 1]
            }
#pragma unroll
            for (uint32_t m = 0; m < 2; ++m) {
                y[m] += dpp_f32<0x124>(y[m]);  // row_ror:4
                y[m] += dpp_f32<0x128>(y[m]);  // row_ror:8
            }
            const float s8 = row_sum16(v[8]);
            if (j < kTilePixels) {
                if ((lane & 12u) == 0u) {
                    const uint32_t q0 = (b0 ? 4u : 0u) + (b1 ? 2u : 0u);
                    if (y[0] != 0.0f) atomicAdd(&L.acc[j][q0], (double)y[0]);
                    if (y[1] != 0.0f) atomicAdd(&L.acc[j][q0 + 1u], (double)y[1]);
                }
                if ((lane & 15u) == 15u && s8 != 0.0f) atomicAdd(&L.acc[j][8], (double)s8);
            }
        };
        for (int32_t i = (int32_t)walk - 1; i >= 0; i -= 2) {
            const uint32_t o1 = list[i], o2 = i >= 1 ? list[i - 1] : kSentinel;
            // lanes whose last blended splat lies at or behind the entry (256 = the sentinel)
            const uint64_t mine1 = ballot((int32_t)(base + o1 / kSlotBytes) <= last);
            const uint64_t mine2 = ballot((int32_t)(base + o2 / kSlotBytes) <= last);
            if ((mine1 | mine2) == 0) continue;
            const float4 A1 = slot4(s.a, o1), B1 = slot4(s.b, o1);
            const float4 A2 = slot4(s.a, o2), B2 = slot4(s.b, o2);
            const float dx1 = A1.x - fx, dy1 = A1.y - fy, dx2 = A2.x - fx, dy2 = A2.y - fy;
            const float p1 = -0.5f * ((A1.z * dx1) * dx1 + (B1.x * dy1) * dy1) - (A1.w * dx1) * dy1;
            const float p2 = -0.5f * ((A2.z * dx2) * dx2 + (B2.x * dy2) * dy2) - (A2.w * dx2) * dy2;
            const uint64_t act1 = ballot(p1 <= 0.0f) & ballot(!(p1 < B1.y)) & mine1;
            const uint64_t act2 = ballot(p2 <= 0.0f) & ballot(!(p2 < B2.y)) & mine2;
            if ((act1 | act2) == 0) continue;
            const float G1 = soft_expf_inrange(p1), G2 = soft_expf_inrange(p2);
            const float raw1 = B1.z * G1, raw2 = B2.z * G2;
            const float alpha1 = raw1 < 0.99f ? raw1 : 0.99f, alpha2 = raw2 < 0.99f ? raw2 : 0.99f;
            const uint64_t updm1 = act1 & ballot(!(alpha1 < 1.0f / 255.0f));
            const uint64_t updm2 = act2 & ballot(!(alpha2 < 1.0f / 255.0f));
            if (updm1 != 0) entry_update(o1, A1, B1, dx1, dy1, G1, raw1, alpha1, updm1);
            if (updm2 != 0) entry_update(o2, A2, B2, dx2, dy2, G2, raw2, alpha2, updm2);
        }
        __syncthreads();
        if (cur.valid) {
            float out[kG];
#pragma unroll
            for (uint32_t q = 0; q < kG; ++q) {
                out[q] = (float)L.acc[tid][q];
                L.acc[tid][q] = 0.0;
            }
            store_partial_at(pos, partials, out);
        }
    }
}

}  // namespace tpd

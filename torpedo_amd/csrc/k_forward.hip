// Forward path kernels: entity matrices, fused preprocess + look-back scan + duplicate,
// tile ranges, and the front-to-back blend.
//
// Reference: torpedo/volumetric/assets/gaussian/{project,prefix,keygen,range,blend}.slang and the
// dispatch sequence of GaussianEngine::recordSplat/recordBlend (GaussianEngine.cpp:777-863).
#include "tpd_kernels.hpp"

namespace tpd {

// ------------------------------------------------------------------ entity matrices

// VM = view * model, PVM = projView * model for every entity (Slang mul order: each entry summed
// k = 0..3 left to right; project.slang:48/57 via common.slang:102,106).
__global__ void k_entity_mats(const float* __restrict__ models, uint32_t n_ent, CameraParams cam,
                              EntityMats* __restrict__ out) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_ent) return;
    float M[16];
    if (models) {
        for (int i = 0; i < 16; ++i) M[i] = models[16 * e + i];
    } else {
        for (int i = 0; i < 16; ++i) M[i] = (i % 5 == 0) ? 1.0f : 0.0f;
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            out[e].vm[i * 4 + j] = ((cam.view[i * 4 + 0] * M[0 * 4 + j] + cam.view[i * 4 + 1] * M[1 * 4 + j])
                                    + cam.view[i * 4 + 2] * M[2 * 4 + j]) + cam.view[i * 4 + 3] * M[3 * 4 + j];
            out[e].pvm[i * 4 + j] = ((cam.proj_view[i * 4 + 0] * M[0 * 4 + j] + cam.proj_view[i * 4 + 1] * M[1 * 4 + j])
                                     + cam.proj_view[i * 4 + 2] * M[2 * 4 + j]) + cam.proj_view[i * 4 + 3] * M[3 * 4 + j];
        }
}

void launch_entity_mats(const float* models, uint32_t n_ent, const CameraParams& cam, EntityMats* out,
                        hipStream_t s) {
    const uint32_t ne = n_ent ? n_ent : 1u;
    hipLaunchKernelGGL(k_entity_mats, dim3((ne + 63) / 64), dim3(64), 0, s, models, ne, cam, out);
}

// ------------------------------------------------------------------ preprocess + scan + duplicate

// Scan look-back descriptor: bits 63..62 = state (0 none, 1 aggregate, 2 inclusive prefix),
// bits 61..0 = value.  One descriptor per partition of kPreBlock Gaussians.
constexpr uint64_t kDescA = 1ull << 62;
constexpr uint64_t kDescP = 2ull << 62;
constexpr uint64_t kDescMask = (1ull << 62) - 1;

__device__ __forceinline__ uint64_t scan_lookback(uint64_t* desc, uint32_t part) {
    // Wave 0 inspects up to 64 predecessors at a time (lane l looks at partition part-1-l).
    uint64_t excl = 0;
    int64_t j = (int64_t)part - 1;
    const uint32_t l = __lane_id();
    while (true) {
        const int64_t idx = j - (int64_t)l;
        const uint64_t d = idx >= 0 ? load_relaxed(&desc[idx]) : kDescP;
        const uint32_t st = (uint32_t)(d >> 62);
        const uint64_t mP = __ballot(st == 2);
        const uint64_t mX = __ballot(st == 0);
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



__global__ __launch_bounds__(kPreBlock) void k_preprocess_dup(const ForwardArgs a) {
    __shared__ uint32_t s_part;
    __shared__ uint32_t s_wave_sum[kPreBlock / kWave];
    __shared__ uint32_t s_incl[kPreBlock];
    __shared__ int32_t s_x0[kPreBlock], s_y0[kPreBlock], s_w[kPreBlock];
    __shared__ uint32_t s_depth[kPreBlock];
    __shared__ uint64_t s_base;

    const uint32_t tid = threadIdx.x;
    if (tid == 0) s_part = atomicAdd(&a.counters[kCtrScanTicket], 1u);
    // Tile ranges are cleared here (every block strides over them) so that the ranges kernel,
    // which runs after the sort, only writes the non-empty tiles.
    for (uint32_t t = blockIdx.x * kPreBlock + tid; t < a.ntiles; t += gridDim.x * kPreBlock)
        a.ranges[t] = make_uint2(0u, 0u);
    __syncthreads();
    const uint32_t part = s_part;
    if (part >= gridDim.x) return;  // tickets are reset every frame; never true in a consistent state
    const uint32_t i = part * kPreBlock + tid;

    Projected p;
    p.tiles = 0;
    p.radius = 0.0f;
    if (i < a.n) {
        const uint32_t e = a.ent_idx ? a.ent_idx[i] : 0u;
        const float* g = a.gauss + (size_t)i * 60;
        const bool vis = project_gaussian(g, a.ent[e].vm, a.ent[e].pvm, a.cam.view,
                                          (0.5f * (float)a.cam.width) * a.cam.focal_ndc[0],
                                          (0.5f * (float)a.cam.height) * a.cam.focal_ndc[1], a.cam.width,
                                          a.cam.height, a.gx, a.gy, a.sh_degree, p);
        if (vis) {
            a.xyzr[i] = make_float4(p.px, p.py, p.vz, p.radius);
            a.conop[i] = make_float4(p.con_a, p.con_b, p.con_c, p.opacity);
            a.rgb[i] = make_float4(p.r, p.g, p.b, 0.0f);
        } else {
            a.xyzr[i] = make_float4(0.f, 0.f, 0.f, 0.f);
            a.conop[i] = make_float4(0.f, 0.f, 0.f, 0.f);
            a.rgb[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        a.tiles[i] = p.tiles;
    }

    // Block-wide inclusive scan of tile counts.
    const uint32_t wave = tid / kWave, lane = __lane_id();
    const uint32_t incl_w = wave_incl_scan(p.tiles);
    if (lane == kWave - 1) s_wave_sum[wave] = incl_w;
    __syncthreads();
    uint32_t wave_off = 0, block_total = 0;
#pragma unroll
    for (uint32_t w = 0; w < kPreBlock / kWave; ++w) {
        const uint32_t v = s_wave_sum[w];
        wave_off += (w < wave) ? v : 0u;
        block_total += v;
    }
    const uint32_t incl = wave_off + incl_w;
    s_incl[tid] = incl;
    if (p.tiles) {
        s_x0[tid] = p.x0;
        s_y0[tid] = p.y0;
        s_w[tid] = p.x1 - p.x0;
        s_depth[tid] = as_uint(p.vz);
    }

    // Decoupled look-back (atomic ticket order guarantees the predecessors are running).
    if (wave == 0) {
        if (lane == 0) store_relaxed(&a.scan_desc[part], (part == 0 ? kDescP : kDescA) | (uint64_t)block_total);
        uint64_t excl = 0;
        if (part > 0) excl = scan_lookback(a.scan_desc, part);
        if (lane == 0) {
            if (part > 0) store_relaxed(&a.scan_desc[part], kDescP | (excl + block_total));
            s_base = excl;
            const uint32_t nparts = (a.n + kPreBlock - 1) / kPreBlock;
            if (part == nparts - 1) {
                const uint64_t total = excl + block_total;
                a.counters[kCtrPairs] = total > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)total;
                a.counters[kCtrOverflow] = total > a.cap ? 1u : 0u;
            }
        }
    }
    __syncthreads();
    const uint64_t base = s_base;
    if (i < a.n) a.offsets[i] = (uint32_t)(base + incl - p.tiles);

    // Duplicate (keygen.slang:21-54): the block emits its pairs cooperatively, pair j of the block
    // belongs to the thread whose inclusive offset first exceeds j; row-major over the rectangle.
    for (uint32_t j = tid; j < block_total; j += kPreBlock) {
        uint32_t lo = 0, hi = kPreBlock - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_incl[mid] > j) hi = mid; else lo = mid + 1;
        }
        const uint32_t owner = lo;
        const uint32_t own_tiles = s_incl[owner] - (owner ? s_incl[owner - 1] : 0u);
        const uint32_t k = j - (s_incl[owner] - own_tiles);
        const uint32_t w = (uint32_t)s_w[owner];
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
        const uint32_t tile = ((uint32_t)s_y0[owner] + ry) * a.gx + ((uint32_t)s_x0[owner] + rx);
        const uint64_t gpos = base + j;
        if (gpos < a.cap) {
            a.keys[gpos] = ((uint64_t)tile << 32) | s_depth[owner];
            a.vals[gpos] = a.val_base + part * kPreBlock + owner;
        }
    }
}

void launch_preprocess_dup(const ForwardArgs& a, hipStream_t s) {
    const uint32_t nb = (a.n + kPreBlock - 1) / kPreBlock;
    if (nb == 0) return;
    hipLaunchKernelGGL(k_preprocess_dup, dim3(nb), dim3(kPreBlock), 0, s, a);
}

// ------------------------------------------------------------------ ranges

// range.slang:16-35.  Empty tiles keep the (0, 0) written by the preprocess kernel (the
// reference clears with fillBuffer, GaussianEngine.cpp:843-848).  The kernel also re-zeroes the
// scan descriptors and the scan ticket for the next frame.
__global__ void k_ranges(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ counters,
                         uint2* __restrict__ ranges, uint64_t* scan_desc, uint32_t n_scan_desc, uint32_t* ticket) {
    const uint32_t P = counters[kCtrPairs];
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nthreads = gridDim.x * blockDim.x;
    for (uint32_t d = gid; d < n_scan_desc; d += nthreads) scan_desc[d] = 0;
    if (gid == 0) *ticket = 0;
    uint32_t* r = reinterpret_cast<uint32_t*>(ranges);
    for (uint32_t idx = gid; idx < P; idx += nthreads) {
        const uint32_t cur = (uint32_t)(keys[idx] >> 32);
        if (idx == 0) {
            r[2 * cur] = 0;
        } else {
            const uint32_t prev = (uint32_t)(keys[idx - 1] >> 32);
            if (cur != prev) {
                r[2 * prev + 1] = idx;
                r[2 * cur] = idx;
            }
        }
        if (idx == P - 1) r[2 * cur + 1] = P;
    }
}

void launch_ranges(const uint64_t* keys, const uint32_t* counters, uint32_t P_hint, uint32_t ntiles, uint2* ranges,
                   uint64_t* scan_desc, uint32_t n_scan_desc, uint32_t* ticket, hipStream_t s) {
    uint32_t work = P_hint > ntiles ? P_hint : ntiles;
    work = work > n_scan_desc ? work : n_scan_desc;
    uint32_t nb = (work + 255) / 256;
    nb = nb < 1 ? 1 : (nb > 8192 ? 8192 : nb);
    hipLaunchKernelGGL(k_ranges, dim3(nb), dim3(256), 0, s, keys, counters, ranges, scan_desc, n_scan_desc, ticket);
}

// ------------------------------------------------------------------ blend

// blend.slang:22-105.  One 16x16 tile per 256-thread block; wave w owns the 8x8 quadrant
// (w & 1, w >> 1).  Gaussians are staged in LDS 256 at a time, the block stops when every pixel
// has terminated.  Per pixel the operations and their order are the reference's; the exact
// shortcuts around them change no pixel:
//   - each staged splat is tested against each quadrant (splat_touches_rect, after a cheap
//     bounding-box test): only a quadrant in which some pixel centre may see it with
//     alpha >= 1/255 gets it in its list (compacted in sorted order);
//   - skip = log(1/(255 o)) - 0.01: a power below it gives o * exp(power) < 0.99/255, so the
//     alpha < 1/255 test would discard the splat anyway; exp() runs only if some lane needs it;
//   - a wave leaves the round as soon as all its pixels have terminated.
// COUNT adds up the pixel-splat pairs whose power is evaluated (for the FLOP roofline).
template <bool COUNT>
__global__ __launch_bounds__(kTilePixels) void k_blend(const BlendArgs a) {
    constexpr uint32_t kWaves = kTilePixels / kWave;
    constexpr uint32_t kQuad = 8;
    __shared__ float4 s_a[kTilePixels];  // x, y, conic a, conic b
    __shared__ float4 s_b[kTilePixels];  // conic c, opacity, skip threshold, -
    __shared__ float4 s_c[kTilePixels];  // colour
    __shared__ uint8_t s_list[kWaves][kTilePixels];
    __shared__ uint32_t s_cnt[kWaves][kWaves];  // [staging wave][quadrant]

    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t tx = tile % a.gx, ty = tile / a.gx;
    const uint32_t tid = threadIdx.x, wave = tid / kWave, lane = __lane_id();
    const uint32_t pxl = tx * kTile + (wave & 1u) * kQuad + (lane % kQuad);
    const uint32_t pyl = ty * kTile + (wave >> 1) * kQuad + (lane / kQuad);
    const bool inside = pxl < a.width && pyl < a.height;
    const float fx = (float)pxl, fy = (float)pyl;
    const uint2 range = a.ranges[tile];
    const uint64_t lt = lanemask_lt();

    float T = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
    bool done = !inside;
    uint32_t evals = 0;
    for (uint32_t base = range.x; base < range.y; base += kTilePixels) {
        if (__syncthreads_count(done) == (int)kTilePixels) break;
        const uint32_t k = base + tid;
        uint32_t mask = 0;
        if (k < range.y) {
            const uint32_t g = a.vals[k];
            const float4 p = a.xyzr[g];
            const float4 c = a.conop[g];
            s_a[tid] = make_float4(p.x, p.y, c.x, c.y);
            s_b[tid] = make_float4(c.z, c.w, __logf(1.0f / (255.0f * c.w)) - 0.01f, 0.0f);
            s_c[tid] = a.rgb[g];
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

        if (__ballot(!done) == 0) continue;
        for (uint32_t i = 0; i < cnt; ++i) {
            const uint32_t j = s_list[wave][i];
            const float4 A = s_a[j];
            const float4 B = s_b[j];
            if (COUNT) evals += done ? 0u : 1u;
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
            const float4 C = s_c[j];
            cr = upd ? cr + (C.x * alpha) * T : cr;
            cg = upd ? cg + (C.y * alpha) * T : cg;
            cb = upd ? cb + (C.z * alpha) * T : cb;
            T = upd ? tn : T;
            done = done || stop;
            if (__ballot(stop) != 0 && __ballot(!done) == 0) break;
        }
    }
    if (COUNT) {
        const uint64_t s = wave_sum_u64(evals);
        if (lane == 0) atomicAdd(reinterpret_cast<unsigned long long*>(a.eval_count), (unsigned long long)s);
    }
    if (inside) {
        const size_t o = (size_t)pyl * a.width + pxl;
        if (a.image) {
            a.image[o * 3 + 0] = cr;
            a.image[o * 3 + 1] = cg;
            a.image[o * 3 + 2] = cb;
        }
        if (a.rgba8) {
            uchar4 q;
            q.x = unorm8(cr); q.y = unorm8(cg); q.z = unorm8(cb); q.w = 255;
            reinterpret_cast<uchar4*>(a.rgba8)[o] = q;
        }
    }
}

void launch_blend(const BlendArgs& a, hipStream_t s) {
    if (a.ntiles == 0) return;
    if (a.eval_count)
        hipLaunchKernelGGL(k_blend<true>, dim3(a.ntiles), dim3(kTilePixels), 0, s, a);
    else
        hipLaunchKernelGGL(k_blend<false>, dim3(a.ntiles), dim3(kTilePixels), 0, s, a);
}

}  // namespace tpd

// Stable LSD radix sort of (64-bit key, 32-bit value) pairs, 8-bit digits, one pass per digit.
//
// Replaces the reference's 2-bit "4-way" sorter (radix-shuffle/-prefixA/-prefixB/-mapping.slang,
// 4 dispatches per 2 bits, GaussianEngine.cpp:819-841) with:
//   k_sort_hist + k_sort_offsets  one read of the keys computes the digit histograms of every pass
//                                 and their global exclusive offsets;
//   k_onesweep (once per pass)    each 256-thread block takes the next 4096-key partition from an
//                                 atomic ticket, ranks its keys by digit with wave64 ballot
//                                 matching (stable: element order inside a wave, then waves in
//                                 order), publishes its per-digit counts with a decoupled
//                                 look-back, reorders the partition in LDS and writes every digit
//                                 run out contiguously.
// The result equals a stable sort by key, i.e. the reference's order (tile, depth, index).
#include "tpd_kernels.hpp"

namespace tpd {

constexpr uint32_t kFlagA = 1u << 30;
constexpr uint32_t kFlagP = 2u << 30;
constexpr uint32_t kValMask = (1u << 30) - 1u;

__device__ __forceinline__ uint32_t digit_of(uint64_t k, uint32_t shift) {
    return (uint32_t)(k >> shift) & (kRadix - 1u);
}

// ------------------------------------------------------------------ histograms

__global__ __launch_bounds__(256) void k_sort_hist(const uint64_t* __restrict__ keys, const uint32_t* count_ptr,
                                                   uint32_t n_imm, uint32_t passes, uint32_t* __restrict__ partials,
                                                   uint32_t* __restrict__ desc0, uint32_t* __restrict__ tickets) {
    __shared__ uint32_t h[kMaxPasses][kRadix];
    const uint32_t n = count_ptr ? *count_ptr : n_imm;
    const uint32_t tid = threadIdx.x;
#pragma unroll
    for (uint32_t p = 0; p < kMaxPasses; ++p) h[p][tid] = 0;
    // The first onesweep pass needs region 0 of the look-back descriptors and all tickets at zero.
    const uint32_t nparts = sort_parts(n);
    for (uint32_t x = blockIdx.x * 256 + tid; x < nparts * kRadix; x += gridDim.x * 256) desc0[x] = 0;
    if (blockIdx.x == 0 && tid < kMaxPasses) tickets[tid] = 0;
    __syncthreads();

    const uint32_t per = ((n + gridDim.x - 1) / gridDim.x + 63u) & ~63u;
    const uint32_t lo = blockIdx.x * per;
    const uint32_t hi = min(n, lo + per);
    for (uint32_t i = lo + tid; i < hi; i += 256) {
        const uint64_t k = keys[i];
#pragma unroll
        for (uint32_t p = 0; p < kMaxPasses; ++p)
            if (p < passes) atomicAdd(&h[p][digit_of(k, p * kRadixBits)], 1u);
    }
    __syncthreads();
    for (uint32_t p = 0; p < passes; ++p) partials[((size_t)blockIdx.x * kMaxPasses + p) * kRadix + tid] = h[p][tid];
}

// One block per pass: sum the block histograms, exclusive scan over digits.
__global__ __launch_bounds__(256) void k_sort_offsets(const uint32_t* __restrict__ partials, uint32_t nblocks,
                                                      uint32_t* __restrict__ offsets) {
    __shared__ uint32_t s_w[4];
    const uint32_t p = blockIdx.x, d = threadIdx.x;
    uint32_t s = 0;
    for (uint32_t b = 0; b < nblocks; ++b) s += partials[((size_t)b * kMaxPasses + p) * kRadix + d];
    const uint32_t incl = wave_incl_scan(s);
    const uint32_t w = d / kWave;
    if (__lane_id() == kWave - 1) s_w[w] = incl;
    __syncthreads();
    uint32_t off = 0;
    for (uint32_t q = 0; q < w; ++q) off += s_w[q];
    offsets[p * kRadix + d] = off + incl - s;
}

// ------------------------------------------------------------------ one pass

struct OnesweepArgs {
    const uint64_t* keys_in;
    const uint32_t* vals_in;
    uint64_t* keys_out;
    uint32_t* vals_out;
    const uint32_t* count_ptr;
    uint32_t n_imm;
    uint32_t shift;
    const uint32_t* offsets;   // kRadix global exclusive offsets of this pass's digits
    uint32_t* desc;            // this pass's look-back region (parts x kRadix), zero on entry
    uint32_t* desc_next;       // the other region, cleared here for the next pass
    uint32_t* ticket;
};

__global__ __launch_bounds__(kSortBlock) void k_onesweep(const OnesweepArgs a) {
    constexpr uint32_t kWaves = kSortBlock / kWave;
    __shared__ uint32_t s_part;
    __shared__ uint32_t s_whist[kWaves][kRadix];
    __shared__ uint32_t s_bstart[kRadix];
    __shared__ uint32_t s_adj[kRadix];
    __shared__ uint32_t s_wtot[kWaves];
    __shared__ uint64_t s_keys[kSortTile];
    __shared__ uint8_t s_dig[kSortTile];

    const uint32_t n = a.count_ptr ? *a.count_ptr : a.n_imm;
    const uint32_t nparts = sort_parts(n);
    const uint32_t tid = threadIdx.x, wave = tid / kWave, lane = __lane_id();
    if (tid == 0) s_part = atomicAdd(a.ticket, 1u);
#pragma unroll
    for (uint32_t w = 0; w < kWaves; ++w) s_whist[w][tid] = 0;
    __syncthreads();
    const uint32_t part = s_part;
    if (part >= nparts) return;
    if (a.desc_next) a.desc_next[part * kRadix + tid] = 0;

    // Load 16 keys per thread, wave-striped: element e = wave*1024 + j*64 + lane.
    const uint32_t base = part * kSortTile;
    uint64_t k[kSortItems];
    uint32_t v[kSortItems];
#pragma unroll
    for (uint32_t j = 0; j < kSortItems; ++j) {
        const uint32_t idx = base + wave * (kSortItems * kWave) + j * kWave + lane;
        const bool ok = idx < n;
        k[j] = ok ? a.keys_in[idx] : ~0ull;
        v[j] = ok ? a.vals_in[idx] : 0u;
    }

    // Rank within the wave: peers with the same digit found with 8 ballots; the lowest lane of
    // each group advances the wave's running digit counter.
    const uint64_t lt = lanemask_lt();
    uint32_t rank[kSortItems];
#pragma unroll
    for (uint32_t j = 0; j < kSortItems; ++j) {
        const uint32_t d = digit_of(k[j], a.shift);
        uint64_t peers = ~0ull;
#pragma unroll
        for (uint32_t b = 0; b < kRadixBits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t below = (uint32_t)__popcll(peers & lt);
        const uint32_t pre = s_whist[wave][d];
        if (below == 0) s_whist[wave][d] = pre + (uint32_t)__popcll(peers);
        rank[j] = pre + below;
    }
    __syncthreads();

    // Per digit (thread = digit): wave exclusive offsets and the partition's digit count.
    const uint32_t c0 = s_whist[0][tid], c1 = s_whist[1][tid], c2 = s_whist[2][tid], c3 = s_whist[3][tid];
    const uint32_t tot = c0 + c1 + c2 + c3;
    store_relaxed(&a.desc[part * kRadix + tid], (part == 0 ? kFlagP : kFlagA) | tot);
    s_whist[0][tid] = 0;
    s_whist[1][tid] = c0;
    s_whist[2][tid] = c0 + c1;
    s_whist[3][tid] = c0 + c1 + c2;
    const uint32_t incl = wave_incl_scan(tot);
    if (lane == kWave - 1) s_wtot[wave] = incl;
    __syncthreads();
    uint32_t bstart = incl - tot;
    for (uint32_t w = 0; w < wave; ++w) bstart += s_wtot[w];
    s_bstart[tid] = bstart;
    __syncthreads();

    // Reorder the partition by digit in LDS (local position = digit start + wave offset + rank).
#pragma unroll
    for (uint32_t j = 0; j < kSortItems; ++j) {
        const uint32_t d = digit_of(k[j], a.shift);
        const uint32_t lp = s_bstart[d] + s_whist[wave][d] + rank[j];
        rank[j] = lp;
        s_keys[lp] = k[j];
        s_dig[lp] = (uint8_t)d;
    }

    // Decoupled look-back for digit tid.
    uint32_t excl = 0;
    if (part > 0) {
        int32_t jj = (int32_t)part - 1;
        while (true) {
            const uint32_t dsc = load_relaxed(&a.desc[(uint32_t)jj * kRadix + tid]);
            const uint32_t f = dsc & ~kValMask;
            if (f == 0) {
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            excl += dsc & kValMask;
            if (f == kFlagP) break;
            --jj;
        }
        store_relaxed(&a.desc[part * kRadix + tid], kFlagP | (excl + tot));
    }
    s_adj[tid] = a.offsets[tid] + excl - bstart;
    __syncthreads();

    const uint32_t nvalid = min(kSortTile, n - base);
#pragma unroll 4
    for (uint32_t t = 0; t < kSortItems; ++t) {
        const uint32_t i = t * kSortBlock + tid;
        if (i < nvalid) a.keys_out[s_adj[s_dig[i]] + i] = s_keys[i];
    }
    __syncthreads();
    uint32_t* s_vals = reinterpret_cast<uint32_t*>(s_keys);
#pragma unroll
    for (uint32_t j = 0; j < kSortItems; ++j) s_vals[rank[j]] = v[j];
    __syncthreads();
#pragma unroll 4
    for (uint32_t t = 0; t < kSortItems; ++t) {
        const uint32_t i = t * kSortBlock + tid;
        if (i < nvalid) a.vals_out[s_adj[s_dig[i]] + i] = s_vals[i];
    }
}

bool launch_sort_pairs(SortWorkspace& ws, uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                       const uint32_t* count_ptr, uint32_t n_max, uint32_t key_bits, hipStream_t s,
                       hipEvent_t after_hist) {
    const uint32_t passes = sort_passes(key_bits);
    const uint32_t parts = sort_parts(n_max);
    if (parts == 0 || passes == 0) return false;
    uint32_t hb = (n_max + 8191u) / 8192u;
    hb = hb < 1 ? 1 : (hb > kHistBlocks ? kHistBlocks : hb);
    hipLaunchKernelGGL(k_sort_hist, dim3(hb), dim3(256), 0, s, keys, count_ptr, n_max, passes, ws.partials, ws.desc,
                       ws.tickets);
    hipLaunchKernelGGL(k_sort_offsets, dim3(passes), dim3(256), 0, s, ws.partials, hb, ws.offsets);
    if (after_hist) hipEventRecord(after_hist, s);
    const uint64_t* kin = keys;
    const uint32_t* vin = vals;
    uint64_t* kout = keys_alt;
    uint32_t* vout = vals_alt;
    for (uint32_t p = 0; p < passes; ++p) {
        OnesweepArgs o;
        o.keys_in = kin; o.vals_in = vin; o.keys_out = kout; o.vals_out = vout;
        o.count_ptr = count_ptr; o.n_imm = n_max;
        o.shift = p * kRadixBits;
        o.offsets = ws.offsets + p * kRadix;
        o.desc = ws.desc + (size_t)(p & 1u) * ws.parts_cap * kRadix;
        o.desc_next = ws.desc + (size_t)((p + 1) & 1u) * ws.parts_cap * kRadix;
        o.ticket = ws.tickets + p;
        hipLaunchKernelGGL(k_onesweep, dim3(parts), dim3(kSortBlock), 0, s, o);
        const uint64_t* tk = kout; kout = const_cast<uint64_t*>(kin); kin = tk;
        const uint32_t* tv = vout; vout = const_cast<uint32_t*>(vin); vin = tv;
    }
    return (passes & 1u) != 0;
}

}  // namespace tpd

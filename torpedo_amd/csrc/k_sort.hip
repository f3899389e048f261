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

// Keys are (high word << 32) | low word.  When the low word is known to use only lo_bits bits
// (a positive float depth has bit 31 clear), the sort runs on the compacted key
// (high << lo_bits) | low, which orders exactly as the full key and can need one pass less.
__device__ __forceinline__ uint64_t compact_key(uint64_t k, uint32_t lo_bits) {
    return ((k >> 32) << lo_bits) | (k & ((1ull << lo_bits) - 1ull));
}

// ------------------------------------------------------------------ histograms

// Digit histograms of every pass in one read of the keys.  Consecutive keys often share a digit
// (all pairs of one Gaussian share its depth bits; neighbouring tiles share the high tile bits),
// so each wave first collapses runs of equal digits (run heads found with one shuffle and one
// ballot) and only run heads add their run length to the block's LDS histogram.  Each thread
// keeps kHistUnroll loads in flight; blocks add their non-zero bins to the global counts
// (kMaxPasses x kRadix, zero on entry; k_sort_offsets re-zeroes them after use).
constexpr uint32_t kHistUnroll = 8;

__global__ __launch_bounds__(256) void k_sort_hist(const uint64_t* __restrict__ keys, const uint32_t* count_ptr,
                                                   uint32_t n_imm, uint32_t passes, uint32_t lo_bits, uint32_t tile_keys,
                                                   uint32_t* __restrict__ counts, uint32_t* __restrict__ desc0,
                                                   uint32_t* __restrict__ tickets) {
    __shared__ uint32_t h[kMaxPasses][kRadix];
    const uint32_t n = count_ptr ? min(*count_ptr, n_imm) : n_imm;  // device count, capped by the capacity
    const uint32_t tid = threadIdx.x, lane = __lane_id();
#pragma unroll
    for (uint32_t p = 0; p < kMaxPasses; ++p) h[p][tid] = 0;
    // The first onesweep pass needs region 0 of the look-back descriptors and all tickets at zero.
    const uint32_t nparts = (n + tile_keys - 1) / tile_keys;
    for (uint32_t x = blockIdx.x * 256 + tid; x < nparts * kRadix; x += gridDim.x * 256) desc0[x] = 0;
    if (blockIdx.x == 0 && tid < kMaxPasses) tickets[tid] = 0;
    __syncthreads();

    constexpr uint32_t kChunk = 256 * kHistUnroll;
    const uint64_t above = lane == 63 ? 0ull : (~0ull << (lane + 1));
    for (uint32_t c0 = blockIdx.x * kChunk; c0 < n; c0 += gridDim.x * kChunk) {
        uint64_t k[kHistUnroll];
#pragma unroll
        for (uint32_t u = 0; u < kHistUnroll; ++u) {
            const uint32_t i = c0 + u * 256 + tid;
            k[u] = i < n ? keys[i] : 0ull;
        }
#pragma unroll
        for (uint32_t u = 0; u < kHistUnroll; ++u) k[u] = compact_key(k[u], lo_bits);
#pragma unroll
        for (uint32_t u = 0; u < kHistUnroll; ++u) {
            const uint32_t i = c0 + u * 256 + tid;
            const bool valid = i < n;
            const uint64_t vmask = ballot(valid);
#pragma unroll
            for (uint32_t p = 0; p < kMaxPasses; ++p) {
                if (p >= passes) break;
                const uint32_t d = digit_of(k[u], p * kRadixBits);
                const uint32_t prev = __shfl_up(d, 1, 64);
                const bool head = valid && (lane == 0 || prev != d);
                const uint64_t heads = ballot(head) | ~vmask;  // invalid lanes end runs
                if (head) {
                    const uint64_t nxt = heads & above;
                    const uint32_t end = nxt ? (uint32_t)__builtin_ctzll(nxt) : 64u;
                    atomicAdd(&h[p][d], end - lane);
                }
            }
        }
    }
    __syncthreads();
    for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t v = h[p][tid];
        if (v) atomicAdd(&counts[p * kRadix + tid], v);
    }
}

// One block per pass: exclusive scan of the digit counts, which are then re-zeroed.
__global__ __launch_bounds__(256) void k_sort_offsets(uint32_t* __restrict__ counts, uint32_t* __restrict__ offsets) {
    __shared__ uint32_t s_w[4];
    const uint32_t p = blockIdx.x, d = threadIdx.x;
    const uint32_t s = counts[p * kRadix + d];
    counts[p * kRadix + d] = 0;
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
    uint32_t lo_bits;          // bits of the key's low word (32, or 31 for positive depths)
    const uint32_t* offsets;   // kRadix global exclusive offsets of this pass's digits
    uint32_t* desc;            // this pass's look-back region (parts x kRadix), zero on entry
    uint32_t* desc_next;       // the other region, cleared here for the next pass
    uint32_t* ticket;
};

// BLOCK threads (256 or 512) x ITEMS keys per partition.  Threads 0..255 own one digit each for
// the per-digit bookkeeping and the look-back.
template <uint32_t BLOCK, uint32_t ITEMS>
__global__ __launch_bounds__(BLOCK) void k_onesweep(const OnesweepArgs a) {
    constexpr uint32_t kWaves = BLOCK / kWave;
    constexpr uint32_t kTileKeys = BLOCK * ITEMS;
    static_assert(BLOCK == 256 || BLOCK == 512, "one or two threads per digit");
    __shared__ uint32_t s_part;
    __shared__ uint32_t s_whist[kWaves][kRadix];
    __shared__ uint32_t s_bstart[kRadix];
    __shared__ uint32_t s_adj[kRadix];
    __shared__ uint32_t s_wtot[4];
    __shared__ uint64_t s_keys[kTileKeys];
    __shared__ uint8_t s_dig[kTileKeys];

    const uint32_t n = a.count_ptr ? min(*a.count_ptr, a.n_imm) : a.n_imm;
    const uint32_t nparts = (n + kTileKeys - 1) / kTileKeys;
    const uint32_t tid = threadIdx.x, wave = tid / kWave, lane = __lane_id();
    if (tid == 0) s_part = atomicAdd(a.ticket, 1u);
    for (uint32_t x = tid; x < kWaves * kRadix; x += BLOCK) (&s_whist[0][0])[x] = 0;
    __syncthreads();
    const uint32_t part = __builtin_amdgcn_readfirstlane(s_part);  // block-uniform: keep it scalar
    if (part >= nparts) return;
    if (a.desc_next && tid < kRadix) a.desc_next[part * kRadix + tid] = 0;

    // Load ITEMS keys per thread, wave-striped: element e = wave*(ITEMS*64) + j*64 + lane.
    const uint32_t base = part * kTileKeys;
    const uint32_t wbase = base + wave * (ITEMS * kWave) + lane;
    uint64_t k[ITEMS];
    uint32_t v[ITEMS];
    // Every partition but the last is full: its loads and stores need no bounds checks.
    const bool full = base + kTileKeys <= n;
    if (full) {
#pragma unroll
        for (uint32_t j = 0; j < ITEMS; ++j) k[j] = a.keys_in[wbase + j * kWave];
#pragma unroll
        for (uint32_t j = 0; j < ITEMS; ++j) v[j] = a.vals_in[wbase + j * kWave];
    } else {
#pragma unroll
        for (uint32_t j = 0; j < ITEMS; ++j) {
            const uint32_t idx = wbase + j * kWave;
            k[j] = idx < n ? a.keys_in[idx] : ~0ull;
        }
#pragma unroll
        for (uint32_t j = 0; j < ITEMS; ++j) {
            const uint32_t idx = wbase + j * kWave;
            v[j] = idx < n ? a.vals_in[idx] : 0u;
        }
    }

    // Rank within the wave: peers with the same digit found with 8 ballots; the lowest lane of
    // each group advances the wave's running digit counter.
    const uint64_t lt = lanemask_lt();
    uint32_t rank[ITEMS], dig[ITEMS];
#pragma unroll
    for (uint32_t j = 0; j < ITEMS; ++j) dig[j] = digit_of(compact_key(k[j], a.lo_bits), a.shift);
#pragma unroll
    for (uint32_t j = 0; j < ITEMS; ++j) {
        const uint32_t d = dig[j];
        uint64_t peers = ~0ull;
#pragma unroll
        for (uint32_t b = 0; b < kRadixBits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t below = (uint32_t)__popcll(peers & lt);
        const uint32_t pre = s_whist[wave][d];
        if (below == 0) s_whist[wave][d] = pre + (uint32_t)__popcll(peers);
        rank[j] = pre + below;
    }
    __syncthreads();

    // Per digit (thread = digit): wave exclusive offsets and the partition's digit count.
    uint32_t tot = 0, bstart = 0;
    if (tid < kRadix) {
#pragma unroll
        for (uint32_t w = 0; w < kWaves; ++w) {
            const uint32_t c = s_whist[w][tid];
            s_whist[w][tid] = tot;
            tot += c;
        }
        store_relaxed(&a.desc[part * kRadix + tid], (part == 0 ? kFlagP : kFlagA) | tot);
        const uint32_t incl = wave_incl_scan(tot);
        if (lane == kWave - 1) s_wtot[wave] = incl;
        bstart = incl - tot;
    }
    __syncthreads();
    if (tid < kRadix) {
        for (uint32_t w = 0; w < wave; ++w) bstart += s_wtot[w];
        s_bstart[tid] = bstart;
    }
    __syncthreads();

    // Reorder the partition by digit in LDS (local position = digit start + wave offset + rank).
#pragma unroll
    for (uint32_t j = 0; j < ITEMS; ++j) {
        const uint32_t d = dig[j];
        const uint32_t lp = s_bstart[d] + s_whist[wave][d] + rank[j];
        rank[j] = lp;
        s_keys[lp] = k[j];
        s_dig[lp] = (uint8_t)d;
    }

    // Decoupled look-back for digit tid: kLookback predecessor descriptors are loaded at once
    // (before partition 0 a virtual inclusive prefix of 0), consumed in order up to the first
    // inclusive prefix; an unpublished one is re-read after a short sleep.
    if (tid < kRadix) {
        constexpr int32_t kLookback = 8;
        uint32_t excl = 0;
        if (part > 0) {
            int32_t jj = (int32_t)part - 1;
            while (true) {
                uint32_t dsc[kLookback];
#pragma unroll
                for (int32_t u = 0; u < kLookback; ++u)
                    dsc[u] = jj - u >= 0 ? load_relaxed(&a.desc[(uint32_t)(jj - u) * kRadix + tid]) : kFlagP;
                uint32_t add = 0;
                int32_t used = 0;
                bool found = false, blocked = false;
#pragma unroll
                for (int32_t u = 0; u < kLookback; ++u) {
                    if (!found && !blocked) {
                        const uint32_t f = dsc[u] & ~kValMask;
                        if (f == 0) {
                            blocked = true;
                        } else {
                            add += dsc[u] & kValMask;
                            used = u + 1;
                            found = f == kFlagP;
                        }
                    }
                }
                excl += add;
                jj -= used;
                if (found) break;
                if (blocked) __builtin_amdgcn_s_sleep(1);
            }
            store_relaxed(&a.desc[part * kRadix + tid], kFlagP | (excl + tot));
        }
        s_adj[tid] = a.offsets[tid] + excl - bstart;
    }
    __syncthreads();

    const uint32_t nvalid = min(kTileKeys, n - base);
#pragma unroll 4
    for (uint32_t t = 0; t < ITEMS; ++t) {
        const uint32_t i = t * BLOCK + tid;
        if (full || i < nvalid) a.keys_out[s_adj[s_dig[i]] + i] = s_keys[i];
    }
    __syncthreads();
    uint32_t* s_vals = reinterpret_cast<uint32_t*>(s_keys);
#pragma unroll
    for (uint32_t j = 0; j < ITEMS; ++j) s_vals[rank[j]] = v[j];
    __syncthreads();
#pragma unroll 4
    for (uint32_t t = 0; t < ITEMS; ++t) {
        const uint32_t i = t * BLOCK + tid;
        if (full || i < nvalid) a.vals_out[s_adj[s_dig[i]] + i] = s_vals[i];
    }
}

// Block shapes the tuner chooses from: {threads, keys per thread}.
struct SortShape {
    uint32_t block, items;
    void (*kernel)(const OnesweepArgs);
};
static const SortShape kShapes[] = {
    {256, 16, k_onesweep<256, 16>}, {256, 8, k_onesweep<256, 8>},   {256, 12, k_onesweep<256, 12>},
    {512, 8, k_onesweep<512, 8>},   {512, 12, k_onesweep<512, 12>}, {256, 20, k_onesweep<256, 20>},
};
constexpr uint32_t kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

uint32_t sort_num_shapes() { return kNumShapes; }
uint32_t sort_shape_keys(uint32_t shape) { return kShapes[shape % kNumShapes].block * kShapes[shape % kNumShapes].items; }

bool launch_sort_pairs(SortWorkspace& ws, uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                       const uint32_t* count_ptr, uint32_t n_max, uint32_t key_bits, hipStream_t s,
                       hipEvent_t after_hist, uint32_t lo_bits) {
    const SortShape& sh = kShapes[ws.shape % kNumShapes];
    const uint32_t tile = sh.block * sh.items;
    const uint32_t passes = sort_passes(key_bits);
    const uint32_t parts = (n_max + tile - 1) / tile;
    if (parts == 0 || passes == 0) return false;
    uint32_t hb = (n_max + 256 * kHistUnroll * 4 - 1) / (256 * kHistUnroll * 4);
    hb = hb < 1 ? 1 : (hb > kHistBlocks ? kHistBlocks : hb);
    hipLaunchKernelGGL(k_sort_hist, dim3(hb), dim3(256), 0, s, keys, count_ptr, n_max, passes, lo_bits, tile, ws.partials,
                       ws.desc, ws.tickets);
    hipLaunchKernelGGL(k_sort_offsets, dim3(passes), dim3(256), 0, s, ws.partials, ws.offsets);
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
        o.lo_bits = lo_bits;
        o.offsets = ws.offsets + p * kRadix;
        o.desc = ws.desc + (size_t)(p & 1u) * ws.parts_cap * kRadix;
        o.desc_next = ws.desc + (size_t)((p + 1) & 1u) * ws.parts_cap * kRadix;
        o.ticket = ws.tickets + p;
        hipLaunchKernelGGL(sh.kernel, dim3(parts), dim3(sh.block), 0, s, o);
        const uint64_t* tk = kout; kout = const_cast<uint64_t*>(kin); kin = tk;
        const uint32_t* tv = vout; vout = const_cast<uint32_t*>(vin); vin = tv;
    }
    return (passes & 1u) != 0;
}

}  // namespace tpd

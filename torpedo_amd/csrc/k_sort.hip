// Stable LSD radix sort of (64-bit key, 32-bit value) pairs, one pass per digit of BITS bits
// (8 or 9; the autotuner picks the shape, digit width included).
//
// Replaces the reference's 2-bit "4-way" sorter (radix-shuffle/-prefixA/-prefixB/-mapping.slang,
// 4 dispatches per 2 bits, GaussianEngine.cpp:819-841) with:
//   k_sort_hist + k_sort_offsets  one read of the keys computes the digit histograms of every pass
//                                 and their global exclusive offsets;
//   k_onesweep (once per pass)    each block takes the next partition of BLOCK x ITEMS keys from an
//                                 atomic ticket, ranks its keys by digit with wave64 ballot
//                                 matching (stable: element order inside a wave, then waves in
//                                 order), publishes its per-digit counts with a decoupled
//                                 look-back, reorders the partition in LDS and writes every digit
//                                 run out contiguously.
// The result equals a stable sort by key, i.e. the reference's order (tile, depth, index).
#include "tpd_kernels.hpp"

#include <type_traits>

namespace tpd {

constexpr uint32_t kFlagA = 1u << 30;
constexpr uint32_t kFlagP = 2u << 30;
constexpr uint32_t kValMask = (1u << 30) - 1u;
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;  // kWideDrop: the rank of a dropped key

template <uint32_t BITS, typename KT>
__device__ __forceinline__ uint32_t digit_of(KT k, uint32_t shift) {
    return (uint32_t)(k >> shift) & ((1u << BITS) - 1u);
}

// Keys are (high word << 32) | low word.  When every low word is known to lie in
// [lo_base, lo_base + 2^lo_bits) (a positive float depth has bit 31 clear; a depth in the window
// has bits in [kDepthBase, kDepthBase + 2^27)), the sort runs on the compacted key
// (high << lo_bits) | (low - lo_base), which orders exactly as the full key and can need fewer
// passes.  The difference is masked to lo_bits bits, so the high word's bits survive any key.
__device__ __forceinline__ uint64_t compact_key(uint64_t k, uint32_t lo_bits, uint32_t lo_base) {
    return ((k >> 32) << lo_bits) | ((uint64_t)((uint32_t)k - lo_base) & ((1ull << lo_bits) - 1ull));
}

// ------------------------------------------------------------------ histograms

// Digit histograms of every pass in one read of the keys.  Consecutive keys often share a digit
// (all pairs of one Gaussian share its depth bits; neighbouring tiles share the high tile bits),
// so each wave first collapses runs of equal digits (run heads found with one shuffle and one
// ballot) and only run heads add their run length to the block's LDS histogram.  Each thread
// keeps kHistUnroll loads in flight; blocks add their non-zero bins to the global counts, block b
// into copy b mod kCountCopies (kCountCopies x kMaxPasses x kMaxRadix, zero on entry), so that
// each counter takes 1/kCountCopies of the blocks' atomics; k_sort_offsets sums the copies and
// re-zeroes them.
constexpr uint32_t kHistUnroll = 8;

// The key of the lane below (lane 0 keeps its own) by a wave-wide DPP shift: register moves, no
// LDS traffic (a shuffle per digit and pass goes through the LDS crossbar).
constexpr int kWaveShr1 = 0x138;  // DPP control wave_shr:1
__device__ __forceinline__ uint64_t key_of_lane_below(uint64_t k) {
    const int lo = (int)(uint32_t)k, hi = (int)(uint32_t)(k >> 32);
    const uint32_t plo = (uint32_t)__builtin_amdgcn_update_dpp(lo, lo, kWaveShr1, 0xF, 0xF, false);
    const uint32_t phi = (uint32_t)__builtin_amdgcn_update_dpp(hi, hi, kWaveShr1, 0xF, 0xF, false);
    return ((uint64_t)phi << 32) | plo;
}
__device__ __forceinline__ uint32_t key_of_lane_below(uint32_t k) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)k, (int)k, kWaveShr1, 0xF, 0xF, false);
}

// KT = uint64_t: (tile | depth) keys, histogrammed on their compacted form; KT = uint32_t: keys
// sorted as they are (lo_bits and lo_base unused).
// Only the key's low key_bits bits (key_mask) order it; bits above ride along.
template <uint32_t BITS, typename KT>
__global__ __launch_bounds__(256) void k_sort_hist(const KT* __restrict__ keys, const uint32_t* count_ptr,
                                                   uint32_t n_imm, uint32_t passes, uint32_t lo_bits, uint32_t lo_base,
                                                   KT key_mask, uint32_t tile_keys,
                                                   uint32_t* __restrict__ counts, uint32_t* __restrict__ desc0,
                                                   uint32_t* __restrict__ tickets) {
    constexpr uint32_t R = 1u << BITS;
    __shared__ uint32_t h[kMaxPasses][R];
    const uint32_t n = count_ptr ? min(*count_ptr, n_imm) : n_imm;  // device count, capped by the capacity
    const uint32_t tid = threadIdx.x, lane = __lane_id();
    for (uint32_t x = tid; x < kMaxPasses * R; x += 256) (&h[0][0])[x] = 0;
    // The first onesweep pass needs region 0 of the look-back descriptors and all tickets at zero.
    const uint32_t nparts = (n + tile_keys - 1) / tile_keys;
    for (uint32_t x = blockIdx.x * 256 + tid; x < nparts * R; x += gridDim.x * 256) desc0[x] = 0;
    if (blockIdx.x == 0 && tid < kMaxPasses) tickets[tid] = 0;
    __syncthreads();

    constexpr uint32_t kChunk = 256 * kHistUnroll;
    const uint64_t above = lane == 63 ? 0ull : (~0ull << (lane + 1));
    for (uint32_t c0 = blockIdx.x * kChunk; c0 < n; c0 += gridDim.x * kChunk) {
        KT k[kHistUnroll];
#pragma unroll
        for (uint32_t u = 0; u < kHistUnroll; ++u) {
            const uint32_t i = c0 + u * 256 + tid;
            k[u] = i < n ? keys[i] : KT(0);
        }
        if constexpr (sizeof(KT) == 8) {
#pragma unroll
            for (uint32_t u = 0; u < kHistUnroll; ++u) k[u] = compact_key(k[u], lo_bits, lo_base);
        }
#pragma unroll
        for (uint32_t u = 0; u < kHistUnroll; ++u) k[u] &= key_mask;
#pragma unroll
        for (uint32_t u = 0; u < kHistUnroll; ++u) {
            const uint32_t i = c0 + u * 256 + tid;
            const bool valid = i < n;
            const uint64_t vmask = ballot(valid);
            const KT below = key_of_lane_below(k[u]);
#pragma unroll
            for (uint32_t p = 0; p < kMaxPasses; ++p) {
                if (p >= passes) break;
                const uint32_t d = digit_of<BITS>(k[u], p * BITS);
                const uint32_t prev = digit_of<BITS>(below, p * BITS);
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
    for (uint32_t p = 0; p < passes; ++p)
        for (uint32_t d = tid; d < R; d += 256) {
            const uint32_t v = h[p][d];
            if (v) atomicAdd(&counts[(blockIdx.x % kCountCopies) * (kMaxPasses * kMaxRadix) + p * R + d], v);
        }
}

// One block per pass: exclusive scan of the 2^BITS digit counts summed over their kCountCopies
// copies (each thread takes a contiguous run of digits); every copy is then re-zeroed.  With
// total_out, pass 0's block stores the number of keys counted there.
template <uint32_t BITS>
__global__ __launch_bounds__(256) void k_sort_offsets(uint32_t* __restrict__ counts, uint32_t* __restrict__ offsets,
                                                      uint32_t* __restrict__ total_out) {
    constexpr uint32_t R = 1u << BITS;
    constexpr uint32_t DPT = R / 256;
    __shared__ uint32_t s_w[4];
    const uint32_t p = blockIdx.x, t = threadIdx.x;
    uint32_t c[DPT], local = 0;
#pragma unroll
    for (uint32_t i = 0; i < DPT; ++i) {
        c[i] = 0;
#pragma unroll
        for (uint32_t k = 0; k < kCountCopies; ++k) {
            uint32_t* q = counts + k * (kMaxPasses * kMaxRadix) + p * R + t * DPT + i;
            c[i] += *q;
            *q = 0;
        }
        local += c[i];
    }
    const uint32_t incl = wave_incl_scan(local);
    const uint32_t w = t / kWave;
    if (__lane_id() == kWave - 1) s_w[w] = incl;
    __syncthreads();
    uint32_t off = incl - local;
    for (uint32_t q = 0; q < w; ++q) off += s_w[q];
#pragma unroll
    for (uint32_t i = 0; i < DPT; ++i) {
        offsets[p * R + t * DPT + i] = off;
        off += c[i];
    }
    if (total_out && p == 0 && t == 255) *total_out = off;
}

// ------------------------------------------------------------------ one pass

// Key widths of a pass: kWide reads and writes 64-bit keys; kNarrow reads 64-bit keys and writes
// the 32-bit key (compacted key >> drop), possible once the passes before it have consumed the bits
// below `drop` and the rest fit in 32 bits; kNarrowed reads and writes such 32-bit keys.
// kWideDrop is kWide that also drops the keys whose high word is zero (the batch's culled items,
// which the digit counts leave out): they take no part in the ranking, get no LDS slot and are
// never written.  Its values are the keys' indices (the batch's items), generated rather than read.
enum : uint32_t { kWide = 0, kNarrow = 1, kNarrowed = 2, kWideDrop = 3 };

struct OnesweepArgs {
    const void* keys_in;
    const uint32_t* vals_in;
    void* keys_out;
    uint32_t* vals_out;
    const uint32_t* count_ptr;
    uint32_t n_imm;
    uint32_t shift;
    uint32_t lo_bits;          // bits of the key's low word (32; 31 for positive depths; 27 in the depth window)
    uint32_t lo_base;          // subtracted from the low word before compaction
    uint32_t drop;             // kNarrow / kNarrowed: low bits of the compacted key left out of the 32-bit key
    uint32_t digit_mask;       // this pass's digit bits below key_bits (bits at and above it do not order)
    const uint32_t* offsets;   // 2^BITS global exclusive offsets of this pass's digits
    uint32_t* desc;            // this pass's look-back region (parts x 2^BITS), zero on entry
    uint32_t* desc_next;       // the other region, cleared here for the next pass
    uint32_t* ticket;
};

// LDS of one partition of BLOCK x ITEMS keys (of type KT) with 2^BITS digits.
template <uint32_t BITS, uint32_t BLOCK, uint32_t ITEMS, typename KT>
struct OnesweepLds {
    static constexpr uint32_t R = 1u << BITS;
    static constexpr uint32_t kWaves = BLOCK / kWave;
    static constexpr uint32_t kTileKeys = BLOCK * ITEMS;
    uint16_t whist[kWaves][R];  // per wave digit counters (then wave offsets), zero on entry
    uint32_t bstart[R];
    uint32_t adj[R];
    uint32_t wtot[8];
    uint32_t nkeep;             // kWideDrop: the partition's kept keys
    uint32_t part;
    KT keys[kTileKeys];         // the partition reordered by digit (then its values)
};

// Decoupled look-back of one digit: kLookback predecessor descriptors are loaded at once (before
// partition 0 a virtual inclusive prefix of 0), consumed in order up to the first inclusive
// prefix; an unpublished one is re-read after a short sleep.  Returns the exclusive prefix.
template <uint32_t R>
__device__ __forceinline__ uint32_t onesweep_lookback(const uint32_t* desc, uint32_t part, uint32_t d) {
    constexpr int32_t kLookback = 8;
    uint32_t excl = 0;
    int32_t jj = (int32_t)part - 1;
    while (true) {
        uint32_t dsc[kLookback];
#pragma unroll
        for (int32_t u = 0; u < kLookback; ++u)
            dsc[u] = jj - u >= 0 ? load_relaxed(&desc[(uint32_t)(jj - u) * R + d]) : kFlagP;
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
        if (found) return excl;
        if (blocked) __builtin_amdgcn_s_sleep(1);
    }
}

// The digits are owned by DTH threads, each taking DPT consecutive digits for the per-digit
// bookkeeping and the look-back.
template <uint32_t BITS, uint32_t BLOCK, uint32_t ITEMS, uint32_t MODE>
__global__ __launch_bounds__(BLOCK) void k_onesweep(const OnesweepArgs a) {
    constexpr bool DROP = MODE == kWideDrop;
    constexpr bool WIDE = MODE == kWide || DROP;
    using KT = typename std::conditional<WIDE, uint64_t, uint32_t>::type;  // key type after the load
    using Lds = OnesweepLds<BITS, BLOCK, ITEMS, KT>;
    constexpr uint32_t R = Lds::R;
    constexpr uint32_t kWaves = Lds::kWaves;
    constexpr uint32_t kTileKeys = Lds::kTileKeys;
    constexpr uint32_t DTH = R < BLOCK ? R : BLOCK;  // threads owning digits
    constexpr uint32_t DPT = R / DTH;                // digits per owning thread
    static_assert(R >= 256 && (R % BLOCK == 0 || BLOCK % R == 0), "digits and threads must tile");
    static_assert(DTH / kWave <= 8, "at most 8 digit-owning waves");
    __shared__ Lds L;

    const uint32_t n = a.count_ptr ? min(*a.count_ptr, a.n_imm) : a.n_imm;
    const uint32_t nparts = (n + kTileKeys - 1) / kTileKeys;
    const uint32_t tid = threadIdx.x, wave = tid / kWave, lane = __lane_id();
    if (tid == 0) L.part = atomicAdd(a.ticket, 1u);
    for (uint32_t x = tid; x < kWaves * R / 2; x += BLOCK) reinterpret_cast<uint32_t*>(&L.whist[0][0])[x] = 0;
    __syncthreads();
    const uint32_t part = __builtin_amdgcn_readfirstlane(L.part);  // block-uniform: keep it scalar
    if (part >= nparts) return;
    if (a.desc_next)
        for (uint32_t d = tid; d < R; d += BLOCK) a.desc_next[part * R + d] = 0;

    // Load ITEMS keys and values per thread, wave-striped: element e = wave*(ITEMS*64) + j*64 +
    // lane; past the end keys are all ones (kWideDrop: zero, i.e. dropped).  Every partition but the
    // last is full: its loads (and, without dropped keys, its stores) need no bounds checks.
    const uint32_t base = part * kTileKeys;
    const uint32_t wbase = base + wave * (ITEMS * kWave) + lane;
    const bool full = base + kTileKeys <= n;
    const auto load_key = [&](uint32_t idx) -> KT {
        if constexpr (WIDE) return static_cast<const uint64_t*>(a.keys_in)[idx];
        if constexpr (MODE == kNarrow)
            return (uint32_t)(compact_key(static_cast<const uint64_t*>(a.keys_in)[idx], a.lo_bits, a.lo_base) >> a.drop);
        return static_cast<const uint32_t*>(a.keys_in)[idx];
    };
    KT k[ITEMS];
    uint32_t v[ITEMS];
    if (full) {
#pragma unroll
        for (uint32_t j = 0; j < ITEMS; ++j) k[j] = load_key(wbase + j * kWave);
#pragma unroll
        for (uint32_t j = 0; j < ITEMS; ++j) v[j] = DROP ? wbase + j * kWave : a.vals_in[wbase + j * kWave];
    } else {
#pragma unroll
        for (uint32_t j = 0; j < ITEMS; ++j) {
            const uint32_t idx = wbase + j * kWave;
            k[j] = idx < n ? load_key(idx) : (DROP ? KT(0) : ~KT(0));
        }
#pragma unroll
        for (uint32_t j = 0; j < ITEMS; ++j) {
            const uint32_t idx = wbase + j * kWave;
            v[j] = idx < n ? (DROP ? idx : a.vals_in[idx]) : 0u;
        }
    }

    // Rank within the wave: peers with the same digit found with BITS ballots; the lowest lane of
    // each group advances the wave's running digit counter.  Digits are recomputed from the keys
    // wherever they are needed (a few instructions) rather than kept in registers or LDS.
    const auto digit = [&](KT key) -> uint32_t {
        if constexpr (WIDE) return digit_of<BITS>(compact_key(key, a.lo_bits, a.lo_base), a.shift) & a.digit_mask;
        else return (uint32_t)(key >> (a.shift - a.drop)) & a.digit_mask;
    };
    const uint64_t lt = lanemask_lt();
    const auto dropped = [&](KT key) { return DROP && (uint32_t)((uint64_t)key >> 32) == 0u; };
    uint32_t rank[ITEMS];
#pragma unroll
    for (uint32_t j = 0; j < ITEMS; ++j) {
        const uint32_t d = digit(k[j]);
        uint64_t peers = ~0ull;
#pragma unroll
        for (uint32_t b = 0; b < BITS; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        if constexpr (DROP) {
            // a dropped key is nobody's peer and gets no rank (kNoSlot)
            const bool cull = dropped(k[j]);
            peers &= ~ballot(cull);
            const uint32_t below = (uint32_t)__popcll(peers & lt);
            const uint32_t pre = L.whist[wave][d];
            if (!cull && below == 0) L.whist[wave][d] = (uint16_t)(pre + (uint32_t)__popcll(peers));
            rank[j] = cull ? kNoSlot : pre + below;
        } else {
            const uint32_t below = (uint32_t)__popcll(peers & lt);
            const uint32_t pre = L.whist[wave][d];
            if (below == 0) L.whist[wave][d] = (uint16_t)(pre + (uint32_t)__popcll(peers));
            rank[j] = pre + below;
        }
    }
    __syncthreads();

    // Per digit: wave exclusive offsets and the partition's digit counts (published); then the
    // digits' block-local starts by an exclusive scan over the digit-owning threads.
    uint32_t tot[DPT];
    uint32_t tstart = 0;
    if (tid < DTH) {
        uint32_t local = 0;
#pragma unroll
        for (uint32_t i = 0; i < DPT; ++i) {
            const uint32_t d = tid * DPT + i;
            uint32_t acc = 0;
#pragma unroll
            for (uint32_t w = 0; w < kWaves; ++w) {
                const uint32_t c = L.whist[w][d];
                L.whist[w][d] = (uint16_t)acc;
                acc += c;
            }
            tot[i] = acc;
            local += acc;
            store_relaxed(&a.desc[part * R + d], (part == 0 ? kFlagP : kFlagA) | acc);
        }
        const uint32_t incl = wave_incl_scan(local);
        if (lane == kWave - 1) L.wtot[wave] = incl;
        tstart = incl - local;
    }
    __syncthreads();
    if (tid < DTH) {
        for (uint32_t w = 0; w < wave; ++w) tstart += L.wtot[w];
#pragma unroll
        for (uint32_t i = 0; i < DPT; ++i) {
            L.bstart[tid * DPT + i] = tstart;
            tstart += tot[i];
        }
        if (DROP && tid == DTH - 1) L.nkeep = tstart;  // the kept keys of all digits
    }
    __syncthreads();

    // Reorder the partition by digit in LDS (local position = digit start + wave offset + rank).
#pragma unroll
    for (uint32_t j = 0; j < ITEMS; ++j) {
        const uint32_t d = digit(k[j]);
        if (DROP && rank[j] == kNoSlot) continue;
        const uint32_t lp = L.bstart[d] + L.whist[wave][d] + rank[j];
        rank[j] = lp;
        L.keys[lp] = k[j];
    }

    // Look-back per digit, then the global position adjustment of each digit run.
    if (tid < DTH) {
#pragma unroll
        for (uint32_t i = 0; i < DPT; ++i) {
            const uint32_t d = tid * DPT + i;
            uint32_t excl = 0;
            if (part > 0) {
                excl = onesweep_lookback<R>(a.desc, part, d);
                store_relaxed(&a.desc[part * R + d], kFlagP | (excl + tot[i]));
            }
            L.adj[d] = a.offsets[d] + excl - L.bstart[d];
        }
    }
    __syncthreads();

    // Keys out, each to its digit's global start + its place in the partition; the global position
    // of each output slot is kept for the values (the keys' LDS is reused for them).
    // (the past-the-end keys are in the last digit, after the others; dropped keys have no slot)
    const uint32_t nvalid = DROP ? L.nkeep : min(kTileKeys, n - base);
    const bool all_valid = full && !DROP;
    uint32_t gpos[ITEMS];
#pragma unroll
    for (uint32_t t = 0; t < ITEMS; ++t) {
        const uint32_t i = t * BLOCK + tid;
        const KT key = L.keys[i];
        gpos[t] = L.adj[digit(key)] + i;
        if (all_valid || i < nvalid) static_cast<KT*>(a.keys_out)[gpos[t]] = key;
    }
    __syncthreads();
    uint32_t* s_vals = reinterpret_cast<uint32_t*>(L.keys);
#pragma unroll
    for (uint32_t j = 0; j < ITEMS; ++j)
        if (!DROP || rank[j] != kNoSlot) s_vals[rank[j]] = v[j];
    __syncthreads();
#pragma unroll
    for (uint32_t t = 0; t < ITEMS; ++t) {
        const uint32_t i = t * BLOCK + tid;
        if (all_valid || i < nvalid) a.vals_out[gpos[t]] = s_vals[i];
    }
}

// Shapes the tuner chooses from: {digit bits, threads, keys per thread}.
// kernel[mode] for the key widths kWide, kNarrow, kNarrowed and kWideDrop.
struct SortShape {
    uint32_t bits, block, items;
    void (*kernel[4])(const OnesweepArgs);
};
#define TPD_SHAPE(B, T, I)                                                                              \
    {                                                                                                   \
        B, T, I, {                                                                                      \
            k_onesweep<B, T, I, kWide>, k_onesweep<B, T, I, kNarrow>, k_onesweep<B, T, I, kNarrowed>, \
                k_onesweep<B, T, I, kWideDrop>                                                          \
        }                                                                                               \
    }
static const SortShape kShapes[] = {
    TPD_SHAPE(8, 256, 16), TPD_SHAPE(8, 256, 8),  TPD_SHAPE(8, 256, 12), TPD_SHAPE(8, 512, 8),
    TPD_SHAPE(8, 512, 12), TPD_SHAPE(8, 256, 20), TPD_SHAPE(9, 512, 12), TPD_SHAPE(9, 512, 8),
    TPD_SHAPE(9, 512, 16), TPD_SHAPE(9, 1024, 16), TPD_SHAPE(9, 1024, 12), TPD_SHAPE(8, 1024, 12),
};
#undef TPD_SHAPE
constexpr uint32_t kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

uint32_t sort_num_shapes() { return kNumShapes; }
uint32_t sort_shape_keys(uint32_t shape) { return kShapes[shape % kNumShapes].block * kShapes[shape % kNumShapes].items; }
uint32_t sort_shape_bits(uint32_t shape) { return kShapes[shape % kNumShapes].bits; }
uint32_t sort_shape_passes(uint32_t shape, uint32_t key_bits) {
    const uint32_t b = sort_shape_bits(shape);
    return (key_bits + b - 1) / b;
}

template <uint32_t BITS, typename KT>
static void launch_hist(uint32_t hb, uint32_t passes, hipStream_t s, const KT* keys, const uint32_t* count_ptr,
                        uint32_t n_max, uint32_t lo_bits, uint32_t lo_base, uint32_t key_bits, uint32_t tile,
                        SortWorkspace& ws) {
    const KT key_mask = key_bits >= 8 * sizeof(KT) ? ~KT(0) : (KT(1) << key_bits) - KT(1);
    hipLaunchKernelGGL((k_sort_hist<BITS, KT>), dim3(hb), dim3(256), 0, s, keys, count_ptr, n_max, passes, lo_bits,
                       lo_base, key_mask, tile, ws.partials, ws.desc, ws.tickets);
    hipLaunchKernelGGL(k_sort_offsets<BITS>, dim3(passes), dim3(256), 0, s, ws.partials, ws.offsets, nullptr);
}

// The digit bits of the pass at `shift` that lie below key_bits.
static uint32_t pass_digit_mask(uint32_t bits, uint32_t key_bits, uint32_t shift) {
    const uint32_t w = key_bits - shift < bits ? key_bits - shift : bits;
    return (1u << w) - 1u;
}

static uint32_t hist_blocks(uint32_t n_max) {
    uint32_t hb = (n_max + 256 * kHistUnroll * 4 - 1) / (256 * kHistUnroll * 4);
    return hb < 1 ? 1 : (hb > kHistBlocks ? kHistBlocks : hb);
}

bool launch_sort_pairs(SortWorkspace& ws, uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                       const uint32_t* count_ptr, uint32_t n_max, uint32_t key_bits, hipStream_t s,
                       hipEvent_t after_hist, uint32_t lo_bits, uint32_t* narrow_drop, uint32_t lo_base, bool counted,
                       uint32_t* kept) {
    const SortShape& sh = kShapes[ws.shape % kNumShapes];
    const uint32_t R = 1u << sh.bits;
    const uint32_t tile = sh.block * sh.items;
    const uint32_t passes = (key_bits + sh.bits - 1) / sh.bits;
    const uint32_t parts = (n_max + tile - 1) / tile;
    if (parts == 0 || passes == 0 || passes > kMaxPasses) return false;

    const uint32_t hb = hist_blocks(n_max);
    if (!counted) kept = nullptr;  // the histogram pass counts every key
    if (counted) {
        if (sh.bits == 8) hipLaunchKernelGGL(k_sort_offsets<8>, dim3(passes), dim3(256), 0, s, ws.partials, ws.offsets, kept);
        else hipLaunchKernelGGL(k_sort_offsets<9>, dim3(passes), dim3(256), 0, s, ws.partials, ws.offsets, kept);
    } else if (sh.bits == 8) {
        launch_hist<8>(hb, passes, s, keys, count_ptr, n_max, lo_bits, lo_base, key_bits, tile, ws);
    } else {
        launch_hist<9>(hb, passes, s, keys, count_ptr, n_max, lo_bits, lo_base, key_bits, tile, ws);
    }
    if (after_hist) hipEventRecord(after_hist, s);
    // With narrow_drop, the first pass whose digit starts at or above key_bits - 32 writes 32-bit
    // keys (the bits from its digit up) and the later passes move 8 instead of 12 bytes per pair.
    uint32_t narrow_pass = passes;
    if (narrow_drop) {
        for (uint32_t p = 0; p < passes; ++p)
            if (key_bits <= p * sh.bits + 32) {
                narrow_pass = p;
                break;
            }
        *narrow_drop = narrow_pass < passes ? narrow_pass * sh.bits : kKeysWide;
    }
    if (narrow_pass == 0) kept = nullptr;  // the dropping pass reads and writes 64-bit keys
    const uint64_t* kin = keys;
    const uint32_t* vin = vals;
    uint64_t* kout = keys_alt;
    uint32_t* vout = vals_alt;
    for (uint32_t p = 0; p < passes; ++p) {
        OnesweepArgs o;
        o.keys_in = kin; o.vals_in = vin; o.keys_out = kout; o.vals_out = vout;
        o.count_ptr = (kept && p > 0) ? kept : count_ptr;
        o.n_imm = n_max;
        o.shift = p * sh.bits;
        o.lo_bits = lo_bits;
        o.lo_base = lo_base;
        o.drop = narrow_pass < passes ? narrow_pass * sh.bits : 0u;
        o.digit_mask = pass_digit_mask(sh.bits, key_bits, o.shift);
        const uint32_t mode = p < narrow_pass ? ((kept && p == 0) ? kWideDrop : kWide) : (p == narrow_pass ? kNarrow : kNarrowed);
        o.offsets = ws.offsets + p * R;
        o.desc = ws.desc + (size_t)(p & 1u) * ws.parts_cap * R;
        o.desc_next = ws.desc + (size_t)((p + 1) & 1u) * ws.parts_cap * R;
        o.ticket = ws.tickets + p;
        hipLaunchKernelGGL(sh.kernel[mode], dim3(parts), dim3(sh.block), 0, s, o);
        const uint64_t* tk = kout; kout = const_cast<uint64_t*>(kin); kin = tk;
        const uint32_t* tv = vout; vout = const_cast<uint32_t*>(vin); vin = tv;
    }
    return (passes & 1u) != 0;
}

bool launch_sort_pairs32(SortWorkspace& ws, uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                         const uint32_t* count_ptr, uint32_t n_max, uint32_t key_bits, hipStream_t s,
                         hipEvent_t after_hist, bool counted) {
    const SortShape& sh = kShapes[ws.shape % kNumShapes];
    const uint32_t R = 1u << sh.bits;
    const uint32_t tile = sh.block * sh.items;
    const uint32_t passes = (key_bits + sh.bits - 1) / sh.bits;
    const uint32_t parts = (n_max + tile - 1) / tile;
    if (parts == 0 || passes == 0 || passes > kMaxPasses) return false;
    const uint32_t hb = hist_blocks(n_max);
    if (counted) {
        uint32_t* counts = ws.partials + kPairCountRow;
        if (sh.bits == 8) hipLaunchKernelGGL(k_sort_offsets<8>, dim3(passes), dim3(256), 0, s, counts, ws.offsets, nullptr);
        else hipLaunchKernelGGL(k_sort_offsets<9>, dim3(passes), dim3(256), 0, s, counts, ws.offsets, nullptr);
    } else if (sh.bits == 8) {
        launch_hist<8>(hb, passes, s, (const uint32_t*)keys, count_ptr, n_max, 32u, 0u, key_bits, tile, ws);
    } else {
        launch_hist<9>(hb, passes, s, (const uint32_t*)keys, count_ptr, n_max, 32u, 0u, key_bits, tile, ws);
    }
    if (after_hist) hipEventRecord(after_hist, s);
    const uint32_t* kin = keys;
    const uint32_t* vin = vals;
    uint32_t* kout = keys_alt;
    uint32_t* vout = vals_alt;
    for (uint32_t p = 0; p < passes; ++p) {
        OnesweepArgs o;
        o.keys_in = kin; o.vals_in = vin; o.keys_out = kout; o.vals_out = vout;
        o.count_ptr = count_ptr; o.n_imm = n_max;
        o.shift = p * sh.bits;
        o.lo_bits = 32u;
        o.lo_base = 0u;
        o.drop = 0u;
        o.digit_mask = pass_digit_mask(sh.bits, key_bits, o.shift);
        o.offsets = ws.offsets + p * R;
        o.desc = ws.desc + (size_t)(p & 1u) * ws.parts_cap * R;
        o.desc_next = ws.desc + (size_t)((p + 1) & 1u) * ws.parts_cap * R;
        o.ticket = ws.tickets + p;
        hipLaunchKernelGGL(sh.kernel[kNarrowed], dim3(parts), dim3(sh.block), 0, s, o);
        const uint32_t* tk = kout; kout = const_cast<uint32_t*>(kin); kin = tk;
        const uint32_t* tv = vout; vout = const_cast<uint32_t*>(vin); vin = tv;
    }
    return (passes & 1u) != 0;
}

}  // namespace tpd

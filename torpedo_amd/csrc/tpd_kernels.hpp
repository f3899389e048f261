// Kernel argument blocks and launchers shared by the kernels and the C-ABI layer.
#pragma once

#include "tpd_device.hpp"

namespace tpd {

// Indices into the per-context device counter block.
enum : uint32_t {
    kCtrScanTicket = 0,   // preprocess partition ticket
    kCtrPairs = 1,        // pair count P written by the last preprocess partition
    kCtrOverflow = 2,     // 1 when P exceeded the key/value capacity
    kCtrDepthOut = 3,     // non-zero when an emitted depth lies outside the sort's depth window
    kCtrScanTicket2 = 4,  // partition ticket of a second scan in the frame (the batch's item offsets)
    kCtrSortPairs = 5,    // the batch's pair count for its pair sort and ranges: P, or 0 when P exceeded
                          // the capacity (that frame is rendered again)
    kCtrItemsKept = 32,   // the batch's visible items: those the item sort keeps (its later passes and
                          // the emission run on them).  128 bytes past the tickets: every emission
                          // block reads it right after its ticket atomic, and on the tickets' L2
                          // line that read would queue behind the other blocks' atomics
    kCtrCount = 64,
};

// The sort's depth window.  Culling keeps view depths inside [near, far], and depths in
// [2^-8, 2^8) have float bit patterns in [kDepthBase, kDepthBase + 2^27): for such frames the sort
// may key on (tile << 27) | (depth bits - kDepthBase), in the same order as (tile << 32) | depth
// bits and with 4 bits less than the positive-float form (tile << 31) | depth bits.  A frame with
// any depth outside the window (flagged by its preprocess in kCtrDepthOut) sorts that 31-bit form.
constexpr uint32_t kDepthBase = 0x3B800000u;  // floatbits(2^-8)
constexpr uint32_t kDepthWindowBits = 27;
constexpr uint32_t kDepthWideBits = 31;
__device__ __forceinline__ bool depth_outside_window(float vz) {
    return __float_as_uint(vz) - kDepthBase >= (1u << kDepthWindowBits);
}

struct ForwardArgs {
    const float* gauss;          // N x 60 floats (GaussianPoint records)
    const uint32_t* ent_idx;     // N entity indices, or null (entity 0 for all)
    const EntityMats* ent;       // per-entity matrices
    CameraParams cam;
    uint32_t n, sh_degree, gx, gy;
    uint32_t cap;                // pair capacity of keys/vals
    Splat* splats;               // per Gaussian (zeros when culled)
    uint32_t* tiles;             // tile count per Gaussian
    uint32_t* offsets;           // exclusive scan of tiles
    uint64_t* scan_desc;         // look-back descriptors, one per partition of the preprocess (the
                                 // buffer holds ceil(N/256)), zero on entry
    uint32_t* counters;          // kCtr* block
    uint64_t* keys;
    uint32_t* vals;
    uint32_t val_base;           // value written for Gaussian i is val_base + i
    uint2* ranges;               // zeroed here for the ranges kernel
    uint32_t ntiles;
};

struct BlendArgs {
    const uint2* ranges;
    const uint32_t* tile_order;  // block -> tile (heaviest first, launch_tile_order), or null
    const uint32_t* vals;
    const Splat* splats;
    uint32_t width, height, gx, ntiles;
    float* image;            // H x W x 3 or null
    uint8_t* rgba8;          // H x W x 4 or null
    uint64_t* eval_count;    // when non-null, incremented by the number of evaluated pixel-splat pairs
    float* final_T;          // H x W final transmittance, for the backward (or null)
    int32_t* final_last;     // H x W sorted index of the last blended splat, -1 for none (or null)
};

struct BackwardArgs {
    const uint2* ranges;
    const uint32_t* tile_order;  // the forward frame's block -> tile order, or null
    const uint32_t* vals;
    const Splat* splats;
    uint32_t width, height, gx, gy, ntiles;
    const float* dL_dimg;    // H x W x 3
    const uint32_t* offsets; // per-Gaussian exclusive scan of tile counts (pre-sort pair position)
    float* partials;         // P x 9 per (tile, splat) pair at its pre-sort position: dL/d(image x, y,
                             // conic a, b, c, r, g, b, opacity); every pair is written
    const float* final_T;    // the forward frame's per-pixel final transmittance
    const int32_t* final_last;  // and last blended splat
};

struct PreBackwardArgs {
    const float* gauss;
    const uint32_t* ent_idx;
    const EntityMats* ent;
    CameraParams cam;
    uint32_t n, sh_degree;
    const uint32_t* tiles;
    const uint32_t* offsets;
    const float* partials;       // P x 9, summed per Gaussian over its tiles in tile order
    float* g2d;                  // N x 9 per-Gaussian sums (written, for inspection)
    float* grad_means;           // N x 3
    float* grad_log_scales;      // N x 3
    float* grad_quats;           // N x 4
    float* grad_opacity_logits;  // N
    float* grad_sh;              // N x 48
    uint32_t accumulate;         // add to the outputs instead of overwriting them
};

void launch_blend_bwd(const BackwardArgs& a, hipStream_t s);
void launch_preprocess_bwd(const PreBackwardArgs& a, hipStream_t s);

void launch_entity_mats(const float* models, uint32_t n_ent, const CameraParams& cam, EntityMats* out, hipStream_t s);
void launch_preprocess_dup(const ForwardArgs& a, hipStream_t s);
// Sorted keys are 64-bit (tile << 32) | depth, or, after a sort that narrowed them, 32-bit
// (compacted key (tile << lo_bits) | (depth - lo_base)) >> drop; key_drop = kKeysWide marks the
// first form.
constexpr uint32_t kKeysWide = 0xFFFFFFFFu;
// Tile ranges of the sorted keys; P = min(counters[count_slot], cap).  Also re-zeroes the scan
// descriptors and the scan ticket for the next frame, and sets kCtrOverflow when a frame sorted
// on the depth-window key (lo_bits = kDepthWindowBits) has kCtrDepthOut set.
void launch_ranges(const void* keys, uint32_t* counters, uint32_t P_hint, uint32_t ntiles, uint2* ranges,
                   uint64_t* scan_desc, uint32_t n_scan_desc, uint32_t* ticket, hipStream_t s,
                   uint32_t cap = 0xFFFFFFFFu, uint32_t key_drop = kKeysWide, uint32_t lo_bits = kDepthWideBits,
                   uint32_t count_slot = kCtrPairs);
// The 64-bit keys (tile << 32) | floatbits(view depth) of n sorted 32-bit keys, the tile from the
// key and the depth from the splat of each value (for readers of the sorted pairs).
void launch_full_keys(const uint32_t* keys32, uint32_t drop, uint32_t lo_bits, const uint32_t* vals,
                      const Splat* splats, uint32_t n, uint64_t* out, hipStream_t s);
void launch_blend(const BlendArgs& a, hipStream_t s);
// The tiles in launch order for the blend: classes of pair count (bit length) from the heaviest
// down, empty tiles last, so that the longest tiles start first and the short ones fill the tail.
// counters (optional): nothing is done after a pair overflow.  `order` holds
// tile_order_words(ntiles) words: the order, then the per-chunk class counts.
constexpr uint32_t kOrderChunk = 1024;  // tiles per block
constexpr uint32_t kTileClasses = 33;   // pair-count bit lengths 0 .. 32
inline size_t tile_order_words(size_t ntiles) {
    return ntiles + kTileClasses * ((ntiles + kOrderChunk - 1) / kOrderChunk);
}
void launch_tile_order(const uint2* ranges, uint32_t ntiles, const uint32_t* counters, uint32_t* order, hipStream_t s);

// ------------------------------------------------------------------ batched rendering
//
// Many views of many scenes in one launch per stage.  Layout:
//   items   view-major: Gaussian g of view v is item views[v].item_base + g; per-item splats,
//           tile counts and pair offsets are indexed by item, pair values are items;
//   tiles   global tile t of view v is views[v].tile_base + (local tile); keys carry global
//           tiles, so one sort orders every view's pairs and the views' ranges are disjoint;
//   pixels  view v's image starts at pixel views[v].pix_base of the concatenated images;
//   blocks  preprocess block b covers Gaussians [256 (b - first_block), +256) of scene
//           block_scene[b] and projects them into every view of that scene, reading each
//           Gaussian record once;
//   order   items are sorted by depth key (stable: equal depths keep item order, which within a
//           view is Gaussian order), their pairs are emitted in that order, and the pairs are
//           then sorted by global tile alone (stable): within each tile they stay in (depth,
//           Gaussian) order, the order of the reference's (tile << 32 | depth) sort.

struct BatchSceneDev {
    const float* gauss;          // N x 60 records
    const uint32_t* ent_idx;     // N entity indices, or null
    const float* models;         // n_ent x 16 model matrices, or null (identity)
    uint32_t n, sh_degree, n_ent;
    uint32_t first_block;        // first preprocess block of the scene
    uint32_t view_begin, nviews; // the scene's views: scene_views[view_begin, view_begin + nviews)
    uint32_t pad;
};

struct BatchViewDev {
    CameraParams cam;
    uint32_t scene, gx, gy, tile_base;
    uint32_t item_base;          // item of Gaussian g = item_base + g
    uint32_t mat_base;           // entity matrices of this view at mats[mat_base + e]
    uint64_t pix_base;           // first pixel of this view in the concatenated images
};

struct BatchArgs {
    const BatchSceneDev* scenes;
    const BatchViewDev* views;
    const uint32_t* scene_views;  // view indices grouped by scene
    const uint32_t* block_scene;  // scene of each preprocess block
    uint32_t nblocks, nviews, nitems;
    const EntityMats* mats;
    uint64_t* scan_desc;          // 2 x ceil(nitems / kEmitBlockItems) look-back descriptors, zero on entry: the
                                  // emission's, then the item offsets
    uint32_t* counters;           // kCtr* block
    uint32_t cap;                 // pair capacity
    Splat* splats;                // per item (written for visible items only)
    uint32_t* tiles;              // per item tile counts, or null (kept for the backward)
    uint32_t* offsets;            // per item exclusive scan of the tile counts in item order, or null:
                                  // where the backward keeps the item's per-pair partial sums
    uint64_t* rects;              // per item emission record (pack_emit_rect), written only where the key
                                  // carries kRectLookup
    uint64_t* item_keys;          // per item: low word the depth key ((depth bits - lo_base) in lo_bits
                                  // bits, all ones within lo_bits when culled), high word the emission
                                  // record as inline_rect encodes it
    uint32_t* item_vals;          // per item: the item itself (null when the item sort generates them)
    uint32_t lo_bits, lo_base;
    // the item sort's digit histograms, counted by the item kernel (the sort then skips its own):
    // counts as k_sort_hist leaves them (copy blockIdx mod kCountCopies), digit width and passes of
    // the item sort's shape; the first pass's look-back descriptors (desc0_words) and the tickets
    // are cleared here too
    uint32_t* hist_counts;
    uint32_t* hist_desc0;
    uint32_t* hist_tickets;
    uint32_t hist_bits, hist_passes, hist_desc0_words;
    // the pair sort's digit histograms, counted by the item kernel from each visible item's tile
    // rectangle (needs hist_counts): counts from row kPairCountRow of the sort's digit counts, digit
    // width and passes of the pair sort's shape; the emission clears the pair sort's first-pass
    // descriptors (pair_desc0_words, for the pair capacity) and the tickets
    uint32_t* pair_counts;
    uint32_t pair_bits, pair_passes, pair_desc0_words;
    const uint32_t* kept_items;     // emission input: the number of sorted items (the visible ones) ...
    const uint64_t* sorted_keys64;  // ... the item keys in depth order ...
    const uint32_t* sorted_items;   // ... and their items
    uint32_t* keys;               // pairs: global tile
    uint32_t* vals;               //        item
    uint2* ranges;                // all views' tiles, cleared by the item kernel
    uint32_t ntiles;
};

struct BatchBlendArgs {
    const BatchViewDev* views;
    const uint32_t* tile_view;    // view of each global tile
    const uint32_t* tile_order;   // block -> tile (heaviest first), or null
    const uint2* ranges;
    const uint32_t* vals;
    const Splat* splats;
    uint32_t ntiles;
    const uint32_t* counters;     // the frame's kCtr* block: nothing is done after an overflow
    float* image;                 // concatenated H x W x 3 images, or null
    uint8_t* rgba8;               // concatenated H x W x 4 images, or null
    const float* dL_dimg;         // backward: concatenated H x W x 3 upstream gradients
    const uint32_t* offsets;      // backward: per item pre-sort pair offsets
    float* partials;              // backward: P x 9 per pair
    float* final_T;               // per pixel final transmittance: written by a forward kept for the
    int32_t* final_last;          // backward, read by it (with the last blended splat); or null
    uint64_t* eval_count;         // forward: evaluated pixel-splat pairs are added here, or null
};

struct BatchGradsDev {            // one scene's parameter-gradient buffers
    float* means;                 // N x 3
    float* log_scales;            // N x 3
    float* quats;                 // N x 4
    float* opacity_logits;        // N
    float* sh;                    // N x 48
};

struct BatchPreBwdArgs {
    const BatchSceneDev* scenes;
    const BatchViewDev* views;
    const uint32_t* scene_views;
    const uint32_t* block_scene;
    uint32_t nblocks;
    const EntityMats* mats;
    const uint32_t* tiles;
    const uint32_t* offsets;
    const float* partials;
    const BatchGradsDev* grads;   // per scene
    const uint32_t* counters;     // nothing is written after an overflow
    uint32_t accumulate;
};

void launch_batch_mats(const BatchSceneDev* scenes, const BatchViewDev* views, uint32_t nviews, uint32_t max_ent,
                       EntityMats* mats, uint32_t* counters, hipStream_t s);
void launch_batch_items(const BatchArgs& a, hipStream_t s);
void launch_batch_emit(const BatchArgs& a, hipStream_t s);
void launch_batch_item_offsets(const BatchArgs& a, hipStream_t s);
constexpr uint32_t kEmitBlockItems = 4096;  // sorted items per emission partition (k_batch_emit)
void launch_batch_blend(const BatchBlendArgs& a, hipStream_t s);
void launch_batch_blend_bwd(const BatchBlendArgs& a, hipStream_t s);
void launch_batch_preprocess_bwd(const BatchPreBwdArgs& a, hipStream_t s);

// ------------------------------------------------------------------ radix sort

constexpr uint32_t kMaxRadixBits = 9;                      // digit widths 8 and 9 bits
constexpr uint32_t kMaxRadix = 1u << kMaxRadixBits;
constexpr uint32_t kSortMinTile = 2048;                    // smallest partition of any block shape
constexpr uint32_t kMaxPasses = 8;
constexpr uint32_t kCountCopies = 8;  // copies of the digit counts, block b adds into copy b mod 8
constexpr uint32_t kHistBlocks = 1024;                      // blocks of the digit-histogram kernel
constexpr uint32_t kMaxSortKeys = (1u << 30) - 1;          // 30-bit look-back values

struct SortWorkspace {
    uint32_t* desc = nullptr;      // 2 regions x parts_cap x kMaxRadix look-back descriptors
    uint32_t parts_cap = 0;
    uint32_t* partials = nullptr;  // kCountCopies x kMaxPasses x kMaxRadix digit counts (zero between sorts)
    uint32_t* offsets = nullptr;   // kMaxPasses x kMaxRadix global exclusive digit offsets
    uint32_t* tickets = nullptr;   // kMaxPasses partition tickets
    uint32_t shape = 0;            // onesweep block shape (index into the tuner's table)
};

// Partitions for n keys at the smallest partition size (descriptor capacity for every shape).
__host__ __device__ inline uint32_t sort_parts(uint32_t n) { return (n + kSortMinTile - 1) / kSortMinTile; }
uint32_t sort_num_shapes();
uint32_t sort_shape_keys(uint32_t shape);
uint32_t sort_shape_bits(uint32_t shape);                      // digit width of a shape
uint32_t sort_shape_passes(uint32_t shape, uint32_t key_bits);  // passes for key_bits bits

// Stable LSD sort of (keys, vals) by the low key_bits bits of the keys.  n comes from *count_ptr
// when count_ptr is non-null (device-resident count capped at n_max, grids sized by n_max), else
// from n_max.
// Returns true when the sorted data ended in the alternate buffers.
// after_hist (optional) is recorded between the histogram kernels and the first onesweep pass.
// lo_bits < 32 sorts the compacted key (high word << lo_bits) | (low word - lo_base), valid when
// every key's low word lies in [lo_base, lo_base + 2^lo_bits); key_bits then counts the compacted
// key's bits.  A low word outside that range cannot change the high word's bits in the compacted
// key (the difference is masked to lo_bits bits), only the order among keys of equal high word.
// narrow_drop (optional) lets the later passes move 32-bit keys: *narrow_drop receives the number
// of low compacted-key bits the sorted 32-bit keys leave out, or kKeysWide if the keys stayed
// 64-bit (see launch_ranges).
bool launch_sort_pairs(SortWorkspace& ws, uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                       const uint32_t* count_ptr, uint32_t n_max, uint32_t key_bits, hipStream_t s,
                       hipEvent_t after_hist = nullptr, uint32_t lo_bits = 32, uint32_t* narrow_drop = nullptr,
                       uint32_t lo_base = 0, bool counted = false, uint32_t* kept = nullptr);
// With counted, the digit counts are already in ws.partials and the first pass's descriptors and the
// tickets already cleared (by the kernel that made the keys): the histogram kernel is not launched.
// With kept as well (and 64-bit keys in the first pass), the counts leave out the keys whose high
// word is zero and the first pass drops those keys: *kept receives the number of keys kept, the
// later passes sort those, and the result holds the kept keys in order (stable), nothing after.
// The first pass then also takes each key's index as its value: vals is not read, only used as
// the later passes' buffer.
constexpr uint32_t kMaxCountedPasses = 4;  // passes whose counts the item kernel can make
// The same sort of 32-bit keys (key bits at and above key_bits zero) with 32-bit values: every pass
// moves 8 bytes per pair.  Returns true when the result is in keys_alt / vals_alt.  With counted,
// the digit counts are in ws.partials from word kPairCountRow of each copy on, and the first pass's
// descriptors and the tickets are already cleared.
bool launch_sort_pairs32(SortWorkspace& ws, uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                         const uint32_t* count_ptr, uint32_t n_max, uint32_t key_bits, hipStream_t s,
                         hipEvent_t after_hist = nullptr, bool counted = false);
constexpr uint32_t kPairCountRow = kMaxCountedPasses * kMaxRadix;  // after the item sort's counted rows
constexpr uint32_t kMaxCountedPairPasses = 3;
static_assert(kPairCountRow + kMaxCountedPairPasses * kMaxRadix <= kMaxPasses * kMaxRadix, "pair count rows");

}  // namespace tpd

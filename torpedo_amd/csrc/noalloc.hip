// The stand-alone sort for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc): tpd_sort_pairs_noalloc runs the
// same kernels as tpd_sort_pairs, with the sort workspace (digit counts and offsets, partition
// tickets, look-back descriptors) in one file-scope device array instead of context allocations.
//
// The array (8,536,064 bytes) is held by every GPU the library runs on until the process ends; every
// call on one GPU shares that GPU's copy: calls that overlap on one GPU (two streams, two threads)
// race and must not be issued.  A call leaves the workspace as it found it except for the digit
// counts, which it clears itself before sorting (with a fill kernel, so that the call can be
// captured into a graph).
#include "tpd_ctx.hpp"

#include <mutex>

namespace tpd {

constexpr uint32_t kNaMaxKeys = 1u << 22;                                      // 4,194,304 pairs
constexpr uint32_t kNaParts = (kNaMaxKeys + kSortMinTile - 1) / kSortMinTile;  // = sort_parts(kNaMaxKeys)
constexpr size_t kNaCountWords = (size_t)kCountCopies * kMaxPasses * kMaxRadix;
constexpr size_t kNaOffsetWords = (size_t)kMaxPasses * kMaxRadix;
constexpr size_t kNaTicketWords = 64;  // kMaxPasses tickets, padded to 256 bytes
constexpr size_t kNaDescWords = 2ull * kNaParts * kMaxRadix;
constexpr size_t kNaWords = kNaCountWords + kNaOffsetWords + kNaTicketWords + kNaDescWords;
static_assert(kNaTicketWords >= kMaxPasses, "ticket region too small");
static_assert(kNaCountWords % 64 == 0 && kNaOffsetWords % 64 == 0, "regions stay 256-byte aligned");

// The scratch array.  Host code never passes this name to a kernel or a copy (on the host it is a
// stand-in, not a device address): it uses the per-device address below.
static __device__ __align__(256) uint32_t scratch[kNaWords];

constexpr int kNaMaxDevices = 64;
static std::mutex g_na_mutex;
static uint32_t* g_na_scratch[kNaMaxDevices] = {};  // device address of `scratch` on each device

__global__ __launch_bounds__(256) void k_clear_words(uint32_t* __restrict__ p, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Looks up this device's copy of the scratch array once (the device must be current).  Called by
// tpd_create for the context's device, and by the entry point if a device has none yet.
int noalloc_setup(int device) {
    if (device < 0 || device >= kNaMaxDevices) return fail(TPD_ERR_INVALID, "noalloc: device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_na_mutex);
    if (g_na_scratch[device]) return TPD_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e, "noalloc: hipGetSymbolAddress(scratch)");
    g_na_scratch[device] = static_cast<uint32_t*>(p);
    return TPD_OK;
}

}  // namespace tpd

using namespace tpd;

extern "C" int tpd_sort_pairs_noalloc(tpd_ctx* c, uint64_t* d_keys, uint32_t* d_vals, uint64_t* d_keys_alt,
                                      uint32_t* d_vals_alt, uint32_t n, uint32_t key_bits, void* stream, int* in_alt) {
    if (!c) return fail(TPD_ERR_INVALID, "null context");
    if (!in_alt) return fail(TPD_ERR_INVALID, "tpd_sort_pairs_noalloc: null in_alt");
    *in_alt = 0;
    if (key_bits == 0 || key_bits > 64) return fail(TPD_ERR_INVALID, "tpd_sort_pairs_noalloc: key_bits must be 1..64");
    if (n > kNaMaxKeys)
        return fail(TPD_ERR_CAPACITY, "tpd_sort_pairs_noalloc: more than 4,194,304 pairs (the fixed scratch limit); "
                                      "use tpd_sort_pairs");
    if (n == 0) return TPD_OK;
    if (!d_keys || !d_vals || !d_keys_alt || !d_vals_alt)
        return fail(TPD_ERR_INVALID, "tpd_sort_pairs_noalloc: null pointer");
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    int rc = noalloc_setup(c->device);
    if (rc != TPD_OK) return rc;
    uint32_t* base = g_na_scratch[c->device];
    SortWorkspace ws;
    ws.partials = base;
    ws.offsets = base + kNaCountWords;
    ws.tickets = base + kNaCountWords + kNaOffsetWords;
    ws.desc = base + kNaCountWords + kNaOffsetWords + kNaTicketWords;
    ws.parts_cap = kNaParts;
    ws.shape = c->ws.shape;
    hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : c->stream;
    // The histogram kernel adds into the digit counts, which must start at zero.
    hipLaunchKernelGGL(k_clear_words, dim3((uint32_t)((kNaCountWords + 255) / 256)), dim3(256), 0, s, ws.partials,
                       (uint32_t)kNaCountWords);
    *in_alt = launch_sort_pairs(ws, d_keys, d_vals, d_keys_alt, d_vals_alt, nullptr, n, key_bits, s) ? 1 : 0;
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "tpd_sort_pairs_noalloc: launch");
    return TPD_OK;
}

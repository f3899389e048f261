// The library's stand-alone radix sort (tpd_sort_pairs) against rocPRIM's radix_sort_pairs on the
// same GPU and the same input: n (64-bit key, 32-bit value) pairs with random keys below 2^key_bits,
// both sorting by the key's low key_bits bits into ping-pong buffers, the library under each of its
// radix block shapes.  Prints one JSON line per case (mean milliseconds over 10 sorts after one
// warm-up, each sort from the same unsorted input) and checks that every result is sorted and equal
// to rocPRIM's.  Run by `bench.py --only sort-vs-rocprim`.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "tpd_hip.h"

namespace {

bool ok(hipError_t e, const char* what) {
    if (e != hipSuccess) std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    return e == hipSuccess;
}

struct Buffers {
    uint64_t *k = nullptr, *k_alt = nullptr;
    uint32_t *v = nullptr, *v_alt = nullptr;
    ~Buffers() {
        (void)hipFree(k); (void)hipFree(k_alt); (void)hipFree(v); (void)hipFree(v_alt);
    }
};

// Mean ms of `sort` over iters runs after one warm-up, each from the input copied back in first.
template <typename F>
float time_sort(F sort, Buffers& b, const std::vector<uint64_t>& keys, const std::vector<uint32_t>& vals,
                hipStream_t s, int iters) {
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    float total = 0.0f;
    for (int it = 0; it <= iters; ++it) {
        (void)hipMemcpyAsync(b.k, keys.data(), keys.size() * 8, hipMemcpyHostToDevice, s);
        (void)hipMemcpyAsync(b.v, vals.data(), vals.size() * 4, hipMemcpyHostToDevice, s);
        (void)hipEventRecord(e0, s);
        sort();
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (it > 0) total += ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return total / (float)iters;
}

}  // namespace

int main() {
    tpd_ctx* ctx = nullptr;
    tpd_config cfg{};
    if (tpd_create(&ctx, &cfg) != TPD_OK) {
        std::fprintf(stderr, "tpd_create: %s\n", tpd_last_error());
        return 1;
    }
    hipStream_t s;
    if (!ok(hipStreamCreate(&s), "hipStreamCreate")) return 1;
    // (pairs, key bits): the batched frame's two sorts (58.3M pairs by 17 tile bits, 32M items by 27
    // depth bits), the plain path's 44-bit (tile, depth) key, and full 64-bit keys
    const uint32_t cases[][2] = {{58288090u, 17u}, {32000000u, 27u}, {58288090u, 44u}, {8000000u, 64u}};
    std::mt19937_64 rng(7);
    int status = 0;
    for (const auto& c : cases) {
        const uint32_t n = c[0], bits = c[1];
        std::vector<uint64_t> keys(n);
        std::vector<uint32_t> vals(n);
        const uint64_t mask = bits >= 64 ? ~0ull : (1ull << bits) - 1ull;
        for (uint32_t i = 0; i < n; ++i) {
            keys[i] = rng() & mask;
            vals[i] = i;
        }
        Buffers b;
        if (!ok(hipMalloc(&b.k, (size_t)n * 8), "hipMalloc") || !ok(hipMalloc(&b.k_alt, (size_t)n * 8), "hipMalloc") ||
            !ok(hipMalloc(&b.v, (size_t)n * 4), "hipMalloc") || !ok(hipMalloc(&b.v_alt, (size_t)n * 4), "hipMalloc"))
            return 1;

        // rocPRIM: temporary storage sized once, double buffers
        size_t tmp_bytes = 0;
        rocprim::double_buffer<uint64_t> dk(b.k, b.k_alt);
        rocprim::double_buffer<uint32_t> dv(b.v, b.v_alt);
        if (!ok(rocprim::radix_sort_pairs(nullptr, tmp_bytes, dk, dv, n, 0, bits, s), "rocprim size")) return 1;
        void* tmp = nullptr;
        if (!ok(hipMalloc(&tmp, tmp_bytes), "hipMalloc")) return 1;
        const uint64_t* roc_keys = nullptr;
        const float ms_roc = time_sort(
            [&]() {
                rocprim::double_buffer<uint64_t> k2(b.k, b.k_alt);
                rocprim::double_buffer<uint32_t> v2(b.v, b.v_alt);
                (void)rocprim::radix_sort_pairs(tmp, tmp_bytes, k2, v2, n, 0, bits, s);
                roc_keys = k2.current();
            },
            b, keys, vals, s, 10);
        std::vector<uint64_t> out_roc(n), out_tpd(n);
        (void)hipMemcpy(out_roc.data(), roc_keys, (size_t)n * 8, hipMemcpyDeviceToHost);
        (void)hipFree(tmp);

        // every radix block shape of the library; each result checked against rocPRIM's
        std::printf("{\"pairs\": %u, \"key_bits\": %u, \"rocprim_ms\": %.4f, \"tpd_ms_per_shape\": [", n, bits, ms_roc);
        bool all_same = true;
        for (uint32_t sh = 0;; ++sh) {
            if (tpd_set_sort_shape(ctx, sh) != TPD_OK) break;  // past the last shape
            int in_alt = 0;
            int rc = TPD_OK;
            const float ms_tpd = time_sort(
                [&]() { rc = tpd_sort_pairs(ctx, b.k, b.v, b.k_alt, b.v_alt, n, bits, s, &in_alt); }, b, keys, vals, s,
                10);
            if (rc != TPD_OK) {
                std::fprintf(stderr, "tpd_sort_pairs: %s\n", tpd_last_error());
                return 1;
            }
            (void)hipMemcpy(out_tpd.data(), in_alt ? b.k_alt : b.k, (size_t)n * 8, hipMemcpyDeviceToHost);
            all_same = all_same && out_tpd == out_roc;
            std::printf("%s%.4f", sh ? ", " : "", ms_tpd);
        }
        bool sorted = true;
        for (uint32_t i = 1; i < n && sorted; ++i) sorted = out_roc[i - 1] <= out_roc[i];
        if (!sorted || !all_same) status = 1;
        std::printf("], \"sorted\": %s, \"same_as_rocprim\": %s}\n", sorted ? "true" : "false",
                    all_same ? "true" : "false");
        std::fflush(stdout);
    }
    (void)hipStreamDestroy(s);
    tpd_destroy(ctx);
    return status;
}

// Host-only driver for the PLY reader (torpedo_amd/csrc/ply.cpp) under AddressSanitizer and
// UndefinedBehaviorSanitizer: loads every file named on the command line with tpd_load_ply and
// prints "<status> <vertices>" per file.  The reader's error helpers normally live in capi.cpp
// (with the HIP runtime); the two it calls are defined here so that nothing else is linked.
#include "tpd_ctx.hpp"

#include <cstdio>
#include <cstdlib>
#include <string>

namespace tpd {
thread_local std::string g_last_error;
int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}
int hip_fail(hipError_t, const char* what) { return fail(TPD_ERR_HIP, what); }
}  // namespace tpd

int main(int argc, char** argv) {
    for (int i = 1; i < argc; ++i) {
        float* rec = nullptr;
        uint32_t n = 0;
        const int rc = tpd_load_ply(argv[i], &rec, &n);
        double sum = 0.0;  // touch every returned value
        if (rc == TPD_OK)
            for (size_t k = 0; k < (size_t)n * 60; ++k) sum += rec[k];
        std::printf("%d %u %g\n", rc, n, sum);
        tpd_free_host(rec);
    }
    return 0;
}

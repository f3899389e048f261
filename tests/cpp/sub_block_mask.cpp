// Host-side check of tpd::splat_sub_block_mask (torpedo_amd/csrc/tpd_device.hpp): for random
// splats and tiles, every pixel centre that could pass the blend's alpha >= 1/255 test (power and
// alpha evaluated as the blend does, exp in double with a 1e-6 relative slack for the float exp's
// error) must lie in a sub-block whose bit is set.  Usage: sub_block_mask [splats] [relative
// error of the mask's sqrt / reciprocal / log].  Prints counts; exit status 1 on a violation.
#include "tpd_device.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

int main(int argc, char** argv) {
    const int n = argc > 1 ? atoi(argv[1]) : 200000;
    // relative error given to the mask's square roots, reciprocals and logarithm (the device uses
    // approximate instructions): 0, or +-2^-21, larger than theirs
    tpd::mask_host_error = argc > 2 ? (float)atof(argv[2]) : 0.0f;
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    long long violations = 0, bits = 0, needed = 0, splats_some = 0;
    for (int it = 0; it < n; ++it) {
        // tile origin: small and large coordinates (float spacing up to 2^-9 near 16384)
        const float ox = 16.0f * (float)(int)(U(rng) < 0.5 ? U(rng) * 8 : U(rng) * 1020);
        const float oy = 16.0f * (float)(int)(U(rng) < 0.5 ? U(rng) * 8 : U(rng) * 1020);
        float gx, gy, a, b, c, o;
        o = (float)(U(rng) < 0.1 ? U(rng) * 0.01 : std::exp(std::log(0.003) * U(rng)));
        gx = ox + (float)(-40.0 + 96.0 * U(rng));
        gy = oy + (float)(-40.0 + 96.0 * U(rng));
        if (U(rng) < 0.8) {
            // conic = inverse of R diag(s1^2, s2^2) R^T + 0.3 I, as the projection builds it
            const double s1 = std::exp(std::log(0.05) + (std::log(40.0) - std::log(0.05)) * U(rng));
            const double s2 = std::exp(std::log(0.05) + (std::log(40.0) - std::log(0.05)) * U(rng));
            const double th = 6.283185307179586 * U(rng), cs = std::cos(th), sn = std::sin(th);
            const double xx = cs * cs * s1 * s1 + sn * sn * s2 * s2 + 0.3, yy = sn * sn * s1 * s1 + cs * cs * s2 * s2 + 0.3;
            const double xy = cs * sn * (s1 * s1 - s2 * s2);
            const double det = xx * yy - xy * xy;
            a = (float)(yy / det); b = (float)(-xy / det); c = (float)(xx / det);
        } else {
            // raw conics, including nearly degenerate and indefinite ones
            a = (float)(std::exp(-6.0 + 8.0 * U(rng)));
            c = (float)(std::exp(-6.0 + 8.0 * U(rng)));
            b = (float)((2.0 * U(rng) - 1.0) * 1.05 * std::sqrt((double)a * c));
        }
        const uint32_t m = tpd::splat_sub_block_mask(gx, gy, a, b, c, o, ox, oy);
        uint32_t need = 0;
        for (int py = 0; py < 16; ++py) {
            for (int px = 0; px < 16; ++px) {
                const float fx = ox + (float)px, fy = oy + (float)py;
                const float dx = gx - fx, dy = gy - fy;
                const float power = -0.5f * ((a * dx) * dx + (c * dy) * dy) - (b * dx) * dy;
                if (!(power <= 0.0f)) continue;
                const double alpha = std::fmin(0.99, (double)o * std::exp((double)power) * (1.0 + 1e-6));
                if (alpha < 1.0 / 255.0) continue;
                need |= 1u << ((py / 4) * 4 + px / 4);
            }
        }
        if (need & ~m) {
            if (violations < 10)
                printf("violation: g=(%.9g, %.9g) conic=(%.9g, %.9g, %.9g) o=%.9g tile=(%g, %g) mask=%04x need=%04x\n",
                       gx, gy, a, b, c, o, ox, oy, m, need);
            ++violations;
        }
        bits += __builtin_popcount(m);
        needed += __builtin_popcount(need);
        splats_some += need != 0;
    }
    printf("splats %d reaching %lld violations %lld mask_bits %lld needed_bits %lld\n", n, splats_some, violations,
           bits, needed);
    return violations ? 1 : 0;
}

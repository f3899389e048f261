// GaussianPoint and its helpers (torpedo/volumetric/include/torpedo/volumetric/GaussianGeometry.h:10-52)
// for C++ callers of libtpd_hip.so.  Header-only, C++20.
#pragma once

#include <torpedo/math.hpp>
#include <tpd_hip.h>

#include <array>
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace tpd {

// 240-byte record: the layout every entry point of tpd_hip.h takes (TPD_GAUSSIAN_FLOATS floats).
struct GaussianPoint {
    static constexpr uint32_t MAX_SH_FLOATS = 48;

    vec3 position;
    float opacity;   // activated (sigmoid of the logit)
    vec4 quaternion; // x, y, z, w; fed to the kernels unnormalised (volume.slang:28-35)
    vec4 scale;      // activated scale xyz, w = modifier
    std::array<float, MAX_SH_FLOATS> sh;  // DC rgb, then 15 coefficients per channel

    // GaussianGeometry.cpp:8-32: positions uniform in the cube of half-size radius around center,
    // identity rotation, uniform scale / opacity, a random DC colour.  The reference seeds from
    // std::random_device; seed makes the draw reproducible.
    [[nodiscard]] static std::vector<GaussianPoint> random(uint32_t count, float radius = 1.0f,
                                                          const vec3& center = {0.f, 0.f, 0.f},
                                                          float minScale = 0.1f, float maxScale = 1.0f,
                                                          float minOpacity = 0.1f, float maxOpacity = 1.0f,
                                                          uint32_t seed = std::random_device{}());

    // GaussianGeometry.cpp:59-127 (tpd_load_ply).  Throws std::runtime_error on failure.
    [[nodiscard]] static std::vector<GaussianPoint> fromModel(const std::filesystem::path& plyFile);
};
static_assert(sizeof(GaussianPoint) == TPD_GAUSSIAN_FLOATS * sizeof(float), "GaussianPoint is 240 bytes");

namespace utils {
// GaussianGeometry.h:38-45
[[nodiscard]] constexpr std::array<float, GaussianPoint::MAX_SH_FLOATS> rgb2sh(float r, float g, float b) noexcept {
    constexpr float C0 = 0.28209479177387814f;
    std::array<float, GaussianPoint::MAX_SH_FLOATS> sh{};
    sh[0] = (r - 0.5f) / C0;
    sh[1] = (g - 0.5f) / C0;
    sh[2] = (b - 0.5f) / C0;
    return sh;
}
// GaussianGeometry.h:47-52
[[nodiscard]] constexpr vec3 sh2rgb(const std::array<float, GaussianPoint::MAX_SH_FLOATS>& sh) noexcept {
    constexpr float C0 = 0.28209479177387814f;
    return vec3{sh[0] * C0 + 0.5f, sh[1] * C0 + 0.5f, sh[2] * C0 + 0.5f};
}
}  // namespace utils

inline std::vector<GaussianPoint> GaussianPoint::random(uint32_t count, float radius, const vec3& center,
                                                        float minScale, float maxScale, float minOpacity,
                                                        float maxOpacity, uint32_t seed) {
    std::mt19937 rng{seed};
    std::uniform_real_distribution<float> dist_pos(-1.0f, 1.0f);
    std::uniform_real_distribution<float> dist_opacity(minOpacity, maxOpacity);
    std::uniform_real_distribution<float> dist_scale(minScale, maxScale);
    std::uniform_real_distribution<float> dist_sh(0.0f, 1.0f);
    std::vector<GaussianPoint> points(count);
    for (auto& p : points) {
        const float x = dist_pos(rng), y = dist_pos(rng), z = dist_pos(rng);
        p.position = vec3{x, y, z} * radius + center;
        p.opacity = dist_opacity(rng);
        p.quaternion = {0.0f, 0.0f, 0.0f, 1.0f};
        const float sx = dist_scale(rng), sy = dist_scale(rng), sz = dist_scale(rng);
        p.scale = {sx, sy, sz, 1.0f};
        const float r = dist_sh(rng), g = dist_sh(rng), b = dist_sh(rng);
        p.sh = utils::rgb2sh(r, g, b);
    }
    return points;
}

inline std::vector<GaussianPoint> GaussianPoint::fromModel(const std::filesystem::path& plyFile) {
    float* records = nullptr;
    uint32_t n = 0;
    if (tpd_load_ply(plyFile.string().c_str(), &records, &n) != TPD_OK)
        throw std::runtime_error(tpd_last_error());
    std::vector<GaussianPoint> points(n);
    if (n) std::memcpy(static_cast<void*>(points.data()), records, sizeof(GaussianPoint) * n);
    tpd_free_host(records);
    return points;
}

}  // namespace tpd

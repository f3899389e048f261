// Scene (torpedo/rendering/include/torpedo/rendering/Scene.h:17-162) for C++ callers of
// libtpd_hip.so: groups of GaussianPoints (borrowed spans, as the reference's EntityGroup) and
// single points, concatenated groups first, then singles.  The reference keeps them in an EnTT
// registry (EnTT v3.15.0, a build-time dependency: cmake/Dependencies.cmake:9-11) and walks each
// kind with a single-storage view; EnTT's sparse set iterates its packed array from the last
// element to the first, and a scene only ever appends, so within each kind the most recently added
// entity comes first.  Every listing below walks the stored vectors backwards to give that order.
// Header-only, C++20.
#pragma once

#include <torpedo/GaussianGeometry.hpp>

#include <concepts>
#include <cstddef>
#include <cstring>
#include <map>
#include <ranges>
#include <span>
#include <type_traits>
#include <utility>
#include <vector>

namespace tpd {

enum class Entity : uint32_t {};

template <typename T>
using EntityGroup = std::span<const T>;

class Scene final {
public:
    template <typename T, typename... Args>
        requires(std::constructible_from<T, Args...>)
    Entity add(Args&&... args) {
        return add(T{std::forward<Args>(args)...});
    }

    Entity add(const GaussianPoint& point) {
        const Entity e{_next++};
        _singles.emplace_back(e, point);
        return e;
    }

    Entity add(EntityGroup<GaussianPoint> points) {
        const Entity e{_next++};
        _groups.emplace_back(e, points);
        return e;
    }

    template <typename T>
    [[nodiscard]] uint32_t count() const noexcept {
        static_assert(std::is_same_v<T, GaussianPoint>, "scenes hold GaussianPoints");
        return static_cast<uint32_t>(_singles.size());
    }

    template <typename T>
    [[nodiscard]] uint32_t countGroup() const noexcept {
        static_assert(std::is_same_v<T, GaussianPoint>, "scenes hold GaussianPoints");
        return static_cast<uint32_t>(_groups.size());
    }

    template <typename T>
    [[nodiscard]] uint32_t countAll() const noexcept {
        uint32_t n = count<T>();
        for (const auto& g : _groups) n += static_cast<uint32_t>(g.second.size());
        return n;
    }

    // Scene.h:130-146: every group's points, then the single points, each kind last added first.
    template <typename T>
    [[nodiscard]] std::vector<std::byte> dataAll() const {
        std::vector<std::byte> bytes(sizeof(GaussianPoint) * countAll<T>());
        size_t off = 0;
        for (const auto& g : std::views::reverse(_groups)) {
            if (!g.second.empty()) std::memcpy(bytes.data() + off, g.second.data(), g.second.size_bytes());
            off += g.second.size_bytes();
        }
        for (const auto& s : std::views::reverse(_singles)) {
            std::memcpy(bytes.data() + off, &s.second, sizeof(GaussianPoint));
            off += sizeof(GaussianPoint);
        }
        return bytes;
    }

    template <typename T>
    [[nodiscard]] std::vector<uint32_t> groupSizes() const noexcept {
        std::vector<uint32_t> sizes;
        for (const auto& g : std::views::reverse(_groups)) sizes.push_back(static_cast<uint32_t>(g.second.size()));
        return sizes;
    }

    // Scene.h:154-162: groups, then singles (each kind last added first), numbered 0, 1, ...
    template <typename T>
    [[nodiscard]] std::map<Entity, uint32_t> buildEntityMap() const {
        std::map<Entity, uint32_t> m;
        uint32_t i = 0;
        for (const auto& g : std::views::reverse(_groups)) m.emplace(g.first, i++);
        for (const auto& s : std::views::reverse(_singles)) m.emplace(s.first, i++);
        return m;
    }

    // Per-point entity index in dataAll order (GaussianEngine.cpp:383-388).
    template <typename T>
    [[nodiscard]] std::vector<uint32_t> entityIndices() const {
        std::vector<uint32_t> idx;
        idx.reserve(countAll<T>());
        uint32_t k = 0;
        for (const auto& g : std::views::reverse(_groups)) idx.insert(idx.end(), g.second.size(), k++);
        for (size_t i = 0; i < _singles.size(); ++i) idx.push_back(k++);
        return idx;
    }

private:
    uint32_t _next = 0;
    std::vector<std::pair<Entity, EntityGroup<GaussianPoint>>> _groups;
    std::vector<std::pair<Entity, GaussianPoint>> _singles;
};

namespace ent {
template <typename T>
[[nodiscard]] EntityGroup<T> group(const std::vector<T>& elements) {
    return EntityGroup<T>{elements.data(), elements.size()};
}
}  // namespace ent

}  // namespace tpd

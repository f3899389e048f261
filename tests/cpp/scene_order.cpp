// Host-only check of tpd::Scene's listing order (include/torpedo/Scene.hpp): two groups and two
// single points added interleaved; prints the first float (position x) of every dataAll record,
// the group sizes, each entity's buildEntityMap index in creation order, and the per-point entity
// indices, one line each.
#include <torpedo/Scene.hpp>

#include <cstdio>
#include <vector>

int main() {
    std::vector<tpd::GaussianPoint> g1(2), g2(3);
    g1[0].position = {10.f, 0.f, 0.f};
    g1[1].position = {11.f, 0.f, 0.f};
    for (int i = 0; i < 3; ++i) g2[i].position = {20.f + i, 0.f, 0.f};
    tpd::GaussianPoint s1{}, s2{};
    s1.position = {7.f, 0.f, 0.f};
    s2.position = {8.f, 0.f, 0.f};

    tpd::Scene scene;
    const tpd::Entity e[4] = {scene.add(tpd::ent::group(g1)), scene.add(s1), scene.add(tpd::ent::group(g2)),
                              scene.add(s2)};

    const auto bytes = scene.dataAll<tpd::GaussianPoint>();
    const auto* rec = reinterpret_cast<const tpd::GaussianPoint*>(bytes.data());
    for (size_t i = 0; i < bytes.size() / sizeof(tpd::GaussianPoint); ++i) std::printf("%g ", rec[i].position.x);
    std::printf("\n");
    for (const auto n : scene.groupSizes<tpd::GaussianPoint>()) std::printf("%u ", n);
    std::printf("\n");
    const auto map = scene.buildEntityMap<tpd::GaussianPoint>();
    for (const auto x : e) std::printf("%u ", map.at(x));
    std::printf("\n");
    for (const auto k : scene.entityIndices<tpd::GaussianPoint>()) std::printf("%u ", k);
    std::printf("\n");
    return 0;
}

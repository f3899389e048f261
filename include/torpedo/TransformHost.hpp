// TransformHost (torpedo/rendering/include/torpedo/rendering/TransformHost.h:10-28,
// rendering/src/TransformHost.cpp:3-12) for C++ callers of libtpd_hip.so: per-entity model
// matrices of a compiled scene.  Header-only.
#pragma once

#include <torpedo/Scene.hpp>
#include <tpd_hip.h>

#include <map>
#include <stdexcept>
#include <string>

namespace tpd {

class TransformHost final {
public:
    // scene < 0: the engine's plain scene (tpd_set_transform); else a batch scene (tpd_scene_set_transform).
    TransformHost(tpd_ctx* ctx, std::map<Entity, uint32_t> entityMap, int scene = -1)
        : _ctx{ctx}, _map{std::move(entityMap)}, _scene{scene} {}

    TransformHost(const TransformHost&) = delete;
    TransformHost& operator=(const TransformHost&) = delete;

    // TransformHost.cpp:3-12: unknown entities are ignored.
    void transform(Entity entity, const mat4& transform) const {
        const auto it = _map.find(entity);
        if (it == _map.end()) return;
        const int rc = _scene < 0 ? tpd_set_transform(_ctx, it->second, transform.data_ptr())
                                  : tpd_scene_set_transform(_ctx, (uint32_t)_scene, it->second, transform.data_ptr());
        if (rc != TPD_OK) throw std::runtime_error(std::string("TransformHost::transform: ") + tpd_last_error());
    }

private:
    tpd_ctx* _ctx;
    std::map<Entity, uint32_t> _map;
    int _scene;
};

}  // namespace tpd

// Camera and PerspectiveCamera (torpedo/rendering/include/torpedo/rendering/Camera.h:11-38,
// torpedo/extension/include/torpedo/extension/PerspectiveCamera.h:8-31) for C++ callers of
// libtpd_hip.so.  The matrices come from the library's host camera math (tpd_look_at,
// tpd_perspective), which restates Camera.cpp:3-16 and PerspectiveCamera.cpp:3-23.  Header-only.
#pragma once

#include <torpedo/math.hpp>
#include <tpd_hip.h>

#include <cstdint>

namespace tpd {

class Camera {
public:
    explicit Camera(const mat4& w2c = mat4{1.0f}) noexcept : _view{w2c} {}
    virtual ~Camera() = default;

    // Camera.cpp:3-16: rows x = normalize(cross(z, up)), y = normalize(cross(z, x)) (y down), z = forward.
    void lookAt(const vec3& eye, const vec3& center, const vec3& up) noexcept {
        const float e[3] = {eye.x, eye.y, eye.z}, c[3] = {center.x, center.y, center.z}, u[3] = {up.x, up.y, up.z};
        tpd_look_at(e, c, u, _view.data.data());
    }
    // Camera.cpp:18-20: view = [R | t] with R row-major 3x3.
    void lookAt(const float R[9], const vec3& t) noexcept {
        _view = mat4{R[0], R[1], R[2], t.x, R[3], R[4], R[5], t.y, R[6], R[7], R[8], t.z, 0.f, 0.f, 0.f, 1.f};
    }

    void setNear(float near) noexcept { _near = near; }
    void setFar(float far) noexcept { _far = far; }

    [[nodiscard]] const mat4& getViewMatrix() const noexcept { return _view; }
    [[nodiscard]] const float* getViewMatrixData() const noexcept { return _view.data_ptr(); }

    [[nodiscard]] virtual const float* getProjectionData() const noexcept = 0;
    [[nodiscard]] virtual uint32_t getProjectionByteSize() const noexcept = 0;
    virtual void onImageSizeChange(uint32_t, uint32_t) noexcept {}

protected:
    float _near{0.01f};
    float _far{100.0f};

private:
    mat4 _view;
};

class PerspectiveCamera final : public Camera {
public:
    // PerspectiveCamera.h:26-31: vertical fov 60 degrees (fy = sqrt(3)).
    PerspectiveCamera(uint32_t imageWidth, uint32_t imageHeight) : _w{imageWidth}, _h{imageHeight} {
        updateProjectionMatrix(1.7320508075688772f);
    }

    // PerspectiveCamera.cpp:3-7
    void setVerticalFov(float degrees) noexcept { updateProjectionMatrix(tpd_fy_from_fov(degrees)); }

    [[nodiscard]] const float* getProjectionData() const noexcept override { return _projection.data_ptr(); }
    [[nodiscard]] uint32_t getProjectionByteSize() const noexcept override { return sizeof(float) * 16; }

    // PerspectiveCamera.h:39-42: new aspect, same fy.
    void onImageSizeChange(uint32_t w, uint32_t h) noexcept override {
        _w = w;
        _h = h;
        updateProjectionMatrix(_projection(1, 1));
    }

    [[nodiscard]] uint32_t width() const noexcept { return _w; }
    [[nodiscard]] uint32_t height() const noexcept { return _h; }

private:
    // PerspectiveCamera.cpp:9-23: reversed-z projection with fx = fy / aspect.
    void updateProjectionMatrix(float fy) noexcept { tpd_perspective(_w, _h, fy, _near, _far, _projection.data.data()); }

    uint32_t _w, _h;
    mat4 _projection;
};

}  // namespace tpd

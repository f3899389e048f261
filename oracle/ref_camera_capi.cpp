// C entry points over the reference's own camera and math code (Camera.cpp, PerspectiveCamera.{h,cpp},
// torpedo/math headers), compiled from /root/reference by ref_camera.sh into oracle/_ref/.
// TEST INFRASTRUCTURE ONLY: used by tests to pin the oracle's camera restatement bit for bit.
#include <torpedo/rendering/Camera.h>
#include <torpedo/extension/PerspectiveCamera.h>

#include <cstring>

namespace {
// Camera is abstract (getProjectionData is pure virtual); this concrete view-only camera exposes lookAt.
class ViewOnly final : public tpd::Camera {
public:
    const float* getProjectionData() const noexcept override { return nullptr; }
    uint32_t getProjectionByteSize() const noexcept override { return 0; }
};
}

extern "C" {
void tpr_look_at(const float* eye, const float* center, const float* up, float* view) {
    ViewOnly cam;
    cam.lookAt(tpd::vec3{eye[0], eye[1], eye[2]}, tpd::vec3{center[0], center[1], center[2]},
               tpd::vec3{up[0], up[1], up[2]});
    std::memcpy(view, cam.getViewMatrixData(), 16 * sizeof(float));
}

// fov_deg <= 0 keeps the constructor's default fy = sqrt(3).
void tpr_perspective(uint32_t w, uint32_t h, float fov_deg, float* proj) {
    tpd::PerspectiveCamera cam(w, h);
    if (fov_deg > 0.f) cam.setVerticalFov(fov_deg);
    std::memcpy(proj, cam.getProjectionData(), 16 * sizeof(float));
}

void tpr_mat4_mul(const float* a, const float* b, float* out) {
    const auto r = tpd::math::mul(tpd::mat4{a}, tpd::mat4{b});
    std::memcpy(out, r.data_ptr(), 16 * sizeof(float));
}

void tpr_dot3(const float* a, const float* b, float* out) {
    *out = tpd::math::dot(tpd::vec3{a[0], a[1], a[2]}, tpd::vec3{b[0], b[1], b[2]});
}

void tpr_cross3(const float* a, const float* b, float* out) {
    const auto r = tpd::math::cross(tpd::vec3{a[0], a[1], a[2]}, tpd::vec3{b[0], b[1], b[2]});
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
}

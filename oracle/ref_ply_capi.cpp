// C entry point over the reference's own PLY loader, GaussianPoint::fromModel
// (volumetric/src/GaussianGeometry.cpp:59-127, with the reference's vendored miniply), compiled
// from /root/reference by ref_ply.sh into oracle/_ref/.
// TEST INFRASTRUCTURE ONLY: used by tests to pin the library's PLY reader bit for bit.
#include <torpedo/volumetric/GaussianGeometry.h>

#include <cstring>
#include <exception>

extern "C" {
// Returns the point count (records written when it fits capacity), or -1 when fromModel throws.
long tpr_from_model(const char* path, float* out, unsigned long capacity) {
    try {
        const auto points = tpd::GaussianPoint::fromModel(path);
        if (points.size() <= capacity && !points.empty())
            std::memcpy(out, points.data(), points.size() * sizeof(tpd::GaussianPoint));
        return static_cast<long>(points.size());
    } catch (const std::exception&) {
        return -1;
    }
}
}

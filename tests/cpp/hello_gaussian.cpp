// The reference's demo/HelloGaussian/main.cpp against this build's C++ API: the same includes,
// construction sequence, scene and render loop, with the lines that need a window replaced
// (SurfaceRenderer -> HeadlessRenderer; no window title; a fixed orbit of the camera stands in for
// OrbitControl, whose eye, target and up the loop would otherwise take).  After the loop it writes
// what the last frame drew (the swap image read back), the scene records and the camera into the
// directory argv[1]; tests/test_gpu_cpp_api.py checks the frame against the oracle.
#include <torpedo/rendering/Context.h>
#include <torpedo/rendering/Scene.h>
#include <torpedo/rendering/HeadlessRenderer.h>  // the demo includes SurfaceRenderer.h
#include <torpedo/rendering/LogUtils.h>

#include <torpedo/volumetric/GaussianEngine.h>
#include <torpedo/volumetric/GaussianGeometry.h>

#include <torpedo/extension/PerspectiveCamera.h>  // the demo also includes OrbitControl.h

#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

namespace {
template <typename T>
void dump(const std::string& path, const T* data, size_t count) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(data), static_cast<std::streamsize>(count * sizeof(T)));
}
}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: hello_gaussian OUTDIR\n");
        return 2;
    }
    const std::string out = argv[1];

    tpd::utils::plantConsoleLogger();
    const auto context = tpd::Context<tpd::HeadlessRenderer>::create();

    const auto renderer = context->initRenderer(1280, 720);
    renderer->setFrameCount(4);  // replaces renderer->getWindow()->setTitle("Hello, Gaussian!")

    const auto engine = context->bindEngine<tpd::GaussianEngine>();
    const auto camera = context->createCamera<tpd::PerspectiveCamera>();

    // Generate random points within a cube of size 10 centered at the origin
    auto points = tpd::GaussianPoint::random(8192, 10.f, {}, 0.005f, 0.2f);

    // A big, white, uniform Gaussian at the center of the scene
    auto gaussian = tpd::GaussianPoint{
        .position = { 0.f, 0.f, 0.f },
        .opacity = 1.f,
        .quaternion = { 0.f, 0.f, 0.f, 1.f },
        .scale = { 2.f, 2.f, 2.f, 1.0f },
        .sh = tpd::utils::rgb2sh(1.0f, 1.0f, 1.0f),
    };

    auto scene = tpd::Scene{};
    scene.add(tpd::ent::group(points));
    scene.add(std::move(gaussian));

    auto settings = tpd::GaussianEngine::Settings::getDefault();
    settings.sphericalHarmonicsDegree = 0;

    engine->compile(scene, settings);
    const auto records = scene.dataAll<tpd::GaussianPoint>();  // kept for the check below
    points.clear(); // all data has been transferred to the GPU

    int frame = 0;
    renderer->loop([&](const float deltaTimeMillis) {
        (void)deltaTimeMillis;
        // in place of control->getCameraUpdate(deltaTimeMillis): an orbit of radius 8
        const float a = 0.4f * static_cast<float>(frame++);
        const tpd::vec3 eye{ 8.f * std::sin(a), -1.f, -8.f * std::cos(a) }, tar{ 0.f, 0.f, 0.f };
        camera->lookAt(eye, tar, { 0.f, -1.f, 0.f });

        engine->rasterFrame(*camera);
        if (const auto swapImage = renderer->launchFrame(); swapImage) {
            engine->draw(swapImage); // mainly to copy rastered image to swap image
            renderer->submitFrame(swapImage.index);
        }
    });

    const auto last = renderer->lastSubmitted();
    if (last < 0) {
        std::fprintf(stderr, "hello_gaussian: no frame was submitted\n");
        return 1;
    }
    const auto image = renderer->readFrame(static_cast<uint32_t>(last));
    dump(out + "/frame.rgba8", image.data(), image.size());
    dump(out + "/records.f32", records.data(), records.size());  // the 240-byte records, as bytes
    dump(out + "/view.f32", camera->getViewMatrixData(), 16);
    dump(out + "/proj.f32", camera->getProjectionData(), 16);
    const uint32_t info[2] = { static_cast<uint32_t>(frame), static_cast<uint32_t>(last) };
    dump(out + "/frames.u32", info, 2);
    std::printf("hello_gaussian: %d frames, last in swap image %d\n", frame, last);
    return 0;
}

// C++ caller of libtpd_hip.so through include/torpedo/*.hpp, written the way a caller of the
// reference's engine is (demo/HelloGaussian/main.cpp): build a Scene of GaussianPoint groups,
// compile it, rasterFrame a PerspectiveCamera, draw it into a device surface and a PNG file,
// transform an entity, render a batch.  Writes what it rendered into the directory argv[1];
// tests/test_gpu_cpp_api.py checks it against the oracle.
#include <torpedo/Context.hpp>
#include <torpedo/GaussianEngine.hpp>

#include <cstdint>
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {
template <typename T>
void dump(const std::string& path, const T* data, size_t count) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(data), static_cast<std::streamsize>(count * sizeof(T)));
}
template <typename T>
void dump(const std::string& path, const std::vector<T>& v) {
    dump(path, v.data(), v.size());
}
}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: engine_api OUTDIR\n");
        return 2;
    }
    const std::string out = argv[1];
    try {
        const auto a = tpd::GaussianPoint::random(3000, 1.0f, {0.f, 0.f, 0.f}, 0.01f, 0.05f, 0.2f, 0.9f, 11);
        const auto b = tpd::GaussianPoint::random(2000, 0.6f, {0.3f, -0.2f, 0.1f}, 0.01f, 0.03f, 0.3f, 1.0f, 12);
        tpd::GaussianPoint single{};
        single.position = {0.0f, 0.0f, 0.5f};
        single.opacity = 0.8f;
        single.quaternion = {0.0f, 0.0f, 0.0f, 1.0f};
        single.scale = {0.1f, 0.05f, 0.02f, 1.0f};
        single.sh = tpd::utils::rgb2sh(0.9f, 0.2f, 0.1f);

        tpd::Scene scene;
        scene.add(tpd::ent::group(a));
        const auto eb = scene.add(tpd::ent::group(b));
        scene.add(single);

        tpd::GaussianEngine engine(320, 200);
        engine.compile(scene);
        tpd::PerspectiveCamera cam(320, 200);
        cam.lookAt({0.2f, -0.1f, -3.0f}, {0.f, 0.f, 0.f}, {0.f, -1.f, 0.f});
        engine.rasterFrame(cam);

        dump(out + "/records.f32", scene.dataAll<tpd::GaussianPoint>());
        dump(out + "/entity_indices.u32", scene.entityIndices<tpd::GaussianPoint>());
        dump(out + "/group_sizes.u32", scene.groupSizes<tpd::GaussianPoint>());
        {
            const uint32_t eb_index = scene.buildEntityMap<tpd::GaussianPoint>().at(eb);
            dump(out + "/eb_index.u32", &eb_index, 1);
        }
        dump(out + "/view.f32", cam.getViewMatrixData(), 16);
        dump(out + "/proj.f32", cam.getProjectionData(), 16);
        dump(out + "/image.f32", engine.readImage());
        dump(out + "/image.rgba8", engine.readRGBA8());
        // draw(): the RGBA8 target into a device surface with a padded row pitch, as into a swap
        // image; the padding must keep the fill byte.  savePNG(): the same image as a file.
        {
            const uint64_t pitch = 4 * 320 + 64, bytes = pitch * 200;
            void* surface = nullptr;
            if (tpd_alloc(engine.native(), bytes, &surface) != TPD_OK) throw std::runtime_error(tpd_last_error());
            std::vector<uint8_t> fill(bytes, 0xA5);
            if (tpd_memcpy(engine.native(), surface, fill.data(), bytes, 0) != TPD_OK)
                throw std::runtime_error(tpd_last_error());
            engine.draw(surface, pitch);
            if (tpd_memcpy(engine.native(), fill.data(), surface, bytes, 1) != TPD_OK)
                throw std::runtime_error(tpd_last_error());
            (void)tpd_free(engine.native(), surface);
            dump(out + "/surface.rgba8", fill);
            engine.savePNG(out + "/frame.png");
        }
        std::vector<uint64_t> keys;
        std::vector<uint32_t> vals;
        engine.readPairs(keys, vals);
        dump(out + "/keys.u64", keys);
        dump(out + "/vals.u32", vals);

        // TransformHost::transform of the second group, then the same camera again
        const tpd::mat4 M{0.9f, -0.2f, 0.0f, 0.4f, 0.2f, 0.9f, 0.1f, -0.3f, 0.0f, -0.1f, 1.1f, 0.0f, 0.f, 0.f, 0.f, 1.f};
        engine.getTransformHost()->transform(eb, M);
        engine.rasterFrame(cam);
        dump(out + "/image_transformed.f32", engine.readImage());
        dump(out + "/model.f32", M.data_ptr(), 16);

        // batch of two views of the (untransformed) scene
        const uint32_t sid = engine.addScene(scene);
        tpd::PerspectiveCamera cam2(128, 96);
        cam2.setVerticalFov(45.0f);
        cam2.lookAt({-1.0f, 0.5f, -2.5f}, {0.f, 0.f, 0.f}, {0.f, -1.f, 0.f});
        engine.setViews({{sid, &cam, 320, 200}, {sid, &cam2, 128, 96}});
        engine.rasterBatch();
        dump(out + "/batch0.f32", engine.readBatchImage(0));
        dump(out + "/batch1.f32", engine.readBatchImage(1));
        dump(out + "/view2.f32", cam2.getViewMatrixData(), 16);
        dump(out + "/proj2.f32", cam2.getProjectionData(), 16);

        // compile of an empty scene warns and keeps the compiled one (GaussianEngine.cpp:362-365)
        tpd::Scene empty;
        engine.compile(empty);
        engine.rasterFrame(cam);
        dump(out + "/image_after_empty.f32", engine.readImage());

        // a bad view is reported as std::runtime_error carrying the library's message
        std::string msg = "no exception";
        try {
            engine.setViews({{7, &cam, 10, 10}});
        } catch (const std::runtime_error& e) {
            msg = e.what();
        }
        std::ofstream(out + "/error.txt") << msg;
        // Context (Context.h:15-56): a camera before any engine throws; an engine bound before the
        // renderer takes its frame size from initRenderer; a second initRenderer or bindEngine
        // answers nullptr; the engine outlives the context and still draws.
        {
            std::string order;
            auto context = tpd::Context<tpd::HeadlessRenderer>::create();
            try {
                (void)context->createCamera<tpd::PerspectiveCamera>();
                order += "camera-before-engine:no-throw ";
            } catch (const std::runtime_error&) {
                order += "camera-before-engine:throws ";
            }
            auto eng = context->bindEngine<tpd::GaussianEngine>();
            auto* r = context->initRenderer(96, 64);
            order += context->initRenderer(10, 10) == nullptr ? "second-init:null " : "second-init:renderer ";
            order += context->bindEngine<tpd::GaussianEngine>() == nullptr ? "second-bind:null " : "second-bind:engine ";
            auto cam3 = context->createCamera<tpd::PerspectiveCamera>();
            cam3->lookAt({0.2f, -0.1f, -3.0f}, {0.f, 0.f, 0.f}, {0.f, -1.f, 0.f});
            eng->compile(scene);
            eng->rasterFrame(*cam3);
            const auto swap = r->launchFrame();
            eng->draw(swap);
            r->submitFrame(swap.index);
            const auto drawn = r->readFrame(swap.index);
            dump(out + "/ctx_frame.rgba8", drawn);
            dump(out + "/ctx_view.f32", cam3->getViewMatrixData(), 16);
            dump(out + "/ctx_proj.f32", cam3->getProjectionData(), 16);
            context.reset();  // the engine outlives its context
            eng->rasterFrame(*cam3);
            dump(out + "/ctx_after.rgba8", eng->readRGBA8());
            std::ofstream(out + "/ctx_order.txt") << order;
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "engine_api: %s\n", e.what());
        return 1;
    }
    std::printf("engine_api: ok\n");
    return 0;
}

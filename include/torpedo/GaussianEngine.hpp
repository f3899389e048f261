// GaussianEngine (torpedo/volumetric/include/torpedo/volumetric/GaussianEngine.h:15-27) for C++
// callers of libtpd_hip.so: compile a Scene, rasterFrame a Camera, read the frame back, and the
// two added forms, the backward pass and batched rendering.  Every call goes through the C ABI
// of tpd_hip.h; failures of the device layer throw std::runtime_error carrying tpd_last_error(),
// misuse that the reference only logs is logged to stderr and ignored the same way
// (compile of an empty scene, GaussianEngine.cpp:362-365; SH degree above 3, :366-370).
// Header-only, C++20.
#pragma once

#include <torpedo/Camera.hpp>
#include <torpedo/GaussianGeometry.hpp>
#include <torpedo/HeadlessRenderer.hpp>
#include <torpedo/Scene.hpp>
#include <torpedo/TransformHost.hpp>
#include <tpd_hip.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace tpd {

class GaussianEngine final {
public:
    struct Settings {
        uint32_t sphericalHarmonicsDegree{3};
        [[nodiscard]] static constexpr Settings getDefault() { return {}; }
    };

    // Device-side gradient buffers of one scene (float, device memory).
    struct Gradients {
        float* means;           // N x 3
        float* logScales;       // N x 3
        float* quaternions;     // N x 4
        float* opacityLogits;   // N
        float* sh;              // N x 48
    };

    // One view of a batch.
    struct View {
        uint32_t scene;
        const Camera* camera;
        uint32_t width, height;
    };

    // Context::create + initRenderer(w, h) + bindEngine<GaussianEngine> (Context.h:18-36).
    GaussianEngine(uint32_t imageWidth, uint32_t imageHeight, int device = 0, bool keepUnsorted = false)
        : _w{imageWidth}, _h{imageHeight} {
        tpd_config cfg{};
        cfg.device = device;
        cfg.keep_unsorted = keepUnsorted ? 1u : 0u;
        check(tpd_create(&_ctx, &cfg), "GaussianEngine");
    }
    ~GaussianEngine() {
        if (_renderer && _renderer->boundEngine() == this) _renderer->resetEngine();  // swap images on _ctx
        tpd_destroy(_ctx);
    }
    GaussianEngine(const GaussianEngine&) = delete;
    GaussianEngine& operator=(const GaussianEngine&) = delete;

    // GaussianEngine.cpp:359-399
    void compile(const Scene& scene, const Settings& settings = Settings::getDefault()) {
        const uint32_t n = scene.countAll<GaussianPoint>();
        if (n == 0) {
            std::fprintf(stderr, "GaussianEngine - Scene compilation warning: Could NOT find a single "
                                 "tpd::GaussianPoint in the scene!\n");
            return;
        }
        uint32_t deg = settings.sphericalHarmonicsDegree;
        if (deg > 3) {
            std::fprintf(stderr, "GaussianEngine - Clamping to the maximum degrees of SH support: up to 3 degrees only\n");
            deg = 3;
        }
        const auto data = scene.dataAll<GaussianPoint>();
        auto map = scene.buildEntityMap<GaussianPoint>();
        const auto idx = scene.entityIndices<GaussianPoint>();
        const uint32_t ne = static_cast<uint32_t>(map.size());
        std::vector<float> models(16 * (size_t)ne, 0.0f);
        for (uint32_t e = 0; e < ne; ++e)
            for (int i = 0; i < 4; ++i) models[16 * (size_t)e + 5 * i] = 1.0f;
        check(tpd_upload(_ctx, reinterpret_cast<const float*>(data.data()), n, idx.data(), models.data(), ne, deg),
              "GaussianEngine::compile");
        _n = n;
        _transformHost = std::make_unique<TransformHost>(_ctx, std::move(map));
    }

    [[nodiscard]] const std::unique_ptr<TransformHost>& getTransformHost() const noexcept { return _transformHost; }

    // onFramebufferResize (GaussianEngine.cpp:206-220)
    void resize(uint32_t imageWidth, uint32_t imageHeight) {
        _w = imageWidth;
        _h = imageHeight;
    }

    // GaussianEngine.cpp:621-712 (+ updateCameraBuffer :764-775)
    void rasterFrame(const Camera& camera, void* stream = nullptr) {
        float pv[16], focal[2];
        tpd_camera_upload(camera.getViewMatrixData(), camera.getProjectionData(), pv, focal);
        check(tpd_set_camera(_ctx, camera.getViewMatrixData(), pv, focal, _w, _h), "GaussianEngine::rasterFrame");
        check(tpd_render(_ctx, stream), "GaussianEngine::rasterFrame");
    }

    void waitIdle() { check(tpd_synchronize(_ctx), "GaussianEngine::waitIdle"); }

    // The C-ABI context, for the tpd_* calls this class does not wrap (device memory, timing).
    [[nodiscard]] tpd_ctx* native() const noexcept { return _ctx; }

    // ---- readbacks (the reference cannot read its frame back)
    [[nodiscard]] uint32_t pairCount() const {
        uint32_t p = 0;
        check(tpd_get_pair_count(_ctx, &p), "GaussianEngine::pairCount");
        return p;
    }
    [[nodiscard]] std::vector<float> readImage() const {
        std::vector<float> img(3 * (size_t)_w * _h);
        check(tpd_read_image_f32(_ctx, img.data()), "GaussianEngine::readImage");
        return img;
    }
    [[nodiscard]] std::vector<uint8_t> readRGBA8() const {
        std::vector<uint8_t> img(4 * (size_t)_w * _h);
        check(tpd_read_image_rgba8(_ctx, img.data()), "GaussianEngine::readRGBA8");
        return img;
    }
    void readPairs(std::vector<uint64_t>& keys, std::vector<uint32_t>& vals) const {
        const uint32_t p = pairCount();
        keys.resize(p);
        vals.resize(p);
        if (p) check(tpd_read_pairs(_ctx, keys.data(), vals.data(), p, 0), "GaussianEngine::readPairs");
    }
    [[nodiscard]] std::vector<uint32_t> readRanges() const {
        std::vector<uint32_t> r(2 * (size_t)((_w + 15) / 16) * ((_h + 15) / 16));
        check(tpd_read_ranges(_ctx, r.data()), "GaussianEngine::readRanges");
        return r;
    }
    // draw(SwapImage) (GaussianEngine.h:27, GaussianEngine.cpp:714-762, 865-875): the RGBA8 target
    // into a device surface of row pitch `pitch` bytes (0: 4 * width), ordered on `stream`.
    void draw(void* surface, uint64_t pitch = 0, void* stream = nullptr) const {
        check(tpd_present_rgba8(_ctx, surface, pitch, stream), "GaussianEngine::draw");
    }
    // draw(SwapImage) as the reference's render loop calls it: into a renderer's swap image.
    void draw(const SwapImage& swapImage, void* stream = nullptr) const {
        if (!swapImage || swapImage.width != _w || swapImage.height != _h)
            throw std::runtime_error("GaussianEngine::draw: the swap image does not match the frame size");
        check(tpd_present_rgba8(_ctx, swapImage.image, swapImage.pitch, stream), "GaussianEngine::draw");
    }
    // Context::bindEngine: the renderer whose swap images live on this engine's context (shared, so
    // that engine and context may be destroyed in either order).
    void attachRenderer(std::shared_ptr<HeadlessRenderer> renderer) noexcept {
        if (_renderer && _renderer != renderer) _renderer->resetEngine();
        _renderer = std::move(renderer);
        if (_renderer) _renderer->setEngine(this, _ctx);
    }
    void savePNG(const std::string& path) const { check(tpd_save_png(_ctx, path.c_str()), "GaussianEngine::savePNG"); }
    [[nodiscard]] float* deviceImage() const {
        float* p = nullptr;
        check(tpd_device_image(_ctx, &p), "GaussianEngine::deviceImage");
        return p;
    }

    // ---- backward of the last frame (the reference's differentiable rasterizer is a stub)
    void backward(const float* dLdImage, const Gradients& g, bool accumulate = false, void* stream = nullptr) {
        check(tpd_backward(_ctx, dLdImage, g.means, g.logScales, g.quaternions, g.opacityLogits, g.sh,
                           accumulate ? 1 : 0, stream),
              "GaussianEngine::backward");
    }

    // ---- batched rendering: several scenes, many views per launch
    uint32_t addScene(const Scene& scene, const Settings& settings = Settings::getDefault()) {
        const uint32_t n = scene.countAll<GaussianPoint>();
        const auto data = scene.dataAll<GaussianPoint>();
        auto map = scene.buildEntityMap<GaussianPoint>();
        const auto idx = scene.entityIndices<GaussianPoint>();
        const uint32_t ne = std::max<uint32_t>(static_cast<uint32_t>(map.size()), 1u);
        std::vector<float> models(16 * (size_t)ne, 0.0f);
        for (uint32_t e = 0; e < ne; ++e)
            for (int i = 0; i < 4; ++i) models[16 * (size_t)e + 5 * i] = 1.0f;
        uint32_t id = 0;
        check(tpd_scene_add(_ctx, n ? data.data() : nullptr, n, n ? idx.data() : nullptr, models.data(), ne,
                            std::min<uint32_t>(settings.sphericalHarmonicsDegree, 3u), &id),
              "GaussianEngine::addScene");
        _batchHosts.push_back(std::make_unique<TransformHost>(_ctx, std::move(map), static_cast<int>(id)));
        return id;
    }
    [[nodiscard]] const TransformHost& getSceneTransformHost(uint32_t scene) const { return *_batchHosts.at(scene); }

    void setViews(const std::vector<View>& views) {
        std::vector<tpd_view> tv(views.size());
        std::vector<std::pair<uint32_t, uint32_t>> sizes;
        for (size_t i = 0; i < views.size(); ++i) {
            tpd_view& t = tv[i];
            t.scene = views[i].scene;
            t.width = views[i].width;
            t.height = views[i].height;
            std::memcpy(t.view, views[i].camera->getViewMatrixData(), sizeof(t.view));
            tpd_camera_upload(t.view, views[i].camera->getProjectionData(), t.proj_view, t.focal_ndc);
            sizes.emplace_back(views[i].width, views[i].height);
        }
        check(tpd_batch_set_views(_ctx, tv.data(), static_cast<uint32_t>(tv.size())), "GaussianEngine::setViews");
        _viewSizes = std::move(sizes);
    }
    void rasterBatch(bool forBackward = false, void* stream = nullptr) {
        check(tpd_batch_render(_ctx, forBackward ? 1 : 0, stream), "GaussianEngine::rasterBatch");
    }
    void syncBatch() { check(tpd_batch_synchronize(_ctx), "GaussianEngine::syncBatch"); }
    void batchBackward(const float* dLdImages, const std::vector<Gradients>& grads, bool accumulate = false,
                       void* stream = nullptr) {
        const auto t = table(grads);
        check(tpd_batch_backward(_ctx, dLdImages, t.data(), accumulate ? 1 : 0, stream), "GaussianEngine::batchBackward");
    }
    void batchStep(const float* dLdImages, const std::vector<Gradients>& grads, bool accumulate = false,
                   void* stream = nullptr) {
        const auto t = table(grads);
        check(tpd_batch_step(_ctx, dLdImages, t.data(), accumulate ? 1 : 0, stream), "GaussianEngine::batchStep");
    }
    [[nodiscard]] std::vector<float> readBatchImage(uint32_t view) const {
        const auto& s = _viewSizes.at(view);
        std::vector<float> img(3 * (size_t)s.first * s.second);
        check(tpd_batch_read_image_f32(_ctx, view, img.data()), "GaussianEngine::readBatchImage");
        return img;
    }
    void saveBatchPNG(uint32_t view, const std::string& path) const {
        check(tpd_batch_save_png(_ctx, view, path.c_str()), "GaussianEngine::saveBatchPNG");
    }
    [[nodiscard]] uint32_t batchPairCount() const {
        uint32_t p = 0;
        check(tpd_batch_get_pair_count(_ctx, &p), "GaussianEngine::batchPairCount");
        return p;
    }

    [[nodiscard]] tpd_ctx* handle() const noexcept { return _ctx; }
    [[nodiscard]] uint32_t width() const noexcept { return _w; }
    [[nodiscard]] uint32_t height() const noexcept { return _h; }

private:
    static void check(int rc, const char* what) {
        if (rc != TPD_OK) throw std::runtime_error(std::string(what) + ": " + tpd_last_error());
    }
    static std::vector<tpd_scene_grads> table(const std::vector<Gradients>& grads) {
        std::vector<tpd_scene_grads> t(grads.size());
        for (size_t i = 0; i < grads.size(); ++i)
            t[i] = tpd_scene_grads{grads[i].means, grads[i].logScales, grads[i].quaternions, grads[i].opacityLogits,
                                   grads[i].sh};
        return t;
    }

    tpd_ctx* _ctx = nullptr;
    std::shared_ptr<HeadlessRenderer> _renderer;
    uint32_t _w, _h;
    uint32_t _n = 0;
    std::unique_ptr<TransformHost> _transformHost;
    std::vector<std::unique_ptr<TransformHost>> _batchHosts;
    std::vector<std::pair<uint32_t, uint32_t>> _viewSizes;
};

}  // namespace tpd

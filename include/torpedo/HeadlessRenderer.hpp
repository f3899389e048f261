// HeadlessRenderer (torpedo/rendering/include/torpedo/rendering/HeadlessRenderer.h) for C++ callers
// of libtpd_hip.so: the frame target without a window.  It keeps the frame loop of the reference's
// SurfaceRenderer (SurfaceRenderer.h: loop, launchFrame, submitFrame; demo/HelloGaussian/main.cpp)
// over a ring of device RGBA8 images ("swap images", row pitch rounded up to 256 bytes) that the
// engine draws into, so that the reference's render loop runs unchanged without a surface:
//
//     renderer->loop([&](float dt) {
//         engine->rasterFrame(*camera);
//         if (const auto swapImage = renderer->launchFrame(); swapImage) {
//             engine->draw(swapImage);
//             renderer->submitFrame(swapImage.index);
//         }
//     });
//
// loop() runs the callback for setFrameCount() frames (a window would run until it closes);
// readFrame() copies a submitted swap image back to the host.  Header-only, C++20.
#pragma once

#include <tpd_hip.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace tpd {

class GaussianEngine;

// One device image of the renderer's ring: RGBA8 rows of `pitch` bytes.
struct SwapImage {
    void* image{nullptr};
    uint64_t pitch{0};
    uint32_t index{0};
    uint32_t width{0}, height{0};
    explicit operator bool() const noexcept { return image != nullptr; }
};

class HeadlessRenderer final {
public:
    static constexpr uint32_t IN_FLIGHT_FRAME_COUNT = 2;

    HeadlessRenderer() = default;
    HeadlessRenderer(const HeadlessRenderer&) = delete;
    HeadlessRenderer& operator=(const HeadlessRenderer&) = delete;
    ~HeadlessRenderer() { release(); }

    [[nodiscard]] std::pair<uint32_t, uint32_t> getFramebufferSize() const noexcept { return {_w, _h}; }
    [[nodiscard]] uint32_t getInFlightFrameCount() const noexcept { return IN_FLIGHT_FRAME_COUNT; }
    [[nodiscard]] uint32_t getCurrentFrameIndex() const noexcept { return _current; }
    [[nodiscard]] bool supportSurfaceRendering() const noexcept { return false; }
    [[nodiscard]] bool initialized() const noexcept { return _w > 0 && _h > 0; }

    // Frames loop() runs (default 1).
    void setFrameCount(uint32_t frames) noexcept { _frames = frames; }

    void loop(const std::function<void()>& onRender) const {
        for (uint32_t i = 0; i < _frames; ++i) onRender();
    }
    // The callback receives the time since the previous frame started, in milliseconds.
    void loop(const std::function<void(float)>& onRender) const {
        auto last = std::chrono::steady_clock::now();
        for (uint32_t i = 0; i < _frames; ++i) {
            const auto now = std::chrono::steady_clock::now();
            onRender(std::chrono::duration<float, std::milli>(now - last).count());
            last = now;
        }
    }

    // The next swap image of the ring, or an empty one before an engine was bound.
    [[nodiscard]] SwapImage launchFrame() {
        if (!_engineInitialized) return {};
        return SwapImage{_images[_current], _pitch, _current, _w, _h};
    }
    // Presents the image: waits for the engine's work on it (the copy of the draw call).
    void submitFrame(uint32_t imageIndex) {
        if (!_engineInitialized || imageIndex >= IN_FLIGHT_FRAME_COUNT)
            throw std::runtime_error("HeadlessRenderer::submitFrame: bad image");
        check(tpd_synchronize(_ctx), "HeadlessRenderer::submitFrame");
        _presented = static_cast<int>(imageIndex);
        _current = (imageIndex + 1) % IN_FLIGHT_FRAME_COUNT;
    }
    [[nodiscard]] int lastSubmitted() const noexcept { return _presented; }

    // Host copy of a swap image: height rows of 4 * width bytes (the pitch padding dropped).
    [[nodiscard]] std::vector<uint8_t> readFrame(uint32_t imageIndex) const {
        if (!_engineInitialized || imageIndex >= IN_FLIGHT_FRAME_COUNT)
            throw std::runtime_error("HeadlessRenderer::readFrame: bad image");
        std::vector<uint8_t> pitched(_pitch * _h), out(4ull * _w * _h);
        check(tpd_memcpy(_ctx, pitched.data(), _images[imageIndex], pitched.size(), 1), "HeadlessRenderer::readFrame");
        for (uint32_t y = 0; y < _h; ++y)
            std::copy_n(pitched.data() + y * _pitch, 4ull * _w, out.data() + 4ull * _w * y);
        return out;
    }

    // Context::initRenderer (Context.h:152-167): the frame size.
    void init(uint32_t frameWidth, uint32_t frameHeight) {
        if (frameWidth == 0 || frameHeight == 0) throw std::runtime_error("HeadlessRenderer: empty frame size");
        _w = frameWidth;
        _h = frameHeight;
    }
    // Context::bindEngine (Context.h:192-238): the engine bound to this renderer and its device context.
    void setEngine(GaussianEngine* engine, tpd_ctx* engineContext) noexcept {
        release();
        _engine = engine;
        _ctx = engineContext;
    }
    [[nodiscard]] GaussianEngine* boundEngine() const noexcept { return _engine; }
    // Context::engineInitRenderer (Context.h:241-257): the ring of swap images, allocated on the bound
    // engine's device context.
    void engineInit() {
        if (!_ctx || !initialized()) throw std::runtime_error("HeadlessRenderer: bind an engine and set a frame size first");
        freeImages();
        _pitch = (4ull * _w + 255) / 256 * 256;
        for (auto& img : _images) check(tpd_alloc(_ctx, _pitch * _h, &img), "HeadlessRenderer: swap image");
        _current = 0;
        _presented = -1;
        _engineInitialized = true;
    }
    [[nodiscard]] bool engineInitialized() const noexcept { return _engineInitialized; }
    // Renderer::resetEngine: the engine goes away (its destructor calls this) or another is bound.
    void resetEngine() noexcept { release(); }

private:
    static void check(int rc, const char* what) {
        if (rc != TPD_OK) throw std::runtime_error(std::string(what) + ": " + tpd_last_error());
    }
    void freeImages() noexcept {
        if (_ctx)
            for (auto& img : _images)
                if (img) (void)tpd_free(_ctx, img);
        _images.fill(nullptr);
        _engineInitialized = false;
    }
    void release() noexcept {
        freeImages();
        _ctx = nullptr;
        _engine = nullptr;
    }

    uint32_t _w{0}, _h{0};
    uint32_t _frames{1};
    tpd_ctx* _ctx{nullptr};
    GaussianEngine* _engine{nullptr};
    bool _engineInitialized{false};
    std::array<void*, IN_FLIGHT_FRAME_COUNT> _images{};
    uint64_t _pitch{0};
    uint32_t _current{0};
    int _presented{-1};
};

}  // namespace tpd

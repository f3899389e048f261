// Context (torpedo/rendering/include/torpedo/rendering/Context.h:15-56) for C++ callers of
// libtpd_hip.so: the construction sequence of the reference's demos,
//
//     const auto context  = tpd::Context<tpd::HeadlessRenderer>::create();
//     const auto renderer = context->initRenderer(1280, 720);
//     const auto engine   = context->bindEngine<tpd::GaussianEngine>();
//     const auto camera   = context->createCamera<tpd::PerspectiveCamera>();
//
// over one HIP device (create(device), 0 by default) instead of a Vulkan instance.  The renderer is
// held by the context and shared with the engine bound to it (so the two may be destroyed in either
// order); engines and cameras are owned by the unique_ptrs returned to the caller, as
// in the reference (which allocates them from a pool behind a custom deleter; here the default
// deleter frees them).  Misuse the reference logs and answers with nullptr (a second initRenderer,
// binding a second engine of the same type) does the same here, logged to stderr; what it throws
// for (a camera before any engine) throws std::runtime_error.  Header-only, C++20.
#pragma once

#include <torpedo/Camera.hpp>
#include <torpedo/GaussianEngine.hpp>
#include <torpedo/HeadlessRenderer.hpp>

#include <cstdio>
#include <memory>
#include <stdexcept>
#include <type_traits>

namespace tpd {

template <typename R>
class Context final {
    static_assert(std::is_same_v<R, HeadlessRenderer>, "this build renders without a window: use HeadlessRenderer");

public:
    // Context.h:63-70
    [[nodiscard]] static std::unique_ptr<Context> create(int device = 0) {
        return std::unique_ptr<Context>(new Context(device));
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;

    // Context.h:152-167
    [[nodiscard]] R* initRenderer(uint32_t frameWidth, uint32_t frameHeight) {
        if (_renderer->initialized()) {
            std::fprintf(stderr, "Context - A Renderer has already been initialized with the current Context: "
                                 "create a new Context if you want to have another Renderer, returning nullptr\n");
            return nullptr;
        }
        _renderer->init(frameWidth, frameHeight);
        if (_renderer->boundEngine()) engineInitRenderer();
        return _renderer.get();
    }

    // Context.h:192-238.  A headless renderer may be initialised after the engine is bound
    // (Context.h:224-229): the engine then takes the frame size from initRenderer.
    template <typename E>
    [[nodiscard]] std::unique_ptr<E> bindEngine() {
        static_assert(std::is_same_v<E, GaussianEngine>, "GaussianEngine is the engine this build provides");
        if (_renderer->boundEngine()) {
            std::fprintf(stderr, "Context - An Engine of the same type has already been bound with the current "
                                 "Context: returning null\n");
            return nullptr;
        }
        const auto [w, h] = _renderer->getFramebufferSize();
        auto engine = std::make_unique<E>(w, h, _device);
        engine->attachRenderer(_renderer);
        if (_renderer->initialized()) engineInitRenderer();
        return engine;
    }

    // Context.h:259-273
    template <typename E>
    void destroyEngine(std::unique_ptr<E> engine) noexcept {
        if (!engine) {
            std::fprintf(stderr, "Context - Destroying an Engine that has already been destroyed\n");
            return;
        }
        if (engine.get() != _renderer->boundEngine()) {
            std::fprintf(stderr, "Context - Destroying an Engine that has not been bound to this Context\n");
            return;
        }
        engine.reset();  // frees the renderer's swap images first
    }

    // Context.h:275-292: the camera takes the renderer's frame size.
    template <typename C>
    [[nodiscard]] std::unique_ptr<C> createCamera() {
        if (!_renderer->engineInitialized())
            throw std::runtime_error("Context - Create Camera before binding an Engine");
        const auto [w, h] = _renderer->getFramebufferSize();
        return std::make_unique<C>(w, h);
    }
    template <typename C>
    void destroyCamera(std::unique_ptr<C> camera) noexcept {
        camera.reset();
    }

    // Context.h:305-314.  The renderer is shared with the bound engine, which may outlive the
    // context (its swap images then stay valid until the engine goes).
    ~Context() noexcept = default;

private:
    explicit Context(int device) : _device{device}, _renderer{std::make_shared<R>()} {}

    // Context.h:241-257: the engine draws into the renderer's swap images, allocated on its device.
    void engineInitRenderer() {
        const auto [w, h] = _renderer->getFramebufferSize();
        _renderer->boundEngine()->resize(w, h);
        _renderer->engineInit();
    }

    int _device;
    std::shared_ptr<R> _renderer;
};

}  // namespace tpd

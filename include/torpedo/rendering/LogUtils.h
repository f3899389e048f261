// The reference's include path torpedo/rendering/LogUtils.h
// (torpedo/rendering/include/torpedo/rendering/LogUtils.h), so that its callers' #include lines compile
// unchanged against this build.  plantConsoleLogger() starts the reference's console logger; the
// C++ layer here writes its warnings to stderr without one, so it has nothing to start.
#pragma once

namespace tpd::utils {
inline void plantConsoleLogger() noexcept {}
}  // namespace tpd::utils

// The reference's include path torpedo/rendering/HeadlessRenderer.h (torpedo/rendering/include/torpedo/rendering/HeadlessRenderer.h),
// so that its callers' #include lines compile unchanged against this build.
#pragma once

#include <torpedo/HeadlessRenderer.hpp>

// The reference's include path torpedo/rendering/Camera.h (torpedo/rendering/include/torpedo/rendering/Camera.h),
// so that its callers' #include lines compile unchanged against this build.
#pragma once

#include <torpedo/Camera.hpp>

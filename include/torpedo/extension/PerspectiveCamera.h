// The reference's include path torpedo/extension/PerspectiveCamera.h (torpedo/extension/include/torpedo/extension/PerspectiveCamera.h),
// so that its callers' #include lines compile unchanged against this build.
#pragma once

#include <torpedo/Camera.hpp>

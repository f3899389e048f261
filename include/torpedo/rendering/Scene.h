// The reference's include path torpedo/rendering/Scene.h (torpedo/rendering/include/torpedo/rendering/Scene.h),
// so that its callers' #include lines compile unchanged against this build.
#pragma once

#include <torpedo/Scene.hpp>

// The reference's include path torpedo/rendering/TransformHost.h (torpedo/rendering/include/torpedo/rendering/TransformHost.h),
// so that its callers' #include lines compile unchanged against this build.
#pragma once

#include <torpedo/TransformHost.hpp>

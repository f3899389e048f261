// The reference's include path torpedo/rendering/Context.h (torpedo/rendering/include/torpedo/rendering/Context.h),
// so that its callers' #include lines compile unchanged against this build.
#pragma once

#include <torpedo/Context.hpp>

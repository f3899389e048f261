// The reference's include path torpedo/volumetric/GaussianEngine.h (torpedo/volumetric/include/torpedo/volumetric/GaussianEngine.h),
// so that its callers' #include lines compile unchanged against this build.
#pragma once

#include <torpedo/GaussianEngine.hpp>

// The reference's include path torpedo/volumetric/GaussianGeometry.h (torpedo/volumetric/include/torpedo/volumetric/GaussianGeometry.h),
// so that its callers' #include lines compile unchanged against this build.
#pragma once

#include <torpedo/GaussianGeometry.hpp>

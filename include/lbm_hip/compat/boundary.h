// Compatibility header: the reference's include/boundary.h name, served by the
// GPU path's C++ front-end. Put include/lbm_hip/compat ahead of (or instead
// of) the reference's include directory.
#pragma once
#include "lbm_hip/lbm.hpp"

// Compatibility header: the reference's include/io/vtk.h name, served by the
// GPU path's front-end (config file and command line, scenario XML, legacy
// VTK mask, .vts output).
#pragma once
#include "lbm_hip/io.hpp"

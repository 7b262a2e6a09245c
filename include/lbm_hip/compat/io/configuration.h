// Compatibility header: the reference's include/io/configuration.h name, served by the
// GPU path's front-end (config file and command line, scenario XML, legacy
// VTK mask, .vts output).
#pragma once
#include <omp.h>  // omp_get_wtime, which the reference driver times with

#include "lbm_hip/io.hpp"

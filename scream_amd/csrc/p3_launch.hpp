// p3_launch.hpp -- host-side entry points of the compiled kernel variants.
#pragma once

#include <hip/hip_runtime.h>

#include "p3_kernel.hpp"

namespace p3d {

// Block shapes compiled into the library: columns per block x threads per block.
// Returns hipErrorInvalidValue for a shape that is not compiled in.
hipError_t launch_p3_fp64(const KParams& p, int columns_per_block, int threads_per_block, hipStream_t s);
hipError_t launch_p3_mixed(const KParams& p, int columns_per_block, int threads_per_block, hipStream_t s);

}  // namespace p3d

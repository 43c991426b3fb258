// FP64 drop-in variants of the fused P3 kernel.
#include "p3_launch.hpp"
namespace p3d {
#define REAL double
#define LAUNCH_NAME launch_p3_fp64
#define LAUNCH_BATCH_NAME launch_p3_batch_fp64
#define OCC_NAME occupancy_p3_fp64
#include "p3_kernel_variants.inc"
}  // namespace p3d

// Mixed-precision variants: FP32 bulk process rates, FP64 saturation,
// limiters, updates and sedimentation fluxes.
#include "p3_launch.hpp"
namespace p3d {
#define REAL float
#define LAUNCH_NAME launch_p3_mixed
#define LAUNCH_BATCH_NAME launch_p3_batch_mixed
#define OCC_NAME occupancy_p3_mixed
#include "p3_kernel_variants.inc"
}  // namespace p3d

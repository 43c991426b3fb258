// p3_launch.hpp -- host-side entry points of the compiled kernel variants.
#pragma once

#include <hip/hip_runtime.h>

#include "p3_batch.hpp"

namespace p3d {

// Shape of one step's launches.
//   part 1: p1lvl = 1, one thread per (level, column) on the pointwise tiles
//     (pw_tiles); else cb columns per block x nt threads per block (compiled:
//     16x128, 16x256, 32x256, 64x256, 32x512, 64x512);
//   part 2: one thread per (level, column) on the pointwise tiles, register
//     allocation for wpe = 1 or 2 wavefronts per SIMD (2: fewer registers,
//     with spills in FP64);
//   sedimentation, two kernels: the liquid group (cloud, rain) on lpc lanes
//     per column and the ice group on lpc_ice (16, 32 or 64 each), each lane
//     holding L = ceil(nlev / lanes) levels (at most 8), 4 wavefronts per
//     block (the batched form runs both groups on lpc);
//   part 3: one thread per (level, column) on the pointwise tiles.
struct Shape {
  int cb, nt, lpc, wpe;
  int p1lvl;    // 1: part 1 one thread per (level, column); 0: the cb x nt block kernel
  int lpc_ice;  // lanes per column of the ice group's sedimentation kernel
};

// Largest levels-per-lane count of the sedimentation kernel.
constexpr int kMaxLevelsPerLane = 8;
// Wavefronts per block of the sedimentation kernel.
#ifndef P3_SED_WAVES
#define P3_SED_WAVES 4
#endif
constexpr int kSedWaves = P3_SED_WAVES;

// Levels per lane the sedimentation kernel is instantiated with for nlev
// levels on lpc lanes (0 if nlev does not fit).
inline int sed_levels_per_lane(int nlev, int lpc) {
  const int l = (nlev + lpc - 1) / lpc;
  if (l <= 6) return l < 1 ? 1 : l;
  return l <= kMaxLevelsPerLane ? kMaxLevelsPerLane : 0;
}

// A second stream for the liquid group's sedimentation kernel: with one, that
// kernel runs beside the ice group's (they touch disjoint fields), forked from
// s after part 2 (event fork) and joined back into s before part 3 (event join).
struct SedFork {
  hipStream_t side;
  hipEvent_t fork, join;
};

// hipErrorInvalidValue for a shape that is not compiled in.
// `between` (optional) is recorded on s after part 2, before sedimentation;
// `fork` (optional) runs the two sedimentation kernels concurrently.
hipError_t launch_p3_fp64(const KParams& p, const Shape& sh, hipStream_t s, hipEvent_t between = nullptr,
                          const SedFork* fork = nullptr);
hipError_t launch_p3_mixed(const KParams& p, const Shape& sh, hipStream_t s, hipEvent_t between = nullptr,
                           const SedFork* fork = nullptr);

// ---- batched form (p3_batch.hpp)
// Sets whose sedimentation kernel uses the same levels-per-lane count.
struct SedClass {
  int L;
  const KParams* sets;  // device array
  const int* first;     // device array, nsets + 1 entries
  int nsets;
  int grid;             // blocks (one per virtual block, or a persistent grid)
};
struct BatchPlan {
  Shape shape;
  const KParams* sets;  // device array, all sets (part 1 and part 2 launches)
  const int* first;     // device array, nsets + 1 entries: part 1 blocks
  const int* first2;    // device array, nsets + 1 entries: pointwise tiles (parts 1 to 3)
  int nsets;
  int grid, grid2;      // blocks of the part 1 block kernel and of the pointwise launches
  int total2;           // pointwise virtual blocks (first2[nsets]); grid2 >= total2: one block each, no loop
  int nclass;
  SedClass cls[kMaxLevelsPerLane];
};
hipError_t launch_p3_batch_fp64(const BatchPlan& b, hipStream_t s);
hipError_t launch_p3_batch_mixed(const BatchPlan& b, hipStream_t s);
// Resident blocks per CU of the batched kernels for this shape (occupancy API).
hipError_t occupancy_p3_fp64(const Shape& sh, int L, int* blocks1, int* blocks2, int* blocks_sed);
hipError_t occupancy_p3_mixed(const Shape& sh, int L, int* blocks1, int* blocks2, int* blocks_sed);

}  // namespace p3d

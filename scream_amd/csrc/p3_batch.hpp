// p3_batch.hpp -- kernels that run one P3 step of many independent column
// sets (each with its own ncol, nlev, dt and flags) in one launch each.
//
// A set's work is cut into the same blocks its single-set launch would use
// ("virtual blocks").  first[s] is the index of set s's first virtual block in
// the concatenation of all sets (first[nsets] = total).  The grid has one block
// per virtual block, or (a persistent grid) gridDim.x blocks walk the virtual
// blocks round-robin; a block finds the set of its virtual block by binary
// search over first[].
//
// The set descriptors are read through the constant address space: the
// compiler then loads them with scalar loads and knows that the field pointers
// inside them address global memory (global rather than flat loads).
#pragma once

#include "p3_sed.hpp"

namespace p3d {

// The virtual blocks of this block: blockIdx.x, blockIdx.x + gridDim.x, ...
// (PERSIST), or blockIdx.x alone on a grid of one block per virtual block.
// Without the loop the compiler keeps nothing live from one virtual block to
// the next: the pointwise kernels then hold 60-100 fewer VGPRs (FP64 part 2:
// 20 spilled VGPRs instead of 102).
template <bool PERSIST, class F>
__device__ __forceinline__ void for_virtual_blocks(int total, F&& body) {
  if constexpr (PERSIST) {
    for (int vb = blockIdx.x; vb < total; vb += gridDim.x) body(vb);
  } else {
    if ((int)blockIdx.x < total) body((int)blockIdx.x);
  }
}

__device__ __forceinline__ int find_set(const int* __restrict__ first, int nsets, int vb) {
  int lo = 0, hi = nsets - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= vb) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <class R, int CB, int NT>
__global__ __launch_bounds__(NT) void p3_part1_batch(const KParams* __restrict__ sets,
                                                     const int* __restrict__ first, int nsets) {
  __shared__ int s_i[NT];
  const auto sets4 = (const __attribute__((address_space(4))) KParams*)sets;
  const int total = first[nsets];
  for (int vb = blockIdx.x; vb < total; vb += gridDim.x) {
    const int s = find_set(first, nsets, vb);
    part1_block<R, CB, NT>(*(const KParams*)(sets4 + s), vb - first[s], s_i);
  }
}

// Zeroes every set's column status words before p3_part1_lvl_batch: block b
// takes sets b, b + gridDim.x, ...
template <class R>
__global__ __launch_bounds__(256) void p3_colstat_clear_batch(const KParams* __restrict__ sets, int nsets) {
  const auto sets4 = (const __attribute__((address_space(4))) KParams*)sets;
  for (int s = blockIdx.x; s < nsets; s += gridDim.x) {
    const KParams& p = *(const KParams*)(sets4 + s);
    for (int i = threadIdx.x; i < p.ncol; i += blockDim.x) p.colstat[i] = 0;
  }
}

// Part 1 one thread per (level, column); first: the pointwise tile offsets.
template <class R, bool PERSIST>
__global__ __launch_bounds__(kTileThreads) void p3_part1_lvl_batch(const KParams* __restrict__ sets,
                                                                   const int* __restrict__ first, int nsets) {
  const auto sets4 = (const __attribute__((address_space(4))) KParams*)sets;
  for_virtual_blocks<PERSIST>(first[nsets], [&](int vb) {
    const int s = find_set(first, nsets, vb);
    part1_tile<R>(*(const KParams*)(sets4 + s), vb - first[s]);
  });
}

// first: the pointwise tile offsets (pw_tiles per set).
template <class R, int WPE, bool PERSIST>
__global__ __launch_bounds__(kPart2Cols * kPart2Levels, WPE) void p3_part2_batch(const KParams* __restrict__ sets,
                                                                                const int* __restrict__ first,
                                                                                int nsets) {
  const auto sets4 = (const __attribute__((address_space(4))) KParams*)sets;
  for_virtual_blocks<PERSIST>(first[nsets], [&](int vb) {
    const int s = find_set(first, nsets, vb);
    part2_tile<R>(*(const KParams*)(sets4 + s), vb - first[s]);
  });
}

// Part 3 one thread per (level, column), after the sedimentation launches;
// first: the pointwise tile offsets.
template <class R, bool PERSIST>
__global__ __launch_bounds__(kTileThreads) void p3_part3_lvl_batch(const KParams* __restrict__ sets,
                                                                  const int* __restrict__ first, int nsets) {
  const auto sets4 = (const __attribute__((address_space(4))) KParams*)sets;
  for_virtual_blocks<PERSIST>(first[nsets], [&](int vb) {
    const int s = find_set(first, nsets, vb);
    part3_tile<R>(*(const KParams*)(sets4 + s), vb - first[s]);
  });
}

// Sedimentation of species group G (SED_LIQ or SED_ICE).  The sets of a batch
// share their context's tables, so the ice group stages the first set's ice
// records in LDS once per block.
template <class R, int LPC, int L, int WPB, int G>
__global__ __launch_bounds__(64 * WPB, (sed_waves_per_simd<R, G, L>())) void p3_sed_batch(const KParams* __restrict__ sets,
                                                         const int* __restrict__ first, int nsets) {
  const auto sets4 = (const __attribute__((address_space(4))) KParams*)sets;
  const int total = first[nsets];
  if constexpr (G == SED_ICE) {
    static_assert(WPB == 4 || !ice_rec_dif<R, L>(), "the LDS budget assumes blocks of 4 wavefronts");
    __shared__ __align__(16) R s_tab[ice_lds_len<R, L>()];
    if ((int)blockIdx.x >= total) return;  // the whole block leaves before the barrier
    stage_ice_table<R, ice_lds_len<R, L>()>(ice_lds_source<R, L>(*(const KParams*)sets4), s_tab);
    for (int vb = blockIdx.x; vb < total; vb += gridDim.x) {
      const int s = find_set(first, nsets, vb);
      sed_block<R, LPC, L, WPB, G>(*(const KParams*)(sets4 + s), vb - first[s], s_tab);
    }
  } else {
    for (int vb = blockIdx.x; vb < total; vb += gridDim.x) {
      const int s = find_set(first, nsets, vb);
      sed_block<R, LPC, L, WPB, G>(*(const KParams*)(sets4 + s), vb - first[s], nullptr);
    }
  }
}

}  // namespace p3d

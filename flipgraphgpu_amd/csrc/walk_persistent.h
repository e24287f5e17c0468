// The persistent loop shared by every walk-family kernel (walk.hip,
// walk_solo.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "lane_common.h"
#include "walk_args.h"

namespace fgk {

// Persistent walk: each workgroup (one wave) claims blocks of at most 32
// walkers from the queue until none is left: one block per wave in the
// lane-pair layout, two in the one-lane layout (lanes 0-31 serve the first,
// lanes 32-63 the second; the second may be missing at the queue's end).
// Blocks never straddle two formats and are queued largest format first, so
// uneven formats share one launch with little tail.  Every wave exits once
// the queue is exhausted.
// Inlined into each kernel so the argument block is read straight from the
// kernel-argument segment (an out-of-line call would copy it to scratch).
template <class Lane, int kMode>
__device__ __attribute__((always_inline)) void walk_persistent(const WalkArgs &A, uint8_t *lds) {
    const int lane = threadIdx.x;
    if (A.args_bytes != (int32_t)sizeof(WalkArgs)) {
        if (lane == 0) atomicOr(A.err, kErrArgs);
        return;
    }
    constexpr int kItems = Lane::kWS / kWalkersPerGroup;
    for (;;) {
        int item = 0;
        if (lane == 0) item = atomicAdd(A.queue, kItems);
        item = __builtin_amdgcn_readfirstlane(item);
        if (item >= A.nitems) break;
        const int mine = item + (kItems > 1 ? lane / kWalkersPerGroup : 0);
        const bool has = mine < A.nitems;
        const int wb = has ? A.item_begin[mine] + lane % kWalkersPerGroup : 0;
        const int we = has ? A.item_end[mine] : 0;
        lane::walk_lane<Lane, kMode>(A, lds, lane, wb, we);
    }
}

// Walk-family kernels: one wave per workgroup.  With 32-bit factor words a
// workgroup's LDS block is at least 18.8 KB (factor planes of 33 rows), so at
// most 8 fit a CU: 2 waves per SIMD.  Telling the compiler so lets it use up
// to 256 VGPRs (3x3x3 Z2 launch 35.9 -> 33.7 ms, integer ring 29.3 -> 25.3
// ms); the grid follows the occupancy the runtime queries, registers included.
#define FGK_WALK_KERNEL __global__ void __launch_bounds__(kThreadsPerGroup) __attribute__((amdgpu_waves_per_eu(1, 2)))
// One-lane layout: a workgroup's LDS block (64 walkers) admits at most 4 per
// CU, one wave per SIMD, so the compiler may use up to 512 VGPRs.
#define FGK_SOLO_KERNEL __global__ void __launch_bounds__(kThreadsPerGroup) __attribute__((amdgpu_waves_per_eu(1, 1)))

}  // namespace fgk

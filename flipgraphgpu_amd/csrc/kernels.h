// Host launchers of the HIP kernels (defined in walk.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "walk_args.h"

namespace fgk {

// ring 0 = Z2, 1 = integer ring with packed ternary words
size_t walk_lds_bytes(int ring, int max_m);
// Sets the kernel's LDS limit and reports how many workgroups fit per CU
// (minimize selects the complexity-minimizer kernel).
hipError_t prepare_walk(int ring, int max_m, bool minimize, int *groups_per_cu);
// Resets a.queue and launches `grid` persistent workgroups.
hipError_t launch_walk(int ring, bool minimize, const WalkArgs &a, int grid, hipStream_t stream);
// Current scheme of walkers dst[0..count) := best scheme of walker src.
hipError_t launch_broadcast_best(uint32_t *cur, int32_t *m, const uint32_t *best, const int32_t *best_m, int walkers,
                                 int max_m, int src, const int32_t *dst, int count, hipStream_t stream);
hipError_t launch_verify(int ring, const VerifyArgs &v, hipStream_t stream);
hipError_t launch_fill(int32_t *p, int32_t v, int n, hipStream_t stream);

}  // namespace fgk

// Host launchers of the HIP kernels (defined in walk.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "walk_args.h"

namespace fgk {

// ring 0 = Z2, 1 = integer ring with packed ternary words
hipError_t launch_walk(int ring, const WalkArgs &a, hipStream_t stream);
hipError_t launch_verify(int ring, const VerifyArgs &v, hipStream_t stream);

}  // namespace fgk

// Host launchers of the HIP kernels (defined in the .hip files).
#pragma once

#include <hip/hip_runtime.h>

#include "walk_args.h"

namespace fgk {

hipError_t launch_walk_z2(const WalkArgs &a, hipStream_t stream);
hipError_t launch_verify_z2(const VerifyArgs &v, hipStream_t stream);

}  // namespace fgk

// Walk kernels of the one-lane class-mask layout (Z2 schemes of at most 32
// terms with factors of up to 16 coefficients): each lane runs one walker
// and both halves of its flips, a workgroup holds 64 walkers.  The lane
// logic is that of the lane-pair kernels in walk.hip (walk_z2_lane.h); only
// the LDS stride and the split of a flip differ.
#include <hip/hip_runtime.h>
#include <stdint.h>

#define FGK_PHASE_ARRAY g_phase_solo
#include "kernels.h"
#include "walk_args.h"
#include "walk_z2_lane.h"
#include "resize_ops.h"
#include "walk_persistent.h"

namespace fgk {
#if defined(FGK_PHASE_CLOCKS)
__device__ unsigned long long g_phase_solo[16];
#endif

FGK_SOLO_KERNEL walk_z2s_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::Z2LaneS, lane::kModeWalk>(A, lds);
}

FGK_SOLO_KERNEL minimize_z2s_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::Z2LaneS, lane::kModeMinimize>(A, lds);
}

FGK_SOLO_KERNEL learn_z2s_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::Z2LaneS, lane::kModeLearn>(A, lds);
}

FGK_SOLO_KERNEL look_z2s_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::Z2LaneS, lane::kModeLook>(A, lds);
}

FGK_SOLO_KERNEL trace_z2s_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::Z2LaneS, lane::kModeTrace>(A, lds);
}

// Kernel of the one-lane layout for a walk-family mode (kernels.h), or null.
const void *solo_walk_kernel(int mode) {
    switch (mode) {
    case lane::kModeWalk: return reinterpret_cast<const void *>(walk_z2s_kernel);
    case lane::kModeMinimize: return reinterpret_cast<const void *>(minimize_z2s_kernel);
    case lane::kModeLearn: return reinterpret_cast<const void *>(learn_z2s_kernel);
    case lane::kModeLook: return reinterpret_cast<const void *>(look_z2s_kernel);
    case lane::kModeTrace: return reinterpret_cast<const void *>(trace_z2s_kernel);
    default: return nullptr;
    }
}

}  // namespace fgk

#if defined(FGK_PHASE_CLOCKS)
// Phase profile build only: the one-lane kernels' per-phase cycle sums.
extern "C" __attribute__((visibility("default"))) int fg_phase_read_solo(unsigned long long *out16) {
    return hipMemcpyFromSymbol(out16, HIP_SYMBOL(fgk::g_phase_solo), 16 * 8) == hipSuccess ? 0 : -1;
}
extern "C" __attribute__((visibility("default"))) int fg_phase_reset_solo(void) {
    unsigned long long z[16] = {0};
    return hipMemcpyToSymbol(HIP_SYMBOL(fgk::g_phase_solo), z, 16 * 8) == hipSuccess ? 0 : -1;
}
#endif

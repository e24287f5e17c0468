// Host launchers of the HIP kernels (defined in walk.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "walk_args.h"

namespace fgk {

// Walk kernel variants: Z2 with 32-bit factor words, the integer ring
// (packed ternary words), Z2 with 16-bit factor words in LDS (all three with
// partner lists), Z2 with 16-bit words and class masks (at most 32 terms),
// and the last with one lane per walker (64 walkers per workgroup).
constexpr int kVarZ2 = 0, kVarZT = 1, kVarZ2h = 2, kVarZ2m = 3, kVarZ2s = 4;
// Class-mask layouts (kVarZ2m, kVarZ2s).
inline bool mask_variant(int variant) { return variant == kVarZ2m || variant == kVarZ2s; }
// The one-lane layout's kernel for a walk-family mode (walk_solo.hip), or null.
const void *solo_walk_kernel(int mode);

size_t walk_lds_bytes(int variant, int max_m);
// Sets the kernel's LDS limit and reports how many workgroups fit per CU.
// mode: 0 walk, 1 complexity minimizer, 2 walk with proposals and logging (Z2 only),
// 3 the additions reducer's scheme flips (integer ring only), 4 resize, 5 walk with sandwiching
// and basis changes (4 and 5: Z2 with 32-bit words or the integer ring), 6 walk with lookahead
// proposals (Z2), 7 walk that records its operators (every layout).
hipError_t prepare_walk(int variant, int max_m, int mode, int *groups_per_cu);
// Resets a.queue and launches `grid` persistent workgroups.
hipError_t launch_walk(int variant, int mode, const WalkArgs &a, int grid, hipStream_t stream);
// Current scheme of walkers dst[0..count) := best scheme of walker src.
hipError_t launch_broadcast_best(uint32_t *cur, int32_t *m, const uint32_t *best, const int32_t *best_m, int walkers,
                                 int max_m, int src, const int32_t *dst, int count, hipStream_t stream);
// Best scheme, rank and format of walkers list[0..count) := their current ones.
hipError_t launch_snapshot_best(const uint32_t *cur, uint32_t *best, const int32_t *m, int32_t *best_m,
                                const int32_t *fmt, int32_t *bfmt, int walkers, int max_m, const int32_t *list,
                                int count, hipStream_t stream);
hipError_t launch_verify(int ring, const VerifyArgs &v, hipStream_t stream);
hipError_t launch_fill(int32_t *p, int32_t v, int n, hipStream_t stream);
// Labels one logged batch of the trajectory ring into the log rows.
hipError_t launch_log_label(const uint32_t *rwords, const int32_t *rmeta, const int32_t *rminm, int horizon, int W,
                            int every, int nslot, int maxm, uint32_t *words, int32_t *m, int32_t *pairs, float *y,
                            hipStream_t stream);

}  // namespace fgk

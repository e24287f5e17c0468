// Views of a walker pool (fg_ctx) for other parts of the library: the
// learned scorer reads the pool's schemes and trajectory log and writes its
// proposals.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "fg.h"

namespace fgrt {

struct PoolView {
    int ring, W, maxm;
    const uint32_t *cur;   // cur[(c * maxm + t) * W + w]
    const int32_t *m;      // [W]
    const int32_t *fmt;    // [W][4]
    int32_t *forced;       // [W][4] proposals read by the next forced launch
    hipStream_t stream;
    uint32_t key0, key1;
};

struct LogView {
    const uint32_t *words;  // [rows][3][maxm]
    const int32_t *m;       // [rows]
    const int32_t *pairs;   // [rows][3]
    const float *y;         // [rows] 1 rank fell within the horizon, 0 not, -1 no flip logged
    const std::vector<int64_t> *valid;  // rows holding a labelled flip
    const std::vector<int32_t> *vpairs; // their pairs (3 per row), host copy
    const std::vector<float> *vy;       // their labels, host copy
    int maxm;
};

bool pool_view(fg_ctx *ctx, PoolView *v);
bool log_view(fg_ctx *ctx, LogView *v);
// One launch whose first iteration flips ctx's proposed pairs (mode 2).
int walk_forced(fg_ctx *ctx, int iterations);
void set_error(const std::string &msg);

}  // namespace fgrt

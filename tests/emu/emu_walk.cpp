// Host emulation of the walk kernels for the CPU test suite: runs the
// kernel's own per-lane code (flipgraphgpu_amd/csrc/walk_z2_lane.h) one lane
// at a time over a host LDS buffer.  Lanes of the kernel never communicate,
// so this reproduces the GPU result exactly.  Test infrastructure only.
#include <cstring>
#include <vector>

#include "fg.h"
#include "walk_args.h"
#include "walk_z2_lane.h"
#include "walk_zt_lane.h"
#include "resize_ops.h"

static uint32_t pack(int ring, uint64_t v, uint64_t s) {
    return ring ? (uint32_t)((v & 0xFFFFu) | ((s & v & 0xFFFFu) << 16)) : (uint32_t)v;
}

extern "C" int emu_walk_records(int ring, fg_walker_record *recs, int count, uint64_t seed, int iterations,
                                   int max_iterations, int random_iterations, int plus_iterations,
                                   int use_reduce, uint32_t thr_reduce, int use_expand, uint32_t thr_expand,
                                   uint32_t thr_best_eq, int check_reduce, int launches, int visit_cap,
                                   int mode, const int32_t *forced, int32_t *log_meta, int32_t *minm,
                                   int force_wide, uint64_t thr_resize, int resize_max_rank, int resize_max_coef,
                                   int use_sandwich, uint32_t thr_sandwich, int use_basis, uint32_t thr_basis,
                                   int look_period, int look_candidates, int32_t *trace, int32_t *trace_n,
                                   int trace_cap, int32_t *err_out) {
    int maxm = 1;
    for (int q = 0; q < count; ++q) {
        maxm = std::max(maxm, std::max(recs[q].m, recs[q].best_m));
        maxm = std::max(maxm, recs[q].n[0] * recs[q].n[1] * recs[q].n[2]);
    }
    if (mode == fgk::lane::kModeResize) maxm = std::max(maxm, resize_max_rank);  // room for grown schemes
    const int W = count;
    const size_t plane = (size_t)3 * maxm * W;
    std::vector<uint32_t> cur(plane), best(plane);
    std::vector<int32_t> m(W), bm(W), br(W), fc(W), ai(W), fmt((size_t)4 * W), bfmt((size_t)4 * W);
    std::vector<uint64_t> it(W), aux(W), tf(W), sm(W);
    std::vector<int32_t> cx(W);
    for (int w = 0; w < W; ++w) {
        const fg_walker_record &r = recs[w];
        for (int c = 0; c < 3; ++c) fmt[4 * w + c] = r.n[c];
        for (int c = 0; c < 3; ++c) bfmt[4 * w + c] = r.best_n[0] > 0 ? r.best_n[c] : r.n[c];
        m[w] = r.m; bm[w] = r.best_m; br[w] = r.best_rank; fc[w] = r.flips_count; ai[w] = r.aux_idx;
        it[w] = r.it; aux[w] = r.aux; tf[w] = r.total_flips; sm[w] = r.sum_m;
        for (int c = 0; c < 3; ++c) {
            for (int t = 0; t < r.m; ++t) cur[((size_t)c * maxm + t) * W + w] = pack(ring, r.cur_v[c][t], r.cur_s[c][t]);
            for (int t = 0; t < r.best_m; ++t) best[((size_t)w * 3 + c) * maxm + t] = pack(ring, r.best_v[c][t], r.best_s[c][t]);
        }
    }
    int32_t err = 0;
    fgk::WalkArgs a{};
    a.cur = cur.data(); a.best = best.data(); a.m = m.data(); a.best_m = bm.data(); a.best_rank = br.data();
    a.flips_count = fc.data(); a.aux_idx = ai.data(); a.it = it.data(); a.aux = aux.data();
    a.total_flips = tf.data(); a.sum_m = sm.data(); a.fmt = fmt.data();
    a.walkers = W; a.max_m = maxm; a.key0 = (uint32_t)seed; a.key1 = (uint32_t)(seed >> 32);
    a.iterations = iterations; a.max_iterations = max_iterations; a.random_iterations = random_iterations;
    a.plus_iterations = plus_iterations; a.use_reduce = use_reduce; a.thr_reduce = thr_reduce;
    a.use_expand = use_expand; a.thr_expand = thr_expand; a.thr_best_eq = thr_best_eq;
    a.check_reduce = check_reduce; a.zt_visit_cap = visit_cap; a.err = &err;
    a.cx = cx.data();
    a.bfmt = bfmt.data();
    a.thr_resize = thr_resize;
    a.resize_max_rank = resize_max_rank;
    a.resize_max_coef = resize_max_coef;
    a.use_sandwich = use_sandwich;
    a.thr_sandwich = thr_sandwich;
    a.use_basis = use_basis;
    a.thr_basis = thr_basis;
    a.look_period = look_period;
    a.look_candidates = look_candidates;
    a.trace = trace;  // the trace mode (7) writes [W][trace_cap][8] operators and their counts
    a.trace_n = trace_n;
    a.trace_cap = trace_cap;
    std::vector<uint32_t> log_words((size_t)W * 3 * maxm);
    std::vector<int32_t> lm((size_t)W * 4, -1), mm(W, 0);
    if (mode == fgk::lane::kModeLearn) {
        a.forced = forced;
        a.log_every = 1;
        a.log_words = log_words.data();
        a.log_meta = lm.data();
        a.minm = mm.data();
    }
    // same word choice as the runtime: Z2 factors of up to 16 coefficients use 16-bit LDS words
    int nn_max = 0;
    for (int q = 0; q < count; ++q)
        for (int c = 0; c < 3; ++c) nn_max = std::max(nn_max, recs[q].n[c] * recs[q].n[(c + 1) % 3]);
    // force_wide: 0 the runtime's automatic choice, 1 32-bit words with partner
    // lists, 2 16-bit words with partner lists, 3 class masks, 4 class masks
    // in the one-lane layout (packed rows, 64 walkers per workgroup)
    const bool small = !ring && nn_max <= 16 && mode != fgk::lane::kModeResize && mode != fgk::lane::kModeMeta;
    const bool solo = small && maxm <= fgk::kMaskMaxTerms && force_wide == 4;
    const bool masks = small && maxm <= fgk::kMaskMaxTerms && (force_wide == 0 || force_wide == 3);
    const bool half = small && !masks && !solo && force_wide != 1;
    std::vector<uint8_t> lds(ring ? fgk::zt_lds_bytes(maxm)
                                  : (solo ? fgk::z2s_lds_bytes(maxm)
                                          : (masks ? fgk::z2m_lds_bytes(maxm) : fgk::z2_lds_bytes(maxm, half ? 2 : 4))),
                             0xCD);
    const int WPG = solo ? fgk::kSoloWalkersPerGroup : fgk::kWalkersPerGroup;
    const int G = (W + WPG - 1) / WPG;
    for (int l = 0; l < launches; ++l) {
        if (l > 0) a.forced = nullptr;  // proposals apply to the first launch
        for (int g = 0; g < G; ++g)
            for (int slot = 0; slot < WPG; ++slot) {
                // one call per walker: the host side of the lane code does both flip halves
                const int lane = slot;  // the side-0 lane of the slot's pair (the slot's lane when solo)
                const int w = g * WPG + slot;
                using namespace fgk::lane;
                uint8_t *L = lds.data();
                if (solo) {
                    if (mode == kModeMinimize) walk_lane<Z2LaneS, kModeMinimize>(a, L, lane, w, W);
                    else if (mode == kModeLearn) walk_lane<Z2LaneS, kModeLearn>(a, L, lane, w, W);
                    else if (mode == kModeTrace) walk_lane<Z2LaneS, kModeTrace>(a, L, lane, w, W);
                    else if (mode == kModeLook) walk_lane<Z2LaneS, kModeLook>(a, L, lane, w, W);
                    else walk_lane<Z2LaneS, kModeWalk>(a, L, lane, w, W);
                } else if (ring) {
                    if (mode == kModeMinimize) walk_lane<ZTLane, kModeMinimize>(a, L, lane, w, W);
                    else if (mode == kModeFlips) walk_lane<ZTLane, kModeFlips>(a, L, lane, w, W);
                    else if (mode == kModeResize) walk_lane<ZTLane, kModeResize>(a, L, lane, w, W);
                    else if (mode == kModeMeta) walk_lane<ZTLane, kModeMeta>(a, L, lane, w, W);
                    else if (mode == kModeTrace) walk_lane<ZTLane, kModeTrace>(a, L, lane, w, W);
                    else walk_lane<ZTLane, kModeWalk>(a, L, lane, w, W);
                } else if (mode == kModeResize) {
                    walk_lane<Z2Lane, kModeResize>(a, L, lane, w, W);
                } else if (mode == kModeMeta) {
                    walk_lane<Z2Lane, kModeMeta>(a, L, lane, w, W);
                } else if (mode == kModeLook) {
                    if (masks) walk_lane<Z2LaneM, kModeLook>(a, L, lane, w, W);
                    else if (half) walk_lane<Z2Lane16, kModeLook>(a, L, lane, w, W);
                    else walk_lane<Z2Lane, kModeLook>(a, L, lane, w, W);
                } else if (masks) {
                    if (mode == kModeMinimize) walk_lane<Z2LaneM, kModeMinimize>(a, L, lane, w, W);
                    else if (mode == kModeLearn) walk_lane<Z2LaneM, kModeLearn>(a, L, lane, w, W);
                    else if (mode == kModeTrace) walk_lane<Z2LaneM, kModeTrace>(a, L, lane, w, W);
                    else walk_lane<Z2LaneM, kModeWalk>(a, L, lane, w, W);
                } else if (half) {
                    if (mode == kModeMinimize) walk_lane<Z2Lane16, kModeMinimize>(a, L, lane, w, W);
                    else if (mode == kModeLearn) walk_lane<Z2Lane16, kModeLearn>(a, L, lane, w, W);
                    else if (mode == kModeTrace) walk_lane<Z2Lane16, kModeTrace>(a, L, lane, w, W);
                    else walk_lane<Z2Lane16, kModeWalk>(a, L, lane, w, W);
                } else {
                    if (mode == kModeMinimize) walk_lane<Z2Lane, kModeMinimize>(a, L, lane, w, W);
                    else if (mode == kModeLearn) walk_lane<Z2Lane, kModeLearn>(a, L, lane, w, W);
                    else if (mode == kModeTrace) walk_lane<Z2Lane, kModeTrace>(a, L, lane, w, W);
                    else walk_lane<Z2Lane, kModeWalk>(a, L, lane, w, W);
                }
            }
    }
    if (log_meta) std::copy(lm.begin(), lm.end(), log_meta);
    if (minm) std::copy(mm.begin(), mm.end(), minm);
    for (int w = 0; w < W; ++w) {
        fg_walker_record &r = recs[w];
        std::memset(r.cur_v, 0, sizeof r.cur_v);
        std::memset(r.best_v, 0, sizeof r.best_v);
        std::memset(r.cur_s, 0, sizeof r.cur_s);
        std::memset(r.best_s, 0, sizeof r.best_s);
        for (int c = 0; c < 3; ++c) r.n[c] = fmt[4 * w + c];
        for (int c = 0; c < 3; ++c) r.best_n[c] = bfmt[4 * w + c];
        r.m = m[w]; r.best_m = bm[w]; r.best_rank = br[w]; r.flips_count = fc[w]; r.aux_idx = ai[w];
        r.it = it[w]; r.aux = aux[w]; r.total_flips = tf[w]; r.sum_m = sm[w];
        for (int c = 0; c < 3; ++c) {
            for (int t = 0; t < r.m; ++t) {
                const uint32_t x = cur[((size_t)c * maxm + t) * W + w];
                r.cur_v[c][t] = ring ? (x & 0xFFFFu) : x;
                r.cur_s[c][t] = ring ? (x >> 16) : 0;
            }
            for (int t = 0; t < r.best_m; ++t) {
                const uint32_t x = best[((size_t)w * 3 + c) * maxm + t];
                r.best_v[c][t] = ring ? (x & 0xFFFFu) : x;
                r.best_s[c][t] = ring ? (x >> 16) : 0;
            }
        }
    }
    *err_out = err;
    return 0;
}

// Kernel argument block shared by the host runtime and the HIP walk kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fgk {

// One wave64 per workgroup; walker slot s is served by the lane pair (s, s + 32)
// (lane s handles the index1 side of a flip, lane s + 32 the index2 side), so
// a workgroup holds 32 walkers.
constexpr int kThreadsPerGroup = 64;
constexpr int kLanesPerWalker = 2;
constexpr int kWalkersPerGroup = kThreadsPerGroup / kLanesPerWalker;
// The one-lane layout (class masks only): every lane runs a walker of its
// own, so a workgroup holds 64 walkers and takes two queue items at a time.
constexpr int kSoloWalkersPerGroup = kThreadsPerGroup;

// Device-resident walker pool.  Factor planes are structure-of-arrays,
// word(c, t, w) = words[(c * max_m + t) * walkers + w], so a wave touching
// term t of 64 consecutive walkers reads 64 consecutive words.
struct WalkArgs {
    // The first two fields keep their offsets whatever is added below: the
    // kernel compares args_bytes with its own sizeof(WalkArgs) and, on a
    // mismatch (a stale object of a different layout), only raises
    // kErrArgs through err and returns.
    int32_t args_bytes;
    int32_t *err;         // invariant violations (bit flags), host checks after each launch
    uint32_t *cur;        // Z2 factor masks / ZT coefficient magnitudes
    uint32_t *cur_s;      // ZT signs (null for Z2)
    uint32_t *best;       // best schemes, per walker [c][t]: best[(w * 3 + c) * max_m + t]
    uint32_t *best_s;
    int32_t *m;
    int32_t *best_m;
    int32_t *best_rank;
    int32_t *flips_count;
    int32_t *aux_idx;
    uint64_t *it;
    uint64_t *aux;
    uint64_t *total_flips;
    uint64_t *sum_m;
    int32_t *fmt;         // [walkers][4]: n1 n2 n3 unused (written by resize launches)
    int32_t walkers;
    int32_t max_m;
    uint32_t key0;
    uint32_t key1;
    int32_t iterations;
    int32_t max_iterations;
    int32_t random_iterations;
    int32_t plus_iterations;
    int32_t use_reduce;
    uint32_t thr_reduce;
    int32_t use_expand;
    uint32_t thr_expand;
    uint32_t thr_best_eq;
    int32_t check_reduce;
    int32_t zt_visit_cap;
    // persistent launch: workgroups take blocks [item_begin[q], item_end[q])
    // of at most kWalkersPerGroup walkers from a queue until it runs dry
    const int32_t *item_begin;
    const int32_t *item_end;
    int32_t nitems;
    int32_t *queue;       // next unclaimed item; reset to 0 before each launch
    int32_t *cx;          // [walkers] complexity of the current scheme after the launch (may be null)
    // learned-proposal launches (Z2 only)
    const int32_t *forced;  // [walkers][4]: component, index1, index2, 1 of a proposed first flip (component -1: none)
    int32_t log_every;      // walkers w with w % log_every == 0 log their first flip (0: none)
    uint32_t *log_words;    // [walkers / log_every][3][max_m] scheme before that flip
    int32_t *log_meta;      // [walkers / log_every][4]: component (-1: no flip), index1, index2, rank
    int32_t *minm;          // [walkers] lowest rank during the launch (or its first min_window iterations)
    int32_t min_window;     // iterations whose ranks count toward minm (0: the whole launch)
    // lookahead launches (Z2): every look_period-th iteration of a launch first
    // tries look_candidates extra pairs and applies the first that reduces
    int32_t look_period;
    int32_t look_candidates;
    // resize launches (Z2, resizeKernel)
    int32_t *bfmt;            // [walkers][4] format of each walker's best scheme (null: same as fmt)
    uint64_t thr_resize;      // a walker resizes when its first aux draw is below this (at most 2^32)
    int32_t resize_max_rank;  // rank limit of a resized scheme (the reference's MAX_RANK)
    int32_t resize_max_coef;  // coefficients per factor (the reference's MAX_MATRIX_ELEMENTS)
    // basis-change edges of the walk (Z2, meta-walk kernel only)
    int32_t use_sandwich;
    uint32_t thr_sandwich;
    int32_t use_basis;
    uint32_t thr_basis;
    // flip-trace launches: walker w writes its operators to
    // trace[w][trace_cap][8] and their count (stored or not) to trace_n[w]
    int32_t *trace;
    int32_t *trace_n;
    int32_t trace_cap;
};

// Error bits written by the walk kernels.
constexpr int32_t kErrSelect = 1;   // pair draw found no term (partner counts inconsistent)
constexpr int32_t kErrWalk = 2;     // class list walk exceeded the term count
constexpr int32_t kErrCapacity = 4; // a term would exceed the LDS capacity
constexpr int32_t kErrArgs = 8;     // the argument block's layout differs from the kernel's
constexpr int32_t kErrCommit = 16;  // host check: an incremental class update differs from a rebuild

// Padding of the packed H / N planes: a multiple of 32 terms (8 words of 4
// terms) so the selection scan reads whole 8-word blocks.
constexpr int kHTermsPad = 32;
__host__ __device__ inline int h_words(int max_m) { return ((max_m + kHTermsPad - 1) / kHTermsPad) * (kHTermsPad / 4); }

// Rows of a factor plane compared per batch by the class-member scan.
constexpr int kScanRows = 32;

// LDS term stride of the factor planes F.  It holds the term capacity
// rounded up to whole scan batches, so a scan never reads past its own
// component's rows (the rows from the rank up are kept zero).  The two lanes
// of a walker scan two different components j != k in step; padding the
// stride makes their rows fall in different LDS bank groups: with 16-bit
// words a row of 32 walkers covers a quarter of the 64 banks, so the stride
// is made 1 mod 4 (component offsets 1, 2 or 3 quarters apart); with 32-bit
// words a row covers half the banks and the stride is made odd.
__host__ __device__ inline int lds_term_stride(int max_m, int word_bytes) {
    int fs = (max_m + kScanRows - 1) / kScanRows * kScanRows;
    if (word_bytes == 2) {
        while (fs % 4 != 1) ++fs;
    } else if (fs % 2 == 0) {
        ++fs;
    }
    return fs;
}

// LDS bytes per workgroup of the factor planes F (word_bytes per factor,
// wpg walkers per workgroup, one word column per walker).
__host__ __device__ inline size_t walk_lds_f_bytes(int max_m, int word_bytes, int wpg = kWalkersPerGroup) {
    return ((size_t)wpg * 3 * lds_term_stride(max_m, word_bytes) * word_bytes + 3) & ~(size_t)3;
}

// LDS bytes per workgroup of the common planes: factor words F and the
// packed H and N words, for every lane.
__host__ __device__ inline size_t walk_lds_base_bytes(int max_m, int word_bytes = 4) {
    return walk_lds_f_bytes(max_m, word_bytes) + (size_t)kWalkersPerGroup * 2 * 3 * (size_t)h_words(max_m) * 4;
}
__host__ __device__ inline size_t z2_lds_bytes(int max_m, int word_bytes = 4) { return walk_lds_base_bytes(max_m, word_bytes); }

// Class-mask layout (Z2, at most kMaskMaxTerms terms, 16-bit factor words):
// F, H, then CM[c][t] (32-bit class masks).  The CM term stride is odd, so
// the rows of two components that the lanes of a pair touch in step fall in
// different halves of the banks.
constexpr int kMaskMaxTerms = 32;
__host__ __device__ inline int lds_mask_stride(int max_m) { return max_m | 1; }
__host__ __device__ inline size_t z2m_lds_bytes(int max_m) {
    return walk_lds_f_bytes(max_m, 2) + (size_t)kWalkersPerGroup * 3 * (size_t)h_words(max_m) * 4 +
           (size_t)kWalkersPerGroup * 3 * (size_t)lds_mask_stride(max_m) * 4;
}
// One-lane class-mask layout (64 walkers): F holds 32 rows of 16-bit words
// per component packed two per dword (one dword column per walker), then H
// and CM as above.
__host__ __device__ inline size_t z2s_lds_f_bytes() { return (size_t)kSoloWalkersPerGroup * 3 * 16 * 4; }
__host__ __device__ inline size_t z2s_lds_bytes(int max_m) {
    return z2s_lds_f_bytes() + (size_t)kSoloWalkersPerGroup * 3 * (size_t)h_words(max_m) * 4 +
           (size_t)kSoloWalkersPerGroup * 3 * (size_t)lds_mask_stride(max_m) * 4;
}

// Integer ring: the candidate pairs one tryFlip visits one at a time
// (before its exhaustive pass) are remembered in LDS (16-bit positions,
// kZtMaxVisits per walker, kept ascending).
constexpr int kZtMaxVisits = 16;
__host__ __device__ inline size_t zt_lds_bytes(int max_m) {
    return walk_lds_base_bytes(max_m, 4) + (size_t)kWalkersPerGroup * kZtMaxVisits * 2;
}

// Brent check of schemes held in the same plane layout.
struct VerifyArgs {
    const uint32_t *words;
    const uint32_t *signs;  // null for Z2
    const int32_t *m;
    const int32_t *fmt;     // [count][4]
    int32_t count;
    int32_t max_m;
    int32_t *ok;            // [count], host sets 1, kernel clears on failure
};

}  // namespace fgk

// Kernel argument block and LDS budget of the batched additions reducer
// (src/entities/additions_reducer.cuh, src/entities/scheme_additions_reducer.cu).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fgk {

// Per-component capacities of the reference configuration (src/config.cuh:17-31).
constexpr int kReducerMaxFresh[3] = {250, 250, 500};
constexpr int kReducerMaxPairs[3] = {64 * 63 / 2, 64 * 63 / 2, 350 * 349 / 2};

// One workgroup (one wave) runs one component (U, V or W) of one reducer.
struct ReduceArgs {
    const int8_t *coef;      // [schemes][3][max_m][max_nn] coefficients of the U, V, W factors per term
    const int32_t *sm;       // [schemes] rank of each scheme
    int32_t schemes;
    int32_t n1, n2, n3;
    int32_t max_m, max_nn;
    int32_t count;           // reducers
    uint32_t key0, key1;     // Philox key (seed)
    int32_t independent;     // 1: per-component best replay draws, 0: one shared decision
    uint64_t decision_word;  // word index of the shared decision stream
    const int32_t *best_fresh;  // [3] fresh variables of the best reducer per component
    const int16_t *best_vars;   // [3][f_cap][2]
    uint64_t *words;            // [count][4] words consumed per stream (components 0-2 advanced here)
    int32_t *info;   // [count * 3][8]: expressions, real, fresh, naive, mode, scale bits, additions, 0
    int16_t *size;   // [count * 3][e_cap]
    int16_t *expr;   // [count * 3][e_cap][l_cap]
    int16_t *vars;   // [count * 3][f_cap][2]
    int32_t e_cap, l_cap, f_cap, p_cap, h_bits;
    int32_t *err;
};

constexpr int32_t kRedErrCapacity = 1;  // an expression, pair list or hash table would overflow
constexpr int32_t kRedErrInput = 2;     // a coefficient outside {-1, 0, 1}

__host__ __device__ inline size_t reducer_align(size_t x) { return (x + 15) & ~(size_t)15; }

// LDS bytes of one workgroup; reducer.hip carves the block in this order.
__host__ __device__ inline size_t reducer_lds_bytes(int e_cap, int l_cap, int f_cap, int p_cap, int h_bits) {
    const size_t H = (size_t)1 << h_bits, pw = ((size_t)p_cap + 31) / 32;
    size_t b = 0;
    b += reducer_align((size_t)e_cap * l_cap * 2);       // expressions
    b += reducer_align((size_t)e_cap * 4);               // sizes
    b += reducer_align(((size_t)e_cap + 1) * 4);         // pair offsets
    b += reducer_align(H * 4) * 3;                       // hash keys, counts, first occurrences
    b += reducer_align((size_t)p_cap * 4) * 5;           // used slots, ordered keys, counts, chain, final list
    b += reducer_align(pw * 4) * 2;                      // first-occurrence bitmap and its prefix
    b += reducer_align((size_t)f_cap * 4);               // fresh variables
    b += 64;                                             // counters
    return b;
}

}  // namespace fgk

// Batched additions reducer for CDNA4 (gfx950).  Each workgroup is one wave
// running the greedy common-subexpression search of one component (U rows,
// V rows or W columns) of one reducer (AdditionsReducer::reduce,
// src/entities/additions_reducer.cuh:89-105) with its state in LDS.
//
// One step of the reference (updateSubexpressions + sort + select +
// replaceSubexpression) is done by the 64 lanes together:
//  - every (expression, a < b) occurrence is inserted into an LDS hash table
//    with atomics, keeping the count and the first occurrence of each pair;
//  - a bitmap over occurrence indices ranks the pairs by first occurrence,
//    which restores the reference's insertion order (pairs_counter.cuh:55-66);
//  - the reference's swap partition (pairs_counter.cuh:68-101) is evaluated
//    in closed form: each swap moves the front of the window of non-top
//    pairs to its back, so the final order is a queue with rotations whose
//    entries are resolved by following rotation links;
//  - selection (pairs_counter.cuh:130-196) runs lane-parallel, including the
//    intersections score whose random draws are assigned by prefix counts;
//  - the replacement runs one expression per lane.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lane_common.h"
#include "reducer_args.h"

// The intersections score rounds every float and double operation on its own
// as the reference does; a fused multiply-add would move near-tied scores by
// an ulp and change the chosen pair.  The Makefile also builds this file with
// -ffp-contract=off, which covers the HIP header helpers inlined here.
#pragma clang fp contract(off)

namespace fgk {
namespace red {

enum { kGreedy = 0, kGreedyAlt = 1, kGreedyRandom = 2, kGreedyInter = 3, kWeighted = 4, kMix = 5, kRandom = 6 };

// modes of runReducersKernel (src/entities/scheme_additions_reducer.cu:495-503)
__constant__ int kModes[15] = {3, 3, 3, 3, 3, 3, 3, 3, 1, 1, 1, 1, 2, 2, 5};

__device__ __forceinline__ uint32_t stream_word(uint32_t k0, uint32_t k1, uint32_t r, uint32_t sid, uint64_t w) {
    uint32_t o0, o1, o2, o3;
    lane::philox10((uint32_t)(w >> 2), (uint32_t)(w >> 34), r, sid, k0, k1, o0, o1, o2, o3);
    const uint32_t q = (uint32_t)w & 3u;
    return q == 0 ? o0 : (q == 1 ? o1 : (q == 2 ? o2 : o3));
}

__device__ __forceinline__ float uniform01(uint32_t x) { return __fmul_rn((float)((x >> 8) + 1u), 1.0f / 16777216.0f); }

__device__ __forceinline__ int wave_excl_sum(int v, int lane, int &total) {
    int x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    total = __shfl(x, 63);
    return x - v;
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int u = __shfl_xor(v, o);
        v = u < v ? u : v;
    }
    return v;
}

__device__ __forceinline__ int iabs(int x) { return x < 0 ? -x : x; }

__device__ __forceinline__ uint32_t pair_key(int i, int j) { return ((uint32_t)i << 12) | (uint32_t)(j + 2048); }
__device__ __forceinline__ int key_i(uint32_t k) { return (int)(k >> 12); }
__device__ __forceinline__ int key_j(uint32_t k) { return (int)(k & 4095u) - 2048; }

// Pair::intersects (pairs_counter.cuh:13-15)
__device__ __forceinline__ bool intersects(uint32_t a, uint32_t b) {
    const int ai = key_i(a), aj = iabs(key_j(a)), bi = key_i(b), bj = iabs(key_j(b));
    return ai == bi || ai == bj || aj == bi || aj == bj;
}

// binarySearch (additions_reducer.cuh:241-262) over |values|
__device__ __forceinline__ int bsearch_abs(const int16_t *e, int size, int value, int start) {
    int lo = start, hi = size - 1;
    const int v = iabs(value);
    while (lo <= hi) {
        const int mid = lo + (hi - lo) / 2, a = iabs(e[mid]);
        if (a < v) lo = mid + 1;
        else if (a > v) hi = mid - 1;
        else return mid;
    }
    return -1;
}

// start offset of row a in the (a < b) enumeration of an expression of size s
__device__ __forceinline__ int pair_row_start(int a, int s) { return a * (2 * s - a - 1) / 2; }

struct Lds {
    int16_t *E;
    int32_t *sz, *off;
    uint32_t *hkey, *hcnt, *hfirst;
    int32_t *used;
    uint32_t *okey, *ocnt;
    int32_t *sx, *fin;
    uint32_t *bits, *bpre;
    int16_t *vars;
    int32_t *misc;
};

__device__ __forceinline__ Lds carve(uint8_t *lds, const ReduceArgs &A) {
    const size_t H = (size_t)1 << A.h_bits, pw = ((size_t)A.p_cap + 31) / 32;
    Lds L;
    size_t b = 0;
    auto take = [&](size_t bytes) {
        uint8_t *p = lds + b;
        b += reducer_align(bytes);
        return p;
    };
    L.E = reinterpret_cast<int16_t *>(take((size_t)A.e_cap * A.l_cap * 2));
    L.sz = reinterpret_cast<int32_t *>(take((size_t)A.e_cap * 4));
    L.off = reinterpret_cast<int32_t *>(take(((size_t)A.e_cap + 1) * 4));
    L.hkey = reinterpret_cast<uint32_t *>(take(H * 4));
    L.hcnt = reinterpret_cast<uint32_t *>(take(H * 4));
    L.hfirst = reinterpret_cast<uint32_t *>(take(H * 4));
    L.used = reinterpret_cast<int32_t *>(take((size_t)A.p_cap * 4));
    L.okey = reinterpret_cast<uint32_t *>(take((size_t)A.p_cap * 4));
    L.ocnt = reinterpret_cast<uint32_t *>(take((size_t)A.p_cap * 4));
    L.sx = reinterpret_cast<int32_t *>(take((size_t)A.p_cap * 4));
    L.fin = reinterpret_cast<int32_t *>(take((size_t)A.p_cap * 4));
    L.bits = reinterpret_cast<uint32_t *>(take(pw * 4));
    L.bpre = reinterpret_cast<uint32_t *>(take(pw * 4));
    L.vars = reinterpret_cast<int16_t *>(take((size_t)A.f_cap * 4));
    L.misc = reinterpret_cast<int32_t *>(lds + b);
    return L;
}

__global__ void __launch_bounds__(64) reduce_kernel(ReduceArgs A) {
    extern __shared__ uint8_t lds[];
    const int lane = threadIdx.x;
    const int r = blockIdx.x / 3, c = blockIdx.x % 3;
    if (r >= A.count) return;
    Lds L = carve(lds, A);
    const int H = 1 << A.h_bits;
    const uint32_t hmask = (uint32_t)H - 1u;
    const uint64_t lt_mask = lane ? (~0ull >> (64 - lane)) : 0ull;

    // copySchemeToReducers (scheme_additions_reducer.cu:461-483)
    const int s = r % A.schemes;
    const int m = A.sm[s];
    const int nn[3] = {A.n1 * A.n2, A.n2 * A.n3, A.n3 * A.n1};
    const int nexpr = c < 2 ? m : nn[2];
    const int nvars = c < 2 ? nn[c] : m;
    const int max_fresh = kReducerMaxFresh[c] < A.f_cap ? kReducerMaxFresh[c] : A.f_cap;
    const uint32_t max_pairs = (uint32_t)kReducerMaxPairs[c];
    bool bad = nexpr > A.e_cap;
    int naive_part = 0;
    if (!bad) {
        for (int e = lane; e < nexpr; e += 64) {
            int len = 0;
            for (int v = 0; v < nvars; ++v) {
                const int x = c < 2 ? A.coef[(((size_t)s * 3 + c) * A.max_m + e) * A.max_nn + v]
                                    : A.coef[(((size_t)s * 3 + 2) * A.max_m + v) * A.max_nn + e];
                if (x != 0) {
                    if (x != 1 && x != -1) atomicOr(A.err, kRedErrInput);
                    if (len < A.l_cap) L.E[(size_t)e * A.l_cap + len] = (int16_t)(x > 0 ? v + 1 : -(v + 1));
                    ++len;
                }
            }
            if (len > A.l_cap) {
                atomicOr(A.err, kRedErrCapacity);
                len = A.l_cap;
            }
            L.sz[e] = len;
            naive_part += len - 1;
        }
    }
    for (int q = lane; q < H; q += 64) {
        L.hkey[q] = 0u;
        L.hcnt[q] = 0u;
        L.hfirst[q] = 0xFFFFFFFFu;
    }
    for (int q = lane; q < (A.p_cap + 31) / 32; q += 64) L.bits[q] = 0u;
    if (lane == 0) L.misc[0] = 0;
    int naive_tot;
    wave_excl_sum(naive_part, lane, naive_tot);
    if (bad) atomicOr(A.err, kRedErrCapacity);
    __syncthreads();

    const int real = nvars;
    int fresh = 0;
    uint64_t w = A.words[(size_t)r * 4 + c];
    const uint32_t sid = 2u + (uint32_t)c;
    auto next = [&]() -> uint32_t { return stream_word(A.key0, A.key1, (uint32_t)r, sid, w++); };

    // replaceSubexpression (additions_reducer.cuh:217-239), one expression per lane
    auto replace = [&](int pi, int pj) {
        const int var = real + fresh + 1;
        for (int e = lane; e < nexpr; e += 64) {
            const int sze = L.sz[e];
            if (sze < 2) continue;
            int16_t *row = L.E + (size_t)e * A.l_cap;
            const int a = bsearch_abs(row, sze, pi, 0);
            if (a < 0) continue;
            const int b = bsearch_abs(row, sze, pj, a + 1);
            if (b < 0) continue;
            int nv;
            if (row[a] == pi && row[b] == pj) nv = var;
            else if (row[a] == -pi && row[b] == -pj) nv = -var;
            else continue;
            int j = a;
            for (int i = a + 1; i < sze; ++i)
                if (i != b) row[j++] = row[i];
            row[j] = (int16_t)nv;
            L.sz[e] = sze - 1;
        }
        if (lane == 0) {
            L.vars[2 * fresh] = (int16_t)pi;
            L.vars[2 * fresh + 1] = (int16_t)pj;
        }
        ++fresh;
        __syncthreads();
    };

    // setMode + best replay (scheme_additions_reducer.cu:505-528)
    const int mode = bad ? 0 : (r == 0 ? kGreedy : kModes[next() % 15u]);
    const int bf = A.best_fresh[c];
    int replay = 0;
    if (!bad) {
        if (A.independent) {
            const float u = uniform01(next());
            if ((double)u < 0.3 && bf > 0) replay = 1 + (int)(next() % (uint32_t)bf);
        } else {
            const float u = uniform01(stream_word(A.key0, A.key1, (uint32_t)r, 5u, A.decision_word));
            if ((double)u < 0.3 && bf > 0) replay = 1 + (int)(next() % (uint32_t)bf);
        }
    }
    for (int q = 0; q < replay && q < bf; ++q)  // partialInitialize (additions_reducer.cuh:134-142)
        replace(A.best_vars[((size_t)c * A.f_cap + q) * 2], A.best_vars[((size_t)c * A.f_cap + q) * 2 + 1]);

    // reduce (additions_reducer.cuh:89-105)
    const float scale = bad ? 0.0f : uniform01(next()) * 0.5f;
    while (!bad && fresh < max_fresh) {
        // ---- updateSubexpressions: count pairs ----
        int carry = 0;
        for (int base = 0; base < nexpr; base += 64) {
            const int e = base + lane;
            const int sze = e < nexpr ? L.sz[e] : 0;
            int tot;
            const int ex = wave_excl_sum(sze * (sze - 1) / 2, lane, tot);
            if (e < nexpr) L.off[e] = carry + ex;
            carry += tot;
        }
        const int P = carry;
        if (P > A.p_cap) {
            if (lane == 0) atomicOr(A.err, kRedErrCapacity);
            bad = true;
            break;
        }
        __syncthreads();
        for (int o = lane; o < P; o += 64) {
            int lo = 0, hi = nexpr - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (L.off[mid] <= o) lo = mid;
                else hi = mid - 1;
            }
            const int e = lo, sze = L.sz[e], k = o - L.off[e];
            const float t = (float)(2 * sze - 1);
            int a = (int)((t - sqrtf(t * t - 8.0f * (float)k)) * 0.5f);
            a = a < 0 ? 0 : (a > sze - 2 ? sze - 2 : a);
            while (a > 0 && pair_row_start(a, sze) > k) --a;
            while (a < sze - 2 && pair_row_start(a + 1, sze) <= k) ++a;
            const int b = a + 1 + (k - pair_row_start(a, sze));
            int x = L.E[(size_t)e * A.l_cap + a], y = L.E[(size_t)e * A.l_cap + b];
            if (iabs(x) > iabs(y)) {
                const int tt = x;
                x = y;
                y = tt;
            }
            if (x < 0) {
                x = -x;
                y = -y;
            }
            const uint32_t key = pair_key(x, y);
            uint32_t h = (key * 2654435761u) >> (32 - A.h_bits);
            for (int probe = 0; probe < H; ++probe) {
                const uint32_t old = atomicCAS(&L.hkey[h], 0u, key);
                if (old == 0u) {
                    const int u = atomicAdd(&L.misc[0], 1);
                    if (u < A.p_cap) L.used[u] = (int32_t)h;
                }
                if (old == 0u || old == key) {
                    atomicAdd(&L.hcnt[h], 1u);
                    atomicMin(&L.hfirst[h], (uint32_t)o);
                    break;
                }
                h = (h + 1u) & hmask;
            }
        }
        __syncthreads();
        const int D = L.misc[0];
        for (int u = lane; u < D; u += 64) {
            const uint32_t f = L.hfirst[L.used[u]];
            atomicOr(&L.bits[f >> 5], 1u << (f & 31u));
        }
        __syncthreads();
        const int pw = (P + 31) / 32;
        carry = 0;
        for (int base = 0; base < pw; base += 64) {
            const int q = base + lane;
            int tot;
            const int ex = wave_excl_sum(q < pw ? __popc(L.bits[q]) : 0, lane, tot);
            if (q < pw) L.bpre[q] = (uint32_t)(carry + ex);
            carry += tot;
        }
        __syncthreads();
        // insertion order = order of first occurrences; pairs beyond the
        // counter's capacity are dropped as in findOrCreateSlot
        int maxc = 0;
        for (int u = lane; u < D; u += 64) {
            const int h = L.used[u];
            const uint32_t f = L.hfirst[h];
            const uint32_t rank = L.bpre[f >> 5] + __popc(L.bits[f >> 5] & ((1u << (f & 31u)) - 1u));
            if (rank < max_pairs) {
                L.okey[rank] = L.hkey[h];
                L.ocnt[rank] = L.hcnt[h];
                maxc = (int)L.hcnt[h] > maxc ? (int)L.hcnt[h] : maxc;
            }
        }
        maxc = lane::wave_max(maxc);
        const int Dk = D < (int)max_pairs ? D : (int)max_pairs;
        __syncthreads();
        for (int u = lane; u < D; u += 64) {
            const int h = L.used[u];
            L.hkey[h] = 0u;
            L.hcnt[h] = 0u;
            L.hfirst[h] = 0xFFFFFFFFu;
        }
        for (int q = lane; q < pw; q += 64) L.bits[q] = 0u;
        if (lane == 0) L.misc[0] = 0;
        __syncthreads();
        if (Dk == 0 || maxc < 2) break;  // PairsCounter::operator bool

        // ---- sort: the swap partition in closed form ----
        int T = 0, f0 = Dk;
        for (int base = 0; base < Dk; base += 64) {
            const int i = base + lane;
            const bool isM = i < Dk && (int)L.ocnt[i] == maxc;
            const bool isN = i < Dk && !isM;
            const uint64_t mm = __ballot(isM), mn = __ballot(isN);
            if (f0 == Dk && mn) f0 = base + __builtin_ctzll(mn);
            T += __popcll(mm);
        }
        const int t0 = f0 < Dk ? f0 : T;
        const int K = Dk - T, R = K > 0 ? T - t0 : 0;
        int ct = 0, cp = 0;
        for (int base = 0; base < Dk; base += 64) {
            const int i = base + lane;
            const bool isM = i < Dk && (int)L.ocnt[i] == maxc;
            const bool isN = i < Dk && !isM;
            const uint64_t mm = __ballot(isM), mn = __ballot(isN);
            const int ti = ct + __popcll(mm & lt_mask), pi = cp + __popcll(mn & lt_mask);
            if (isM) {
                L.fin[ti] = i;
                if (pi > 0) L.sx[pi + (ti - t0)] = -(ti - t0 + 1);
            } else if (isN) {
                const int rot = ti - t0 > 0 ? ti - t0 : 0;
                L.sx[pi + rot] = i;
            }
            ct += __popcll(mm);
            cp += __popcll(mn);
        }
        __syncthreads();
        int kept = 0;
        for (int base = 0; base < K; base += 64) {
            const int q = base + lane;
            int pos = -1;
            if (q < K) {
                int v = L.sx[R + q];
                for (int guard = 0; v < 0 && guard <= Dk; ++guard) v = L.sx[-v - 1];
                pos = v;
            }
            const bool keep = pos >= 0 && L.ocnt[pos] > 1u;
            const uint64_t mk = __ballot(keep);
            if (keep) L.fin[T + kept + __popcll(mk & lt_mask)] = pos;
            kept += __popcll(mk);
        }
        const int size = T + kept;
        __syncthreads();

        // ---- selectSubexpression (additions_reducer.cuh:264-282) ----
        const int step_mode = mode == kMix ? (int)(next() % (uint32_t)kMix) : mode;
        int sel = 0;
        if (step_mode == kGreedy) {
            sel = 0;
        } else if (step_mode == kGreedyAlt) {
            sel = (int)(next() % (uint32_t)T);
        } else if (step_mode == kGreedyRandom) {
            if (uniform01(next()) < scale) sel = (int)(next() % (uint32_t)size);
            else sel = (int)(next() % (uint32_t)T);
        } else if (step_mode == kGreedyInter) {
            // draws: one per intersecting (i, j), in i-major order
            int best_i = 0x7FFFFFFF;
            float best_s = 0.0f;
            int cbase = 0;
            for (int base = 0; base < size; base += 64) {
                const int i = base + lane;
                int X = 0;
                uint32_t ki = 0;
                if (i < size) {
                    ki = L.okey[L.fin[i]];
                    for (int j = 0; j < size; ++j)
                        if (j != i && intersects(ki, L.okey[L.fin[j]])) ++X;
                }
                int tot;
                const int ex = wave_excl_sum(X, lane, tot);
                if (i < size) {
                    uint64_t wi = w + (uint64_t)(cbase + ex);
                    uint64_t blk = ~0ull;
                    uint32_t o0 = 0, o1 = 0, o2 = 0, o3 = 0;
                    float inter = 0.0f;
                    for (int j = 0; j < size; ++j) {
                        if (j == i) continue;
                        const int pj = L.fin[j];
                        const uint32_t kj = L.okey[pj];
                        const int cj = (int)L.ocnt[pj];
                        if (!intersects(ki, kj)) {
                            inter = __fadd_rn(inter, (float)(cj - 1));
                        } else {
                            if ((wi >> 2) != blk) {
                                blk = wi >> 2;
                                lane::philox10((uint32_t)blk, (uint32_t)(blk >> 32), (uint32_t)r, sid, A.key0, A.key1,
                                               o0, o1, o2, o3);
                            }
                            const uint32_t q = (uint32_t)wi & 3u;
                            const uint32_t x = q == 0 ? o0 : (q == 1 ? o1 : (q == 2 ? o2 : o3));
                            ++wi;
                            if ((x % 2u) == 0u) inter = (float)__dadd_rn((double)inter, __dmul_rn(0.69, (double)cj));
                        }
                    }
                    const float score = __fadd_rn((float)((int)L.ocnt[L.fin[i]] - 1), __fmul_rn(scale, inter));
                    if (score > best_s) {
                        best_s = score;
                        best_i = i;
                    }
                }
                cbase += tot;
            }
            // first index of the largest score above 0, else 0
            for (int o = 32; o > 0; o >>= 1) {
                const float os = __shfl_xor(best_s, o);
                const int oi = __shfl_xor(best_i, o);
                if (os > best_s || (os == best_s && oi < best_i)) {
                    best_s = os;
                    best_i = oi;
                }
            }
            sel = best_s > 0.0f ? best_i : 0;
            w += (uint64_t)cbase;
        } else if (step_mode == kWeighted) {
            int total_cnt = 0;
            for (int i = lane; i < size; i += 64) total_cnt += (int)L.ocnt[L.fin[i]];
            int total;
            wave_excl_sum(total_cnt, lane, total);
            const float x = __fmul_rn(uniform01(next()), (float)total);
            int found = 0x7FFFFFFF, cbase = 0;
            for (int base = 0; base < size; base += 64) {
                const int i = base + lane;
                const int ci = i < size ? (int)L.ocnt[L.fin[i]] : 0;
                int tot;
                const int ex = wave_excl_sum(ci, lane, tot);
                if (i < size && x <= (float)(cbase + ex + ci) && i < found) found = i;
                cbase += tot;
            }
            found = wave_min(found);
            sel = found < size ? found : size - 1;
        } else {
            sel = (int)(next() % (uint32_t)size);
        }
        const uint32_t kk = L.okey[L.fin[sel]];
        replace(key_i(kk), key_j(kk));
    }

    // outputs
    int add_part = 0;
    for (int e = lane; e < nexpr; e += 64) add_part += L.sz[e] - 1;
    int add_tot;
    wave_excl_sum(add_part, lane, add_tot);
    const size_t rc = (size_t)r * 3 + c;
    if (lane == 0) {
        int32_t *inf = A.info + rc * 8;
        inf[0] = bad ? 0 : nexpr;
        inf[1] = real;
        inf[2] = fresh;
        inf[3] = naive_tot;
        inf[4] = mode;
        inf[5] = (int32_t)__float_as_uint(scale);
        inf[6] = fresh + add_tot;
        inf[7] = 0;
        A.words[(size_t)r * 4 + c] = w;
    }
    if (!bad) {
        for (int e = lane; e < nexpr; e += 64) A.size[rc * A.e_cap + e] = (int16_t)L.sz[e];
        for (int q = lane; q < nexpr * A.l_cap; q += 64) {
            const int e = q / A.l_cap, x = q % A.l_cap;
            A.expr[(rc * A.e_cap + e) * A.l_cap + x] = x < L.sz[e] ? L.E[(size_t)e * A.l_cap + x] : (int16_t)0;
        }
        for (int q = lane; q < 2 * fresh; q += 64) A.vars[rc * A.f_cap * 2 + q] = L.vars[q];
    }
}

}  // namespace red

hipError_t launch_reduce(const ReduceArgs &a, hipStream_t stream) {
    const size_t lds = reducer_lds_bytes(a.e_cap, a.l_cap, a.f_cap, a.p_cap, a.h_bits);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(red::reduce_kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(red::reduce_kernel, dim3((unsigned)a.count * 3), dim3(64), lds, stream, a);
    return hipGetLastError();
}

}  // namespace fgk

// Ring-independent per-lane machinery of the walk kernels, shared by the
// HIP kernels and the host emulation used by the tests.  Everything here is
// __host__ __device__; lanes never communicate, so running lanes one after
// another on the host reproduces a kernel exactly.
//
// A lane pair (s, s + 32) owns walker slot s (the reference's randomWalkKernel
// runs one thread per scheme, src/entities/flip_graph.cu:520-573).  The two
// lanes sit in different half-waves, so their reads of the walker's LDS
// column never compete for a bank within a half-wave.  Both lanes of
// the pair run the same instruction stream on the same walker state; the
// only work they split is the two halves of a flip (the index1 side with
// component j, the index2 side with component k), which exchange their
// results with one permlane32 swap.  The walker's scheme lives in LDS for the
// whole launch as three factor planes F[c][t][slot] (one 32-bit word per
// factor: a Z2 mask, or for the integer
// ring the coefficient magnitudes in bits 0-15 and their signs in bits
// 16-31); two byte planes describe the flip partners:
//
//   N[c][t]  next member (cyclic, ascending term index) of the class of terms
//            whose component-c factors are equal to term t's,
//   H[c][t]  number of class members with an index above t.
//
// Both byte planes hold terms 4q..4q+3 of a walker in one dword, dwords
// slot-interleaved, so a wave touching any term of each walker hits 32
// different banks (the two lanes of a pair share an address).  The best scheme found so far is copied to HBM by the
// whole wave at once (one walker per pass, contiguous per walker).
//
// The reference keeps the partner pairs as an explicit list per component
// (FlipSet, src/entities/flip_set.cu) and rebuilds or scans it on every
// flip.  Here the unordered pair (t1 < t2) at position r of the fresh-list
// order (component, t1, t2 -- the order initFlips() produces,
// src/schemes/scheme_z2.cu:141-150) is found by a prefix scan of H followed
// by r' + 1 steps along N, and a flip updates only the two classes it
// touches.  Rare events (reductions, zero factors, plus / split) rebuild
// N and H from scratch.
//
// Randomness: Philox4x32-10 keyed by the seed, counter (iteration, walker,
// stream).  One block per walk iteration supplies the pair draw, the two
// orientation coins, the 1% equal-rank best-copy draw and the expand
// draw (the two lanes of a pair compute the blocks of two consecutive
// iterations and swap them); rare operators use the aux stream.  The CPU oracle consumes the
// same stream in the same order, so a walker's trajectory is bit-identical.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "walk_args.h"

#define FGK_HD __host__ __device__ __forceinline__

namespace fgk {

FGK_HD void fgk_atomic_or(int32_t *p, int32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
    atomicOr(p, v);
#else
    *p |= v;
#endif
}

// Sum of the four bytes of w (v_sad_u8 on the GPU).
FGK_HD uint32_t byte_sum(uint32_t w) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sad_u8(w, 0u, 0u);
#else
    return (w & 0xFFu) + ((w >> 8) & 0xFFu) + ((w >> 16) & 0xFFu) + (w >> 24);
#endif
}
// acc plus the four bytes of w (one v_sad_u8, its accumulator operand).
FGK_HD uint32_t byte_sum_acc(uint32_t w, uint32_t acc) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sad_u8(w, 0u, acc);
#else
    return acc + byte_sum(w);
#endif
}

#if defined(FGK_PHASE_CLOCKS) && defined(__HIP_DEVICE_COMPILE__)
// phase profile build: cycles per phase, [15] iterations (one array per
// kernel source file, named by FGK_PHASE_ARRAY)
#ifndef FGK_PHASE_ARRAY
#define FGK_PHASE_ARRAY g_phase
#endif
extern __device__ unsigned long long FGK_PHASE_ARRAY[16];
#endif
namespace lane {

#if defined(__HIP_DEVICE_COMPILE__)
constexpr bool kOnDevice = true;
#else
constexpr bool kOnDevice = false;
#endif
constexpr uint32_t kPM0 = 0xD2511F53u, kPM1 = 0xCD9E8D57u;
constexpr uint32_t kPW0 = 0x9E3779B9u, kPW1 = 0xBB67AE85u;
constexpr int kMaxRank = 350;        // src/config.cuh:10
constexpr int kPlusMaxTries = 1000;  // tryPlus / trySplitExisted rejection bound

// a ^ b ^ c (GPU: one v_bitop3_b32 with the three-input XOR table 0x96,
// which the compiler leaves as two XORs here).
FGK_HD uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t r;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
#else
    return a ^ b ^ c;
#endif
}

// Per 16-bit half of x: 1 when the half is non-zero, else 0 (one
// v_pk_min_u16 on the GPU; written out, since the compiler turns a packed
// minimum into two compares and selects per dword).
FGK_HD uint32_t nonzero_halves(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t r;
    asm("v_pk_min_u16 %0, %1, %2" : "=v"(r) : "v"(x), "s"(0x00010001u));
    return r;
#else
    return ((x & 0xFFFFu) ? 1u : 0u) | ((x >> 16) ? 0x10000u : 0u);
#endif
}

// (acc << 1) | x as one v_lshl_or_b32 (the compiler otherwise spreads a
// shift-or chain into a shift per element and three-input ors).
FGK_HD uint32_t shl1_or(uint32_t acc, uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t r;
    asm("v_lshl_or_b32 %0, %1, 1, %2" : "=v"(r) : "v"(acc), "v"(x));
    return r;
#else
    return (acc << 1) | x;
#endif
}

// Rows of a packed component (dword p: row 15 - p low, row 31 - p high)
// that differ from v: bit t set when row t != v.  Shifting the two half
// flags of dword p in at bits 0 and 16 leaves row 15 - p at bit 15 - p and
// row 31 - p at bit 31 - p after the 16 dwords.
FGK_HD uint32_t packed_ne(const uint32_t (&R)[16], uint32_t v) {
    const uint32_t vv = v | (v << 16);
    uint32_t acc = nonzero_halves(R[0] ^ vv);
#pragma unroll
    for (int p = 1; p < 16; ++p) acc = shl1_or(acc, nonzero_halves(R[p] ^ vv));
    return acc;
}

// Philox4x32-10 computed a few rounds at a time: start() sets the counter
// and key, rounds<R0, R1>() applies rounds R0..R1-1 (round r keyed by
// key + r * (kPW0, kPW1)); after rounds 0..9 x0..x3 equal philox10's output.
struct PhiloxRounds {
    uint32_t x0, x1, x2, x3, k0, k1;
    FGK_HD void start(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t key0, uint32_t key1) {
        x0 = c0;
        x1 = c1;
        x2 = c2;
        x3 = c3;
        k0 = key0;
        k1 = key1;
    }
    template <int R0, int R1>
    FGK_HD void rounds() {
#pragma unroll
        for (int r = R0; r < R1; ++r) {
            const uint32_t rk0 = k0 + (uint32_t)r * kPW0, rk1 = k1 + (uint32_t)r * kPW1;
            const uint64_t p0 = (uint64_t)kPM0 * x0;
            const uint64_t p1 = (uint64_t)kPM1 * x2;
            const uint32_t y0 = xor3((uint32_t)(p1 >> 32), x1, rk0);
            const uint32_t y2 = xor3((uint32_t)(p0 >> 32), x3, rk1);
            x0 = y0;
            x1 = (uint32_t)p1;
            x2 = y2;
            x3 = (uint32_t)p0;
        }
    }
};

FGK_HD void philox10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1, uint32_t &o0,
                     uint32_t &o1, uint32_t &o2, uint32_t &o3) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += kPW0;
            k1 += kPW1;
        }
        const uint64_t p0 = (uint64_t)kPM0 * c0;
        const uint64_t p1 = (uint64_t)kPM1 * c2;
        const uint32_t y0 = xor3((uint32_t)(p1 >> 32), c1, k0);
        const uint32_t y2 = xor3((uint32_t)(p0 >> 32), c3, k1);
        c0 = y0;
        c1 = (uint32_t)p1;
        c2 = y2;
        c3 = (uint32_t)p0;
    }
    o0 = c0; o1 = c1; o2 = c2; o3 = c3;
}

// Word: the LDS type of one factor (uint16_t for Z2 factors of up to 16
// coefficients, uint32_t otherwise).  kMaskT selects how class membership is
// kept beside H: false, the N plane (next member of each term's class);
// true, a plane CM of 32-bit class masks (bit u of CM[c][t] set when terms t
// and u have equal component-c factors, t included), for schemes of at most
// 32 terms.  With masks a class is known without walking it: a term leaves
// or joins a class with one store per member and no dependent reads.
//
// kWSv is the number of walkers per workgroup and the LDS stride between
// consecutive terms of one walker: kWalkersPerGroup (32) for the lane-pair
// layout, kSoloWalkersPerGroup (64) for the one-lane layout, where a lane
// runs both halves of a flip itself (the code path the host emulation
// runs) and its walker's column never shares an LDS bank with another lane
// of its half-wave.
template <typename WordT, bool kMaskT = false, int kWSv = kWalkersPerGroup>
struct LaneBaseT {
    using Word = WordT;
    static constexpr bool kMask = kMaskT;
    static constexpr int kWS = kWSv;
    // the two halves of a flip run on the two lanes of a pair (GPU, pair layout)
    static constexpr bool kSplit = kOnDevice && kWSv == kWalkersPerGroup;
    // The one-lane layout packs two 16-bit factor rows per dword: dword p of
    // component c holds row 15 - p in its low half and row 31 - p in its high
    // half (32 rows, kPackedPairs dwords per component, one dword column per
    // walker).  A scan of the 16 dwords then yields a row mask in term order
    // (packed_ne).
    static constexpr bool kPacked = kWSv == kSoloWalkersPerGroup;
    static constexpr int kPackedPairs = 16;
    static_assert(!kPacked || (kMaskT && sizeof(WordT) == 2), "the one-lane layout keeps 16-bit words and class masks");
    // Offset of walker slot `slot` in the F plane, in words.
    FGK_HD static int fslot(int slot) { return kPacked ? 2 * slot : slot; }
    // Address of row t of component c in a walker's F column based at Fb.
    FGK_HD static Word *fat(Word *Fb, int fs, int c, int t) {
        if constexpr (kPacked) return Fb + (c * kPackedPairs + 15 - (t & 15)) * (2 * kWS) + (t >> 4);
        else return Fb + (mul24(c, fs) + t) * kWS;
    }
    // c * x for the small non-negative plane indices (one full-rate
    // v_mul_u32_u24 instead of a quarter-rate 32-bit multiply).
    FGK_HD static int mul24(int c, int x) {
#if defined(__HIP_DEVICE_COMPILE__)
        // in the one-lane layout the compiler turned __umul24 of a value it
        // cannot bound into a quarter-rate v_mul_lo_u32 on the flip's path,
        // so there a product of two variables is written as the instruction
        // itself (the lane-pair layouts keep __umul24, which the compiler
        // folds into their address arithmetic)
        if constexpr (kPacked) {
            if (!__builtin_constant_p(c) && !__builtin_constant_p(x)) {
                int r;
                asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "v"(c), "v"(x));
                return r;
            }
        }
        return (int)__umul24((unsigned)c, (unsigned)x);
#else
        return c * x;
#endif
    }
    // a * b for small signed values (|a|, |b| < 2^23: one v_mul_i32_i24)
    FGK_HD static int smul24(int a, int b) {
#if defined(__HIP_DEVICE_COMPILE__)
        return __mul24(a, b);
#else
        return a * b;
#endif
    }
    FGK_HD static const Word *fat(const Word *Fb, int fs, int c, int t) { return fat(const_cast<Word *>(Fb), fs, c, t); }
    Word *F;      // walker-slot-offset plane bases
    uint8_t *H;   // slot-offset byte base of the packed H plane
    uint8_t *N;
    uint32_t *CM; // kMask: slot-offset base of the class-mask plane
    int cs;       // kMask: term stride of CM
    int side;     // 0: index1 / component-j half of a flip, 1: index2 / component-k half
    int32_t *err;
    int maxm;     // term capacity
    int fs;       // LDS term stride of F (lds_term_stride)
    int hw;       // H words per component
    int n0, n1, n2;
    int nn0, nn1, nn2;
    int m;
    uint32_t S0, S1, S2;  // unordered pair count per component
    uint32_t k0, k1, wid;
    uint64_t aux;
    uint32_t a0, a1, a2, a3;
    int aidx;
    int visit_cap;  // integer ring: candidates one tryFlip may visit
    uint32_t vmask; // coefficient bits of a factor word (all for Z2, bits 0-15 for the integer ring)
    int pop;        // coefficient count of the scheme, kept for getComplexity
    bool track_pop; // keep pop current through flips (the minimizer reads it every flip)
    // One-lane layout: rows changed since CM, H and S0..S2 were last exact
    // (bit t of dirtyC: row t of component C), set by the operators that
    // change a few rows (a reduction after a flip, a flip that leaves a zero
    // factor, plus, split) before they bring the classes up to date; zero
    // otherwise.
    uint32_t dirty0, dirty1, dirty2;
    FGK_HD void mark_row(int c, int t) {
        const uint32_t b = 1u << t;
        dirty0 |= c == 0 ? b : 0u;
        dirty1 |= c == 1 ? b : 0u;
        dirty2 |= c == 2 ? b : 0u;
    }
    int sel_c, sel_t1, sel_t2;  // pair of the last tryFlip (component, index1 < index2)
    // Flip-trace launches only (null in every other kernel, where the
    // recording folds away): the operators applied to this walker's scheme,
    // 8 int32 each, [code, a0..a6] with the codes of an explicit trace:
    // 0 flip(i, j, k, i1, i2, checkReduce), 1 plus(i, j, k, i1, i2, variant),
    // 2 split(i, j, k, index, a, s) (Z2: a = the new factor, s = 0; integer
    // ring: a = magnitudes, s = signs), 6 tryReduce (fresh-list order).  trc_n counts every
    // operator, also those past trc_cap, which are not stored.
    int32_t *trc;
    int trc_cap, trc_n;
    FGK_HD void rec(int code, int x0, int x1, int x2, int x3, int x4, int x5, int x6) {
        if (!trc) return;
        if (side == 0 && trc_n < trc_cap) {
            int32_t *o = trc + (size_t)trc_n * 8;
            o[0] = code;
            o[1] = x0;
            o[2] = x1;
            o[3] = x2;
            o[4] = x3;
            o[5] = x4;
            o[6] = x5;
            o[7] = x6;
        }
        ++trc_n;
    }
#if defined(FGK_PHASE_CLOCKS) && defined(__HIP_DEVICE_COMPILE__)
    // Phase profile build only: cycles (s_memtime) spent since the previous
    // mark, summed per phase: 0 RNG block, 1 pair selection, 2 factor reads
    // and class-member scan, 3 class walk or mask logic, 4 exchange of the
    // flip halves, 5 leave / join, 6 bookkeeping and best copy, 7 loop tail,
    // 8 reduction with rebuild, 9 zero-factor flip.
    uint32_t ph[10];
    uint64_t pht;
    FGK_HD void mark(int k) {
        const uint64_t t = __builtin_readcyclecounter();
        ph[k] += (uint32_t)(t - pht);
        pht = t;
    }
    FGK_HD void mark_start() {
        for (int q = 0; q < 10; ++q) ph[q] = 0;
        pht = __builtin_readcyclecounter();
    }
#else
    FGK_HD void mark(int) {}
    FGK_HD void mark_start() {}
#endif

    // Extra per-lane LDS beyond the common planes (none by default).
    FGK_HD void lane_extra(uint8_t *, int) {}

    FGK_HD Word &f(int c, int t) { return *fat(F, fs, c, t); }
    // Packed layout: dword p of component c's rows (rows 15 - p and 31 - p).
    FGK_HD uint32_t fpair(int c, int p) const {
        return reinterpret_cast<const uint32_t *>(F)[(c * kPackedPairs + p) * kWS];
    }
    // Packed layout: component c's column, and the offset of row t in it
    // (in words), so fcol(c)[packed_off(t)] is f(c, t).
    FGK_HD const Word *fcol(int c) const { return F + c * kPackedPairs * 2 * kWS; }
    FGK_HD static int packed_off(int t) { return (15 - (t & 15)) * (2 * kWS) + (t >> 4); }
    // H packs terms 4q..4q+3 of a walker into one dword: byte (t & 3) of word q.
    // H words per component (a constant 8 in the one-lane layout, schemes of at most 32 terms)
    FGK_HD int hwc() const { return kPacked ? 8 : hw; }
    FGK_HD uint8_t &h(int c, int t) { return H[((mul24(c, hwc()) + (t >> 2)) * kWS) * 4 + (t & 3)]; }
    FGK_HD uint32_t &hword(int c, int q) { return reinterpret_cast<uint32_t *>(H)[(mul24(c, hwc()) + q) * kWS]; }
    FGK_HD uint8_t &nx(int c, int t) { return N[((mul24(c, hw) + (t >> 2)) * kWS) * 4 + (t & 3)]; }
    FGK_HD uint32_t &cm(int c, int t) { return CM[(mul24(c, cs) + t) * kWS]; }
    // H[c][t] += d (d = +-1) as one LDS add on t's dword that returns nothing,
    // so the wave does not wait for it (the byte never wraps: H counts are
    // below 256 and a decremented count is at least 1).
    FGK_HD void h_add(int c, int t, int d) { word_add(&hword(c, t >> 2), (uint32_t)d << (8 * (t & 3))); }
    // *p += v as an LDS add that returns nothing (GPU)
    FGK_HD static void word_add(uint32_t *p, uint32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
        atomicAdd(p, v);
#else
        *p += v;
#endif
    }
    // all ones when bit y of x is set, else 0 (one v_bfe_i32)
    FGK_HD static uint32_t sbit(uint32_t x, int y) {
#if defined(__HIP_DEVICE_COMPILE__)
        return (uint32_t)__builtin_amdgcn_sbfe((int)x, (unsigned)y, 1u);
#else
        return 0u - ((x >> y) & 1u);
#endif
    }
    FGK_HD void fail(int32_t code) { fgk_atomic_or(err, code); }

    // Component-indexed scalars are selected arithmetically: a select
    // between member addresses would keep the lane state out of registers.
    FGK_HD int nn(int c) const {
        return (c == 0 ? 1 : 0) * nn0 + (c == 1 ? 1 : 0) * nn1 + (c == 2 ? 1 : 0) * nn2;
    }
    FGK_HD void addS(int c, int d) {
        S0 += (uint32_t)((c == 0 ? 1 : 0) * d);
        S1 += (uint32_t)((c == 1 ? 1 : 0) * d);
        S2 += (uint32_t)((c == 2 ? 1 : 0) * d);
    }
    FGK_HD uint32_t Sc(int c) const { return (c == 0 ? S0 : 0u) + (c == 1 ? S1 : 0u) + (c == 2 ? S2 : 0u); }
    FGK_HD int naive_rank() const { return n0 * n1 * n2; }
    // getComplexity (scheme_z2.cu:127-139, scheme_integer.cu:139-147)
    FGK_HD int complexity() const { return pop - 2 * m - nn2; }
    FGK_HD static int popc(uint32_t x) { return __builtin_popcount(x); }
    // coefficient count change of a flip that replaces factors a0 -> a1 and b0 -> b1
    FGK_HD void pop_update(uint32_t a0, uint32_t a1, uint32_t b0, uint32_t b1) {
        if (track_pop) pop += popc(a1 & vmask) + popc(b1 & vmask) - popc(a0 & vmask) - popc(b0 & vmask);
    }
    FGK_HD void recount_pop() {
        pop = 0;
        for (int c = 0; c < 3; ++c)
            for (int t = 0; t < m; ++t) pop += popc(f(c, t) & vmask);
    }

    // ---- aux RNG stream ----
    FGK_HD void aux_fill(uint64_t block) {
        philox10((uint32_t)block, (uint32_t)(block >> 32), wid, 1u, k0, k1, a0, a1, a2, a3);
    }
    // a0..a3 hold the unread words of the current aux block, shifted down
    // as they are consumed; aidx counts the words already taken (4 = empty).
    FGK_HD uint32_t aux32() {
        if (aidx >= 4) {
            aux_fill(aux);
            ++aux;
            aidx = 0;
        }
        const uint32_t r = a0;
        a0 = a1;
        a1 = a2;
        a2 = a3;
        ++aidx;
        return r;
    }
    // Reload the unread tail of block aux - 1 after a launch boundary.
    FGK_HD void aux_resume() {
        aux_fill(aux - 1);
        for (int q = 0; q < aidx; ++q) {
            a0 = a1;
            a1 = a2;
            a2 = a3;
        }
    }
    // randomPermutation(p, 3), src/utils/random.cuh:9-22: step i = 2 swaps
    // p[2] with p[rand % 3], step i = 1 swaps p[1] with p[rand % 2].
    FGK_HD void perm3(int &p0, int &p1, int &p2) {
        const uint32_t j2 = aux32() % 3u;
        const int q0 = j2 == 0 ? 2 : 0;
        const int q1 = j2 == 1 ? 2 : 1;
        const int q2 = j2 == 0 ? 0 : (j2 == 1 ? 1 : 2);
        const uint32_t j1 = aux32() % 2u;
        p0 = j1 == 0 ? q1 : q0;
        p1 = j1 == 0 ? q0 : q1;
        p2 = q2;
    }

    // ---- partner classes ----
    // The classes of component c from factors held in registers: rows [0, R)
    // are loaded at once (rows from the rank up are zero, the plane holds at
    // least R rows), then each class is found by comparing its first
    // unassigned member with all R registers, so no comparison waits for an
    // LDS read.  Writes H and, per the layout, the class masks or the next
    // members; returns the component's pair count.
    template <int R>
    FGK_HD uint32_t rebuild_rows(int c) {
        using M = typename std::conditional<(R > 32), uint64_t, uint32_t>::type;
        uint32_t v[R];
#pragma unroll
        for (int e = 0; e < R; ++e) v[e] = f(c, e);
        // the one-lane layout also keeps the rows as packed pairs, so a class
        // is one packed compare (packed_ne) instead of R compares
        uint32_t P[kPackedPairs];
        if constexpr (kPacked && R == 32) {
#pragma unroll
            for (int p = 0; p < kPackedPairs; ++p) P[p] = fpair(c, p);
        }
        const M live = m >= R ? ~M(0) : ((M(1) << m) - M(1));
        M done = ~live;
        uint32_t S = 0;
        while (~done) {
            const int t = ctz_m(~done);
            // the one-lane layout (one wave per SIMD, nothing to hide an LDS
            // round trip behind) takes the leader's factor from the
            // registers by a select tree; the lane pairs read it
            uint32_t vt;
            if constexpr (kPacked) vt = pick_row<R>(v, t);
            else vt = f(c, t);
            M mask = 0;
            if constexpr (kPacked && R == 32) {
                mask = (M)~packed_ne(P, vt);
            } else {
#pragma unroll
                for (int e = 0; e < R; ++e) mask |= M(v[e] == vt ? 1u : 0u) << e;
            }
            mask &= live;
            done |= mask;
            const uint32_t k = popc_m(mask);
            S += k * (k - 1) / 2;
            for (M q = mask; q; q &= q - M(1)) {
                const int y = ctz_m(q);
                const M above = mask & ~((M(2) << y) - M(1));
                h(c, y) = (uint8_t)popc_m(above);
                if constexpr (kMask) cm(c, y) = (uint32_t)mask;
                else nx(c, y) = (uint8_t)(above ? ctz_m(above) : ctz_m(mask));
            }
        }
        return S;
    }
    // v[t] for a lane-varying t < R as a binary select tree over the
    // registers (GPU; plain indexing on the host)
    template <int R>
    FGK_HD static uint32_t pick_row(const uint32_t (&v)[R], int t) {
#if defined(__HIP_DEVICE_COMPILE__)
        uint32_t w[R];
#pragma unroll
        for (int e = 0; e < R; ++e) w[e] = v[e];
#pragma unroll
        for (int half = R / 2, b = 0; half >= 1; half /= 2, ++b) {
            const bool hi = ((t >> b) & 1) != 0;
#pragma unroll
            for (int e = 0; e < half; ++e) w[e] = hi ? w[2 * e + 1] : w[2 * e];
        }
        return w[0];
#else
        return v[t];
#endif
    }
    FGK_HD static int ctz_m(uint32_t x) { return __builtin_ctz(x); }
    FGK_HD static int ctz_m(uint64_t x) { return __builtin_ctzll(x); }
    FGK_HD static uint32_t popc_m(uint32_t x) { return (uint32_t)__builtin_popcount(x); }
    FGK_HD static uint32_t popc_m(uint64_t x) { return (uint32_t)__builtin_popcountll(x); }

    // initFlips (src/schemes/scheme_z2.cu:141-150) for the N/H (or class-mask)
    // representation.  Schemes of up to 64 terms take the register path
    // (rebuild_rows); larger ones walk the plane.
    FGK_HD void rebuild() {
        S0 = S1 = S2 = 0;
        for (int c = 0; c < 3; ++c)
            for (int t = m; t < fs; ++t) f(c, t) = 0;  // the scan's rows past the rank
        if (track_pop) recount_pop();
        for (int c = 0; c < 3; ++c) {
            for (int q = 0; q < hw; ++q) hword(c, q) = 0;
            uint32_t S = 0;
            if (m <= 32) {
                S = rebuild_rows<32>(c);
            } else if constexpr (!kMask) {
                if (m <= 64 && fs >= 64) {
                    S = rebuild_rows<64>(c);
                } else {
                    S = rebuild_walk(c);
                }
            }
            addS(c, (int)S);
        }
    }
    FGK_HD uint32_t rebuild_walk(int c) {
        for (int t = 0; t < m; ++t) nx(c, t) = 0xFF;
        uint32_t S = 0;
        for (int t = 0; t < m; ++t) {
            if (nx(c, t) != 0xFF) continue;
            const uint32_t v = f(c, t);
            int prev = t, cnt = 0;
            for (int u = t + 1; u < m; ++u) {
                if (f(c, u) == v) {
                    nx(c, prev) = (uint8_t)u;
                    prev = u;
                    ++cnt;
                }
            }
            nx(c, prev) = (uint8_t)t;
            int y = t, r = cnt;
            do {
                h(c, y) = (uint8_t)r;
                --r;
                y = nx(c, y);
            } while (y != t && r >= 0);
            S += (uint32_t)(cnt * (cnt + 1) / 2);
        }
        return S;
    }

    // Term x leaves its component-c class (its factor is about to change).
    // Each step issues both loads of a member before using them and writes
    // H back unconditionally, so a step waits for one LDS round trip.
    FGK_HD void leave(int c, int x) {
        int y = nx(c, x);
        if (y == x) return;
        int p = x, cnt = 0;
        while (y != x) {
            const int ny = nx(c, y);
            const uint8_t hy = h(c, y);
            h(c, y) = (uint8_t)(hy - (y < x ? 1 : 0));
            ++cnt;
            p = y;
            y = ny;
            if (cnt > m) {
                fail(kErrWalk);
                return;
            }
        }
        nx(c, p) = nx(c, x);
        nx(c, x) = (uint8_t)x;
        h(c, x) = 0;
        addS(c, -cnt);
    }

    // Walk the component-c class containing y0 that singleton term x is about
    // to join: raise H of the members below x, collect what the join needs,
    // and let `rule(a, b)` test each member for a reduction from its factors
    // a = f(ca, y) and b = f(cb, y) (0 = none); the lowest member with a rule
    // is returned in (best_t, best_rule).  Every load of a member is issued
    // before the first use and H is written back unconditionally, so a step
    // waits for one LDS round trip.  H is touched before the caller knows
    // whether a reduction follows, which is harmless: every reduction
    // rebuilds the classes.
    struct JoinInfo {
        int cnt, gt, p, mx;
    };
    template <class Rule>
    FGK_HD void class_walk(int c, int x, int y0, int ca, int cb, JoinInfo &ji, Rule rule, int &best_t,
                           int &best_rule) {
        int y = y0;
        ji.cnt = 0;
        ji.gt = 0;
        ji.p = -1;
        ji.mx = -1;
        do {
            const int ny = nx(c, y);
            const uint8_t hy = h(c, y);
            const uint32_t a = f(ca, y), b = f(cb, y);
            const bool below = y < x;
            h(c, y) = (uint8_t)(hy + (below ? 1 : 0));
            ++ji.cnt;
            ji.p = below && y > ji.p ? y : ji.p;
            ji.gt += below ? 0 : 1;
            ji.mx = y > ji.mx ? y : ji.mx;
            const int r = rule(a, b);
            if (r && y < best_t) {
                best_t = y;
                best_rule = r;
            }
            y = ny;
            if (ji.cnt > m) {
                fail(kErrWalk);
                return;
            }
        } while (y != y0);
    }
    // Link singleton x into the class walked by class_walk.
    FGK_HD void link_join(int c, int x, const JoinInfo &ji) {
        const int p = ji.p < 0 ? ji.mx : ji.p;
        h(c, x) = (uint8_t)ji.gt;
        nx(c, x) = nx(c, p);
        nx(c, p) = (uint8_t)x;
        addS(c, ji.cnt);
    }

    // The term t1 holding position r of component c's H plane, in blocks of
    // NB words (NB divides hw); r is left as the offset within t1's count.
    template <int NB>
    FGK_HD void select_scan(int c, uint32_t &r, int &t1) {
        for (int q = 0; q < hw; q += NB) {
            uint32_t wd[NB];
#pragma unroll
            for (int e = 0; e < NB; ++e) wd[e] = hword(c, q + e);
            // running sums P[e] are non-decreasing, so the word holding
            // position r is the count of P[e] <= r, and the last such P[e]
            // is what the words before it hold (branch-free, no indexing)
            uint32_t acc = 0, below = 0;
            int cnt = 0;
#pragma unroll
            for (int e = 0; e < NB; ++e) {
                acc = byte_sum_acc(wd[e], acc);
                const bool le = acc <= r;
                below = le ? acc : below;
                cnt += le ? 1 : 0;
            }
            if (cnt == NB) {
                r -= acc;
                continue;
            }
            r -= below;
            uint32_t word = wd[0];
#pragma unroll
            for (int e = 1; e < NB; ++e) word = cnt == e ? wd[e] : word;
            // the same within the word's four term bytes
            const uint32_t c0 = word & 0xFFu, c1 = c0 + ((word >> 8) & 0xFFu), c2 = c1 + ((word >> 16) & 0xFFu);
            const int b = (r >= c0 ? 1 : 0) + (r >= c1 ? 1 : 0) + (r >= c2 ? 1 : 0);
            r -= b == 0 ? 0u : (b == 1 ? c0 : (b == 2 ? c1 : c2));
            t1 = (q + cnt) * 4 + b;
            break;
        }
    }

    // select_scan for planes of exactly NB words, which always hold position
    // r: no loop, and the word holding r is chosen with the compares the
    // running sums already made.
    template <int NB>
    FGK_HD void select_block(int c, uint32_t &r, int &t1) {
        uint32_t wd[NB];
#pragma unroll
        for (int e = 0; e < NB; ++e) wd[e] = hword(c, e);
        select_words<NB>(wd, r, t1);
    }
    // select_block on H words already loaded.
    template <int NB>
    FGK_HD static void select_words(const uint32_t (&wd)[NB], uint32_t &r, int &t1) {
        bool le[NB];
        uint32_t acc = 0, below = 0;
#pragma unroll
        for (int e = 0; e < NB; ++e) {
            acc = byte_sum_acc(wd[e], acc);
            le[e] = acc <= r;
            below = le[e] ? acc : below;
        }
        uint32_t word = wd[NB - 1];
        int idx = NB - 1;
#pragma unroll
        for (int e = NB - 2; e >= 0; --e) {
            word = le[e] ? word : wd[e];
            idx = le[e] ? idx : e;
        }
        r -= below;
        const uint32_t c0 = word & 0xFFu, c1 = c0 + ((word >> 8) & 0xFFu), c2 = c1 + ((word >> 16) & 0xFFu);
        const int b = (r >= c0 ? 1 : 0) + (r >= c1 ? 1 : 0) + (r >= c2 ? 1 : 0);
        // the running sums c0 <= c1 <= c2 below r: the last one is subtracted
        // (selects, not a branch on b)
        uint32_t sub = r >= c0 ? c0 : 0u;
        sub = r >= c1 ? c1 : sub;
        sub = r >= c2 ? c2 : sub;
        r -= sub;
        t1 = idx * 4 + b;
    }

    // The unordered pair at position r (< S0 + S1 + S2) of the fresh-list
    // order: component, then t1, then t2.  H is scanned in blocks of 8 or 16
    // words (32 or 64 terms; the loads of a block are issued together) with
    // byte sums, then t2 is r' + 1 steps along N from t1.
    FGK_HD bool select_pair(uint32_t r, int &c, int &t1, int &t2) {
        c = 0;
        if (r >= S0) {
            r -= S0;
            c = 1;
            if (r >= S1) {
                r -= S1;
                c = 2;
            }
        }
        t1 = -1;
        // planes of 16 H words (64 terms) are scanned in one block, so every
        // H read of the selection goes out in one round trip
        if constexpr (kMask) select_block<8>(c, r, t1);  // at most 32 terms: 8 H words
        else if (hw % 16 == 0) select_scan<16>(c, r, t1);
        else select_scan<8>(c, r, t1);
        if (t1 < 0 || t1 >= m) {
            fail(kErrSelect);
            return false;
        }
        if constexpr (kMask) {
            // the (r + 1)-th member above t1, from t1's class mask
            uint32_t q = cm(c, t1) & ~((2u << t1) - 1u);
            for (uint32_t e = 0; e < r; ++e) q &= q - 1u;
            t2 = q ? __builtin_ctz(q) : -1;
            if (t2 < 0) {
                fail(kErrSelect);
                return false;
            }
            return true;
        }
        t2 = nx(c, t1);
        for (uint32_t q = 1; q <= r; ++q) t2 = nx(c, t2);
        return true;
    }

    // The last term t < m of component c whose factor word equals v (v != 0),
    // or -1.  Rows are loaded kScanBatch at a time with every load of a batch
    // issued before the first compare, so a scan of m <= 32 terms waits for
    // one LDS round trip (a loop over a run-time m would issue its remainder
    // one load at a time).  The rows from m to the end of the plane are zero
    // (rebuild clears them) and the plane holds whole batches
    // (lds_term_stride), so rows past m never match and need no rank test:
    // each row costs a compare and a select of a constant index.
    static constexpr int kScanBatch = kScanRows;
    FGK_HD int find_last(int c, uint32_t v) const {
        int any = -1;
        for (int base = 0; base < m; base += kScanBatch) {
            Word w[kScanBatch];
#pragma unroll
            for (int e = 0; e < kScanBatch; ++e) w[e] = *fat(F, fs, c, base + e);
            // four independent select chains (rows e mod 4), so the compiler
            // can interleave them; the last match is the largest of the four
            int l0 = -1, l1 = -1, l2 = -1, l3 = -1;
#pragma unroll
            for (int e = 0; e < kScanBatch; e += 4) {
                l0 = (uint32_t)w[e] == v ? e : l0;
                l1 = (uint32_t)w[e + 1] == v ? e + 1 : l1;
                l2 = (uint32_t)w[e + 2] == v ? e + 2 : l2;
                l3 = (uint32_t)w[e + 3] == v ? e + 3 : l3;
            }
            const int a01 = l0 > l1 ? l0 : l1, a23 = l2 > l3 ? l2 : l3;
            const int local = a01 > a23 ? a01 : a23;
            any = local >= 0 ? base + local : any;
        }
        return any;
    }

    // ---- term edits (callers rebuild the classes afterwards) ----
    FGK_HD void copy_term(int dst, int src) {
        f(0, dst) = f(0, src);
        f(1, dst) = f(1, src);
        f(2, dst) = f(2, src);
    }
    FGK_HD bool nonzero_term(int t) { return f(0, t) && f(1, t) && f(2, t); }
    // removeAt (src/schemes/scheme_z2.cu:166-175)
    FGK_HD void remove_at(int t) {
        --m;
        if (t != m) copy_term(t, m);
    }
};

FGK_HD uint32_t min_u32(uint32_t a, uint32_t b) { return a < b ? a : b; }

// No instruction is moved across this point by the GPU compiler's
// scheduler (nothing on the host).
FGK_HD void sched_fence() {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_sched_barrier(0);
#endif
}

// x in a vector register (an opaque copy on the GPU), so values derived
// from it stay in vector registers too.
FGK_HD uint32_t vgpr_copy(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t r;
    asm("v_mov_b32 %0, %1" : "=v"(r) : "s"(x));
    return r;
#else
    return x;
#endif
}

// A uniform draw in [0, size): floor(x * size / 2^32) (one v_mul_hi_u32; the
// reference's curand() % size needs an integer division).
FGK_HD uint32_t draw_below(uint32_t x, uint32_t size) { return (uint32_t)(((uint64_t)x * size) >> 32); }

// Orientation of a flip in component c: i = c, (j, k) = the other two in
// increasing order, swapped by coin bit 0 (scheme_z2.cu:988-1016).
FGK_HD void orientation(int c, uint32_t coin, int &j, int &k) {
    j = c == 0 ? 1 : 0;
    k = c == 2 ? 1 : 2;
    if (coin & 1u) {
        const int t = j;
        j = k;
        k = t;
    }
}

// The value held by the other lane of this lane's pair (GPU only): one
// v_permlane32_swap exchanges the two half-waves.
FGK_HD int pair_exchange(int v, int side) {
#if defined(__HIP_DEVICE_COMPILE__)
    const auto r = __builtin_amdgcn_permlane32_swap((unsigned)v, (unsigned)v, false, false);
    return (int)(side ? r[0] : r[1]);
#else
    (void)side;
    return v;
#endif
}

// Maximum of v over the wave (the lane's own value on the host).
FGK_HD int wave_max(int v) {
#if defined(__HIP_DEVICE_COMPILE__)
    for (int o = 32; o > 0; o >>= 1) {
        const int u = __shfl_xor(v, o);
        v = u > v ? u : v;
    }
#endif
    return v;
}

// copyTo(schemesBest[idx]) (flip_graph.cu:547): the current scheme of every
// walker with `want` goes to its slot of A.best, laid out per walker as
// [c][t] (max_m words per component).  On the GPU the whole wave copies one
// walker per pass with coalesced stores; the host emulation copies its own
// walker.
template <class Lane>
FGK_HD void copy_best(Lane &s, const WalkArgs &A, bool want, int lane, int w) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint64_t mask = __ballot(want && s.side == 0);
    if constexpr (Lane::kPacked) {
        // at most 32 terms: two walkers per pass, lanes 0-31 copy the first
        // and lanes 32-63 the second (row t on lane t of its half); a pass
        // with one walker left copies it on both halves
        const int half = lane >> 5, t = lane & 31;
        while (mask) {
            const int sa = __builtin_ctzll(mask);
            mask &= mask - 1;
            const int sb = mask ? __builtin_ctzll(mask) : sa;
            mask &= mask - 1;
            const int msa = __builtin_amdgcn_readlane(s.m, sa), msb = __builtin_amdgcn_readlane(s.m, sb);
            const int wsa = __builtin_amdgcn_readlane(w, sa), wsb = __builtin_amdgcn_readlane(w, sb);
            const int src = half ? sb : sa, ms = half ? msb : msa, ws = half ? wsb : wsa;
            if (A.bfmt) {
                const int f0 = half ? __builtin_amdgcn_readlane(s.n0, sb) : __builtin_amdgcn_readlane(s.n0, sa);
                const int f1 = half ? __builtin_amdgcn_readlane(s.n1, sb) : __builtin_amdgcn_readlane(s.n1, sa);
                const int f2 = half ? __builtin_amdgcn_readlane(s.n2, sb) : __builtin_amdgcn_readlane(s.n2, sa);
                if (t == 0) {
                    int32_t *bf = A.bfmt + 4 * (size_t)ws;
                    bf[0] = f0;
                    bf[1] = f1;
                    bf[2] = f2;
                }
            }
            if (t < ms) {
                const typename Lane::Word *Fs = s.F - Lane::fslot(lane % Lane::kWS) + Lane::fslot(src);
                uint32_t *dst = A.best + (size_t)ws * 3 * A.max_m;
                const int o = Lane::packed_off(t);
                const uint32_t x0 = Fs[o], x1 = Fs[o + 32 * Lane::kWS], x2 = Fs[o + 64 * Lane::kWS];
                dst[t] = x0;
                dst[A.max_m + t] = x1;
                dst[2 * A.max_m + t] = x2;
            }
        }
        return;
    }
    while (mask) {
        const int src = __builtin_ctzll(mask);
        mask &= mask - 1;
        const int ms = __builtin_amdgcn_readlane(s.m, src);
        // the source lane's walker (the one-lane layout takes two queue
        // items per wave, so the walkers of a wave need not be consecutive)
        const int ws = __builtin_amdgcn_readlane(w, src);
        const int slot = lane % Lane::kWS, src_slot = src % Lane::kWS;
        if (A.bfmt) {
            const int f0 = __builtin_amdgcn_readlane(s.n0, src), f1 = __builtin_amdgcn_readlane(s.n1, src),
                      f2 = __builtin_amdgcn_readlane(s.n2, src);
            if (lane == 0) {
                int32_t *bf = A.bfmt + 4 * (size_t)ws;
                bf[0] = f0;
                bf[1] = f1;
                bf[2] = f2;
            }
        }
        const typename Lane::Word *Fs = s.F - Lane::fslot(slot) + Lane::fslot(src_slot);
        uint32_t *dst = A.best + (size_t)ws * 3 * A.max_m;
        // the three components' reads go out together: one LDS round trip per pass
        for (int t = lane; t < ms; t += kThreadsPerGroup) {
            const uint32_t x0 = *Lane::fat(Fs, s.fs, 0, t), x1 = *Lane::fat(Fs, s.fs, 1, t),
                           x2 = *Lane::fat(Fs, s.fs, 2, t);
            dst[t] = x0;
            dst[A.max_m + t] = x1;
            dst[2 * A.max_m + t] = x2;
        }
    }
#else
    if (want) {
        if (A.bfmt) {
            A.bfmt[4 * (size_t)w] = s.n0;
            A.bfmt[4 * (size_t)w + 1] = s.n1;
            A.bfmt[4 * (size_t)w + 2] = s.n2;
        }
        uint32_t *dst = A.best + (size_t)w * 3 * A.max_m;
        for (int c = 0; c < 3; ++c)
            for (int t = 0; t < s.m; ++t) dst[c * A.max_m + t] = s.f(c, t);
    }
#endif
}

// True when `v` holds on any lane of the wave (on the host: for this walker).
FGK_HD bool wave_any(bool v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __ballot(v) != 0;
#else
    return v;
#endif
}

// Modes of walk_lane.
constexpr int kModeWalk = 0;      // randomWalkKernel
constexpr int kModeMinimize = 1;  // minimizeKernel
constexpr int kModeLearn = 2;     // walk + proposed first flip + trajectory logging
constexpr int kModeFlips = 3;     // flipSchemesKernel of the additions reducer
constexpr int kModeResize = 4;    // resizeKernel (Z2)
constexpr int kModeMeta = 5;      // walk with the sandwiching and basis-change edges (Z2)
constexpr int kModeLook = 6;      // walk with reduction-seeking lookahead proposals (Z2)
constexpr int kModeTrace = 7;     // plain walk that records every operator it applies

// Resize draw thresholds: a uniform compare u < q is x < floor(q * 2^32).
constexpr uint32_t kThrHalf = 2147483648u, kThr005 = 214748364u, kThr055 = 2362232012u, kThr085 = 3650722201u;
constexpr int kMinProjectN = 2;    // src/config.cuh:34
constexpr int kMaxExtensionN = 16; // src/config.cuh:37

// The walk of one lane (randomWalkKernel body, src/entities/flip_graph.cu:520-573),
// or in kModeMinimize the complexity minimizer's (minimizeKernel,
// src/entities/complexity_minimizer.cu:190-214): flips without reductions
// until one fails, keeping the scheme of lowest complexity; best_rank then
// holds that complexity.  kModeLearn is the walk with two additions at the
// first iteration of a launch: a proposed pair (A.forced) replaces the
// uniform pair draw when it is still a flip, and logged walkers record their
// scheme and the pair they flip (A.log_words / A.log_meta); the lowest rank
// of the launch goes to A.minm.  kModeFlips is the additions reducer's
// scheme flips (flipSchemesKernel, src/entities/scheme_additions_reducer.cu:448-460):
// every walker but walker 0 draws a count in [0, A.iterations] from its aux
// stream and applies that many flips with reductions, stopping at the first
// that finds no pair.  kModeResize is resizeKernel
// (src/entities/flip_graph.cu:575-613): with probability thr_resize / 2^32 a
// walker swaps two dimensions (half the time), then tries to merge with up
// to three random walkers' best schemes, else projects, multiplies by a
// random best scheme, doubles a dimension or extends one.  kModeMeta is the
// walk with the sandwiching and basis-change checks of randomWalkKernel
// (:558-562), kept out of the plain walk kernel.  kModeTrace is the plain walk
// that also writes every operator it applies to A.trace (LaneBaseT::rec), an
// explicit trace that replays the launch without its random draws.  `lds` is the workgroup's LDS block, `lane`
// the lane in the workgroup (walker slot lane / 2, flip side lane % 2), `w`
// the walker index.  The host emulation calls it once per walker with an
// even lane and does both sides.
template <class Lane, int kMode>
FGK_HD void walk_lane(const WalkArgs &A, uint8_t *lds, int lane, int w, int w_end) {
    const bool active = w < w_end;
    const int W = A.walkers;
    const int maxm = A.max_m;

    using Word = typename Lane::Word;
    Lane s;
    const int hw = h_words(maxm);
    const int fs = Lane::kPacked ? 2 * Lane::kPackedPairs : lds_term_stride(maxm, (int)sizeof(Word));
    constexpr int WS = Lane::kWS;
    const size_t fbytes = Lane::kPacked ? z2s_lds_f_bytes() : walk_lds_f_bytes(maxm, (int)sizeof(Word), WS);
    const int slot = lane % WS;
    s.side = lane / WS;
    s.F = reinterpret_cast<Word *>(lds) + Lane::fslot(slot);
    s.H = lds + fbytes + (size_t)slot * 4;
    if constexpr (Lane::kMask) {
        s.N = nullptr;
        s.cs = lds_mask_stride(maxm);
        s.CM = reinterpret_cast<uint32_t *>(lds + fbytes + (size_t)3 * hw * WS * 4) + slot;
    } else {
        s.CM = nullptr;
        s.cs = 0;
        s.N = lds + fbytes + (size_t)3 * hw * WS * 4 + (size_t)slot * 4;
        s.lane_extra(lds + walk_lds_base_bytes(maxm, (int)sizeof(Word)), slot);
    }
    s.fs = fs;
    s.err = A.err;
    s.maxm = maxm;
    s.hw = hw;
    // one-lane layout: the Philox key is kept in vector registers, so the
    // twenty round keys the compiler hoists out of the loop do not crowd the
    // scalar registers (the kernel's SGPRs spill otherwise)
    s.k0 = Lane::kPacked ? vgpr_copy(A.key0) : A.key0;
    s.k1 = Lane::kPacked ? vgpr_copy(A.key1) : A.key1;
    s.wid = (uint32_t)w;
    s.dirty0 = s.dirty1 = s.dirty2 = 0u;
    s.m = 0;
    s.n0 = s.n1 = s.n2 = 1;
    s.visit_cap = A.zt_visit_cap;
    s.vmask = Lane::kValueMask;
    s.pop = 0;
    s.track_pop = kMode == kModeMinimize;
    s.trc = nullptr;
    s.trc_cap = 0;
    s.trc_n = 0;
    if constexpr (kMode == kModeTrace) {
        if (active) {
            s.trc = A.trace + (size_t)w * A.trace_cap * 8;
            s.trc_cap = A.trace_cap;
        }
    }
    int best_rank = 0, best_m = 0, flips_count = 0;
    uint64_t it = 0, total_flips = 0, sum_m = 0;
    s.aux = 0;
    s.aidx = 4;
    if (active) {
        s.n0 = A.fmt[4 * w];
        s.n1 = A.fmt[4 * w + 1];
        s.n2 = A.fmt[4 * w + 2];
        s.m = A.m[w];
        best_rank = A.best_rank[w];
        best_m = A.best_m[w];
        flips_count = A.flips_count[w];
        it = A.it[w];
        s.aux = A.aux[w];
        s.aidx = A.aux_idx[w];
        total_flips = A.total_flips[w];
        sum_m = A.sum_m[w];
        for (int c = 0; c < 3; ++c)
            for (int t = 0; t < s.m; ++t) s.f(c, t) = A.cur[((size_t)c * maxm + t) * W + w];
        if (s.aidx < 4 && s.aux > 0) s.aux_resume();
    }
    s.nn0 = s.n0 * s.n1;
    s.nn1 = s.n1 * s.n2;
    s.nn2 = s.n2 * s.n0;
    if (active) s.rebuild();

    if constexpr (kMode == kModeMinimize) {
        const int iters = active ? A.iterations : 0;
        const int trips = wave_max(iters);
        bool live = active;
        for (int iter = 0; iter < trips && wave_any(live); ++iter) {
            bool copy = false;
            if (live && iter < iters) {
                uint32_t b0, b1, b2, b3;
                philox10((uint32_t)it, (uint32_t)(it >> 32), (uint32_t)w, 0u, s.k0, s.k1, b0, b1, b2, b3);
                ++it;
                const int rank = s.m;
                if (!s.try_flip(b0, b1, false)) {  // complexity_minimizer.cu:200-201
                    live = false;
                } else {
                    ++total_flips;
                    sum_m += (uint64_t)rank;
                    const int cx = s.complexity();
                    if (cx < best_rank) {  // :203-207
                        best_rank = cx;
                        best_m = s.m;
                        copy = true;
                    }
                }
            }
            copy_best(s, A, copy, lane, w);
        }
    } else if constexpr (kMode == kModeResize) {
        if (active && (uint64_t)s.aux32() < A.thr_resize) {  // :579
            const int R = A.resize_max_rank, C = A.resize_max_coef;
            if (s.aux32() < kThrHalf) rz_swap_size_random(s);  // :584-585
            bool merged = false;
            for (int i = 0; i < 3 && !merged; ++i) {  // :587-592
                const int x = (int)(s.aux32() % (uint32_t)W);
                const int32_t *bf = A.bfmt + 4 * (size_t)x;
                merged = rz_try_merge(s, A.best + (size_t)x * 3 * maxm, maxm, A.best_m[x], bf[0], bf[1], bf[2], R, C);
            }
            if (!merged) {  // :594-609
                const uint32_t u = s.aux32();
                if (u < kThr005) {
                    rz_try_project(s);
                } else if (u < kThr055) {
                    const int x = (int)(s.aux32() % (uint32_t)W);
                    const int32_t *bf = A.bfmt + 4 * (size_t)x;
                    rz_try_product_scheme(s, A.best + (size_t)x * 3 * maxm, maxm, A.best_m[x], bf[0], bf[1], bf[2], R, C);
                } else if (u < kThr085) {
                    rz_try_product_dim(s, R, C);
                } else {
                    rz_try_extend(s, R, C);
                }
            }
        }
    } else if constexpr (kMode == kModeFlips) {
        int count = 0;
        if (active && w != 0) count = (int)(s.aux32() % (uint32_t)(A.iterations + 1));  // :454
        const int trips = wave_max(count);
        bool live = active;
        for (int iter = 0; iter < trips && wave_any(live); ++iter) {
            if (live && iter < count) {
                uint32_t b0, b1, b2, b3;
                philox10((uint32_t)it, (uint32_t)(it >> 32), (uint32_t)w, 0u, s.k0, s.k1, b0, b1, b2, b3);
                ++it;
                const int rank = s.m;
                if (!s.try_flip(b0, b1, true)) {  // :458-459
                    live = false;
                } else {
                    ++total_flips;
                    sum_m += (uint64_t)rank;
                }
            }
        }
    } else {
    int iters = 0;
    if (active) {
        iters = A.iterations;
        if (A.random_iterations) iters = 1 + (int)(s.aux32() % (uint32_t)A.max_iterations);
    }
    const bool check_reduce = A.check_reduce != 0;
    // loop-invariant thresholds (vector registers in the one-lane layout, as the key)
    const uint32_t thr_best_eq = Lane::kPacked ? vgpr_copy(A.thr_best_eq) : A.thr_best_eq;
    const uint32_t thr_reduce = Lane::kPacked ? vgpr_copy(A.thr_reduce) : A.thr_reduce;
    const uint32_t thr_expand = Lane::kPacked ? vgpr_copy(A.thr_expand) : A.thr_expand;
    const int plus_iterations = Lane::kPacked ? (int)vgpr_copy((uint32_t)A.plus_iterations) : A.plus_iterations;
    // every lane runs the same number of loop trips so the wave stays
    // converged at the cooperative best copy
    const int trips = wave_max(iters);
    int min_m = s.m;
    const int log_slot = (kMode == kModeLearn && A.log_every > 0 && w % A.log_every == 0) ? w / A.log_every : -1;

    // On the GPU the two lanes of a pair draw the main blocks of two
    // iterations at once (side 0 the block of `it`, side 1 that of `it + 1`)
    // and swap them, so each lane runs one Philox per two iterations.
    uint32_t nb0 = 0, nb1 = 0, nb2 = 0, nb3 = 0;
    bool have_next = false;
    PhiloxRounds px;
    // one-lane layout, plain and trace walks: every iteration's block is
    // computed during the previous flip (the first one here)
    constexpr bool kPrefetch = Lane::kPacked && (kMode == kModeWalk || kMode == kModeTrace);
    if constexpr (kPrefetch)
        philox10((uint32_t)it, (uint32_t)(it >> 32), (uint32_t)w, 0u, s.k0, s.k1, nb0, nb1, nb2, nb3);
    s.mark_start();
    for (int iter = 0; iter < trips; ++iter) {
        bool flipped = false, copy = false;
        uint32_t b0 = 0, b1 = 0, b2 = 0, b3 = 0;
        if (iter < iters) {
            if constexpr (kPrefetch) {
                b0 = nb0;
                b1 = nb1;
                b2 = nb2;
                b3 = nb3;
            } else if constexpr (!Lane::kSplit) {
                // one-lane layout: the block may have been computed during
                // the previous iteration's flip
                if (have_next) {
                    b0 = nb0;
                    b1 = nb1;
                    b2 = nb2;
                    b3 = nb3;
                    have_next = false;
                } else {
                    philox10((uint32_t)it, (uint32_t)(it >> 32), (uint32_t)w, 0u, s.k0, s.k1, b0, b1, b2, b3);
                }
            } else if (have_next) {
                b0 = nb0;
                b1 = nb1;
                b2 = nb2;
                b3 = nb3;
                have_next = false;
            } else {
                const uint64_t own = it + (uint64_t)s.side;
                uint32_t x0, x1, x2, x3;
                philox10((uint32_t)own, (uint32_t)(own >> 32), (uint32_t)w, 0u, s.k0, s.k1, x0, x1, x2, x3);
                const uint32_t y0 = (uint32_t)pair_exchange((int)x0, s.side),
                               y1 = (uint32_t)pair_exchange((int)x1, s.side),
                               y2 = (uint32_t)pair_exchange((int)x2, s.side),
                               y3 = (uint32_t)pair_exchange((int)x3, s.side);
                b0 = s.side ? y0 : x0;
                b1 = s.side ? y1 : x1;
                b2 = s.side ? y2 : x2;
                b3 = s.side ? y3 : x3;
                nb0 = s.side ? x0 : y0;
                nb1 = s.side ? x1 : y1;
                nb2 = s.side ? x2 : y2;
                nb3 = s.side ? x3 : y3;
                have_next = true;
            }
            ++it;
            s.mark(0);
            const int rank = s.m;
            bool ok = false;
            if constexpr (kMode == kModeLearn) {
                if (iter == 0) {
                    if (log_slot >= 0 && A.log_words && s.side == 0)
                        for (int c = 0; c < 3; ++c)
                            for (int t = 0; t < s.m; ++t)
                                A.log_words[((size_t)log_slot * 3 + c) * maxm + t] = s.f(c, t);
                    if (A.forced && A.forced[4 * w] >= 0)
                        ok = s.try_flip_pair(A.forced[4 * w], A.forced[4 * w + 1], A.forced[4 * w + 2], b1, check_reduce);
                }
                if (!ok) ok = s.try_flip(b0, b1, check_reduce);
                if (iter == 0 && log_slot >= 0 && A.log_meta && s.side == 0) {
                    int32_t *lm = A.log_meta + (size_t)log_slot * 4;
                    lm[0] = ok ? s.sel_c : -1;
                    lm[1] = s.sel_t1;
                    lm[2] = s.sel_t2;
                    lm[3] = rank;
                }
            } else if constexpr (kMode == kModeLook) {
                if (iter % A.look_period == A.look_period - 1) ok = s.try_lookahead(A.look_candidates, check_reduce);
                if (!ok) ok = s.try_flip(b0, b1, check_reduce);
            } else if constexpr (kPrefetch) {
                // the next iteration's block (the counter is already advanced)
                // fills the flip's three read waits (rounds 0-3, 4-6, 7-9);
                // the block is used only if the launch goes on
                ok = s.try_flip(b0, b1, check_reduce, [&](auto phase) {
                    constexpr int P = decltype(phase)::value;
                    if constexpr (P == 0) {
                        px.start((uint32_t)it, (uint32_t)(it >> 32), (uint32_t)w, 0u, s.k0, s.k1);
                        px.template rounds<0, 4>();
                    } else if constexpr (P == 1) {
                        px.template rounds<4, 7>();
                    } else {
                        px.template rounds<7, 10>();
                        nb0 = px.x0;
                        nb1 = px.x1;
                        nb2 = px.x2;
                        nb3 = px.x3;
                        Lane::keep_loaded(nb0, nb1, nb2, nb3);  // computed here, not sunk to its use
                    }
                    Lane::keep_loaded(px.x0, px.x1, px.x2, px.x3);
                });
            } else {
                ok = s.try_flip(b0, b1, check_reduce);
            }
            if (!ok) s.try_expand(1 + (int)(b1 % 2u));  // flip_graph.cu:535-537
            // the bookkeeping of a flip, written as selects (no branches)
            flipped = ok;
            total_flips += ok ? 1u : 0u;
            sum_m += ok ? (uint64_t)rank : 0u;
            flips_count = ok ? (s.m < rank ? 0 : flips_count) + 1 : flips_count;
            copy = ok && (s.m < best_rank || (s.m == best_rank && b2 < thr_best_eq));  // :545-548
            best_rank = copy ? s.m : best_rank;
            best_m = copy ? s.m : best_m;
        }
        copy_best(s, A, copy, lane, w);
        s.mark(6);
        // the switches are tested first, so a walk without these edges
        // branches on uniform values only
        if (A.use_reduce) {
            if (flipped && s.aux32() < thr_reduce) {  // :550-551
                s.rec(6, 0, 0, 0, 0, 0, 0, 0);
                s.try_reduce();
            }
        }
        if (A.use_expand) {
            if (flipped && b3 < thr_expand && s.m <= best_m + 2) {  // :553-556
                if (s.try_expand(1 + (int)(s.aux32() % 2u))) flips_count = 0;
            }
        }
        if constexpr (kMode == kModeMeta) {
            if (flipped) {
                if (A.use_sandwich && s.aux32() < A.thr_sandwich) s.sandwiching();  // :558-559
                if (A.use_basis && s.aux32() < A.thr_basis) rz_swap_basis(s);       // :561-562
            }
        }
        if (flipped && flips_count >= plus_iterations && s.try_plus()) flips_count = 0;  // :564-565
        if (A.min_window <= 0 || iter < A.min_window) min_m = s.m < min_m ? s.m : min_m;
        s.mark(7);
    }
    if (kMode == kModeLearn && active && s.side == 0 && A.minm) A.minm[w] = min_m;
#if defined(FGK_PHASE_CLOCKS) && defined(__HIP_DEVICE_COMPILE__)
    if (active && kMode == kModeWalk) {
        for (int q = 0; q < 10; ++q) atomicAdd(&FGK_PHASE_ARRAY[q], (unsigned long long)s.ph[q]);
        atomicAdd(&FGK_PHASE_ARRAY[15], (unsigned long long)iters);
    }
#endif
    }

    if (active && s.side == 0) {
        if constexpr (kMode == kModeTrace) A.trace_n[w] = s.trc_n;
        A.m[w] = s.m;
        A.best_m[w] = best_m;
        A.best_rank[w] = best_rank;
        A.flips_count[w] = flips_count;
        A.it[w] = it;
        A.aux[w] = s.aux;
        A.aux_idx[w] = s.aidx;
        A.total_flips[w] = total_flips;
        A.sum_m[w] = sum_m;
        if (A.cx) {
            if (!s.track_pop) s.recount_pop();
            A.cx[w] = s.complexity();
        }
        if constexpr (kMode == kModeResize) {
            A.fmt[4 * w] = s.n0;
            A.fmt[4 * w + 1] = s.n1;
            A.fmt[4 * w + 2] = s.n2;
        }
        for (int c = 0; c < 3; ++c)
            for (int t = 0; t < s.m; ++t) A.cur[((size_t)c * maxm + t) * W + w] = s.f(c, t);
    }
}

}  // namespace lane
}  // namespace fgk

// Z2 operators of the walk (src/schemes/scheme_z2.cu) on the LDS lane state
// of lane_common.h.  __host__ __device__ so the tests can run the exact
// kernel code on the host.
#pragma once

#include "lane_common.h"

namespace fgk {
namespace lane {

// GF(2) square matrices of up to 16 x 16 as row bitmasks (bit j of r[i] is
// entry (i, j)), for sandwiching.
struct Gf2Mat {
    uint32_t r[16];
};
// inverseMatrixZ2 (scheme_z2.cu:444-486): Gauss-Jordan on [a | I]; false
// when a is singular.
FGK_HD bool gf2_invert(int n, const Gf2Mat &a, Gf2Mat &inv) {
    uint32_t aug[16];
    const uint32_t mask = (1u << n) - 1u;
    for (int i = 0; i < 16; ++i) aug[i] = i < n ? ((a.r[i] & mask) | (1u << (n + i))) : 0u;
    for (int col = 0; col < n; ++col) {
        int piv = -1;
        for (int row = col; row < n; ++row)
            if (piv < 0 && ((aug[row] >> col) & 1u)) piv = row;
        if (piv < 0) return false;
        const uint32_t t = aug[col];
        aug[col] = aug[piv];
        aug[piv] = t;
        for (int row = 0; row < n; ++row)
            if (row != col && ((aug[row] >> col) & 1u)) aug[row] ^= aug[col];
    }
    for (int i = 0; i < 16; ++i) inv.r[i] = i < n ? (aug[i] >> n) & mask : 0u;
    return true;
}
// L * X for an a x b factor X and an a x a matrix L.
FGK_HD uint32_t gf2_left(const Gf2Mat &L, uint32_t x, int a, int b) {
    const uint32_t rm = (1u << b) - 1u;
    uint32_t y = 0;
    for (int i = 0; i < a; ++i) {
        uint32_t row = 0;
        for (int k = 0; k < a; ++k)
            if ((L.r[i] >> k) & 1u) row ^= (x >> (k * b)) & rm;
        y |= row << (i * b);
    }
    return y;
}
// X * R for an a x b factor X and a b x b matrix R.
FGK_HD uint32_t gf2_right(uint32_t x, const Gf2Mat &R, int a, int b) {
    const uint32_t rm = (1u << b) - 1u;
    uint32_t y = 0;
    for (int i = 0; i < a; ++i) {
        const uint32_t xi = (x >> (i * b)) & rm;
        uint32_t row = 0;
        for (int k = 0; k < b; ++k)
            if ((xi >> k) & 1u) row ^= R.r[k] & rm;
        y |= row << (i * b);
    }
    return y;
}

template <typename WordT, bool kMaskT = false, int kWSv = kWalkersPerGroup>
struct Z2LaneT : LaneBaseT<WordT, kMaskT, kWSv> {
    using B = LaneBaseT<WordT, kMaskT, kWSv>;
    // the two halves of a flip run on the two lanes of a pair
    static constexpr bool kPairSplit = LaneBaseT<WordT, kMaskT, kWSv>::kSplit;
    using typename B::Word;
    using typename B::JoinInfo;
    using B::F;
    using B::m;
    using B::maxm;
    using B::fs;
    using B::side;
    using B::S0;
    using B::S1;
    using B::S2;
    using B::sel_c;
    using B::sel_t1;
    using B::sel_t2;
    using B::f;
    using B::nx;
    using B::copy_term;
    using B::nonzero_term;
    using B::remove_at;
    using B::rebuild;
    using B::dirty0;
    using B::dirty1;
    using B::dirty2;
    using B::mark_row;
    using B::class_walk;
    using B::leave;
    using B::link_join;
    using B::Sc;
    using B::addS;
    using B::pop_update;
    using B::naive_rank;
    using B::aux32;
    using B::perm3;
    using B::nn;
    using B::select_pair;
    using B::cm;
    using B::h_add;
    using B::popc;
    using B::n0;
    using B::n1;
    using B::n2;
    using B::nn0;
    using B::nn1;
    using B::nn2;
    static constexpr uint32_t kValueMask = 0xFFFFFFFFu;
    static constexpr bool kSigned = false;  // factor words are plain masks

    // removeZeroes (scheme_z2.cu:152-164), including its trailing trim and
    // the missing re-check of a swapped-in term.
    FGK_HD void remove_zeroes() {
        while (m > 0 && !nonzero_term(m - 1)) --m;
        for (int t = 0; t < m; ++t) {
            if (!nonzero_term(t)) {
                --m;
                copy_term(t, m);
            }
        }
    }
    // remove_zeroes for the one-lane layout (at most 32 terms), with every
    // row it moves or drops marked in dirty0..2: the terms with a zero factor
    // come from one packed compare per component, then the same trailing
    // trim and swaps, including the unchecked swapped-in term.
    FGK_HD void remove_zero_rows() {
        auto lowmask = [](int n) { return n >= 32 ? 0xFFFFFFFFu : ((1u << n) - 1u); };
        uint32_t Z = 0u;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            uint32_t P[B::kPackedPairs];
#pragma unroll
            for (int p = 0; p < B::kPackedPairs; ++p) P[p] = this->fpair(c, p);
            Z |= ~packed_ne(P, 0u);
        }
        Z &= lowmask(m);
        uint32_t moved = 0u;
        while (m > 0 && ((Z >> (m - 1)) & 1u)) {
            --m;
            moved |= 1u << m;
        }
        for (uint32_t q = Z & lowmask(m); q; q &= lowmask(m)) {
            const int t = __builtin_ctz(q);
            q &= q - 1u;
            --m;
            copy_term(t, m);
            moved |= (1u << t) | (1u << m);
        }
        dirty0 |= moved;
        dirty1 |= moved;
        dirty2 |= moved;
    }

    // reduce (scheme_z2.cu:631-641).  `marked`: every row changed since the
    // classes were last exact is marked in dirty0..2 (a flip's two rows,
    // marked by flip_reduce), so the one-lane layout may update the classes
    // instead of rebuilding them; other callers rebuild.
    FGK_HD void reduce(int c, int a, int b, bool marked = false) {
        f(c, a) ^= f(c, b);
        const bool zero = f(c, a) == 0;
        if constexpr (B::kPacked) {
            if (marked && !zero && !this->track_pop) {
                // term b is replaced by the last term: rows b and the old
                // last change in every component, row a in component c
                const int old_m = m;
                remove_at(b);
                mark_row(c, a);
                const uint32_t moved = (1u << b) | (1u << (old_m - 1));
                dirty0 |= moved;
                dirty1 |= moved;
                dirty2 |= moved;
                commit_rows(old_m);
                return;
            }
            dirty0 = dirty1 = dirty2 = 0u;
        }
        remove_at(b);
        if (zero) remove_zeroes();
        rebuild();
    }

    // ---- one-lane layout: class bookkeeping after a few row changes ----
    // The classes of the scheme after its rows marked in dirty0..2 changed
    // and its rank went from old_m to m, equal to what rebuild() computes.
    // A row y that is neither a changed row D nor in a changed row's old or
    // new class keeps its class and H.  For the others: a changed live row's
    // class E is the packed compare of its new factor; an unchanged row's
    // class is its old class without D plus the changed rows whose E holds
    // it; H(y) = members above y; S changes by the H changes.  Falls back to
    // rebuild() when a component has more than four changed rows.
    FGK_HD void commit_rows(int old_m) {
        const uint32_t live = m >= 32 ? 0xFFFFFFFFu : ((1u << m) - 1u);
        const uint32_t olive = old_m >= 32 ? 0xFFFFFFFFu : ((1u << old_m) - 1u);
        if (popc(dirty0) > 4 || popc(dirty1) > 4 || popc(dirty2) > 4) {
            dirty0 = dirty1 = dirty2 = 0u;
            rebuild();
            return;
        }
        // rows past the rank are zero (the row compares rely on it)
        for (int t = m; t < old_m; ++t) {
            f(0, t) = 0;
            f(1, t) = 0;
            f(2, t) = 0;
        }
        for (int c = 0; c < 3; ++c) {
            const uint32_t D = (c == 0 ? dirty0 : (c == 1 ? dirty1 : dirty2)) & (live | olive);
            if (!D) continue;
            uint32_t P[B::kPackedPairs];
#pragma unroll
            for (int p = 0; p < B::kPackedPairs; ++p) P[p] = this->fpair(c, p);
            // the changed rows (up to four), their new classes and old masks
            int dk[4];
            uint32_t vk[4], ok[4], Ek[4];
            uint32_t q = D;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                dk[e] = q ? __builtin_ctz(q) : -1;
                q &= q - 1u;
                const int d = dk[e] < 0 ? 0 : dk[e];
                vk[e] = f(c, d);
                ok[e] = cm(c, d);
            }
            uint32_t A = D;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const bool has = dk[e] >= 0;
                Ek[e] = has && dk[e] < m ? (~packed_ne(P, vk[e]) & live) : 0u;
                A |= Ek[e] | (has && dk[e] < old_m ? (ok[e] & olive) : 0u);
            }
            int dS = 0;
            for (uint32_t r = A; r; r &= r - 1u) {
                const int y = __builtin_ctz(r);
                const uint32_t above = ~((2u << y) - 1u);
                const uint32_t oldmask = y < old_m ? (cm(c, y) & olive) : 0u;
                uint32_t nw;
                if ((D >> y) & 1u) {
                    nw = 0u;
#pragma unroll
                    for (int e = 0; e < 4; ++e) nw |= dk[e] == y ? Ek[e] : 0u;
                } else {
                    nw = oldmask & ~D;
#pragma unroll
                    for (int e = 0; e < 4; ++e) nw |= dk[e] >= 0 && ((Ek[e] >> y) & 1u) ? (1u << dk[e]) : 0u;
                }
                nw = y < m ? nw : 0u;
                const int hold = popc(oldmask & above), hnew = popc(nw & above);
                cm(c, y) = nw;
                this->h(c, y) = (uint8_t)hnew;
                dS += hnew - hold;
            }
            addS(c, dS);
        }
        dirty0 = dirty1 = dirty2 = 0u;
#if defined(FGK_CHECK_COMMIT) && !defined(__HIP_DEVICE_COMPILE__)
        // host emulation built with FGK_CHECK_COMMIT: the update must equal
        // a full rebuild (masks, H and pair counts of every live row)
        uint32_t cmv[3][32], hv[3][32];
        const uint32_t Sv[3] = {S0, S1, S2};
        for (int c = 0; c < 3; ++c)
            for (int t = 0; t < m; ++t) {
                cmv[c][t] = cm(c, t);
                hv[c][t] = this->h(c, t);
            }
        rebuild();
        bool same = Sv[0] == S0 && Sv[1] == S1 && Sv[2] == S2;
        for (int c = 0; c < 3; ++c)
            for (int t = 0; t < m; ++t) same = same && cmv[c][t] == cm(c, t) && hv[c][t] == this->h(c, t);
        if (!same) this->fail(kErrCommit);
#endif
    }

    // flip (scheme_z2.cu:531-577) taken literally; used when a zero factor is
    // involved.  Partner bookkeeping is rebuilt afterwards.
    FGK_HD void flip_literal(int i, int j, int k, int i1, int i2, bool check_reduce) {
        f(j, i1) ^= f(j, i2);
        f(k, i2) ^= f(k, i1);
        if (!f(j, i1) || !f(k, i2)) {
            if constexpr (B::kPacked) {
                if (!this->track_pop) {
                    // the two changed rows and the rows removeZeroes moves
                    // or drops: the classes are updated, not rebuilt
                    mark_row(j, i1);
                    mark_row(k, i2);
                    const int old_m = m;
                    remove_zero_rows();
                    commit_rows(old_m);
                    return;
                }
                dirty0 = dirty1 = dirty2 = 0u;
            }
            remove_zeroes();
            rebuild();
            return;
        }
        for (int t = 0; t < m; ++t) {
            if (t != i1 && f(j, t) == f(j, i1) && check_reduce) {
                if (f(i, t) == f(i, i1)) { reduce(k, t, i1); return; }
                if (f(k, t) == f(k, i1)) { reduce(i, t, i1); return; }
            }
            if (t != i2 && f(k, t) == f(k, i2) && check_reduce) {
                if (f(i, t) == f(i, i2)) { reduce(j, t, i2); return; }
                if (f(j, t) == f(j, i2)) { reduce(i, t, i2); return; }
            }
        }
        rebuild();
    }

    // One half of a flip with every involved factor non-zero: term x of
    // component mc gets the new factor nv.  The self-matches the reference
    // excludes with `index != index1/2` cannot occur, so one scan of component
    // mc finds a member of the new class; one walk of that class collects the
    // join data and the first reduction candidate (rule 1: component i also
    // equals fi, rule 2: component oc also equals ov).
    struct Half {
        int any, t, rule;
        JoinInfo ji;
    };
    FGK_HD Half flip_half(int mc, int x, uint32_t nv, int i, int oc, uint32_t fi, uint32_t ov, bool check_reduce) {
        Half hf;
        hf.any = -1;
        hf.t = 1 << 30;
        hf.rule = 0;
        hf.any = this->find_last(mc, nv);
        this->mark(2);
        if (hf.any >= 0)
            class_walk(mc, x, hf.any, i, oc, hf.ji,
                       [&](uint32_t a, uint32_t b) { return !check_reduce ? 0 : (a == fi ? 1 : (b == ov ? 2 : 0)); },
                       hf.t, hf.rule);
        this->mark(3);
        return hf;
    }

    // flip (scheme_z2.cu:531-577) when every involved factor is non-zero.
    // Side 0 of the lane pair does the index1 half (component j, new factor
    // A), side 1 the index2 half (component k, new factor B); the host
    // emulation does both.  Reduction order as in the reference: the lowest
    // term first, the index1 half before the index2 half at equal terms.
    FGK_HD void flip(int i, int j, int k, int i1, int i2, bool check_reduce) {
        this->rec(0, i, j, k, i1, i2, check_reduce ? 1 : 0, 0);
        const uint32_t fj1 = f(j, i1), fj2 = f(j, i2);
        const uint32_t fk1 = f(k, i1), fk2 = f(k, i2);
        const uint32_t A = fj1 ^ fj2;
        const uint32_t B = fk2 ^ fk1;
        if (!(fj1 && fj2 && fk1 && fk2 && A && B)) {
            flip_literal(i, j, k, i1, i2, check_reduce);
            this->mark(9);
            return;
        }
        if constexpr (kMaskT) {
            flip_masks(i, j, k, i1, i2, A, B, fj1, fk2, check_reduce);
            return;
        }
        const uint32_t fi = f(i, i1);
        if constexpr (kPairSplit) {
            const int mc = side ? k : j, x = side ? i2 : i1, oc = side ? j : k;
            const uint32_t nv = side ? B : A, ov = side ? fj2 : fk1;
            const Half mine = flip_half(mc, x, nv, i, oc, fi, ov, check_reduce);
            const int ot = pair_exchange(mine.t, side), orule = pair_exchange(mine.rule, side);
            const int tA = side ? ot : mine.t, rA = side ? orule : mine.rule;
            const int tB = side ? mine.t : ot, rB = side ? mine.rule : orule;
            this->mark(4);
            if (rA || rB) {
                flip_reduce(i, j, k, i1, i2, A, B, tA, rA, tB, rB);
                return;
            }
            pop_update(fj1, A, fk2, B);
            const uint32_t before = Sc(mc);
            leave(mc, x);
            if (mine.any >= 0) link_join(mc, x, mine.ji);
            f(mc, x) = nv;
            addS(oc, pair_exchange((int)(Sc(mc) - before), side));
            this->mark(5);
        } else {
            const Half ha = flip_half(j, i1, A, i, k, fi, fk1, check_reduce);
            const Half hb = flip_half(k, i2, B, i, j, fi, fj2, check_reduce);
            if (ha.rule || hb.rule) {
                flip_reduce(i, j, k, i1, i2, A, B, ha.t, ha.rule, hb.t, hb.rule);
                return;
            }
            pop_update(fj1, A, fk2, B);
            leave(j, i1);
            if (ha.any >= 0) link_join(j, i1, ha.ji);
            f(j, i1) = A;
            leave(k, i2);
            if (hb.any >= 0) link_join(k, i2, hb.ji);
            f(k, i2) = B;
        }
    }
    // The reduction a flip found (first match: the lowest term, the index1
    // half before the index2 half at equal terms), after the flip's updates.
    FGK_HD void flip_reduce(int i, int j, int k, int i1, int i2, uint32_t A, uint32_t Bv, int tA, int rA, int tB,
                            int rB) {
        f(j, i1) = A;
        f(k, i2) = Bv;
        if constexpr (B::kPacked) {
            mark_row(j, i1);
            mark_row(k, i2);
        }
        if (rA && tA <= tB) reduce(rA == 1 ? k : i, tA, i1, true);
        else reduce(rB == 1 ? j : i, tB, i2, true);
        this->mark(8);
    }

    // flip with class masks (every involved factor non-zero).  A half of the
    // flip (term x of component mc gets factor nv) reads x's masks before
    // anything changes: its old class in mc and its classes in i and in the
    // other changed component oc; the new class is the mask of any member
    // found by the scan.  The half's reduction candidates are the new class
    // members that share component i with x (rule 1) or component oc (rule 2);
    // the lowest one is its first match (scheme_z2.cu:544-576).  The flip's
    // other changed term is never in either new class (its factor there
    // differs from the new one), so masks read before the flip are the ones
    // the reference's scan sees after it.  Without a reduction, x leaves its
    // old class and joins the new one: one mask store per member and H counts
    // adjusted with stores that return nothing, no walk along the class.
    struct MaskHalf {
        uint32_t old, nw;
        int t, rule;
    };
    FGK_HD MaskHalf mask_half(int mc, int x, uint32_t nv, int i, int oc, bool check_reduce) {
        MaskHalf h;
        h.old = cm(mc, x);
        const uint32_t ci = cm(i, x), co = cm(oc, x);
        const int z = this->find_last(mc, nv);
        this->mark(2);
        h.nw = z >= 0 ? cm(mc, z) : 0u;
        const uint32_t red = check_reduce ? (h.nw & (ci | co)) : 0u;
        h.t = red ? __builtin_ctz(red) : (1 << 30);
        h.rule = red ? (((ci >> h.t) & 1u) ? 1 : 2) : 0;
        this->mark(3);
        return h;
    }
    // x leaves class `old` of component mc and joins `nw`; its factor becomes
    // nv.  Returns the change of component mc's pair count.
    FGK_HD int mask_move(int mc, int x, uint32_t old, uint32_t nw, uint32_t nv) {
        const uint32_t xb = 1u << x, below = xb - 1u;
        const uint32_t rest = old & ~xb, joined = nw | xb;
        for (uint32_t q = rest | nw; q; q &= q - 1u) {
            const int y = __builtin_ctz(q);
            const bool in_new = (nw >> y) & 1u;
            cm(mc, y) = in_new ? joined : rest;
            if ((below >> y) & 1u) h_add(mc, y, in_new ? 1 : -1);
        }
        cm(mc, x) = joined;
        this->h(mc, x) = (uint8_t)popc(nw & ~(below | xb));
        f(mc, x) = (Word)nv;
        return popc(nw) - popc(rest);
    }
    FGK_HD void flip_masks(int i, int j, int k, int i1, int i2, uint32_t A, uint32_t Bv, uint32_t fj1, uint32_t fk2,
                           bool check_reduce) {
        int dj = 0, dk = 0;
        if constexpr (kPairSplit) {
            const int mc = side ? k : j, x = side ? i2 : i1, oc = side ? j : k;
            const uint32_t nv = side ? Bv : A;
            const MaskHalf mine = mask_half(mc, x, nv, i, oc, check_reduce);
            const int ot = pair_exchange(mine.t, side), orule = pair_exchange(mine.rule, side);
            const int tA = side ? ot : mine.t, rA = side ? orule : mine.rule;
            const int tB = side ? mine.t : ot, rB = side ? mine.rule : orule;
            this->mark(4);
            if (rA || rB) {
                flip_reduce(i, j, k, i1, i2, A, Bv, tA, rA, tB, rB);
                return;
            }
            pop_update(fj1, A, fk2, Bv);
            const int d = mask_move(mc, x, mine.old, mine.nw, nv);
            const int od = pair_exchange(d, side);
            dj = side ? od : d;
            dk = side ? d : od;
            this->mark(5);
        } else {
            const MaskHalf ha = mask_half(j, i1, A, i, k, check_reduce);
            const MaskHalf hb = mask_half(k, i2, Bv, i, j, check_reduce);
            if (ha.rule || hb.rule) {
                flip_reduce(i, j, k, i1, i2, A, Bv, ha.t, ha.rule, hb.t, hb.rule);
                return;
            }
            pop_update(fj1, A, fk2, Bv);
            dj = mask_move(j, i1, ha.old, ha.nw, A);
            dk = mask_move(k, i2, hb.old, hb.nw, Bv);
        }
        // component i keeps its classes; j and k change by dj and dk
        S0 += (uint32_t)(j == 0 ? dj : (k == 0 ? dk : 0));
        S1 += (uint32_t)(j == 1 ? dj : (k == 1 ? dk : 0));
        S2 += (uint32_t)(j == 2 ? dj : (k == 2 ? dk : 0));
    }

    // ---- reduction-seeking lookahead proposals (no reference counterpart) ----
    // Whether flip(i, j, k, i1, i2) would lower the rank: a zero factor, or a
    // term that afterwards matches i1 or i2 in two components (the reduction
    // flip() takes).  Read-only.  One half per lane of the pair, as in flip.
    FGK_HD bool half_reduces(int mc, int x, uint32_t nv, int i, int oc, uint32_t fi, uint32_t ov) {
        const int z = this->find_last(mc, nv);
        if (z < 0) return false;
        if constexpr (kMaskT) {
            return (cm(mc, z) & (cm(i, x) | cm(oc, x))) != 0u;
        } else {
            int y = z, steps = 0;
            do {
                if (f(i, y) == fi || f(oc, y) == ov) return true;
                y = nx(mc, y);
            } while (y != z && ++steps <= m);
            return false;
        }
    }
    FGK_HD bool flip_reduces(int i, int j, int k, int i1, int i2) {
        const uint32_t fj1 = f(j, i1), fj2 = f(j, i2), fk1 = f(k, i1), fk2 = f(k, i2);
        const uint32_t A = fj1 ^ fj2, B = fk2 ^ fk1;
        if (!(fj1 && fj2 && fk1 && fk2 && A && B)) return true;
        const uint32_t fi = f(i, i1);
        if constexpr (kPairSplit) {
            const int mc = side ? k : j, x = side ? i2 : i1, oc = side ? j : k;
            const int mine = half_reduces(mc, x, side ? B : A, i, oc, fi, side ? fj2 : fk1) ? 1 : 0;
            return (mine | pair_exchange(mine, side)) != 0;
        } else {
            return half_reduces(j, i1, A, i, k, fi, fk1) || half_reduces(k, i2, B, i, j, fi, fj2);
        }
    }
    // `count` candidates drawn as tryFlip draws them (an aux word for the
    // pair in fresh-list order, an aux word for the orientation coins); the
    // first whose flip lowers the rank is applied.  False when none does.
    FGK_HD bool try_lookahead(int count, bool check_reduce) {
        const uint32_t total = S0 + S1 + S2;
        if (!total) return false;
        for (int q = 0; q < count; ++q) {
            const uint32_t x = aux32(), coin = aux32();
            int c, t1, t2, j, k;
            if (!select_pair(draw_below(x, total), c, t1, t2)) return false;
            orientation(c, coin, j, k);
            if (coin & 2u) {
                const int t = t1;
                t1 = t2;
                t2 = t;
            }
            if (flip_reduces(c, j, k, t1, t2)) {
                sel_c = c;
                sel_t1 = t1 < t2 ? t1 : t2;
                sel_t2 = t1 < t2 ? t2 : t1;
                flip(c, j, k, t1, t2, check_reduce);
                return true;
            }
        }
        return false;
    }

    // plus (scheme_z2.cu:579-612)
    FGK_HD void plus(int i, int j, int k, int i1, int i2, int variant) {
        this->rec(1, i, j, k, i1, i2, variant, 0);
        const uint32_t a1 = f(i, i1), b1 = f(j, i1), c1 = f(k, i1);
        const uint32_t a2 = f(i, i2), b2 = f(j, i2), c2 = f(k, i2);
        const uint32_t a = a1 ^ a2, b = b1 ^ b2, c = c1 ^ c2;
        uint32_t u, v, w;
        if (variant == 0) {
            f(j, i1) = b; f(i, i2) = a;
            u = a1; v = b2; w = c;
        } else if (variant == 1) {
            f(k, i1) = c; f(j, i2) = b;
            u = a; v = b1; w = c2;
        } else {
            f(i, i1) = a; f(k, i2) = c;
            u = a2; v = b; w = c1;
        }
        f(i, m) = u; f(j, m) = v; f(k, m) = w;
        ++m;
        if constexpr (B::kPacked) {
            if (a && b && c && !this->track_pop) {
                // the two changed rows and the new term's row in every component
                mark_row(variant == 0 ? j : (variant == 1 ? k : i), i1);
                mark_row(variant == 0 ? i : (variant == 1 ? j : k), i2);
                const uint32_t nb = 1u << (m - 1);
                dirty0 |= nb;
                dirty1 |= nb;
                dirty2 |= nb;
                commit_rows(m - 1);
                return;
            }
            dirty0 = dirty1 = dirty2 = 0u;
        }
        if (!a || !b || !c) remove_zeroes();
        rebuild();
    }

    // split (scheme_z2.cu:614-619)
    FGK_HD void split(int i, int j, int k, int index, uint32_t x1) {
        this->rec(2, i, j, k, index, (int)x1, 0, 0);
        const uint32_t x2 = f(i, index) ^ x1;
        f(i, index) = x1;
        f(i, m) = x2;
        f(j, m) = f(j, index);
        f(k, m) = f(k, index);
        ++m;
        if constexpr (B::kPacked) {
            if (x1 && x2 && !this->track_pop) {
                mark_row(i, index);
                const uint32_t nb = 1u << (m - 1);
                dirty0 |= nb;
                dirty1 |= nb;
                dirty2 |= nb;
                commit_rows(m - 1);
                return;
            }
            dirty0 = dirty1 = dirty2 = 0u;
        }
        rebuild();
    }

    FGK_HD bool can_grow() const { return m < kMaxRank && m < naive_rank() && m < maxm; }

    FGK_HD bool try_plus() {  // scheme_z2.cu:1029-1050
        if (!can_grow()) return false;
        const uint32_t mm = (uint32_t)m;
        int i1 = (int)(aux32() % mm);
        int i2 = (int)(aux32() % mm);
        int tries = 0;
        while (i1 == i2 || f(0, i1) == f(0, i2) || f(1, i1) == f(1, i2) || f(2, i1) == f(2, i2)) {
            if (++tries > kPlusMaxTries) return false;
            i1 = (int)(aux32() % mm);
            i2 = (int)(aux32() % mm);
        }
        int p0, p1, p2;
        perm3(p0, p1, p2);
        const int variant = (int)(aux32() % 3u);
        plus(p0, p1, p2, i1, i2, variant);
        return true;
    }

    FGK_HD uint32_t rand_factor(int bits) {
        const uint32_t x = aux32();
        return bits >= 32 ? x : (x % (1u << bits));
    }

    FGK_HD bool try_split() {  // scheme_z2.cu:1052-1078
        if (!can_grow()) return false;
        const int index = (int)(aux32() % (uint32_t)m);
        int i, j, k;
        perm3(i, j, k);
        uint32_t x1 = rand_factor(nn(i));
        const uint32_t x2 = f(i, index);
        int loop = 0;
        while (loop < 5 && x1 == x2) {
            x1 = rand_factor(nn(i));
            ++loop;
        }
        if (loop == 5) return false;
        split(i, j, k, index, x1);
        return true;
    }

    FGK_HD bool try_split_existed() {  // scheme_z2.cu:1080-1104
        if (!can_grow()) return false;
        int i, j, k;
        perm3(i, j, k);
        const uint32_t mm = (uint32_t)m;
        const int i1 = (int)(aux32() % mm);
        int i2 = (int)(aux32() % mm);
        int loop = 0;
        while (loop < 5 && (i1 == i2 || f(i, i1) == f(i, i2))) {
            i2 = (int)(aux32() % mm);
            ++loop;
        }
        if (loop == 5) return false;
        split(i, j, k, i1, f(i, i2));
        return true;
    }

    FGK_HD bool try_expand(int count) {  // scheme_z2.cu:1106-1125
        const int maxr = naive_rank();
        bool res = false;
        for (int q = 0; q < count && m < maxr; ++q) {
            const uint32_t v = aux32() % 3u;
            if (v == 0) res |= try_plus();
            else if (v == 1) res |= try_split();
            else res |= try_split_existed();
        }
        return res;
    }

    // tryReduce (scheme_z2.cu:1127-1154) over the fresh-list order.
    FGK_HD bool try_reduce() {
        if constexpr (kMaskT) {
            for (int a = 0; a < m; ++a)
                for (uint32_t q = cm(0, a) & ~((2u << a) - 1u); q; q &= q - 1u) {
                    const int y = __builtin_ctz(q);
                    if (f(1, a) == f(1, y)) { reduce(2, a, y); return true; }
                    if (f(2, a) == f(2, y)) { reduce(1, a, y); return true; }
                }
            for (int a = 0; a < m; ++a)
                for (uint32_t q = cm(1, a) & ~((2u << a) - 1u); q; q &= q - 1u) {
                    const int y = __builtin_ctz(q);
                    if (f(2, a) == f(2, y)) { reduce(0, a, y); return true; }
                }
            return false;
        }
        for (int a = 0; a < m; ++a) {
            int g = 0;
            for (int y = nx(0, a); y > a && g <= m; y = nx(0, y), ++g) {
                if (f(1, a) == f(1, y)) { reduce(2, a, y); return true; }
                if (f(2, a) == f(2, y)) { reduce(1, a, y); return true; }
            }
        }
        for (int a = 0; a < m; ++a) {
            int g = 0;
            for (int y = nx(1, a); y > a && g <= m; y = nx(1, y), ++g) {
                if (f(2, a) == f(2, y)) { reduce(0, a, y); return true; }
            }
        }
        return false;
    }

    // tryFlip of the one-lane layout: the same pair, orientation and flip as
    // the general path below (flip_masks), with the LDS reads of an iteration
    // in four dependent batches: (1) the H words of the pair's component and
    // the packed factor rows of the two components the flip changes, both
    // known from the draw alone; (2) t1's class mask, which gives t2; (3) the
    // factors and class masks of index1 and index2 the halves need; (4) the
    // class stores.  The new classes come from the rows of (1) compared with
    // the new factors (packed_ne) instead of a member's class mask.
    // `pre(phase)` is called three times per call, with phase 0, 1 and 2 as
    // std::integral_constant, once after each of the three batches of reads
    // is issued (on every return path, so all three always run), so the
    // caller's work that needs no scheme state (the next iteration's Philox
    // block, in three parts) fills the waits for them.
    template <class Pre>
    FGK_HD bool try_flip_packed(uint32_t r, uint32_t coin, bool check_reduce, Pre &&pre) {
        const uint32_t total = S0 + S1 + S2;
        if (!total) {
            pre_all(pre);
            return false;
        }
        uint32_t rr = draw_below(r, total);
        int c = 0;
        if (rr >= S0) {
            rr -= S0;
            c = 1;
            if (rr >= S1) {
                rr -= S1;
                c = 2;
            }
        }
        int j, k;
        orientation(c, coin, j, k);
        uint32_t hw8[8], Rj[16], Rk[16];
#pragma unroll
        for (int e = 0; e < 8; ++e) hw8[e] = this->hword(c, e);
#pragma unroll
        for (int p = 0; p < 16; ++p) {
            Rj[p] = this->fpair(j, p);
            Rk[p] = this->fpair(k, p);
        }
        // every read of the first batch is issued before `pre` and the
        // selection (the scheduler would otherwise sink the row reads below
        // the selection, behind the H words' wait)
        sched_fence();
        pre(std::integral_constant<int, 0>{});
        sched_fence();
        int t1;
        B::select_words(hw8, rr, t1);
        // the (rr + 1)-th member above t1 (t1 clamped into the plane for the
        // read; an out-of-range t1 or an empty q is reported below)
        const int t1c = t1 < m ? t1 : 0;
        const uint32_t cmt1 = cm(c, t1c);
        sched_fence();
        pre(std::integral_constant<int, 1>{});
        sched_fence();
        // clear the rr lowest members: two steps as selects, more (rare)
        // in a loop (clearing from an empty q leaves it empty)
        uint32_t q = cmt1 & ~((2u << t1c) - 1u);
        const uint32_t q1 = q & (q - 1u), q2 = q1 & (q1 - 1u);
        q = rr == 0u ? q : (rr == 1u ? q1 : q2);
        if (rr > 2u)
            for (uint32_t e = 2; e < rr && q; ++e) q &= q - 1u;
        if (t1 >= m || !q) {
            this->fail(kErrSelect);
            pre(std::integral_constant<int, 2>{});
            return false;
        }
        const int t2 = __builtin_ctz(q);
        this->mark(1);
        sel_c = c;
        sel_t1 = t1;
        sel_t2 = t2;
        const int i = c;
        const bool sw = (coin & 2u) != 0;
        const int i1 = sw ? t2 : t1, i2 = sw ? t1 : t2;
        this->rec(0, i, j, k, i1, i2, check_reduce ? 1 : 0, 0);
        // the reads of index1 and index2: their factor offsets in a packed
        // component column and their class-mask rows, computed once
        const int o1 = B::packed_off(i1), o2 = B::packed_off(i2);
        const Word *Fj = this->fcol(j), *Fk = this->fcol(k);
        const uint32_t fj1 = Fj[o1], fj2 = Fj[o2], fk1 = Fk[o1], fk2 = Fk[o2];
        const uint32_t *Cj = &cm(j, 0), *Ck = &cm(k, 0);
        const uint32_t oldj = Cj[i1 * B::kWS], ck1 = Ck[i1 * B::kWS];
        const uint32_t oldk = Ck[i2 * B::kWS], cj2 = Cj[i2 * B::kWS];
        // index1 and index2 share their component-i class: t1's mask
        const uint32_t ci1 = cmt1, ci2 = cmt1;
        sched_fence();
        pre(std::integral_constant<int, 2>{});
        sched_fence();
        // one round trip for all eight reads: the masks are read before the
        // zero-factor test, not after it
        keep_loaded(fj1, fj2, fk1, fk2, oldj, ck1, oldk, cj2);
        const uint32_t A = fj1 ^ fj2, Bv = fk2 ^ fk1;
        const uint32_t low = min_u32(min_u32(min_u32(fj1, fj2), min_u32(fk1, fk2)), min_u32(A, Bv));
        if (low == 0u) {
            flip_literal(i, j, k, i1, i2, check_reduce);
            this->mark(9);
            return true;
        }
        // new classes: the rows equal to the new factors (rows from the rank
        // up are zero, new factors are not); neither changed term is in them
        const uint32_t nwj = ~packed_ne(Rj, A), nwk = ~packed_ne(Rk, Bv);
        this->mark(2);
        const uint32_t redj = check_reduce ? (nwj & (ci1 | ck1)) : 0u;
        const uint32_t redk = check_reduce ? (nwk & (ci2 | cj2)) : 0u;
        if (redj || redk) {
            const int tA = redj ? __builtin_ctz(redj) : (1 << 30);
            const int rA = redj ? (((ci1 >> tA) & 1u) ? 1 : 2) : 0;
            const int tB = redk ? __builtin_ctz(redk) : (1 << 30);
            const int rB = redk ? (((ci2 >> tB) & 1u) ? 1 : 2) : 0;
            flip_reduce(i, j, k, i1, i2, A, Bv, tA, rA, tB, rB);
            return true;
        }
        this->mark(3);
        pop_update(fj1, A, fk2, Bv);
        int dj, dk;
        mask_move2(j, i1, oldj, nwj, A, k, i2, oldk, nwk, Bv, dj, dk);
        // the pair-count changes of components j and k: d_lo goes to the
        // lower of the two components other than c (j unless coin bit 0
        // swapped j and k), d_hi to the higher; the choices are bit masks
        // (0 or all ones) from the bits of c and the coin, not selects on
        // compare results
        const uint32_t ms = 0u - (coin & 1u), cu = (uint32_t)c;
        const uint32_t ujd = (uint32_t)dj, ukd = (uint32_t)dk;
        const uint32_t d_lo = (ukd & ms) | (ujd & ~ms), d_hi = (ujd & ms) | (ukd & ~ms);
        const uint32_t m_ne0 = 0u - ((cu | (cu >> 1)) & 1u), m_is2 = 0u - (cu >> 1);
        S0 += d_lo & m_ne0;
        S1 += (d_lo & ~m_ne0) | (d_hi & m_is2);
        S2 += d_hi & ~m_is2;
        this->mark(5);
        return true;
    }

    // Values that must be at hand at this point (GPU: an empty asm that
    // reads them, so their loads or computation are not sunk past it).
    template <class... T>
    FGK_HD static void keep_loaded(T... v) {
#if defined(__HIP_DEVICE_COMPILE__)
        (asm_use(v), ...);
#else
        ((void)v, ...);
#endif
    }
    FGK_HD static void asm_use(uint32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
        asm volatile("" ::"v"(v));
#else
        (void)v;
#endif
    }
    // mask_move for both halves of a flip in one loop: each trip moves one
    // member of each half's list (old class without x, new class); a list
    // already empty repeats its own term x, with a class-mask store the
    // final one overwrites and an H add of 0.  The wave loops over the
    // longer list of each lane instead of both lists one after the other.
    FGK_HD void mask_move2(int j, int xj, uint32_t oldj, uint32_t nwj, uint32_t vj, int k, int xk, uint32_t oldk,
                           uint32_t nwk, uint32_t vk, int &dj, int &dk) {
        const uint32_t bj = 1u << xj, belj = bj - 1u, restj = oldj & ~bj, joinj = nwj | bj;
        const uint32_t bk = 1u << xk, belk = bk - 1u, restk = oldk & ~bk, joink = nwk | bk;
        uint32_t qj = restj | nwj, qk = restk | nwk;
        // per-component bases, so a member's class-mask and H addresses are
        // one shift-add each
        uint32_t *Cj = &cm(j, 0), *Ck = &cm(k, 0);
        uint32_t *Hj = &this->hword(j, 0), *Hk = &this->hword(k, 0);
        while (qj | qk) {
            const int yj = qj ? __builtin_ctz(qj) : xj;
            const int yk = qk ? __builtin_ctz(qk) : xk;
            // all-ones when the member is in the new class / below x
            const uint32_t mnj = B::sbit(nwj, yj), mnk = B::sbit(nwk, yk);
            const uint32_t mbj = B::sbit(belj, yj), mbk = B::sbit(belk, yk);
            Cj[yj * B::kWS] = (mnj & joinj) | (~mnj & restj);
            Ck[yk * B::kWS] = (mnk & joink) | (~mnk & restk);
            // H change of the member: +1 joining below x, -1 leaving below
            // x, 0 above x
            B::word_add(Hj + (yj >> 2) * B::kWS, ((~mnj | 1u) & mbj) << (8 * (yj & 3)));
            B::word_add(Hk + (yk >> 2) * B::kWS, ((~mnk | 1u) & mbk) << (8 * (yk & 3)));
            qj &= qj - 1u;
            qk &= qk - 1u;
        }
        cm(j, xj) = joinj;
        cm(k, xk) = joink;
        this->h(j, xj) = (uint8_t)popc(nwj & ~(belj | bj));
        this->h(k, xk) = (uint8_t)popc(nwk & ~(belk | bk));
        f(j, xj) = (Word)vj;
        f(k, xk) = (Word)vk;
        dj = popc(nwj) - popc(restj);
        dk = popc(nwk) - popc(restk);
    }

    // tryFlip (scheme_z2.cu:982-1027): r picks the pair in fresh-list order,
    // coin bit 0 swaps j/k, coin bit 1 swaps the indices.  `pre` runs once per
    // call (see try_flip_packed).
    template <class Pre>
    FGK_HD bool try_flip(uint32_t r, uint32_t coin, bool check_reduce, Pre &&pre) {
        if constexpr (B::kPacked) {
            return try_flip_packed(r, coin, check_reduce, pre);
        } else {
            pre_all(pre);
            return try_flip(r, coin, check_reduce);
        }
    }
    // the three phases of `pre` at once
    template <class Pre>
    FGK_HD static void pre_all(Pre &&pre) {
        pre(std::integral_constant<int, 0>{});
        pre(std::integral_constant<int, 1>{});
        pre(std::integral_constant<int, 2>{});
    }
    FGK_HD bool try_flip(uint32_t r, uint32_t coin, bool check_reduce) {
        if constexpr (B::kPacked) return try_flip_packed(r, coin, check_reduce, [](auto) {});
        const uint32_t total = S0 + S1 + S2;
        if (!total) return false;
        int c, t1, t2, j, k;
        if (!select_pair(draw_below(r, total), c, t1, t2)) return false;
        this->mark(1);
        sel_c = c;
        sel_t1 = t1;
        sel_t2 = t2;
        orientation(c, coin, j, k);
        if (coin & 2u) {
            const int t = t1;
            t1 = t2;
            t2 = t;
        }
        flip(c, j, k, t1, t2, check_reduce);
        return true;
    }

    FGK_HD void put(int c, int t, uint32_t v) { f(c, t) = (Word)v; }

    // ---- sandwiching (scheme_z2.cu:444-528, 1277-1300) ----
    // invertibleMatrixZ2 (:488-492): a random n x n matrix (row i = the low n
    // bits of one aux draw, random.cuh:24-32) until it is invertible; the
    // redraws are bounded.
    FGK_HD bool random_invertible(int n, Gf2Mat &a, Gf2Mat &inv) {
        const uint32_t mask = (1u << n) - 1u;
        for (int tries = 0; tries < kPlusMaxTries; ++tries) {
            for (int i = 0; i < 16; ++i) a.r[i] = i < n ? (aux32() & mask) : 0u;
            if (gf2_invert(n, a, inv)) return true;
        }
        return false;
    }
    // sandwiching (:1277-1300): U -> A U B^-1, V -> B V C^-1, W -> C W A^-1
    // for random invertible A (n0), B (n1), C (n2), drawn in that order.  Each
    // matrix is applied as soon as it is drawn (the products associate), so
    // only one matrix and its inverse are live.
    FGK_HD void sandwiching() {
        Gf2Mat a, ai;
        if (!random_invertible(n0, a, ai)) return;
        for (int t = 0; t < m; ++t) {
            put(0, t, gf2_left(a, f(0, t), n0, n1));
            put(2, t, gf2_right(f(2, t), ai, n2, n0));
        }
        if (!random_invertible(n1, a, ai)) {
            rebuild();
            return;
        }
        for (int t = 0; t < m; ++t) {
            put(0, t, gf2_right(f(0, t), ai, n0, n1));
            put(1, t, gf2_left(a, f(1, t), n1, n2));
        }
        if (random_invertible(n2, a, ai)) {
            for (int t = 0; t < m; ++t) {
                put(1, t, gf2_right(f(1, t), ai, n1, n2));
                put(2, t, gf2_left(a, f(2, t), n2, n0));
            }
        }
        rebuild();
    }
    // A given pair (component c, index1 t1 < index2 t2) flipped with the
    // orientation coins of tryFlip; false when the factors differ.
    FGK_HD bool try_flip_pair(int c, int t1, int t2, uint32_t coin, bool check_reduce) {
        if (c < 0 || c > 2 || t1 < 0 || t2 <= t1 || t2 >= m || f(c, t1) != f(c, t2)) return false;
        sel_c = c;
        sel_t1 = t1;
        sel_t2 = t2;
        int j, k;
        orientation(c, coin, j, k);
        if (coin & 2u) {
            const int t = t1;
            t1 = t2;
            t2 = t;
        }
        flip(c, j, k, t1, t2, check_reduce);
        return true;
    }
};

using Z2Lane = Z2LaneT<uint32_t>;    // Z2 factors of up to 32 coefficients
using Z2Lane16 = Z2LaneT<uint16_t>;  // Z2 factors of up to 16 coefficients
using Z2LaneM = Z2LaneT<uint16_t, true>;  // up to 16 coefficients and 32 terms, class masks
using Z2LaneS = Z2LaneT<uint16_t, true, kSoloWalkersPerGroup>;  // the same, one lane per walker

}  // namespace lane
}  // namespace fgk

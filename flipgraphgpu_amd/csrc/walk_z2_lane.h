// Z2 operators of the walk (src/schemes/scheme_z2.cu) on the LDS lane state
// of lane_common.h.  __host__ __device__ so the tests can run the exact
// kernel code on the host.
#pragma once

#include "lane_common.h"

namespace fgk {
namespace lane {

struct Z2Lane : LaneBase {
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
    // reduce (scheme_z2.cu:631-641)
    FGK_HD void reduce(int c, int a, int b) {
        f(c, a) ^= f(c, b);
        const bool zero = f(c, a) == 0;
        remove_at(b);
        if (zero) remove_zeroes();
        rebuild();
    }

    // flip (scheme_z2.cu:531-577) taken literally; used when a zero factor is
    // involved.  Partner bookkeeping is rebuilt afterwards.
    FGK_HD void flip_literal(int i, int j, int k, int i1, int i2, bool check_reduce) {
        f(j, i1) ^= f(j, i2);
        f(k, i2) ^= f(k, i1);
        if (!f(j, i1) || !f(k, i2)) {
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

    // flip with every involved factor non-zero: the self-matches the
    // reference excludes with `index != index1/2` cannot occur, so one scan
    // finds a member of each new class and the classes give the rest.
    FGK_HD void flip(int i, int j, int k, int i1, int i2, bool check_reduce) {
        const uint32_t fj1 = f(j, i1), fj2 = f(j, i2);
        const uint32_t fk1 = f(k, i1), fk2 = f(k, i2);
        const uint32_t A = fj1 ^ fj2;
        const uint32_t B = fk2 ^ fk1;
        if (!(fj1 && fj2 && fk1 && fk2 && A && B)) {
            flip_literal(i, j, k, i1, i2, check_reduce);
            return;
        }
        uint32_t *Fj = F + j * maxm * WG;
        uint32_t *Fk = F + k * maxm * WG;
        int anyA = -1, anyB = -1;
#pragma unroll 8
        for (int t = 0; t < m; ++t) {
            const uint32_t a = Fj[t * WG];
            const uint32_t b = Fk[t * WG];
            anyA = a == A ? t : anyA;
            anyB = b == B ? t : anyB;
        }
        // one walk per new class: reduction candidates (reference order:
        // lowest member first, index1's side before index2's) and join data
        const uint32_t fi = f(i, i1);
        int tA = 1 << 30, cA = 0, tB = 1 << 30, cB = 0;
        JoinInfo jA, jB;
        if (anyA >= 0)
            class_walk(j, i1, anyA, jA,
                       [&](int y) { return !check_reduce ? 0 : (f(i, y) == fi ? 1 : (f(k, y) == fk1 ? 2 : 0)); },
                       tA, cA);
        if (anyB >= 0)
            class_walk(k, i2, anyB, jB,
                       [&](int y) { return !check_reduce ? 0 : (f(i, y) == fi ? 1 : (f(j, y) == fj2 ? 2 : 0)); },
                       tB, cB);
        if (cA || cB) {
            f(j, i1) = A;
            f(k, i2) = B;
            if (cA && tA <= tB) reduce(cA == 1 ? k : i, tA, i1);
            else reduce(cB == 1 ? j : i, tB, i2);
            return;
        }
        leave(j, i1);
        if (anyA >= 0) link_join(j, i1, jA);
        leave(k, i2);
        if (anyB >= 0) link_join(k, i2, jB);
        f(j, i1) = A;
        f(k, i2) = B;
    }

    // plus (scheme_z2.cu:579-612)
    FGK_HD void plus(int i, int j, int k, int i1, int i2, int variant) {
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
        if (!a || !b || !c) remove_zeroes();
        rebuild();
    }

    // split (scheme_z2.cu:614-619)
    FGK_HD void split(int i, int j, int k, int index, uint32_t x1) {
        const uint32_t x2 = f(i, index) ^ x1;
        f(i, index) = x1;
        f(i, m) = x2;
        f(j, m) = f(j, index);
        f(k, m) = f(k, index);
        ++m;
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

    // tryFlip (scheme_z2.cu:982-1027): r picks the pair in fresh-list order,
    // coin bit 0 swaps j/k, coin bit 1 swaps the indices.
    FGK_HD bool try_flip(uint32_t r, uint32_t coin, bool check_reduce) {
        const uint32_t total = S0 + S1 + S2;
        if (!total) return false;
        int c, t1, t2, j, k;
        if (!select_pair(draw_below(r, total), c, t1, t2)) return false;
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

}  // namespace lane
}  // namespace fgk

// Integer-ring operators of the walk (src/schemes/scheme_integer.cu over
// src/entities/addition.cu) on the LDS lane state of lane_common.h.  A factor
// with at most 16 coefficients in {-1, 0, 1} is one 32-bit word: magnitudes
// in bits 0-15, signs in bits 16-31 (signs are always a subset of the
// magnitudes, so a zero word is a zero factor and word equality is factor
// equality).
#pragma once

#include "lane_common.h"

namespace fgk {
namespace lane {
namespace tern {

FGK_HD uint32_t val(uint32_t x) { return x & 0xFFFFu; }
FGK_HD uint32_t sgn(uint32_t x) { return x >> 16; }
FGK_HD uint32_t pack(uint32_t v, uint32_t s) { return (v & 0xFFFFu) | ((s & v & 0xFFFFu) << 16); }
// lowest non-zero coefficient is positive (addition.cu:162-170)
FGK_HD bool first_pos(uint32_t v, uint32_t s) { return v == 0 || ((v & (0u - v)) & ~s) != 0; }
FGK_HD bool first_pos(uint32_t x) { return first_pos(val(x), sgn(x)); }

// operator+ / operator- (addition.cu:112-128); ok is Addition::valid
FGK_HD uint32_t add(uint32_t a, uint32_t b, bool &ok) {
    const uint32_t av = val(a), as = sgn(a), bv = val(b), bs = sgn(b);
    const uint32_t v = av ^ bv;
    ok = (av & bv & ~(as ^ bs)) == 0;
    return pack(v, ((as & av) | (bs & bv)) & v);
}
FGK_HD uint32_t sub(uint32_t a, uint32_t b, bool &ok) {
    const uint32_t av = val(a), as = sgn(a), bv = val(b), bs = sgn(b);
    const uint32_t v = av ^ bv;
    ok = (av & bv & (as ^ bs)) == 0;
    return pack(v, ((as & av) | (~bs & bv)) & v);
}
// compare (addition.cu:82-93)
FGK_HD int cmp(uint32_t a, uint32_t b) {
    if (val(a) != val(b)) return 0;
    if (sgn(a) == sgn(b)) return 1;
    if (sgn(a) == (~sgn(b) & val(b))) return -1;
    return 0;
}
// limitSum / limitSub (addition.cu:172-201)
FGK_HD bool limit_sum(uint32_t a, uint32_t b, bool fp) {
    bool ok;
    const uint32_t r = add(a, b, ok);
    return ok && (!fp || first_pos(r));
}
FGK_HD bool limit_sub(uint32_t a, uint32_t b, bool fp) {
    bool ok;
    const uint32_t r = sub(a, b, ok);
    return ok && (!fp || first_pos(r));
}
FGK_HD bool pos_sub(uint32_t a, uint32_t b) {  // positiveFirstNonZeroSub (addition.cu:203-208)
    bool ok;
    return first_pos(sub(a, b, ok));
}
FGK_HD uint32_t inverse(uint32_t a) { return pack(val(a), ~sgn(a) & val(a)); }

}  // namespace tern

struct ZTLane : LaneBaseT<uint32_t> {
    static constexpr uint32_t kValueMask = 0xFFFFu;  // magnitudes; signs sit in bits 16-31
    static constexpr bool kSigned = true;            // a bit map acts on both halves of a word

    FGK_HD void put(int c, int t, uint32_t v) { f(c, t) = v; }
    // sandwiching (scheme_integer.cu:1215-1217) is empty in the reference
    FGK_HD void sandwiching() {}

    uint16_t *V;  // lane-offset base of the visited-candidate list

    FGK_HD void lane_extra(uint8_t *base, int lane) { V = reinterpret_cast<uint16_t *>(base) + lane; }
    FGK_HD uint16_t &visited(int q) { return V[q * kWS]; }

    // fixSigns (scheme_integer.cu:398-425)
    FGK_HD void fix_signs() {
        for (int t = 0; t < m; ++t) {
            const bool pi = tern::first_pos(f(0, t)), pj = tern::first_pos(f(1, t));
            if (pi && pj) continue;
            if (!pi && !pj) {
                f(0, t) = tern::inverse(f(0, t));
                f(1, t) = tern::inverse(f(1, t));
            } else if (!pi) {
                f(0, t) = tern::inverse(f(0, t));
                f(2, t) = tern::inverse(f(2, t));
            } else {
                f(1, t) = tern::inverse(f(1, t));
                f(2, t) = tern::inverse(f(2, t));
            }
        }
    }
    // removeZeroes (scheme_integer.cu:237-249): re-checks the swapped-in term
    FGK_HD void remove_zeroes() {
        for (int t = 0; t < m; ++t) {
            if (nonzero_term(t)) continue;
            --m;
            copy_term(t, m);
            --t;
        }
    }
    // reduceAdd / reduceSub (scheme_integer.cu:549-571)
    FGK_HD void reduce(int c, int a, int b, bool subtract) {
        bool ok;
        f(c, a) = subtract ? tern::sub(f(c, a), f(c, b), ok) : tern::add(f(c, a), f(c, b), ok);
        const bool zero = f(c, a) == 0;
        remove_at(b);
        if (zero) remove_zeroes();
        rebuild();
    }

    // Reduction test of a term y that shares component j exactly with term x
    // after the flip (scheme_integer.cu:447-470 for x = index1, 472-495 for
    // x = index2, where `k` is the other changed component): returns
    // 1..4 for the first applicable rule, 0 for none.
    FGK_HD int reduce_rule(int i, int j, int k, int y, uint32_t fix, uint32_t fkx) {
        return reduce_rule_v(i, k, f(i, y), f(k, y), fix, fkx);
    }
    // The same from term y's factors fiy = f(i, y) and fky = f(k, y).
    FGK_HD static int reduce_rule_v(int i, int k, uint32_t fiy, uint32_t fky, uint32_t fix, uint32_t fkx) {
        const int cmpI = tern::cmp(fiy, fix);
        if (cmpI == 1 && tern::limit_sum(fky, fkx, k != 2)) return 1;
        if (i == 2 && cmpI == -1 && tern::limit_sub(fky, fkx, true)) return 2;
        const int cmpK = tern::cmp(fky, fkx);
        if (cmpK == 1 && tern::limit_sum(fiy, fix, i != 2)) return 3;
        if (k == 2 && cmpK == -1 && tern::limit_sub(fiy, fix, true)) return 4;
        return 0;
    }
    FGK_HD void apply_rule(int rule, int i, int k, int y, int x) {
        if (rule == 1) reduce(k, y, x, false);
        else if (rule == 2) reduce(k, y, x, true);
        else if (rule == 3) reduce(i, y, x, false);
        else reduce(i, y, x, true);
    }

    // flip (scheme_integer.cu:428-501) taken literally; used when a zero
    // factor is involved.  Partner bookkeeping is rebuilt afterwards.
    FGK_HD void flip_literal(int i, int j, int k, int i1, int i2, bool check_reduce) {
        bool ok;
        f(j, i1) = tern::add(f(j, i1), f(j, i2), ok);
        f(k, i2) = tern::sub(f(k, i2), f(k, i1), ok);
        if (!f(j, i1) || !f(k, i2)) {
            remove_zeroes();
            rebuild();
            while (try_reduce()) {
            }
            return;
        }
        for (int t = 0; t < m; ++t) {
            if (t != i1 && f(j, t) == f(j, i1) && check_reduce) {
                const int rule = reduce_rule(i, j, k, t, f(i, i1), f(k, i1));
                if (rule) { apply_rule(rule, i, k, t, i1); return; }
            }
            if (t != i2 && f(k, t) == f(k, i2) && check_reduce) {
                const int rule = reduce_rule(i, k, j, t, f(i, i2), f(j, i2));
                if (rule) { apply_rule(rule, i, j, t, i2); return; }
            }
        }
        rebuild();
    }

    // One half of a flip with every involved factor non-zero (see the Z2
    // lane): term x of component mc gets the new factor nv; reduction rules
    // as in scheme_integer.cu:447-495 with the other changed component oc
    // whose value at x is ov.
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
        if (hf.any >= 0)
            class_walk(mc, x, hf.any, i, oc, hf.ji,
                       [&](uint32_t a, uint32_t b) { return check_reduce ? reduce_rule_v(i, oc, a, b, fi, ov) : 0; },
                       hf.t, hf.rule);
        return hf;
    }

    // flip (scheme_integer.cu:428-501) when every involved factor is
    // non-zero; the two halves run on the two lanes of the pair (both on the
    // host), as in the Z2 lane.
    FGK_HD void flip(int i, int j, int k, int i1, int i2, bool check_reduce) {
        rec(0, i, j, k, i1, i2, check_reduce ? 1 : 0, 0);
        const uint32_t fj1 = f(j, i1), fj2 = f(j, i2);
        const uint32_t fk1 = f(k, i1), fk2 = f(k, i2);
        bool ok;
        const uint32_t A = tern::add(fj1, fj2, ok);
        const uint32_t B = tern::sub(fk2, fk1, ok);
        if (!(fj1 && fj2 && fk1 && fk2 && A && B)) {
            flip_literal(i, j, k, i1, i2, check_reduce);
            return;
        }
        const uint32_t fi = f(i, i1);  // == f(i, i2): the pair shares component i
#if defined(__HIP_DEVICE_COMPILE__)
        const int mc = side ? k : j, x = side ? i2 : i1, oc = side ? j : k;
        const uint32_t nv = side ? B : A, ov = side ? fj2 : fk1;
        const Half mine = flip_half(mc, x, nv, i, oc, fi, ov, check_reduce);
        const int ot = pair_exchange(mine.t, side), orule = pair_exchange(mine.rule, side);
        const int tA = side ? ot : mine.t, rA = side ? orule : mine.rule;
        const int tB = side ? mine.t : ot, rB = side ? mine.rule : orule;
#else
        const Half ha = flip_half(j, i1, A, i, k, fi, fk1, check_reduce);
        const Half hb = flip_half(k, i2, B, i, j, fi, fj2, check_reduce);
        const int tA = ha.t, rA = ha.rule, tB = hb.t, rB = hb.rule;
#endif
        if (rA || rB) {
            f(j, i1) = A;
            f(k, i2) = B;
            if (rA && tA <= tB) apply_rule(rA, i, k, tA, i1);
            else apply_rule(rB, i, j, tB, i2);
            return;
        }
        pop_update(fj1, A, fk2, B);
#if defined(__HIP_DEVICE_COMPILE__)
        const uint32_t before = Sc(mc);
        leave(mc, x);
        if (mine.any >= 0) link_join(mc, x, mine.ji);
        f(mc, x) = nv;
        addS(oc, pair_exchange((int)(Sc(mc) - before), side));
#else
        leave(j, i1);
        if (ha.any >= 0) link_join(j, i1, ha.ji);
        f(j, i1) = A;
        leave(k, i2);
        if (hb.any >= 0) link_join(k, i2, hb.ji);
        f(k, i2) = B;
#endif
    }

    // plus (scheme_integer.cu:503-539)
    FGK_HD void plus(int i, int j, int k, int i1, int i2, int variant) {
        rec(1, i, j, k, i1, i2, variant, 0);
        const uint32_t a1 = f(i, i1), b1 = f(j, i1), c1 = f(k, i1);
        const uint32_t a2 = f(i, i2), b2 = f(j, i2), c2 = f(k, i2);
        bool oaA, obA, ocA, oaS, obS, ocS;
        const uint32_t aAdd = tern::add(a1, a2, oaA), bAdd = tern::add(b1, b2, obA), cAdd = tern::add(c1, c2, ocA);
        const uint32_t aSub = tern::sub(a2, a1, oaS), bSub = tern::sub(b2, b1, obS), cSub = tern::sub(c2, c1, ocS);
        auto lim = [](uint32_t x, bool ok, bool fp) { return ok && (!fp || tern::first_pos(x)); };
        if (variant == 0 && lim(aSub, oaS, i != 2) && lim(bAdd, obA, j != 2) && lim(cSub, ocS, k != 2)) {
            f(j, i1) = bAdd; f(i, i2) = aSub;
            f(i, m) = a1; f(j, m) = b2; f(k, m) = cSub;
            ++m;
        } else if (variant == 1 && lim(aSub, oaS, i != 2) && lim(bSub, obS, j != 2) && lim(cAdd, ocA, k != 2)) {
            f(k, i1) = cAdd; f(j, i2) = bSub;
            f(i, m) = aSub; f(j, m) = b1; f(k, m) = c2;
            ++m;
        } else if (lim(aAdd, oaA, i != 2) && lim(bSub, obS, j != 2) && lim(cSub, ocS, k != 2)) {
            f(i, i1) = aAdd; f(k, i2) = cSub;
            f(i, m) = a2; f(j, m) = bSub; f(k, m) = c1;
            ++m;
        }
        remove_zeroes();
        fix_signs();
        rebuild();
    }

    // split (scheme_integer.cu:541-547)
    FGK_HD void split(int i, int j, int k, int index, uint32_t a) {
        rec(2, i, j, k, index, (int)(a & 0xFFFFu), (int)(a >> 16), 0);
        bool ok;
        f(i, m) = tern::sub(f(i, index), a, ok);
        f(j, m) = f(j, index);
        f(k, m) = f(k, index);
        ++m;
        f(i, index) = a;
        fix_signs();
        rebuild();
    }

    FGK_HD bool can_grow() const { return m < kMaxRank && m < naive_rank() && m < maxm; }

    FGK_HD bool try_plus() {  // scheme_integer.cu:977-998
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

    // trySplit (scheme_integer.cu:1000-1020); the random vector
    // (Addition::random, addition.cu:68-80: sign half-word, then value
    // half-word) is masked to the factor's nn coefficients.
    FGK_HD bool try_split() {
        if (!can_grow()) return false;
        const int index = (int)(aux32() % (uint32_t)m);
        int i, j, k;
        perm3(i, j, k);
        uint32_t av = 0, as = 0;
        const int bits = nn(i);
        for (int x = 0; x < bits; x += 16) {
            const uint32_t sb = aux32() & 0xFFFFu;
            const uint32_t vb = aux32() & 0xFFFFu;
            av |= vb << x;
            as |= sb << x;
        }
        if (bits < 32) av &= (1u << bits) - 1u;
        const uint32_t a = tern::pack(av, as);
        if (!tern::limit_sub(f(i, index), a, i != 2)) return false;
        split(i, j, k, index, a);
        return true;
    }

    FGK_HD bool try_split_existed() {  // scheme_integer.cu:1022-1043
        if (!can_grow()) return false;
        const uint32_t mm = (uint32_t)m;
        int i1, i2, i, tries = 0;
        do {
            if (++tries > kPlusMaxTries) return false;
            i1 = (int)(aux32() % mm);
            i2 = (int)(aux32() % mm);
            i = (int)(aux32() % 3u);
        } while (i1 == i2 || f(i, i1) == f(i, i2));
        if (!tern::limit_sub(f(i, i1), f(i, i2), i != 2)) return false;
        split(i, (i + 1) % 3, (i + 2) % 3, i1, f(i, i2));
        return true;
    }

    FGK_HD bool try_expand(int count) {  // scheme_integer.cu:1045-1060
        bool res = false;
        for (int q = 0; q < count && m < kMaxRank; ++q) {
            const uint32_t v = aux32() % 3u;
            if (v == 0) res |= try_plus();
            else if (v == 1) res |= try_split();
            else res |= try_split_existed();
        }
        return res;
    }

    // tryReduce (scheme_integer.cu:1062-1101) over the fresh-list order.
    FGK_HD bool try_reduce() {
        for (int a = 0; a < m; ++a) {
            int g = 0;
            for (int y = nx(0, a); y > a && g <= m; y = nx(0, y), ++g) {
                if (f(1, a) == f(1, y) && tern::limit_sum(f(2, a), f(2, y), false)) { reduce(2, a, y, false); return true; }
                const int c2 = tern::cmp(f(2, a), f(2, y));
                if (c2 == 1 && tern::limit_sum(f(1, a), f(1, y), true)) { reduce(1, a, y, false); return true; }
                if (c2 == -1 && tern::limit_sub(f(1, a), f(1, y), true)) { reduce(1, a, y, true); return true; }
            }
        }
        for (int a = 0; a < m; ++a) {
            int g = 0;
            for (int y = nx(1, a); y > a && g <= m; y = nx(1, y), ++g) {
                const int c2 = tern::cmp(f(2, a), f(2, y));
                if (c2 == 1 && tern::limit_sum(f(0, a), f(0, y), true)) { reduce(0, a, y, false); return true; }
                if (c2 == -1 && tern::limit_sub(f(0, a), f(0, y), true)) { reduce(0, a, y, true); return true; }
            }
        }
        return false;
    }

    // The two admissibility tests of candidate (component c, t1 < t2) under
    // the orientation coins cn (scheme_integer.cu:944-971); on success op
    // holds the flip to apply: i, j, k, index1, index2.
    struct ZtOp {
        int i, j, k, a, b;
    };
    FGK_HD bool zt_candidate(int c, int t1, int t2, uint32_t cn, ZtOp &op) {
        int j, k;
        orientation(c, cn, j, k);
        if (cn & 2u) {
            const int t = t1;
            t1 = t2;
            t2 = t;
        }
        const uint32_t j1 = f(j, t1), j2 = f(j, t2), k1 = f(k, t1), k2 = f(k, t2);
        if (tern::limit_sum(j1, j2, j != 2) && tern::limit_sub(k2, k1, false)) {
            const bool fwd = k == 2 || tern::pos_sub(k2, k1);
            op = {c, j, k, fwd ? t1 : t2, fwd ? t2 : t1};
            return true;
        }
        if (tern::limit_sum(k1, k2, k != 2) && tern::limit_sub(j2, j1, false)) {
            const bool fwd = j == 2 || tern::pos_sub(j2, j1);
            op = {c, k, j, fwd ? t1 : t2, fwd ? t2 : t1};
            return true;
        }
        return false;
    }

    // tryFlip (scheme_integer.cu:907-975).  The reference visits all
    // candidates in a uniformly random order, each with two fresh coins, and
    // applies the first admissible one; it fails only when none is.  Phase 1
    // visits up to min(visit_cap, kZtMaxVisits) candidates one at a time:
    // visit q draws k uniform in [0, total - q) (visit 0 from the main block
    // r, later visits from the aux stream) and takes the k-th candidate not
    // yet visited (the visited list is kept ascending), with coins from the
    // main block (visit 0) or the aux stream.  If all fail and unvisited
    // candidates remain, phase 2 draws one coin word per unvisited candidate
    // (fresh-list order) and picks uniformly among those admissible under
    // their coins by reservoir sampling (a draw per admissible candidate
    // after the first).  The first admissible candidate of a uniform order
    // with independent coins is a uniform pick among the candidates
    // admissible under their coins, so both phases follow the reference's
    // distribution.
    FGK_HD bool try_flip(uint32_t r, uint32_t coin, bool check_reduce) {
        const uint32_t total = S0 + S1 + S2;
        if (!total) return false;
        if (total > 0xFFFFu) {  // the visited list holds 16-bit positions
            fail(kErrCapacity);
            return false;
        }
        int cap = visit_cap > 0 && visit_cap < kZtMaxVisits ? visit_cap : kZtMaxVisits;
        if ((uint32_t)cap > total) cap = (int)total;
        ZtOp op;
        int nv = 0;
        for (int q = 0; q < cap; ++q) {
            uint32_t cand = draw_below(q == 0 ? r : aux32(), total - (uint32_t)q);
            for (int o = 0; o < nv; ++o) cand += visited(o) <= cand ? 1u : 0u;
            int pos = nv++;
            for (; pos > 0 && visited(pos - 1) > cand; --pos) visited(pos) = visited(pos - 1);
            visited(pos) = (uint16_t)cand;
            const uint32_t cn = q == 0 ? coin : aux32();
            int c, t1, t2;
            if (!select_pair(cand, c, t1, t2)) return false;
            if (zt_candidate(c, t1, t2, cn, op)) {
                flip(op.i, op.j, op.k, op.a, op.b, check_reduce);
                return true;
            }
        }
        if ((uint32_t)nv >= total) return false;
        int n = 0, pv = 0, bc = 0, bt1 = 0, bt2 = 0;
        uint32_t bcoin = 0, idx = 0;
        for (int c = 0; c < 3; ++c) {
            for (int t1 = 0; t1 < m; ++t1) {
                const int above = this->h(c, t1);
                int y = t1;
                for (int q = 0; q < above; ++q) {
                    y = nx(c, y);
                    const uint32_t cur = idx++;
                    if (pv < nv && visited(pv) == cur) {
                        ++pv;
                        continue;
                    }
                    const uint32_t cn = aux32();
                    if (!zt_candidate(c, t1, y, cn, op)) continue;
                    ++n;
                    if (n == 1 || draw_below(aux32(), (uint32_t)n) == 0) {
                        bc = c;
                        bt1 = t1;
                        bt2 = y;
                        bcoin = cn;
                    }
                }
            }
        }
        if (!n) return false;
        zt_candidate(bc, bt1, bt2, bcoin, op);
        flip(op.i, op.j, op.k, op.a, op.b, check_reduce);
        return true;
    }
};

}  // namespace lane
}  // namespace fgk

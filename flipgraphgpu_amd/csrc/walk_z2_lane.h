// Z2 operators of the walk (src/schemes/scheme_z2.cu) on the LDS lane state
// of lane_common.h.  __host__ __device__ so the tests can run the exact
// kernel code on the host.
#pragma once

#include "lane_common.h"

namespace fgk {
namespace lane {

// ---- bit maps of the resize operators (factor (i, j) of an a x b matrix at
// bit i * b + j) ----
FGK_HD uint32_t rz_bit(uint32_t x, int pos) { return (x >> pos) & 1u; }
// excludeRow (scheme_z2.cu:205-224)
FGK_HD uint32_t rz_drop_row(uint32_t x, int a, int b, int row) {
    uint32_t y = 0;
    int ii = 0;
    for (int i = 0; i < a; ++i) {
        if (i == row) continue;
        for (int j = 0; j < b; ++j) y |= rz_bit(x, i * b + j) << (ii * b + j);
        ++ii;
    }
    return y;
}
// excludeColumn (scheme_z2.cu:184-203)
FGK_HD uint32_t rz_drop_col(uint32_t x, int a, int b, int col) {
    uint32_t y = 0;
    for (int i = 0; i < a; ++i) {
        int jj = 0;
        for (int j = 0; j < b; ++j) {
            if (j == col) continue;
            y |= rz_bit(x, i * b + j) << (i * (b - 1) + jj);
            ++jj;
        }
    }
    return y;
}
// An a x b block placed at (di, dj) of a matrix with nb columns (addColumn,
// the diagonal blocks of product and merge).
FGK_HD uint32_t rz_embed(uint32_t x, int a, int b, int nb, int di, int dj) {
    uint32_t y = 0;
    for (int i = 0; i < a; ++i)
        for (int j = 0; j < b; ++j) y |= rz_bit(x, i * b + j) << ((i + di) * nb + j + dj);
    return y;
}
// The a x b factor whose bit (i, j) is bit (j, i) of a b x a factor (swapSize).
FGK_HD uint32_t rz_transpose(uint32_t x, int a, int b) {
    uint32_t y = 0;
    for (int i = 0; i < a; ++i)
        for (int j = 0; j < b; ++j) y |= rz_bit(x, j * a + i) << (i * b + j);
    return y;
}
// Kronecker product of an xa x xb and a ya x yb factor (product of schemes,
// scheme_z2.cu:819-862).
FGK_HD uint32_t rz_kron(uint32_t x, int xa, int xb, uint32_t y, int ya, int yb) {
    uint32_t z = 0;
    const int nb = xb * yb;
    for (int i = 0; i < xa * xb; ++i) {
        if (!rz_bit(x, i)) continue;
        const int r1 = i / xb, c1 = i % xb;
        for (int j = 0; j < ya * yb; ++j)
            if (rz_bit(y, j)) z |= 1u << ((r1 * ya + j / yb) * nb + c1 * yb + j % yb);
    }
    return z;
}


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

template <typename WordT>
struct Z2LaneT : LaneBaseT<WordT> {
    using B = LaneBaseT<WordT>;
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
    using B::n0;
    using B::n1;
    using B::n2;
    using B::nn0;
    using B::nn1;
    using B::nn2;
    static constexpr uint32_t kValueMask = 0xFFFFFFFFu;

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
        if (hf.any >= 0)
            class_walk(mc, x, hf.any, hf.ji,
                       [&](int y) { return !check_reduce ? 0 : (f(i, y) == fi ? 1 : (f(oc, y) == ov ? 2 : 0)); },
                       hf.t, hf.rule);
        return hf;
    }

    // flip (scheme_z2.cu:531-577) when every involved factor is non-zero.
    // Side 0 of the lane pair does the index1 half (component j, new factor
    // A), side 1 the index2 half (component k, new factor B); the host
    // emulation does both.  Reduction order as in the reference: the lowest
    // term first, the index1 half before the index2 half at equal terms.
    FGK_HD void flip(int i, int j, int k, int i1, int i2, bool check_reduce) {
        const uint32_t fj1 = f(j, i1), fj2 = f(j, i2);
        const uint32_t fk1 = f(k, i1), fk2 = f(k, i2);
        const uint32_t A = fj1 ^ fj2;
        const uint32_t B = fk2 ^ fk1;
        if (!(fj1 && fj2 && fk1 && fk2 && A && B)) {
            flip_literal(i, j, k, i1, i2, check_reduce);
            return;
        }
        const uint32_t fi = f(i, i1);
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
            if (rA && tA <= tB) reduce(rA == 1 ? k : i, tA, i1);
            else reduce(rB == 1 ? j : i, tB, i2);
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

    // ---- resize operators (scheme_z2.cu:643-979, 1165-1324) ----
    FGK_HD int dim(int c) const { return (c == 0 ? n0 : 0) + (c == 1 ? n1 : 0) + (c == 2 ? n2 : 0); }
    FGK_HD void set_dims(int a, int b, int c) {
        n0 = a;
        n1 = b;
        n2 = c;
        nn0 = a * b;
        nn1 = b * c;
        nn2 = c * a;
    }
    FGK_HD void set_dim(int p, int v) { set_dims(p == 0 ? v : n0, p == 1 ? v : n1, p == 2 ? v : n2); }
    FGK_HD void put(int c, int t, uint32_t v) { f(c, t) = (Word)v; }

    // project (scheme_z2.cu:643-653): row q of dimension p goes
    FGK_HD void project(int p, int q) {
        const int pc = (p + 2) % 3, a = dim(p), b = dim((p + 1) % 3), c = dim(pc);
        for (int t = 0; t < m; ++t) {
            put(p, t, rz_drop_row(f(p, t), a, b, q));
            put(pc, t, rz_drop_col(f(pc, t), c, a, q));
        }
        set_dim(p, a - 1);
        remove_zeroes();
        rebuild();
    }
    // extend (scheme_z2.cu:655-687)
    FGK_HD void extend(int p) {
        const int a = n0, b = n1, c = n2;
        const int mc = (p + 2) % 3;  // the component that gains a column
        for (int t = 0; t < m; ++t) put(mc, t, rz_embed(f(mc, t), dim(mc), dim(p), dim(p) + 1, 0, 0));
        if (p == 0) {
            for (int i = 0; i < c; ++i)
                for (int j = 0; j < b; ++j) add_term(1u << (a * b + j), 1u << (j * c + i), 1u << (i * (a + 1) + a));
        } else if (p == 1) {
            for (int i = 0; i < a; ++i)
                for (int j = 0; j < c; ++j) add_term(1u << (i * (b + 1) + b), 1u << (b * c + j), 1u << (j * a + i));
        } else {
            for (int i = 0; i < a; ++i)
                for (int j = 0; j < b; ++j) add_term(1u << (i * b + j), 1u << (j * (c + 1) + c), 1u << (c * a + i));
        }
        set_dim(p, dim(p) + 1);
        rebuild();
    }
    FGK_HD void add_term(uint32_t u, uint32_t v, uint32_t w) {
        put(0, m, u);
        put(1, m, v);
        put(2, m, w);
        ++m;
    }
    // product (scheme_z2.cu:689-747): dimension p doubled, the terms repeated
    // in the second diagonal block
    FGK_HD void product_dim(int p) {
        const int d0 = p == 0 ? n0 : 0, d1 = p == 1 ? n1 : 0, d2 = p == 2 ? n2 : 0;
        const int a = n0, b = n1, c = n2;
        const int A2 = a + d0, B2 = b + d1, C2 = c + d2;
        const int m0 = m;
        for (int t = 0; t < m0; ++t) {
            const uint32_t u = f(0, t), v = f(1, t), w = f(2, t);
            put(0, t, rz_embed(u, a, b, B2, 0, 0));
            put(1, t, rz_embed(v, b, c, C2, 0, 0));
            put(2, t, rz_embed(w, c, a, A2, 0, 0));
            put(0, t + m0, rz_embed(u, a, b, B2, d0, d1));
            put(1, t + m0, rz_embed(v, b, c, C2, d1, d2));
            put(2, t + m0, rz_embed(w, c, a, A2, d2, d0));
        }
        m = 2 * m0;
        set_dims(A2, B2, C2);
        rebuild();
    }
    // merge (scheme_z2.cu:749-817): scheme o (words o[c * stride + t], format
    // o0 x o1 x o2, rank om) joins as the second block along p
    FGK_HD void merge(const uint32_t *o, int stride, int om, int o0, int o1, int o2, int p) {
        const int d0 = p == 0 ? n0 : 0, d1 = p == 1 ? n1 : 0, d2 = p == 2 ? n2 : 0;
        const int a = n0, b = n1, c = n2;
        const int A2 = a + (p == 0 ? o0 : 0), B2 = b + (p == 1 ? o1 : 0), C2 = c + (p == 2 ? o2 : 0);
        for (int t = 0; t < m; ++t) {
            put(0, t, rz_embed(f(0, t), a, b, B2, 0, 0));
            put(1, t, rz_embed(f(1, t), b, c, C2, 0, 0));
            put(2, t, rz_embed(f(2, t), c, a, A2, 0, 0));
        }
        for (int t = 0; t < om; ++t) {
            put(0, m + t, rz_embed(o[t], o0, o1, B2, d0, d1));
            put(1, m + t, rz_embed(o[stride + t], o1, o2, C2, d1, d2));
            put(2, m + t, rz_embed(o[2 * stride + t], o2, o0, A2, d2, d0));
        }
        m += om;
        set_dims(A2, B2, C2);
        rebuild();
    }
    // product of schemes (scheme_z2.cu:819-862): term t1 * om + t2.  Terms are
    // rewritten from the last one down, so a term is read before any of its
    // replacements (positions >= t1 * om) can land on an unread term.
    FGK_HD void product_scheme(const uint32_t *o, int stride, int om, int o0, int o1, int o2) {
        const int a = n0, b = n1, c = n2;
        for (int t1 = m - 1; t1 >= 0; --t1) {
            const uint32_t u = f(0, t1), v = f(1, t1), w = f(2, t1);
            for (int t2 = 0; t2 < om; ++t2) {
                const int t = t1 * om + t2;
                put(0, t, rz_kron(u, a, b, o[t2], o0, o1));
                put(1, t, rz_kron(v, b, c, o[stride + t2], o1, o2));
                put(2, t, rz_kron(w, c, a, o[2 * stride + t2], o2, o0));
            }
        }
        m *= om;
        set_dims(a * o0, b * o1, c * o2);
        rebuild();
    }
    // swapSize (scheme_z2.cu:916-979)
    FGK_HD void swap_size(int p1, int p2) {
        if (p1 == p2) return;
        if (p1 > p2) {
            const int t = p1;
            p1 = p2;
            p2 = t;
        }
        // source component of each new component
        const int i0 = (p1 == 0 && p2 == 1) ? 0 : ((p1 == 0 && p2 == 2) ? 1 : 2);
        const int i1 = (p1 == 0 && p2 == 1) ? 2 : ((p1 == 0 && p2 == 2) ? 0 : 1);
        const int i2 = (p1 == 0 && p2 == 1) ? 1 : ((p1 == 0 && p2 == 2) ? 2 : 0);
        const int a = dim((i0 + 1) % 3), b = dim((i1 + 1) % 3), c = dim((i2 + 1) % 3);
        for (int t = 0; t < m; ++t) {
            const uint32_t x0 = f(i0, t), x1 = f(i1, t), x2 = f(i2, t);
            put(0, t, rz_transpose(x0, a, b));
            put(1, t, rz_transpose(x1, b, c));
            put(2, t, rz_transpose(x2, c, a));
        }
        set_dims(a, b, c);
        rebuild();
    }

    // isValidProject / isValidExtension / isValidProduct / tryMerge's and
    // tryProduct's size rules (scheme_z2.cu:256-310, 1225-1275) with the
    // pool's rank limit and factor width for MAX_RANK / MAX_MATRIX_ELEMENTS
    FGK_HD bool valid_project(int i) const {
        return dim(i) - 1 >= kMinProjectN && dim((i + 1) % 3) >= kMinProjectN && dim((i + 2) % 3) >= kMinProjectN;
    }
    FGK_HD static bool valid_sizes(int a, int b, int c, int max_coef) {
        return a * b <= max_coef && b * c <= max_coef && c * a <= max_coef && a <= kMaxExtensionN &&
               b <= kMaxExtensionN && c <= kMaxExtensionN;
    }
    FGK_HD bool valid_extension(int i, int max_rank, int max_coef) const {
        const int a = n0 + (i == 0), b = n1 + (i == 1), c = n2 + (i == 2);
        const int add = i == 0 ? b * c : (i == 1 ? a * c : a * b);
        return m + add <= max_rank && valid_sizes(a, b, c, max_coef);
    }
    FGK_HD bool valid_product(int i, int max_rank, int max_coef) const {
        const int a = n0 * (i == 0 ? 2 : 1), b = n1 * (i == 1 ? 2 : 1), c = n2 * (i == 2 ? 2 : 1);
        return 2 * m <= max_rank && valid_sizes(a, b, c, max_coef);
    }

    FGK_HD bool try_project() {  // scheme_z2.cu:1165-1185
        int idx[3], size = 0;
        for (int i = 0; i < 3; ++i)
            if (valid_project(i)) idx[size++] = i;
        if (!size) return false;
        const int p = pick3(idx, aux32() % (uint32_t)size);
        const int q = (int)(aux32() % (uint32_t)dim(p));
        project(p, q);
        while (try_reduce()) {
        }
        return true;
    }
    FGK_HD bool try_extend(int max_rank, int max_coef) {  // :1187-1206
        int idx[3], size = 0;
        for (int i = 0; i < 3; ++i)
            if (valid_extension(i, max_rank, max_coef)) idx[size++] = i;
        if (!size) return false;
        extend(pick3(idx, aux32() % (uint32_t)size));
        while (try_reduce()) {
        }
        return true;
    }
    FGK_HD bool try_product_dim(int max_rank, int max_coef) {  // :1208-1223
        int idx[3], size = 0;
        for (int i = 0; i < 3; ++i)
            if (valid_product(i, max_rank, max_coef)) idx[size++] = i;
        if (!size) return false;
        product_dim(pick3(idx, aux32() % (uint32_t)size));
        return true;
    }
    FGK_HD bool try_merge(const uint32_t *o, int stride, int om, int o0, int o1, int o2, int max_rank, int max_coef) {
        if (m + om > max_rank) return false;  // :1225-1254
        const bool e0 = n0 == o0, e1 = n1 == o1, e2 = n2 == o2;
        const int a0 = n0 + o0, a1 = n1 + o1, a2 = n2 + o2;
        int p[3], size = 0;
        if (a0 <= kMaxExtensionN && a0 * n1 <= max_coef && a0 * n2 <= max_coef && e1 && e2) p[size++] = 0;
        if (e0 && a1 <= kMaxExtensionN && n0 * a1 <= max_coef && n2 * a1 <= max_coef && e2) p[size++] = 1;
        if (e0 && e1 && a2 <= kMaxExtensionN && n0 * a2 <= max_coef && n1 * a2 <= max_coef) p[size++] = 2;
        if (!size) return false;
        merge(o, stride, om, o0, o1, o2, pick3(p, aux32() % (uint32_t)size));
        return true;
    }
    FGK_HD bool try_product_scheme(const uint32_t *o, int stride, int om, int o0, int o1, int o2, int max_rank,
                                   int max_coef) {  // :1256-1275
        if (m * om > max_rank) return false;
        const int a = n0 * o0, b = n1 * o1, c = n2 * o2;
        if (!valid_sizes(a, b, c, max_coef)) return false;
        product_scheme(o, stride, om, o0, o1, o2);
        return true;
    }
    FGK_HD void swap_size_random() {  // :1315-1324, redraws bounded
        for (int tries = 0; tries < kPlusMaxTries; ++tries) {
            const int p1 = (int)(aux32() % 3u), p2 = (int)(aux32() % 3u);
            if (p1 != p2) {
                swap_size(p1, p2);
                return;
            }
        }
    }
    FGK_HD static int pick3(const int *v, uint32_t q) { return q == 0 ? v[0] : (q == 1 ? v[1] : v[2]); }

    // ---- basis changes (scheme_z2.cu:864-914, 1277-1313) ----
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
    // swapBasis (:1302-1313): swaps two rows of A (rows of U, columns of W)
    // or two columns of B (columns of V, rows of W).  A permutation keeps
    // equal factors equal and coefficient counts unchanged, so the partner
    // classes stay as they are, as the reference keeps its FlipSets.
    FGK_HD void swap_basis() {
        if (aux32() % 2u) {
            const int i1 = (int)(aux32() % (uint32_t)n0), i2 = (int)(aux32() % (uint32_t)n0);
            for (int t = 0; t < m; ++t) {
                put(0, t, swap_rows(f(0, t), n0, n1, i1, i2));
                put(2, t, swap_cols(f(2, t), n2, n0, i1, i2));
            }
        } else {
            const int j1 = (int)(aux32() % (uint32_t)n2), j2 = (int)(aux32() % (uint32_t)n2);
            for (int t = 0; t < m; ++t) {
                put(1, t, swap_cols(f(1, t), n1, n2, j1, j2));
                put(2, t, swap_rows(f(2, t), n2, n0, j1, j2));
            }
        }
    }
    FGK_HD static uint32_t swap_rows(uint32_t x, int a, int b, int i1, int i2) {
        const uint32_t rm = (1u << b) - 1u;
        const uint32_t r1 = (x >> (i1 * b)) & rm, r2 = (x >> (i2 * b)) & rm;
        x &= ~((rm << (i1 * b)) | (rm << (i2 * b)));
        return x | (r2 << (i1 * b)) | (r1 << (i2 * b));
    }
    FGK_HD static uint32_t swap_cols(uint32_t x, int a, int b, int j1, int j2) {
        uint32_t y = x;
        for (int i = 0; i < a; ++i) {
            const uint32_t c1 = (x >> (i * b + j1)) & 1u, c2 = (x >> (i * b + j2)) & 1u;
            y = (y & ~((1u << (i * b + j1)) | (1u << (i * b + j2)))) | (c2 << (i * b + j1)) | (c1 << (i * b + j2));
        }
        return y;
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

}  // namespace lane
}  // namespace fgk

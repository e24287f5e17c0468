// Resize and basis-change operators of the walk (resizeKernel and the
// swapBasis edge) on the LDS lane state, for both rings: src/schemes/scheme_z2.cu
// :184-310,643-979,864-914,1165-1324 and their integer counterparts
// src/schemes/scheme_integer.cu:273-900,1104-1246.  A factor is an a x b
// matrix with entry (i, j) at bit i * b + j; an integer-ring word carries the
// same bit layout twice (magnitudes in bits 0-15, signs in bits 16-31), so every
// bit map acts on both halves.  Lane types provide kSigned, put(), f(),
// remove_zeroes(), rebuild(), try_reduce() and aux32(); the integer lane also
// fix_signs(), which the reference calls after project, swapSize and swapBasis.
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



// A bit map applied to a factor word of either ring.
template <class Lane, class Fn>
FGK_HD uint32_t rz_map(uint32_t x, Fn fn) {
    if constexpr (Lane::kSigned) return fn(x & 0xFFFFu) | (fn(x >> 16) << 16);
    else return fn(x);
}
// Kronecker product of two factor words; an integer coefficient is the
// product of the two: magnitude AND, sign XOR where both are set.
template <class Lane>
FGK_HD uint32_t rz_kron_word(uint32_t x, int xa, int xb, uint32_t y, int ya, int yb) {
    if constexpr (Lane::kSigned) {
        const uint32_t mx = x & 0xFFFFu, sx = x >> 16, my = y & 0xFFFFu, sy = y >> 16;
        const uint32_t mag = rz_kron(mx, xa, xb, my, ya, yb);
        const uint32_t sgn = rz_kron(sx, xa, xb, my, ya, yb) ^ rz_kron(mx, xa, xb, sy, ya, yb);
        return mag | (sgn << 16);
    } else {
        return rz_kron(x, xa, xb, y, ya, yb);
    }
}
template <class Lane>
FGK_HD void rz_fix_signs(Lane &s) {
    if constexpr (Lane::kSigned) s.fix_signs();
}

template <class Lane>
FGK_HD int rz_dim(const Lane &s, int c) {
    return (c == 0 ? s.n0 : 0) + (c == 1 ? s.n1 : 0) + (c == 2 ? s.n2 : 0);
}
template <class Lane>
FGK_HD void rz_set_dims(Lane &s, int a, int b, int c) {
    s.n0 = a;
    s.n1 = b;
    s.n2 = c;
    s.nn0 = a * b;
    s.nn1 = b * c;
    s.nn2 = c * a;
}
template <class Lane>
FGK_HD void rz_set_dim(Lane &s, int p, int v) {
    rz_set_dims(s, p == 0 ? v : s.n0, p == 1 ? v : s.n1, p == 2 ? v : s.n2);
}
template <class Lane>
FGK_HD void rz_add_term(Lane &s, uint32_t u, uint32_t v, uint32_t w) {
    s.put(0, s.m, u);
    s.put(1, s.m, v);
    s.put(2, s.m, w);
    ++s.m;
}

// project (z2 :643-653, integer :573-584): row q of dimension p goes
template <class Lane>
FGK_HD void rz_project(Lane &s, int p, int q) {
    const int pc = (p + 2) % 3, a = rz_dim(s, p), b = rz_dim(s, (p + 1) % 3), c = rz_dim(s, pc);
    for (int t = 0; t < s.m; ++t) {
        s.put(p, t, rz_map<Lane>(s.f(p, t), [&](uint32_t x) { return rz_drop_row(x, a, b, q); }));
        s.put(pc, t, rz_map<Lane>(s.f(pc, t), [&](uint32_t x) { return rz_drop_col(x, c, a, q); }));
    }
    rz_set_dim(s, p, a - 1);
    s.remove_zeroes();
    rz_fix_signs(s);
    s.rebuild();
}
// extend (z2 :655-687, integer :586-618): the new terms have single +1
// coefficients
template <class Lane>
FGK_HD void rz_extend(Lane &s, int p) {
    const int a = s.n0, b = s.n1, c = s.n2;
    const int mc = (p + 2) % 3;  // the component that gains a column
    const int ra = rz_dim(s, mc), rb = rz_dim(s, p);
    for (int t = 0; t < s.m; ++t)
        s.put(mc, t, rz_map<Lane>(s.f(mc, t), [&](uint32_t x) { return rz_embed(x, ra, rb, rb + 1, 0, 0); }));
    if (p == 0) {
        for (int i = 0; i < c; ++i)
            for (int j = 0; j < b; ++j) rz_add_term(s, 1u << (a * b + j), 1u << (j * c + i), 1u << (i * (a + 1) + a));
    } else if (p == 1) {
        for (int i = 0; i < a; ++i)
            for (int j = 0; j < c; ++j) rz_add_term(s, 1u << (i * (b + 1) + b), 1u << (b * c + j), 1u << (j * a + i));
    } else {
        for (int i = 0; i < a; ++i)
            for (int j = 0; j < b; ++j) rz_add_term(s, 1u << (i * b + j), 1u << (j * (c + 1) + c), 1u << (c * a + i));
    }
    rz_set_dim(s, p, rz_dim(s, p) + 1);
    s.rebuild();
}
// product (z2 :689-747, integer :620-678): dimension p doubled, the terms
// repeated in the second diagonal block
template <class Lane>
FGK_HD void rz_product_dim(Lane &s, int p) {
    const int d0 = p == 0 ? s.n0 : 0, d1 = p == 1 ? s.n1 : 0, d2 = p == 2 ? s.n2 : 0;
    const int a = s.n0, b = s.n1, c = s.n2;
    const int A2 = a + d0, B2 = b + d1, C2 = c + d2;
    const int m0 = s.m;
    for (int t = 0; t < m0; ++t) {
        const uint32_t u = s.f(0, t), v = s.f(1, t), w = s.f(2, t);
        s.put(0, t, rz_map<Lane>(u, [&](uint32_t x) { return rz_embed(x, a, b, B2, 0, 0); }));
        s.put(1, t, rz_map<Lane>(v, [&](uint32_t x) { return rz_embed(x, b, c, C2, 0, 0); }));
        s.put(2, t, rz_map<Lane>(w, [&](uint32_t x) { return rz_embed(x, c, a, A2, 0, 0); }));
        s.put(0, t + m0, rz_map<Lane>(u, [&](uint32_t x) { return rz_embed(x, a, b, B2, d0, d1); }));
        s.put(1, t + m0, rz_map<Lane>(v, [&](uint32_t x) { return rz_embed(x, b, c, C2, d1, d2); }));
        s.put(2, t + m0, rz_map<Lane>(w, [&](uint32_t x) { return rz_embed(x, c, a, A2, d2, d0); }));
    }
    s.m = 2 * m0;
    rz_set_dims(s, A2, B2, C2);
    s.rebuild();
}
// merge (z2 :749-817, integer :680-745): scheme o (words o[c * stride + t],
// format o0 x o1 x o2, rank om) joins as the second block along p
template <class Lane>
FGK_HD void rz_merge(Lane &s, const uint32_t *o, int stride, int om, int o0, int o1, int o2, int p) {
    const int d0 = p == 0 ? s.n0 : 0, d1 = p == 1 ? s.n1 : 0, d2 = p == 2 ? s.n2 : 0;
    const int a = s.n0, b = s.n1, c = s.n2;
    const int A2 = a + (p == 0 ? o0 : 0), B2 = b + (p == 1 ? o1 : 0), C2 = c + (p == 2 ? o2 : 0);
    for (int t = 0; t < s.m; ++t) {
        s.put(0, t, rz_map<Lane>(s.f(0, t), [&](uint32_t x) { return rz_embed(x, a, b, B2, 0, 0); }));
        s.put(1, t, rz_map<Lane>(s.f(1, t), [&](uint32_t x) { return rz_embed(x, b, c, C2, 0, 0); }));
        s.put(2, t, rz_map<Lane>(s.f(2, t), [&](uint32_t x) { return rz_embed(x, c, a, A2, 0, 0); }));
    }
    for (int t = 0; t < om; ++t) {
        s.put(0, s.m + t, rz_map<Lane>(o[t], [&](uint32_t x) { return rz_embed(x, o0, o1, B2, d0, d1); }));
        s.put(1, s.m + t, rz_map<Lane>(o[stride + t], [&](uint32_t x) { return rz_embed(x, o1, o2, C2, d1, d2); }));
        s.put(2, s.m + t, rz_map<Lane>(o[2 * stride + t], [&](uint32_t x) { return rz_embed(x, o2, o0, A2, d2, d0); }));
    }
    s.m += om;
    rz_set_dims(s, A2, B2, C2);
    s.rebuild();
}
// product of schemes (z2 :819-862, integer :747-789): term t1 * om + t2.
// Terms are rewritten from the last one down, so a term is read before any
// of its replacements (positions >= t1 * om) can land on an unread term.
template <class Lane>
FGK_HD void rz_product_scheme(Lane &s, const uint32_t *o, int stride, int om, int o0, int o1, int o2) {
    const int a = s.n0, b = s.n1, c = s.n2;
    for (int t1 = s.m - 1; t1 >= 0; --t1) {
        const uint32_t u = s.f(0, t1), v = s.f(1, t1), w = s.f(2, t1);
        for (int t2 = 0; t2 < om; ++t2) {
            const int t = t1 * om + t2;
            s.put(0, t, rz_kron_word<Lane>(u, a, b, o[t2], o0, o1));
            s.put(1, t, rz_kron_word<Lane>(v, b, c, o[stride + t2], o1, o2));
            s.put(2, t, rz_kron_word<Lane>(w, c, a, o[2 * stride + t2], o2, o0));
        }
    }
    s.m *= om;
    rz_set_dims(s, a * o0, b * o1, c * o2);
    s.rebuild();
}
// swapSize (z2 :916-979, integer :843-905)
template <class Lane>
FGK_HD void rz_swap_size(Lane &s, int p1, int p2) {
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
    const int a = rz_dim(s, (i0 + 1) % 3), b = rz_dim(s, (i1 + 1) % 3), c = rz_dim(s, (i2 + 1) % 3);
    for (int t = 0; t < s.m; ++t) {
        const uint32_t x0 = s.f(i0, t), x1 = s.f(i1, t), x2 = s.f(i2, t);
        s.put(0, t, rz_map<Lane>(x0, [&](uint32_t x) { return rz_transpose(x, a, b); }));
        s.put(1, t, rz_map<Lane>(x1, [&](uint32_t x) { return rz_transpose(x, b, c); }));
        s.put(2, t, rz_map<Lane>(x2, [&](uint32_t x) { return rz_transpose(x, c, a); }));
    }
    rz_set_dims(s, a, b, c);
    rz_fix_signs(s);
    s.rebuild();
}

// isValidProject / isValidExtension / isValidProduct / tryMerge's and
// tryProduct's size rules (z2 :256-310, 1225-1275) with the pool's rank limit
// and factor width for MAX_RANK / MAX_MATRIX_ELEMENTS
template <class Lane>
FGK_HD bool rz_valid_project(const Lane &s, int i) {
    return rz_dim(s, i) - 1 >= kMinProjectN && rz_dim(s, (i + 1) % 3) >= kMinProjectN &&
           rz_dim(s, (i + 2) % 3) >= kMinProjectN;
}
FGK_HD bool rz_valid_sizes(int a, int b, int c, int max_coef) {
    return a * b <= max_coef && b * c <= max_coef && c * a <= max_coef && a <= kMaxExtensionN &&
           b <= kMaxExtensionN && c <= kMaxExtensionN;
}
template <class Lane>
FGK_HD bool rz_valid_extension(const Lane &s, int i, int max_rank, int max_coef) {
    const int a = s.n0 + (i == 0), b = s.n1 + (i == 1), c = s.n2 + (i == 2);
    const int add = i == 0 ? b * c : (i == 1 ? a * c : a * b);
    return s.m + add <= max_rank && rz_valid_sizes(a, b, c, max_coef);
}
template <class Lane>
FGK_HD bool rz_valid_product(const Lane &s, int i, int max_rank, int max_coef) {
    const int a = s.n0 * (i == 0 ? 2 : 1), b = s.n1 * (i == 1 ? 2 : 1), c = s.n2 * (i == 2 ? 2 : 1);
    return 2 * s.m <= max_rank && rz_valid_sizes(a, b, c, max_coef);
}
FGK_HD int rz_pick3(const int *v, uint32_t q) { return q == 0 ? v[0] : (q == 1 ? v[1] : v[2]); }

template <class Lane>
FGK_HD bool rz_try_project(Lane &s) {  // z2 :1165-1185
    int idx[3], size = 0;
    for (int i = 0; i < 3; ++i)
        if (rz_valid_project(s, i)) idx[size++] = i;
    if (!size) return false;
    const int p = rz_pick3(idx, s.aux32() % (uint32_t)size);
    const int q = (int)(s.aux32() % (uint32_t)rz_dim(s, p));
    rz_project(s, p, q);
    while (s.try_reduce()) {
    }
    return true;
}
template <class Lane>
FGK_HD bool rz_try_extend(Lane &s, int max_rank, int max_coef) {  // :1187-1206
    int idx[3], size = 0;
    for (int i = 0; i < 3; ++i)
        if (rz_valid_extension(s, i, max_rank, max_coef)) idx[size++] = i;
    if (!size) return false;
    rz_extend(s, rz_pick3(idx, s.aux32() % (uint32_t)size));
    while (s.try_reduce()) {
    }
    return true;
}
template <class Lane>
FGK_HD bool rz_try_product_dim(Lane &s, int max_rank, int max_coef) {  // :1208-1223
    int idx[3], size = 0;
    for (int i = 0; i < 3; ++i)
        if (rz_valid_product(s, i, max_rank, max_coef)) idx[size++] = i;
    if (!size) return false;
    rz_product_dim(s, rz_pick3(idx, s.aux32() % (uint32_t)size));
    return true;
}
template <class Lane>
FGK_HD bool rz_try_merge(Lane &s, const uint32_t *o, int stride, int om, int o0, int o1, int o2, int max_rank,
                         int max_coef) {  // :1225-1254
    if (s.m + om > max_rank) return false;
    const bool e0 = s.n0 == o0, e1 = s.n1 == o1, e2 = s.n2 == o2;
    const int a0 = s.n0 + o0, a1 = s.n1 + o1, a2 = s.n2 + o2;
    int p[3], size = 0;
    if (a0 <= kMaxExtensionN && a0 * s.n1 <= max_coef && a0 * s.n2 <= max_coef && e1 && e2) p[size++] = 0;
    if (e0 && a1 <= kMaxExtensionN && s.n0 * a1 <= max_coef && s.n2 * a1 <= max_coef && e2) p[size++] = 1;
    if (e0 && e1 && a2 <= kMaxExtensionN && s.n0 * a2 <= max_coef && s.n1 * a2 <= max_coef) p[size++] = 2;
    if (!size) return false;
    rz_merge(s, o, stride, om, o0, o1, o2, rz_pick3(p, s.aux32() % (uint32_t)size));
    return true;
}
template <class Lane>
FGK_HD bool rz_try_product_scheme(Lane &s, const uint32_t *o, int stride, int om, int o0, int o1, int o2, int max_rank,
                                  int max_coef) {  // :1256-1275
    if (s.m * om > max_rank) return false;
    if (!rz_valid_sizes(s.n0 * o0, s.n1 * o1, s.n2 * o2, max_coef)) return false;
    rz_product_scheme(s, o, stride, om, o0, o1, o2);
    return true;
}
template <class Lane>
FGK_HD void rz_swap_size_random(Lane &s) {  // :1315-1324, redraws bounded
    for (int tries = 0; tries < kPlusMaxTries; ++tries) {
        const int p1 = (int)(s.aux32() % 3u), p2 = (int)(s.aux32() % 3u);
        if (p1 != p2) {
            rz_swap_size(s, p1, p2);
            return;
        }
    }
}

// swapBasis (z2 :1302-1313, integer :1219-1233): swaps two rows of A (rows
// of U, columns of W) or two columns of B (columns of V, rows of W).  A
// permutation keeps equal factors equal and coefficient counts unchanged, so
// a Z2 walker keeps its partner classes as the reference keeps its FlipSets;
// the integer ring re-canonicalises signs (fixSigns) and rebuilds.
FGK_HD uint32_t rz_swap_rows(uint32_t x, int b, int i1, int i2) {
    const uint32_t rm = (1u << b) - 1u;
    const uint32_t r1 = (x >> (i1 * b)) & rm, r2 = (x >> (i2 * b)) & rm;
    x &= ~((rm << (i1 * b)) | (rm << (i2 * b)));
    return x | (r2 << (i1 * b)) | (r1 << (i2 * b));
}
FGK_HD uint32_t rz_swap_cols(uint32_t x, int a, int b, int j1, int j2) {
    uint32_t y = x;
    for (int i = 0; i < a; ++i) {
        const uint32_t c1 = (x >> (i * b + j1)) & 1u, c2 = (x >> (i * b + j2)) & 1u;
        y = (y & ~((1u << (i * b + j1)) | (1u << (i * b + j2)))) | (c2 << (i * b + j1)) | (c1 << (i * b + j2));
    }
    return y;
}
template <class Lane>
FGK_HD void rz_swap_basis(Lane &s) {
    const int n0 = s.n0, n1 = s.n1, n2 = s.n2;
    if (s.aux32() % 2u) {
        const int i1 = (int)(s.aux32() % (uint32_t)n0), i2 = (int)(s.aux32() % (uint32_t)n0);
        for (int t = 0; t < s.m; ++t) {
            s.put(0, t, rz_map<Lane>(s.f(0, t), [&](uint32_t x) { return rz_swap_rows(x, n1, i1, i2); }));
            s.put(2, t, rz_map<Lane>(s.f(2, t), [&](uint32_t x) { return rz_swap_cols(x, n2, n0, i1, i2); }));
        }
    } else {
        const int j1 = (int)(s.aux32() % (uint32_t)n2), j2 = (int)(s.aux32() % (uint32_t)n2);
        for (int t = 0; t < s.m; ++t) {
            s.put(1, t, rz_map<Lane>(s.f(1, t), [&](uint32_t x) { return rz_swap_cols(x, n1, n2, j1, j2); }));
            s.put(2, t, rz_map<Lane>(s.f(2, t), [&](uint32_t x) { return rz_swap_rows(x, n0, j1, j2); }));
        }
    }
    if constexpr (Lane::kSigned) {
        s.fix_signs();
        s.rebuild();
    }
}

}  // namespace lane
}  // namespace fgk

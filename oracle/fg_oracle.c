/*
 * fg_oracle.c -- CPU oracle for the flip-graph random walk (test
 * infrastructure only; see fg_oracle.h for the contract and the two
 * documented differences from the reference).
 *
 * Citations are to the reference repository dronperminov/FlipGraphGPU.
 */
#include "fg_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: */
/* as easy as 1, 2, 3", SC'11).  Replaces cuRAND (src/utils/random.cuh) */
/* ------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void fgo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r) {
            k0 += PHILOX_W0;
            k1 += PHILOX_W1;
        }
        uint64_t p0 = (uint64_t)PHILOX_M0 * x0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * x2;
        uint32_t y0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0;
        uint32_t y1 = (uint32_t)p1;
        uint32_t y2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1;
        uint32_t y3 = (uint32_t)p0;
        x0 = y0; x1 = y1; x2 = y2; x3 = y3;
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

/* Stream layout: key = seed; main block b of walker w uses counter
 * (b_lo, b_hi, w, 0); aux block a uses (a_lo, a_hi, w, 1). */
void fgo_rng_init(fgo_rng *r, uint64_t seed, uint32_t walker) {
    r->key[0] = (uint32_t)seed;
    r->key[1] = (uint32_t)(seed >> 32);
    r->walker = walker;
    r->it = 0;
    r->aux = 0;
    r->aidx = 4;
    memset(r->abuf, 0, sizeof(r->abuf));
}

void fgo_rng_main_block(fgo_rng *r, uint32_t out[4]) {
    uint32_t ctr[4] = {(uint32_t)r->it, (uint32_t)(r->it >> 32), r->walker, 0u};
    fgo_philox4x32_10(ctr, r->key, out);
    r->it++;
}

static void aux_fill(fgo_rng *r, uint64_t block) {
    uint32_t ctr[4] = {(uint32_t)block, (uint32_t)(block >> 32), r->walker, 1u};
    fgo_philox4x32_10(ctr, r->key, r->abuf);
}

uint32_t fgo_rng_aux32(fgo_rng *r) {
    if (r->aidx >= 4) {
        aux_fill(r, r->aux);
        r->aux++;
        r->aidx = 0;
    }
    return r->abuf[r->aidx++];
}

/* Pair draws map a 32-bit word x onto [0, size) as floor(x * size / 2^32)
 * (multiply-high) where the reference takes curand() % size
 * (src/schemes/scheme_z2.cu:988); both are uniform up to size / 2^32. */
static uint32_t draw_below(uint32_t x, uint32_t size) { return (uint32_t)(((uint64_t)x * size) >> 32); }

/* randint(a, b) = a + rand % (b - a + 1)   (src/utils/random.cuh:5-7) */
static int aux_randint(fgo_rng *r, int a, int b) {
    return a + (int)(fgo_rng_aux32(r) % (uint32_t)(b - a + 1));
}

/* randomPermutation(array, 3) (src/utils/random.cuh:9-22) */
static void aux_perm3(fgo_rng *r, int *p) {
    p[0] = 0; p[1] = 1; p[2] = 2;
    for (int i = 2; i > 0; i--) {
        int j = (int)(fgo_rng_aux32(r) % (uint32_t)(i + 1));
        int t = p[i]; p[i] = p[j]; p[j] = t;
    }
}

/* ------------------------------------------------------------------ */
/* FlipSet (src/entities/flip_set.cu:3-43)                             */
/* ------------------------------------------------------------------ */
static void fs_clear(fgo_flipset *f) { f->size = 0; }

static void fs_add(fgo_flipset *f, uint32_t a, uint32_t b) {
    if (f->size >= FGO_MAX_PAIRS) return; /* silently dropped, flip_set.cu:8-9 */
    f->pairs[f->size++] = (a << 16) | b;
}

static void fs_remove_index(fgo_flipset *f, uint32_t index) {
    /* flip_set.cu:27-31: swap-with-last, re-examine the swapped-in entry */
    for (int i = 0; i < f->size; i++) {
        uint32_t p = f->pairs[i];
        if ((p >> 16) == index || (p & 0xFFFFu) == index) {
            f->pairs[i] = f->pairs[--f->size];
            i--;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Ternary vector arithmetic (src/entities/addition.cu)                */
/* ------------------------------------------------------------------ */
typedef struct { uint64_t v, s; uint8_t ok; } tv;

static tv tv_get(const fgo_scheme *x, int c, int t) { tv a = {x->v[c][t], x->s[c][t], x->ok[c][t]}; return a; }
static void tv_put(fgo_scheme *x, int c, int t, tv a) { x->v[c][t] = a.v; x->s[c][t] = a.s; x->ok[c][t] = a.ok; }

/* operator+ (addition.cu:112-119) */
static tv tv_add(tv a, tv b) {
    tv r;
    r.v = a.v ^ b.v;
    r.s = ((a.s & a.v) | (b.s & b.v)) & r.v;
    r.ok = !(a.v & b.v & ~(a.s ^ b.s));
    return r;
}
/* operator- (addition.cu:121-128) */
static tv tv_sub(tv a, tv b) {
    tv r;
    r.v = a.v ^ b.v;
    r.s = ((a.s & a.v) | (~b.s & b.v)) & r.v;
    r.ok = !(a.v & b.v & (a.s ^ b.s));
    return r;
}
static int tv_eq(tv a, tv b) { return a.v == b.v && a.s == b.s; } /* addition.cu:95-97 */
/* compare (addition.cu:82-93) */
static int tv_cmp(tv a, tv b) {
    if (a.v != b.v) return 0;
    if (a.s == b.s) return 1;
    if (a.s == (~b.s & b.v)) return -1;
    return 0;
}
static int first_pos(uint64_t v, uint64_t s) { return v == 0 || ((v & ~(v & (v - 1))) & ~s) != 0; }
/* limit (addition.cu:162-170) */
static int tv_limit(tv a, int fp) { if (!a.ok) return 0; return fp ? first_pos(a.v, a.s) : 1; }
/* limitSum (addition.cu:172-186) */
static int tv_limit_sum(tv a, tv b, int fp) {
    if (a.v & b.v & ~(a.s ^ b.s)) return 0;
    if (fp) {
        uint64_t sv = a.v ^ b.v;
        uint64_t ss = ((a.s & a.v) | (b.s & b.v)) & sv;
        return first_pos(sv, ss);
    }
    return 1;
}
/* limitSub (addition.cu:188-201) */
static int tv_limit_sub(tv a, tv b, int fp) {
    if (a.v & b.v & (a.s ^ b.s)) return 0;
    if (fp) {
        uint64_t sv = a.v ^ b.v;
        uint64_t ss = ((a.s & a.v) | (~b.s & b.v)) & sv;
        return first_pos(sv, ss);
    }
    return 1;
}
/* positiveFirstNonZeroSub (addition.cu:203-208) */
static int tv_pos_sub(tv a, tv b) {
    uint64_t sv = a.v ^ b.v;
    uint64_t ss = ((a.s & a.v) | (~b.s & b.v)) & sv;
    return first_pos(sv, ss);
}
static tv tv_inverse(tv a) { a.s = ~a.s & a.v; return a; } /* addition.cu:62-64 */

/* ------------------------------------------------------------------ */
/* Scheme basics                                                       */
/* ------------------------------------------------------------------ */
int fgo_scheme_size(void) { return (int)sizeof(fgo_scheme); }

static int fac_eq(const fgo_scheme *x, int c, int a, int b) {
    return x->v[c][a] == x->v[c][b] && (x->ring == FGO_RING_Z2 || x->s[c][a] == x->s[c][b]);
}
static int nz(const fgo_scheme *x, int c, int t) { return x->v[c][t] != 0; }

/* initializeNaive (scheme_z2.cu:32-55, scheme_integer.cu:41-64) */
void fgo_init_naive(fgo_scheme *s, int ring, int n1, int n2, int n3) {
    memset(s, 0, sizeof(*s));
    s->ring = ring;
    s->n[0] = n1; s->n[1] = n2; s->n[2] = n3;
    s->nn[0] = n1 * n2; s->nn[1] = n2 * n3; s->nn[2] = n3 * n1;
    s->m = n1 * n2 * n3;
    for (int i = 0; i < n1; i++)
        for (int j = 0; j < n3; j++)
            for (int k = 0; k < n2; k++) {
                int t = (i * n3 + j) * n2 + k;
                s->v[0][t] = 1ull << (i * n2 + k);
                s->v[1][t] = 1ull << (k * n3 + j);
                s->v[2][t] = 1ull << (j * n1 + i);
                s->ok[0][t] = s->ok[1][t] = s->ok[2][t] = 1;
            }
    fgo_init_flips(s);
}

/* validateEquation + validate (scheme_z2.cu:3-29, scheme_integer.cu:3-38) */
int fgo_validate(const fgo_scheme *s) {
    for (int i = 0; i < s->nn[0]; i++)
        for (int j = 0; j < s->nn[1]; j++)
            for (int k = 0; k < s->nn[2]; k++) {
                int i1 = i / s->n[1], i2 = i % s->n[1];
                int j1 = j / s->n[2], j2 = j % s->n[2];
                int k1 = k / s->n[0], k2 = k % s->n[0];
                int target = (i2 == j1) && (i1 == k2) && (j2 == k1);
                if (s->ring == FGO_RING_Z2) {
                    int e = 0;
                    for (int t = 0; t < s->m; t++)
                        e ^= (int)(((s->v[0][t] >> i) & (s->v[1][t] >> j) & (s->v[2][t] >> k)) & 1);
                    if (e != target) return 0;
                } else {
                    int e = 0;
                    for (int t = 0; t < s->m; t++) {
                        int a = (int)((s->v[0][t] >> i) & 1), b = (int)((s->v[1][t] >> j) & 1), c = (int)((s->v[2][t] >> k) & 1);
                        if (a && b && c) {
                            int sg = (int)(((s->s[0][t] >> i) ^ (s->s[1][t] >> j) ^ (s->s[2][t] >> k)) & 1);
                            e += sg ? -1 : 1;
                        }
                    }
                    if (e != target) return 0;
                }
            }
    if (s->ring == FGO_RING_ZT)
        for (int t = 0; t < s->m; t++)
            for (int c = 0; c < 3; c++)
                if (!s->ok[c][t]) return 0; /* scheme_integer.cu:28-35 */
    return 1;
}

static int popc64(uint64_t x) { return __builtin_popcountll(x); }

/* getComplexity (scheme_z2.cu:127-139, scheme_integer.cu:139-147) */
int fgo_complexity(const fgo_scheme *s) {
    int count = 0;
    for (int t = 0; t < s->m; t++)
        for (int c = 0; c < 3; c++) count += popc64(s->v[c][t]);
    return count - 2 * s->m - s->nn[2];
}

/* initFlips (scheme_z2.cu:141-150, scheme_integer.cu:225-235) */
void fgo_init_flips(fgo_scheme *s) {
    for (int c = 0; c < 3; c++) {
        fs_clear(&s->flips[c]);
        for (int a = 0; a < s->m; a++)
            for (int b = a + 1; b < s->m; b++)
                if (fac_eq(s, c, a, b)) fs_add(&s->flips[c], (uint32_t)a, (uint32_t)b);
    }
}

int fgo_count_pairs(const fgo_scheme *s) {
    int n = 0;
    for (int c = 0; c < 3; c++)
        for (int a = 0; a < s->m; a++)
            for (int b = a + 1; b < s->m; b++) n += fac_eq(s, c, a, b);
    return n;
}

/* r-th unordered equal pair in fresh-initFlips order (component, index1,
 * index2).  Returns 0 when r is out of range. */
int fgo_pick_canonical(const fgo_scheme *s, uint32_t r, int *c, int *t1, int *t2) {
    for (int cc = 0; cc < 3; cc++)
        for (int a = 0; a < s->m; a++)
            for (int b = a + 1; b < s->m; b++)
                if (fac_eq(s, cc, a, b)) {
                    if (r == 0) { *c = cc; *t1 = a; *t2 = b; return 1; }
                    r--;
                }
    return 0;
}

static void copy_term(fgo_scheme *s, int dst, int src) {
    for (int c = 0; c < 3; c++) {
        s->v[c][dst] = s->v[c][src];
        s->s[c][dst] = s->s[c][src];
        s->ok[c][dst] = s->ok[c][src];
    }
}

/* removeZeroes.  Z2 (scheme_z2.cu:152-164) trims trailing zero terms first
 * and does not re-check a swapped-in term; ZT (scheme_integer.cu:237-249)
 * re-checks. */
static void remove_zeroes(fgo_scheme *s) {
    if (s->ring == FGO_RING_Z2) {
        while (s->m > 0 && !(nz(s, 0, s->m - 1) && nz(s, 1, s->m - 1) && nz(s, 2, s->m - 1))) s->m--;
        for (int t = 0; t < s->m; t++)
            if (!(nz(s, 0, t) && nz(s, 1, t) && nz(s, 2, t))) {
                s->m--;
                copy_term(s, t, s->m);
            }
    } else {
        for (int t = 0; t < s->m; t++) {
            if (nz(s, 0, t) && nz(s, 1, t) && nz(s, 2, t)) continue;
            s->m--;
            copy_term(s, t, s->m);
            t--;
        }
    }
}

/* removeAt (scheme_z2.cu:166-175, scheme_integer.cu:251-260) */
static void remove_at(fgo_scheme *s, int t) {
    s->m--;
    if (t == s->m) return;
    copy_term(s, t, s->m);
}

static void add_triplet(fgo_scheme *s, int i, int j, int k, tv u, tv v, tv w) {
    int t = s->m;
    tv_put(s, i, t, u); tv_put(s, j, t, v); tv_put(s, k, t, w);
    s->m++;
}

/* fixSigns (scheme_integer.cu:398-425) */
int fgo_fix_signs(fgo_scheme *s) {
    int changed = 0;
    for (int t = 0; t < s->m; t++) {
        int pi = first_pos(s->v[0][t], s->s[0][t]);
        int pj = first_pos(s->v[1][t], s->s[1][t]);
        if (pi && pj) continue;
        if (!pi && !pj) { tv_put(s, 0, t, tv_inverse(tv_get(s, 0, t))); tv_put(s, 1, t, tv_inverse(tv_get(s, 1, t))); }
        else if (!pi) { tv_put(s, 0, t, tv_inverse(tv_get(s, 0, t))); tv_put(s, 2, t, tv_inverse(tv_get(s, 2, t))); }
        else { tv_put(s, 1, t, tv_inverse(tv_get(s, 1, t))); tv_put(s, 2, t, tv_inverse(tv_get(s, 2, t))); }
        changed = 1;
    }
    return changed;
}

/* reduce (scheme_z2.cu:631-641); reduceAdd/reduceSub (scheme_integer.cu:549-571) */
void fgo_reduce(fgo_scheme *s, int i, int i1, int i2, int sub) {
    int zero;
    if (s->ring == FGO_RING_Z2) {
        s->v[i][i1] ^= s->v[i][i2];
        zero = s->v[i][i1] == 0;
    } else {
        tv a = tv_get(s, i, i1), b = tv_get(s, i, i2);
        tv r = sub ? tv_sub(a, b) : tv_add(a, b);
        tv_put(s, i, i1, r);
        zero = r.v == 0;
    }
    remove_at(s, i2);
    if (zero) remove_zeroes(s);
    fgo_init_flips(s);
}

/* tryReduce (scheme_z2.cu:1127-1154, scheme_integer.cu:1062-1101) over the
 * component-0 pairs, then the component-1 pairs.  The pair lists are taken
 * as stored (REFERENCE order) or in fresh-list order, index1 < index2
 * (CANONICAL).  For the integer ring the orientation of a stored pair matters:
 * limitSub / limitSum with the sign-canonical test are not symmetric. */
static int try_reduce_pair_z2(fgo_scheme *s, int comp, int a, int b) {
    if (comp == 0) {
        if (s->v[1][a] == s->v[1][b]) { fgo_reduce(s, 2, a, b, 0); return 1; }
        if (s->v[2][a] == s->v[2][b]) { fgo_reduce(s, 1, a, b, 0); return 1; }
        return 0;
    }
    if (s->v[2][a] == s->v[2][b]) { fgo_reduce(s, 0, a, b, 0); return 1; }
    return 0;
}

static int try_reduce_pair_zt(fgo_scheme *s, int comp, int a, int b) {
    if (comp == 0) {
        if (tv_eq(tv_get(s, 1, a), tv_get(s, 1, b)) && tv_limit_sum(tv_get(s, 2, a), tv_get(s, 2, b), 0)) { fgo_reduce(s, 2, a, b, 0); return 1; }
        int c2 = tv_cmp(tv_get(s, 2, a), tv_get(s, 2, b));
        if (c2 == 1 && tv_limit_sum(tv_get(s, 1, a), tv_get(s, 1, b), 1)) { fgo_reduce(s, 1, a, b, 0); return 1; }
        if (c2 == -1 && tv_limit_sub(tv_get(s, 1, a), tv_get(s, 1, b), 1)) { fgo_reduce(s, 1, a, b, 1); return 1; }
        return 0;
    }
    int c2 = tv_cmp(tv_get(s, 2, a), tv_get(s, 2, b));
    if (c2 == 1 && tv_limit_sum(tv_get(s, 0, a), tv_get(s, 0, b), 1)) { fgo_reduce(s, 0, a, b, 0); return 1; }
    if (c2 == -1 && tv_limit_sub(tv_get(s, 0, a), tv_get(s, 0, b), 1)) { fgo_reduce(s, 0, a, b, 1); return 1; }
    return 0;
}

static int try_reduce_mode(fgo_scheme *s, int mode) {
    int (*pair)(fgo_scheme *, int, int, int) = s->ring == FGO_RING_Z2 ? try_reduce_pair_z2 : try_reduce_pair_zt;
    for (int comp = 0; comp < 2; comp++) {
        if (mode == FGO_SELECT_REFERENCE) {
            for (int q = 0; q < s->flips[comp].size; q++)
                if (pair(s, comp, (int)(s->flips[comp].pairs[q] >> 16), (int)(s->flips[comp].pairs[q] & 0xFFFF))) return 1;
        } else {
            for (int a = 0; a < s->m; a++)
                for (int b = a + 1; b < s->m; b++)
                    if (fac_eq(s, comp, a, b) && pair(s, comp, a, b)) return 1;
        }
    }
    return 0;
}

/* On a freshly built list both orders coincide. */
int fgo_try_reduce(fgo_scheme *s) { return try_reduce_mode(s, FGO_SELECT_REFERENCE); }

/* ------------------------------------------------------------------ */
/* flip                                                                */
/* ------------------------------------------------------------------ */
/* Z2 flip (scheme_z2.cu:531-577) */
static void flip_z2(fgo_scheme *s, int i, int j, int k, int i1, int i2, int check_reduce) {
    s->v[j][i1] ^= s->v[j][i2];
    s->v[k][i2] ^= s->v[k][i1];
    fs_remove_index(&s->flips[j], (uint32_t)i1);
    fs_remove_index(&s->flips[k], (uint32_t)i2);
    if (!s->v[j][i1] || !s->v[k][i2]) {
        remove_zeroes(s);
        fgo_init_flips(s);
        return;
    }
    for (int t = 0; t < s->m; t++) {
        if (t != i1 && s->v[j][t] == s->v[j][i1]) {
            if (check_reduce) {
                if (s->v[i][t] == s->v[i][i1]) { fgo_reduce(s, k, t, i1, 0); return; }
                if (s->v[k][t] == s->v[k][i1]) { fgo_reduce(s, i, t, i1, 0); return; }
            }
            fs_add(&s->flips[j], (uint32_t)i1, (uint32_t)t);
        }
        if (t != i2 && s->v[k][t] == s->v[k][i2]) {
            if (check_reduce) {
                if (s->v[i][t] == s->v[i][i2]) { fgo_reduce(s, j, t, i2, 0); return; }
                if (s->v[j][t] == s->v[j][i2]) { fgo_reduce(s, i, t, i2, 0); return; }
            }
            fs_add(&s->flips[k], (uint32_t)i2, (uint32_t)t);
        }
    }
}

/* integer flip (scheme_integer.cu:428-501) */
static void flip_zt(fgo_scheme *s, int i, int j, int k, int i1, int i2, int check_reduce) {
    tv_put(s, j, i1, tv_add(tv_get(s, j, i1), tv_get(s, j, i2)));
    tv_put(s, k, i2, tv_sub(tv_get(s, k, i2), tv_get(s, k, i1)));
    fs_remove_index(&s->flips[j], (uint32_t)i1);
    fs_remove_index(&s->flips[k], (uint32_t)i2);
    if (!nz(s, j, i1) || !nz(s, k, i2)) {
        remove_zeroes(s);
        fgo_init_flips(s);
        while (fgo_try_reduce(s)) {}
        return;
    }
    for (int t = 0; t < s->m; t++) {
        if (t != i1 && fac_eq(s, j, t, i1)) {
            if (check_reduce) {
                int cmpI = tv_cmp(tv_get(s, i, t), tv_get(s, i, i1));
                if (cmpI == 1 && tv_limit_sum(tv_get(s, k, t), tv_get(s, k, i1), k != 2)) { fgo_reduce(s, k, t, i1, 0); return; }
                if (i == 2 && cmpI == -1 && tv_limit_sub(tv_get(s, k, t), tv_get(s, k, i1), 1)) { fgo_reduce(s, k, t, i1, 1); return; }
                int cmpK = tv_cmp(tv_get(s, k, t), tv_get(s, k, i1));
                if (cmpK == 1 && tv_limit_sum(tv_get(s, i, t), tv_get(s, i, i1), i != 2)) { fgo_reduce(s, i, t, i1, 0); return; }
                if (k == 2 && cmpK == -1 && tv_limit_sub(tv_get(s, i, t), tv_get(s, i, i1), 1)) { fgo_reduce(s, i, t, i1, 1); return; }
            }
            fs_add(&s->flips[j], (uint32_t)i1, (uint32_t)t);
        }
        if (t != i2 && fac_eq(s, k, t, i2)) {
            if (check_reduce) {
                int cmpI = tv_cmp(tv_get(s, i, t), tv_get(s, i, i2));
                if (cmpI == 1 && tv_limit_sum(tv_get(s, j, t), tv_get(s, j, i2), j != 2)) { fgo_reduce(s, j, t, i2, 0); return; }
                if (i == 2 && cmpI == -1 && tv_limit_sub(tv_get(s, j, t), tv_get(s, j, i2), 1)) { fgo_reduce(s, j, t, i2, 1); return; }
                int cmpJ = tv_cmp(tv_get(s, j, t), tv_get(s, j, i2));
                if (cmpJ == 1 && tv_limit_sum(tv_get(s, i, t), tv_get(s, i, i2), i != 2)) { fgo_reduce(s, i, t, i2, 0); return; }
                if (j == 2 && cmpJ == -1 && tv_limit_sub(tv_get(s, i, t), tv_get(s, i, i2), 1)) { fgo_reduce(s, i, t, i2, 1); return; }
            }
            fs_add(&s->flips[k], (uint32_t)i2, (uint32_t)t);
        }
    }
}

void fgo_flip(fgo_scheme *s, int i, int j, int k, int i1, int i2, int check_reduce) {
    if (s->ring == FGO_RING_Z2) flip_z2(s, i, j, k, i1, i2, check_reduce);
    else flip_zt(s, i, j, k, i1, i2, check_reduce);
}

/* plus (scheme_z2.cu:579-612, scheme_integer.cu:503-539) */
void fgo_plus(fgo_scheme *s, int i, int j, int k, int i1, int i2, int variant) {
    if (s->ring == FGO_RING_Z2) {
        uint64_t a1 = s->v[i][i1], b1 = s->v[j][i1], c1 = s->v[k][i1];
        uint64_t a2 = s->v[i][i2], b2 = s->v[j][i2], c2 = s->v[k][i2];
        uint64_t a = a1 ^ a2, b = b1 ^ b2, c = c1 ^ c2;
        tv T1 = {0, 0, 1}, T2 = {0, 0, 1}, T3 = {0, 0, 1};
        if (variant == 0) {
            s->v[j][i1] = b; s->v[i][i2] = a;
            T1.v = a1; T2.v = b2; T3.v = c;
        } else if (variant == 1) {
            s->v[k][i1] = c; s->v[j][i2] = b;
            T1.v = a; T2.v = b1; T3.v = c2;
        } else {
            s->v[i][i1] = a; s->v[k][i2] = c;
            T1.v = a2; T2.v = b; T3.v = c1;
        }
        add_triplet(s, i, j, k, T1, T2, T3);
        if (!a || !b || !c) remove_zeroes(s);
        fgo_init_flips(s);
        return;
    }
    tv a1 = tv_get(s, i, i1), b1 = tv_get(s, j, i1), c1 = tv_get(s, k, i1);
    tv a2 = tv_get(s, i, i2), b2 = tv_get(s, j, i2), c2 = tv_get(s, k, i2);
    tv aAdd = tv_add(a1, a2), bAdd = tv_add(b1, b2), cAdd = tv_add(c1, c2);
    tv aSub = tv_sub(a2, a1), bSub = tv_sub(b2, b1), cSub = tv_sub(c2, c1);
    if (variant == 0 && tv_limit(aSub, i != 2) && tv_limit(bAdd, j != 2) && tv_limit(cSub, k != 2)) {
        tv_put(s, j, i1, bAdd); tv_put(s, i, i2, aSub);
        add_triplet(s, i, j, k, a1, b2, cSub);
    } else if (variant == 1 && tv_limit(aSub, i != 2) && tv_limit(bSub, j != 2) && tv_limit(cAdd, k != 2)) {
        tv_put(s, k, i1, cAdd); tv_put(s, j, i2, bSub);
        add_triplet(s, i, j, k, aSub, b1, c2);
    } else if (tv_limit(aAdd, i != 2) && tv_limit(bSub, j != 2) && tv_limit(cSub, k != 2)) {
        tv_put(s, i, i1, aAdd); tv_put(s, k, i2, cSub);
        add_triplet(s, i, j, k, a2, bSub, c1);
    }
    remove_zeroes(s);
    fgo_fix_signs(s);
    fgo_init_flips(s);
}

/* split (scheme_z2.cu:614-619) */
void fgo_split_z2(fgo_scheme *s, int i, int j, int k, int index, uint64_t a1) {
    uint64_t a2 = s->v[i][index] ^ a1;
    s->v[i][index] = a1;
    tv u = {a2, 0, 1}, v = {s->v[j][index], 0, 1}, w = {s->v[k][index], 0, 1};
    add_triplet(s, i, j, k, u, v, w);
    fgo_init_flips(s);
}

/* split (scheme_integer.cu:541-547) */
void fgo_split_zt(fgo_scheme *s, int i, int j, int k, int index, uint64_t av, uint64_t as) {
    tv a = {av, as & av, 1};
    tv d = tv_sub(tv_get(s, i, index), a);
    add_triplet(s, i, j, k, d, tv_get(s, j, index), tv_get(s, k, index));
    tv_put(s, i, index, a);
    fgo_fix_signs(s);
    fgo_init_flips(s);
}

/* ------------------------------------------------------------------ */
/* random operators                                                    */
/* ------------------------------------------------------------------ */
static int naive_rank(const fgo_scheme *s) { return s->n[0] * s->n[1] * s->n[2]; }

static void pick_orientation(int c, int coin, int *i, int *j, int *k) {
    *i = c;
    *j = c == 0 ? 1 : 0;
    *k = c == 2 ? 1 : 2;
    if (coin & 1) { int t = *j; *j = *k; *k = t; }
}

/* decode a stored-list index (scheme_z2.cu:990-1011) */
static void decode_reference(const fgo_scheme *s, int idx, int *c, int *t1, int *t2) {
    int cc = 0;
    while (idx >= s->flips[cc].size) { idx -= s->flips[cc].size; cc++; }
    *c = cc;
    *t1 = (int)(s->flips[cc].pairs[idx] >> 16);
    *t2 = (int)(s->flips[cc].pairs[idx] & 0xFFFF);
}

static int total_pairs(const fgo_scheme *s, int mode) {
    if (mode == FGO_SELECT_REFERENCE) return s->flips[0].size + s->flips[1].size + s->flips[2].size;
    return fgo_count_pairs(s);
}

static void select_pair(const fgo_scheme *s, int mode, uint32_t r, int *c, int *t1, int *t2) {
    if (mode == FGO_SELECT_REFERENCE) decode_reference(s, (int)r, c, t1, t2);
    else fgo_pick_canonical(s, r, c, t1, t2);
}

/* Z2 tryFlip (scheme_z2.cu:982-1027): b[0] picks the pair (multiply-high), bit 0 of b[1]
 * swaps j/k, bit 1 swaps index1/index2. */
/* pair (component, index1 < index2) of the last Z2 tryFlip of this thread */
static __thread int tl_sel[3];

static int try_flip_z2(fgo_scheme *s, const uint32_t b[4], const fgo_walk_params *p) {
    int size = total_pairs(s, p->select_mode);
    if (!size) return 0;
    int c, t1, t2, i, j, k;
    select_pair(s, p->select_mode, draw_below(b[0], (uint32_t)size), &c, &t1, &t2);
    tl_sel[0] = c; tl_sel[1] = t1; tl_sel[2] = t2;
    pick_orientation(c, (int)b[1], &i, &j, &k);
    if (b[1] & 2) { int t = t1; t1 = t2; t2 = t; }
    flip_z2(s, i, j, k, t1, t2, p->check_reduce);
    return 1;
}

/* Reduction-seeking lookahead (no reference counterpart): whether flip(i, j,
 * k, i1, i2) of the Z2 scheme would lower the rank -- a zero factor (the term
 * disappears) or a term that afterwards matches i1 or i2 in two components
 * (the reduction flip() takes, scheme_z2.cu:544-576). */
static int z2_flip_reduces(const fgo_scheme *s, int i, int j, int k, int i1, int i2) {
    const uint64_t A = s->v[j][i1] ^ s->v[j][i2], B = s->v[k][i2] ^ s->v[k][i1];
    if (!A || !B || !s->v[j][i1] || !s->v[j][i2] || !s->v[k][i1] || !s->v[k][i2]) return 1;
    for (int t = 0; t < s->m; t++) {
        if (t != i1 && s->v[j][t] == A && (s->v[i][t] == s->v[i][i1] || s->v[k][t] == s->v[k][i1])) return 1;
        if (t != i2 && s->v[k][t] == B && (s->v[i][t] == s->v[i][i2] || s->v[j][t] == s->v[j][i2])) return 1;
    }
    return 0;
}

/* Lookahead proposals: `count` candidates drawn as tryFlip draws them (one
 * aux word picks the pair in fresh-list order, one gives the orientation
 * coins); the first whose flip lowers the rank is applied.  0 when none does. */
static int try_lookahead_z2(fgo_scheme *s, fgo_rng *r, int count, int check_reduce) {
    const int size = fgo_count_pairs(s);
    if (!size) return 0;
    for (int q = 0; q < count; q++) {
        const uint32_t x = fgo_rng_aux32(r), coin = fgo_rng_aux32(r);
        int c, t1, t2, i, j, k;
        fgo_pick_canonical(s, draw_below(x, (uint32_t)size), &c, &t1, &t2);
        pick_orientation(c, (int)coin, &i, &j, &k);
        if (coin & 2) { int t = t1; t1 = t2; t2 = t; }
        if (z2_flip_reduces(s, i, j, k, t1, t2)) {
            flip_z2(s, i, j, k, t1, t2, check_reduce);
            return 1;
        }
    }
    return 0;
}

/* A given pair flipped with tryFlip's orientation coins; 0 when the two
 * factors differ (the pair is no longer a flip). */
static int try_flip_pair_z2(fgo_scheme *s, int c, int t1, int t2, uint32_t coin, int check_reduce) {
    if (c < 0 || c > 2 || t1 < 0 || t2 <= t1 || t2 >= s->m || !fac_eq(s, c, t1, t2)) return 0;
    tl_sel[0] = c; tl_sel[1] = t1; tl_sel[2] = t2;
    int i, j, k;
    pick_orientation(c, (int)coin, &i, &j, &k);
    if (coin & 2) { int t = t1; t1 = t2; t2 = t; }
    flip_z2(s, i, j, k, t1, t2, check_reduce);
    return 1;
}

/* Integer tryFlip (scheme_integer.cu:907-975).  The reference draws a
 * Fisher-Yates permutation of all candidate pairs up front
 * (random.cuh:9-22), draws two orientation coins per visited candidate and
 * applies the first candidate whose coins make it admissible; it fails only
 * when every candidate, with its coins, is inadmissible.  REFERENCE mode
 * does exactly that (stored-list order, one aux word per permutation step,
 * one per coin).
 *
 * CANONICAL mode (the GPU's rule) draws the same process lazily.
 * Phase 1 visits up to min(zt_visit_cap, FGO_ZT_MAX_VISITS) candidates one
 * at a time: visit q draws k uniform in [0, size - q) (visit 0 from the main
 * block b[0], later visits from the aux stream) and takes the k-th candidate
 * not yet visited, so it is uniform over the unvisited ones; its coins are
 * b[1] for visit 0, an aux word later (bit 0 swaps j/k, bit 1 swaps the
 * indices).  When every visited candidate is inadmissible and unvisited ones
 * remain, phase 2 gives each unvisited candidate (fresh-list order) its own
 * coin word from the aux stream and picks uniformly among the admissible
 * ones by reservoir sampling (one aux word per admissible candidate after
 * the first).  The first admissible candidate of a uniformly random order,
 * each candidate with independent coins, is a uniform pick among the
 * candidates admissible under their coins, so both phases follow the
 * reference's distribution, and tryFlip fails exactly when no unvisited
 * candidate is admissible under its coins. */
#define FGO_ZT_MAX_VISITS 16

/* The two admissibility tests of one candidate (component c, indices t1 <
 * t2) under coin word `coin` (scheme_integer.cu:944-971); on success op
 * holds the flip (i, j, k, index1, index2) to apply. */
static int zt_candidate(const fgo_scheme *s, int c, int t1, int t2, uint32_t coin, int op[5]) {
    int i, j, k;
    pick_orientation(c, (int)coin, &i, &j, &k);
    if (coin & 2) { int t = t1; t1 = t2; t2 = t; }
    tv j1 = tv_get(s, j, t1), j2 = tv_get(s, j, t2), k1 = tv_get(s, k, t1), k2 = tv_get(s, k, t2);
    if (tv_limit_sum(j1, j2, j != 2) && tv_limit_sub(k2, k1, 0)) {
        int fwd = k == 2 || tv_pos_sub(k2, k1);
        op[0] = i; op[1] = j; op[2] = k; op[3] = fwd ? t1 : t2; op[4] = fwd ? t2 : t1;
        return 1;
    }
    if (tv_limit_sum(k1, k2, k != 2) && tv_limit_sub(j2, j1, 0)) {
        int fwd = j == 2 || tv_pos_sub(j2, j1);
        op[0] = i; op[1] = k; op[2] = j; op[3] = fwd ? t1 : t2; op[4] = fwd ? t2 : t1;
        return 1;
    }
    return 0;
}

/* flip of the last integer tryFlip of this thread (i, j, k, index1, index2) */
static __thread int tl_zt_op[5];

static void apply_zt(fgo_scheme *s, const int op[5], int check_reduce) {
    for (int q = 0; q < 5; q++) tl_zt_op[q] = op[q];
    flip_zt(s, op[0], op[1], op[2], op[3], op[4], check_reduce);
}

/* REFERENCE mode: the reference's permutation, in stored-list order. */
static int try_flip_zt_reference(fgo_scheme *s, fgo_rng *r, int check_reduce) {
    int size = total_pairs(s, FGO_SELECT_REFERENCE);
    if (!size) return 0;
    int *perm = (int *)malloc(sizeof(int) * (size_t)size);
    for (int q = 0; q < size; q++) perm[q] = q;
    for (int q = size - 1; q > 0; q--) { /* randomPermutation, random.cuh:14-21 */
        int x = (int)(fgo_rng_aux32(r) % (uint32_t)(q + 1));
        int t = perm[q]; perm[q] = perm[x]; perm[x] = t;
    }
    int ok = 0, op[5];
    for (int q = 0; q < size && !ok; q++) {
        int c, t1, t2;
        decode_reference(s, perm[q], &c, &t1, &t2);
        uint32_t coin = fgo_rng_aux32(r) % 2u;           /* :944-948 */
        coin |= (fgo_rng_aux32(r) % 2u) << 1;            /* :950-954 */
        ok = zt_candidate(s, c, t1, t2, coin, op);
    }
    free(perm);
    if (ok) apply_zt(s, op, check_reduce);
    return ok;
}

static int try_flip_zt(fgo_scheme *s, const uint32_t b[4], fgo_rng *r, const fgo_walk_params *p) {
    if (p->select_mode == FGO_SELECT_REFERENCE) return try_flip_zt_reference(s, r, p->check_reduce);
    int size = fgo_count_pairs(s);
    if (!size) return 0;
    int cap = p->zt_visit_cap > 0 && p->zt_visit_cap < FGO_ZT_MAX_VISITS ? p->zt_visit_cap : FGO_ZT_MAX_VISITS;
    if (cap > size) cap = size;
    int visited[FGO_ZT_MAX_VISITS]; /* ascending */
    int nv = 0, op[5], c, t1, t2;
    for (int q = 0; q < cap; q++) {
        int cand = (int)draw_below(q == 0 ? b[0] : fgo_rng_aux32(r), (uint32_t)(size - q));
        for (int o = 0; o < nv; o++) cand += visited[o] <= cand;  /* the cand-th unvisited */
        int pos = nv++;
        while (pos > 0 && visited[pos - 1] > cand) { visited[pos] = visited[pos - 1]; pos--; }
        visited[pos] = cand;
        uint32_t coin = q == 0 ? b[1] : fgo_rng_aux32(r);
        fgo_pick_canonical(s, (uint32_t)cand, &c, &t1, &t2);
        if (zt_candidate(s, c, t1, t2, coin, op)) {
            apply_zt(s, op, p->check_reduce);
            return 1;
        }
    }
    if (nv >= size) return 0;
    /* phase 2: reservoir pick among the unvisited candidates admissible
     * under their own coins */
    int n = 0, pv = 0, idx = 0, bc = 0, bt1 = 0, bt2 = 0;
    uint32_t bcoin = 0;
    for (int cc = 0; cc < 3; cc++)
        for (int a = 0; a < s->m; a++)
            for (int bb = a + 1; bb < s->m; bb++) {
                if (!fac_eq(s, cc, a, bb)) continue;
                int cur = idx++;
                if (pv < nv && visited[pv] == cur) { pv++; continue; }
                uint32_t coin = fgo_rng_aux32(r);
                if (!zt_candidate(s, cc, a, bb, coin, op)) continue;
                n++;
                if (n == 1 || draw_below(fgo_rng_aux32(r), (uint32_t)n) == 0) {
                    bc = cc; bt1 = a; bt2 = bb; bcoin = coin;
                }
            }
    if (!n) return 0;
    zt_candidate(s, bc, bt1, bt2, bcoin, op);
    apply_zt(s, op, p->check_reduce);
    return 1;
}

/* tryPlus (scheme_z2.cu:1029-1050, scheme_integer.cu:977-998).  The
 * reference rejection loop has no bound; here it gives up after
 * FGO_PLUS_MAX_TRIES redraws. */
static int try_plus(fgo_scheme *s, fgo_rng *r) {
    if (s->m >= FGO_MAX_RANK || s->m >= naive_rank(s)) return 0;
    uint32_t m = (uint32_t)s->m;
    int i1 = (int)(fgo_rng_aux32(r) % m);
    int i2 = (int)(fgo_rng_aux32(r) % m);
    int tries = 0;
    while (i1 == i2 || fac_eq(s, 0, i1, i2) || fac_eq(s, 1, i1, i2) || fac_eq(s, 2, i1, i2)) {
        if (++tries > FGO_PLUS_MAX_TRIES) return 0;
        i1 = (int)(fgo_rng_aux32(r) % m);
        i2 = (int)(fgo_rng_aux32(r) % m);
    }
    int p[3];
    aux_perm3(r, p);
    int variant = (int)(fgo_rng_aux32(r) % 3u);
    fgo_plus(s, p[0], p[1], p[2], i1, i2, variant);
    return 1;
}

/* Z2 trySplit (scheme_z2.cu:1052-1078), including the reference's
 * give-up-when-loop-reaches-5 rule. */
static int try_split_z2(fgo_scheme *s, fgo_rng *r) {
    if (s->m >= FGO_MAX_RANK || s->m >= naive_rank(s)) return 0;
    int index = (int)(fgo_rng_aux32(r) % (uint32_t)s->m);
    int p[3];
    aux_perm3(r, p);
    int i = p[0], j = p[1], k = p[2];
    uint64_t mod = s->nn[i] >= 64 ? 0 : (1ull << s->nn[i]);
    uint64_t a1 = mod ? (uint64_t)fgo_rng_aux32(r) % mod : (uint64_t)fgo_rng_aux32(r);
    uint64_t a2 = s->v[i][index];
    int loop = 0;
    while (loop < 5 && a1 == a2) {
        a1 = mod ? (uint64_t)fgo_rng_aux32(r) % mod : (uint64_t)fgo_rng_aux32(r);
        loop++;
    }
    if (loop == 5) return 0;
    fgo_split_z2(s, i, j, k, index, a1);
    return 1;
}

/* Z2 trySplitExisted (scheme_z2.cu:1080-1104) */
static int try_split_existed_z2(fgo_scheme *s, fgo_rng *r) {
    if (s->m >= FGO_MAX_RANK || s->m >= naive_rank(s)) return 0;
    int p[3];
    aux_perm3(r, p);
    int i = p[0], j = p[1], k = p[2];
    uint32_t m = (uint32_t)s->m;
    int i1 = (int)(fgo_rng_aux32(r) % m);
    int i2 = (int)(fgo_rng_aux32(r) % m);
    int loop = 0;
    while (loop < 5 && (i1 == i2 || s->v[i][i1] == s->v[i][i2])) {
        i2 = (int)(fgo_rng_aux32(r) % m);
        loop++;
    }
    if (loop == 5) return 0;
    fgo_split_z2(s, i, j, k, i1, s->v[i][i2]);
    return 1;
}

/* Integer trySplit (scheme_integer.cu:1000-1020).  Addition::random
 * (addition.cu:68-80) draws sign bits then value bits, 16 at a time; the
 * reference does not mask them to nn bits (and sign-extends the second
 * half-word), which leaves coefficients outside the matrix.  Here the
 * random vector is masked to nn bits. */
static int try_split_zt(fgo_scheme *s, fgo_rng *r) {
    if (s->m >= FGO_MAX_RANK || s->m >= naive_rank(s)) return 0;
    int index = (int)(fgo_rng_aux32(r) % (uint32_t)s->m);
    int p[3];
    aux_perm3(r, p);
    int i = p[0], j = p[1], k = p[2];
    uint64_t av = 0, as = 0;
    for (int x = 0; x < s->nn[i]; x += 16) {
        uint64_t sb = fgo_rng_aux32(r) & 0xFFFFu;
        uint64_t vb = fgo_rng_aux32(r) & 0xFFFFu;
        av |= vb << x;
        as |= sb << x;
    }
    uint64_t mask = s->nn[i] >= 64 ? ~0ull : ((1ull << s->nn[i]) - 1);
    av &= mask;
    as &= av;
    tv a = {av, as, 1};
    if (!tv_limit_sub(tv_get(s, i, index), a, i != 2)) return 0;
    fgo_split_zt(s, i, j, k, index, av, as);
    return 1;
}

/* Integer trySplitExisted (scheme_integer.cu:1022-1043); the unbounded
 * do-while is bounded by FGO_PLUS_MAX_TRIES. */
static int try_split_existed_zt(fgo_scheme *s, fgo_rng *r) {
    if (s->m >= FGO_MAX_RANK || s->m >= naive_rank(s)) return 0;
    uint32_t m = (uint32_t)s->m;
    int i1, i2, i, tries = 0;
    do {
        if (++tries > FGO_PLUS_MAX_TRIES) return 0;
        i1 = (int)(fgo_rng_aux32(r) % m);
        i2 = (int)(fgo_rng_aux32(r) % m);
        i = (int)(fgo_rng_aux32(r) % 3u);
    } while (i1 == i2 || fac_eq(s, i, i1, i2));
    if (!tv_limit_sub(tv_get(s, i, i1), tv_get(s, i, i2), i != 2)) return 0;
    int j = (i + 1) % 3, k = (i + 2) % 3;
    tv a = tv_get(s, i, i2);
    fgo_split_zt(s, i, j, k, i1, a.v, a.s);
    return 1;
}

/* tryExpand (scheme_z2.cu:1106-1125, scheme_integer.cu:1045-1060) */
static int try_expand(fgo_scheme *s, int count, fgo_rng *r) {
    int res = 0;
    if (s->ring == FGO_RING_Z2) {
        int maxr = naive_rank(s);
        for (int q = 0; q < count && s->m < maxr; q++) {
            int v = (int)(fgo_rng_aux32(r) % 3u);
            if (v == 0) res |= try_plus(s, r);
            else if (v == 1) res |= try_split_z2(s, r);
            else res |= try_split_existed_z2(s, r);
        }
    } else {
        for (int q = 0; q < count && s->m < FGO_MAX_RANK; q++) {
            int v = (int)(fgo_rng_aux32(r) % 3u);
            if (v == 0) res |= try_plus(s, r);
            else if (v == 1) res |= try_split_zt(s, r);
            else res |= try_split_existed_zt(s, r);
        }
    }
    return res;
}

/* ------------------------------------------------------------------ */
/* walk (src/entities/flip_graph.cu:520-573)                           */
/* ------------------------------------------------------------------ */
void fgo_copy(fgo_scheme *dst, const fgo_scheme *src, int with_flips) {
    dst->ring = src->ring;
    dst->m = src->m;
    for (int c = 0; c < 3; c++) {
        dst->n[c] = src->n[c];
        dst->nn[c] = src->nn[c];
        for (int t = 0; t < src->m; t++) {
            dst->v[c][t] = src->v[c][t];
            dst->s[c][t] = src->s[c][t];
            dst->ok[c][t] = src->ok[c][t];
        }
    }
    if (with_flips) fgo_init_flips(dst);
}

void fgo_walker_init(fgo_walker *w, const fgo_scheme *start, int best_rank, uint64_t seed, uint32_t walker) {
    memset(w, 0, sizeof(*w));
    fgo_copy(&w->cur, start, 1);
    fgo_copy(&w->best, start, 0); /* initializeSchemesKernel, flip_graph.cu:510-518 */
    w->best_rank = best_rank;
    w->flips_count = 0;
    fgo_rng_init(&w->rng, seed, walker);
    w->forced[0] = -1;
    w->log[0] = -1;
}

/* minimizeKernel (src/entities/complexity_minimizer.cu:190-214): flips without
 * reductions until a flip fails; the walker's best_rank holds its best
 * complexity and best the scheme that reached it. */
static void minimize(fgo_walker *w, const fgo_walk_params *p) {
    fgo_scheme *s = &w->cur;
    for (int it = 0; it < p->iterations; it++) {
        uint32_t b[4];
        fgo_rng_main_block(&w->rng, b);
        int rank = s->m;
        int ok = s->ring == FGO_RING_Z2 ? try_flip_z2(s, b, p) : try_flip_zt(s, b, &w->rng, p);
        if (!ok) break;
        w->total_flips++;
        w->sum_m += (uint64_t)rank;
        int cx = fgo_complexity(s);
        if (cx < w->best_rank) {
            w->best_rank = cx;
            fgo_copy(&w->best, s, 0);
        }
    }
}

/* flipSchemesKernel (src/entities/scheme_additions_reducer.cu:448-460):
 * every walker but walker 0 draws a count in [0, iterations] (:454, from
 * the aux stream) and applies that many flips with reductions (:457-459),
 * stopping at the first that finds no pair. */
static void flip_schemes(fgo_walker *w, const fgo_walk_params *p) {
    if (w->rng.walker == 0) return; /* :451 */
    fgo_scheme *s = &w->cur;
    const int count = (int)(fgo_rng_aux32(&w->rng) % (uint32_t)(p->iterations + 1));
    fgo_walk_params q = *p;
    q.check_reduce = 1;
    for (int it = 0; it < count; it++) {
        uint32_t b[4];
        fgo_rng_main_block(&w->rng, b);
        int rank = s->m;
        int ok = s->ring == FGO_RING_Z2 ? try_flip_z2(s, b, &q) : try_flip_zt(s, b, &w->rng, &q);
        if (!ok) break;
        w->total_flips++;
        w->sum_m += (uint64_t)rank;
    }
}

static void sandwiching(fgo_scheme *s, fgo_rng *r);
static void swap_basis(fgo_scheme *s, fgo_rng *r);

void fgo_walk(fgo_walker *w, const fgo_walk_params *p) {
    if (p->mode == 1) {
        minimize(w, p);
        return;
    }
    if (p->mode == 3) {
        flip_schemes(w, p);
        return;
    }
    fgo_scheme *s = &w->cur;
    int iters = p->iterations;
    if (p->random_iterations) iters = aux_randint(&w->rng, 1, p->max_iterations); /* :530 */
    w->min_m = s->m;
    w->log[0] = -1;
    for (int it = 0; it < iters; it++) {
        uint32_t b[4];
        fgo_rng_main_block(&w->rng, b);
        int rank = s->m;
        int ok = 0;
        if (it == 0 && w->forced[0] >= 0 && s->ring == FGO_RING_Z2)
            ok = try_flip_pair_z2(s, w->forced[0], w->forced[1], w->forced[2], b[1], p->check_reduce);
        if (!ok && p->look_candidates > 0 && p->look_period > 0 && s->ring == FGO_RING_Z2 &&
            it % p->look_period == p->look_period - 1)
            ok = try_lookahead_z2(s, &w->rng, p->look_candidates, p->check_reduce);
        if (!ok) ok = s->ring == FGO_RING_Z2 ? try_flip_z2(s, b, p) : try_flip_zt(s, b, &w->rng, p);
        if (it == 0) {
            w->forced[0] = -1;
            w->log[0] = ok && s->ring == FGO_RING_Z2 ? tl_sel[0] : -1;
            w->log[1] = tl_sel[1];
            w->log[2] = tl_sel[2];
            w->log[3] = rank;
        }
        if (!ok) { /* :535-537 */
            try_expand(s, 1 + (int)(b[1] % 2u), &w->rng);
            if ((p->min_window <= 0 || it < p->min_window) && s->m < w->min_m) w->min_m = s->m;
            continue;
        }
        w->total_flips++;
        w->sum_m += (uint64_t)rank;
        if (s->m < rank) w->flips_count = 0; /* :540-543 */
        w->flips_count++;
        if (s->m < w->best_rank || (s->m == w->best_rank && b[2] < p->thr_best_eq)) { /* :545-548 */
            w->best_rank = s->m;
            fgo_copy(&w->best, s, 0);
        }
        if (p->use_reduce && fgo_rng_aux32(&w->rng) < p->thr_reduce) try_reduce_mode(s, p->select_mode); /* :550-551 */
        if (p->use_expand && b[3] < p->thr_expand && s->m <= w->best.m + 2) { /* :553-556 */
            if (try_expand(s, aux_randint(&w->rng, 1, 2), &w->rng)) w->flips_count = 0;
        }
        /* :558-562; the integer ring's sandwiching is empty (scheme_integer.cu:1215-1217) */
        if (p->use_sandwich && fgo_rng_aux32(&w->rng) < p->thr_sandwich && s->ring == FGO_RING_Z2) sandwiching(s, &w->rng);
        if (p->use_basis && fgo_rng_aux32(&w->rng) < p->thr_basis) swap_basis(s, &w->rng);
        if (w->flips_count >= p->plus_iterations && try_plus(s, &w->rng)) w->flips_count = 0; /* :564-565 */
        if ((p->min_window <= 0 || it < p->min_window) && s->m < w->min_m) w->min_m = s->m;
    }
}

/* ------------------------------------------------------------------ */
/* records and blobs                                                   */
/* ------------------------------------------------------------------ */
void fgo_walker_to_record(const fgo_walker *w, fgo_walker_record *rec) {
    memset(rec, 0, sizeof(*rec));
    for (int c = 0; c < 3; c++) rec->n[c] = w->cur.n[c];
    rec->m = w->cur.m;
    rec->best_m = w->best.m;
    rec->best_rank = w->best_rank;
    rec->flips_count = w->flips_count;
    rec->aux_idx = w->rng.aidx;
    rec->it = w->rng.it;
    rec->aux = w->rng.aux;
    rec->total_flips = w->total_flips;
    rec->sum_m = w->sum_m;
    for (int c = 0; c < 3; c++) rec->best_n[c] = w->best.n[c];
    for (int c = 0; c < 3; c++) {
        for (int t = 0; t < w->cur.m && t < FGO_REC_MAX_M; t++) {
            rec->cur_v[c][t] = w->cur.v[c][t];
            rec->cur_s[c][t] = w->cur.ring == FGO_RING_ZT ? w->cur.s[c][t] : 0;
        }
        for (int t = 0; t < w->best.m && t < FGO_REC_MAX_M; t++) {
            rec->best_v[c][t] = w->best.v[c][t];
            rec->best_s[c][t] = w->best.ring == FGO_RING_ZT ? w->best.s[c][t] : 0;
        }
    }
}

static void scheme_from_words(fgo_scheme *s, int ring, const int32_t n[3], int m, const uint64_t v[3][FGO_REC_MAX_M], const uint64_t sg[3][FGO_REC_MAX_M]) {
    memset(s, 0, sizeof(*s));
    s->ring = ring;
    for (int c = 0; c < 3; c++) s->n[c] = n[c];
    for (int c = 0; c < 3; c++) s->nn[c] = s->n[c] * s->n[(c + 1) % 3];
    s->m = m;
    for (int c = 0; c < 3; c++)
        for (int t = 0; t < m; t++) {
            s->v[c][t] = v[c][t];
            s->s[c][t] = ring == FGO_RING_ZT ? sg[c][t] : 0;
            s->ok[c][t] = 1;
        }
}

int fgo_walker_from_record(fgo_walker *w, int ring, const fgo_walker_record *rec, uint64_t seed, uint32_t walker) {
    if (rec->m < 0 || rec->m > FGO_REC_MAX_M || rec->best_m < 0 || rec->best_m > FGO_REC_MAX_M) return -1;
    memset(w, 0, sizeof(*w));
    w->forced[0] = -1;
    w->log[0] = -1;
    scheme_from_words(&w->cur, ring, rec->n, rec->m, rec->cur_v, rec->cur_s);
    fgo_init_flips(&w->cur);
    scheme_from_words(&w->best, ring, rec->best_n[0] > 0 ? rec->best_n : rec->n, rec->best_m, rec->best_v, rec->best_s);
    w->best_rank = rec->best_rank;
    w->flips_count = rec->flips_count;
    fgo_rng_init(&w->rng, seed, walker);
    w->rng.it = rec->it;
    w->rng.aux = rec->aux;
    w->rng.aidx = rec->aux_idx;
    if (w->rng.aidx < 4 && w->rng.aux > 0) aux_fill(&w->rng, w->rng.aux - 1);
    w->total_flips = rec->total_flips;
    w->sum_m = rec->sum_m;
    return 0;
}

typedef struct {
    int ring;
    fgo_walker_record *recs;
    int begin, end;
    uint32_t walker_base;
    uint64_t seed;
    const fgo_walk_params *p;
    int launches;
    int rc;
} walk_job;

static void *walk_job_run(void *arg) {
    walk_job *j = (walk_job *)arg;
    fgo_walker *w = (fgo_walker *)malloc(sizeof(fgo_walker));
    if (!w) { j->rc = -2; return NULL; }
    for (int q = j->begin; q < j->end; q++) {
        if (fgo_walker_from_record(w, j->ring, &j->recs[q], j->seed, j->walker_base + (uint32_t)q)) { j->rc = -1; break; }
        /* REFERENCE mode: the FlipSet history starts from a fresh build at
         * every load, as after the reference's initializeCopyKernel. */
        for (int l = 0; l < j->launches; l++) fgo_walk(w, j->p);
        fgo_walker_to_record(w, &j->recs[q]);
    }
    free(w);
    return NULL;
}

int fgo_walk_records(int ring, fgo_walker_record *recs, int count, uint32_t walker_base,
                     uint64_t seed, const fgo_walk_params *p, int launches, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > count) nthreads = count > 0 ? count : 1;
    walk_job jobs[256];
    pthread_t th[256];
    if (nthreads > 256) nthreads = 256;
    for (int t = 0; t < nthreads; t++) {
        jobs[t].ring = ring; jobs[t].recs = recs;
        jobs[t].begin = (int)((long)count * t / nthreads);
        jobs[t].end = (int)((long)count * (t + 1) / nthreads);
        jobs[t].walker_base = walker_base; jobs[t].seed = seed; jobs[t].p = p;
        jobs[t].launches = launches; jobs[t].rc = 0;
    }
    if (nthreads == 1) {
        walk_job_run(&jobs[0]);
        return jobs[0].rc;
    }
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, walk_job_run, &jobs[t]);
    int rc = 0;
    for (int t = 0; t < nthreads; t++) {
        pthread_join(th[t], NULL);
        if (jobs[t].rc) rc = jobs[t].rc;
    }
    return rc;
}

int fgo_walk_records_forced(fgo_walker_record *recs, int count, uint32_t walker_base, uint64_t seed,
                            const fgo_walk_params *p, const int32_t *forced, int32_t *log, int32_t *minm) {
    fgo_walker *w = (fgo_walker *)malloc(sizeof(fgo_walker));
    if (!w) return -2;
    for (int q = 0; q < count; q++) {
        if (fgo_walker_from_record(w, FGO_RING_Z2, &recs[q], seed, walker_base + (uint32_t)q)) { free(w); return -1; }
        if (forced) { w->forced[0] = forced[4 * q]; w->forced[1] = forced[4 * q + 1]; w->forced[2] = forced[4 * q + 2]; }
        fgo_walk(w, p);
        fgo_walker_to_record(w, &recs[q]);
        if (log) for (int k = 0; k < 4; k++) log[4 * q + k] = w->log[k];
        if (minm) minm[q] = w->min_m;
    }
    free(w);
    return 0;
}

int fgo_naive_records(int ring, int n1, int n2, int n3, fgo_walker_record *recs, int count) {
    fgo_walker *w = (fgo_walker *)malloc(sizeof(fgo_walker));
    fgo_scheme *s = (fgo_scheme *)malloc(sizeof(fgo_scheme));
    if (!w || !s) { free(w); free(s); return -2; }
    fgo_init_naive(s, ring, n1, n2, n3);
    for (int q = 0; q < count; q++) {
        fgo_walker_init(w, s, s->m, 0, (uint32_t)q);
        fgo_walker_to_record(w, &recs[q]);
    }
    free(w); free(s);
    return 0;
}

int fgo_record_validate(int ring, const fgo_walker_record *rec, int best) {
    fgo_scheme *s = (fgo_scheme *)malloc(sizeof(fgo_scheme));
    if (!s) return -2;
    if (best) scheme_from_words(s, ring, rec->n, rec->best_m, rec->best_v, rec->best_s);
    else scheme_from_words(s, ring, rec->n, rec->m, rec->cur_v, rec->cur_s);
    int ok = fgo_validate(s);
    free(s);
    return ok;
}

int fgo_blob_len(const fgo_scheme *s) { return 4 + s->m * (s->nn[0] + s->nn[1] + s->nn[2]); }

/* read (scheme_z2.cu:72-96: |value| mod 2; scheme_integer.cu:81-107:
 * Addition::set, then fixSigns) */
int fgo_from_blob(fgo_scheme *s, int ring, const int32_t *blob, int check_validity) {
    memset(s, 0, sizeof(*s));
    s->ring = ring;
    s->n[0] = blob[0]; s->n[1] = blob[1]; s->n[2] = blob[2]; s->m = blob[3];
    for (int c = 0; c < 3; c++) s->nn[c] = s->n[c] * s->n[(c + 1) % 3];
    if (s->nn[0] > 64 || s->nn[1] > 64 || s->nn[2] > 64 || s->m > FGO_MAX_RANK || s->m < 0) return -1;
    const int32_t *p = blob + 4;
    for (int c = 0; c < 3; c++)
        for (int t = 0; t < s->m; t++) {
            s->ok[c][t] = 1;
            for (int q = 0; q < s->nn[c]; q++) {
                int x = *p++;
                if (ring == FGO_RING_Z2) {
                    s->v[c][t] |= (uint64_t)((x < 0 ? -x : x) % 2) << q;
                } else {
                    uint64_t bit = 1ull << q;
                    if (x == 1) s->v[c][t] |= bit;
                    else if (x == -1) { s->v[c][t] |= bit; s->s[c][t] |= bit; }
                    else if (x != 0) s->ok[c][t] = 0;
                }
            }
        }
    if (ring == FGO_RING_ZT) fgo_fix_signs(s);
    fgo_init_flips(s);
    if (check_validity && !fgo_validate(s)) return 0;
    return 1;
}

int fgo_to_blob(const fgo_scheme *s, int32_t *blob) {
    blob[0] = s->n[0]; blob[1] = s->n[1]; blob[2] = s->n[2]; blob[3] = s->m;
    int32_t *p = blob + 4;
    for (int c = 0; c < 3; c++)
        for (int t = 0; t < s->m; t++)
            for (int q = 0; q < s->nn[c]; q++) {
                int bit = (int)((s->v[c][t] >> q) & 1);
                int neg = (int)((s->s[c][t] >> q) & 1);
                *p++ = s->ring == FGO_RING_ZT && neg ? -bit : bit;
            }
    return (int)(p - blob);
}

int fgo_blob_validate(int ring, const int32_t *blob) {
    fgo_scheme *s = (fgo_scheme *)malloc(sizeof(fgo_scheme));
    if (!s) return -2;
    int rc = fgo_from_blob(s, ring, blob, 1);
    free(s);
    return rc;
}

int fgo_blob_complexity(int ring, const int32_t *blob) {
    fgo_scheme *s = (fgo_scheme *)malloc(sizeof(fgo_scheme));
    if (!s) return -(1 << 30);
    fgo_from_blob(s, ring, blob, 0);
    int c = fgo_complexity(s);
    free(s);
    return c;
}

int fgo_blob_naive(int ring, int n1, int n2, int n3, int32_t *out) {
    fgo_scheme *s = (fgo_scheme *)malloc(sizeof(fgo_scheme));
    if (!s) return -1;
    fgo_init_naive(s, ring, n1, n2, n3);
    int len = fgo_to_blob(s, out);
    free(s);
    return len;
}

int fgo_blob_replay(int ring, const int32_t *in, const int32_t *ops, int nops, int32_t *out) {
    fgo_scheme *s = (fgo_scheme *)malloc(sizeof(fgo_scheme));
    if (!s) return -1;
    if (fgo_from_blob(s, ring, in, 0) < 0) { free(s); return -1; }
    for (int q = 0; q < nops; q++) {
        const int32_t *o = ops + 8 * q;
        switch (o[0]) {
        case 0: fgo_flip(s, o[1], o[2], o[3], o[4], o[5], o[6]); break;
        case 1: fgo_plus(s, o[1], o[2], o[3], o[4], o[5], o[6]); break;
        case 2:
            if (ring == FGO_RING_Z2) fgo_split_z2(s, o[1], o[2], o[3], o[4], (uint64_t)(uint32_t)o[5] | ((uint64_t)(uint32_t)o[6] << 32));
            else fgo_split_zt(s, o[1], o[2], o[3], o[4], (uint64_t)(uint32_t)o[5], (uint64_t)(uint32_t)o[6]);
            break;
        case 3: fgo_reduce(s, o[1], o[2], o[3], o[4]); break;
        case 4: fgo_try_reduce(s); break;
        case 5: fgo_fix_signs(s); break;
        case 6: try_reduce_mode(s, FGO_SELECT_CANONICAL); break;
        default: free(s); return -1;
        }
    }
    int len = fgo_to_blob(s, out);
    free(s);
    return len;
}

/* Pairs as (c << 24 | t1 << 12 | t2).  mode 0: canonical; 1: stored FlipSet
 * order of a freshly read scheme. */
int fgo_blob_pairs(int ring, const int32_t *blob, int mode, uint32_t *out, int cap) {
    fgo_scheme *s = (fgo_scheme *)malloc(sizeof(fgo_scheme));
    if (!s) return -1;
    fgo_from_blob(s, ring, blob, 0);
    int n = 0;
    if (mode == 0) {
        for (int c = 0; c < 3; c++)
            for (int a = 0; a < s->m; a++)
                for (int b = a + 1; b < s->m; b++)
                    if (fac_eq(s, c, a, b)) { if (n < cap) out[n] = ((uint32_t)c << 24) | ((uint32_t)a << 12) | (uint32_t)b; n++; }
    } else {
        for (int c = 0; c < 3; c++)
            for (int q = 0; q < s->flips[c].size; q++) {
                if (n < cap) out[n] = ((uint32_t)c << 24) | ((s->flips[c].pairs[q] >> 16) << 12) | (s->flips[c].pairs[q] & 0xFFFF);
                n++;
            }
    }
    free(s);
    return n;
}

/* Self-check used by tests: walk `walkers` naive walkers in REFERENCE mode
 * and after every iteration compare each stored FlipSet, as a set, with the
 * brute-force set of equal pairs.  Returns the number of mismatching
 * (walker, iteration) states, or -1 on allocation failure. */
static int flipset_matches(const fgo_scheme *s) {
    for (int c = 0; c < 3; c++) {
        int want = 0;
        for (int a = 0; a < s->m; a++)
            for (int b = a + 1; b < s->m; b++) want += fac_eq(s, c, a, b);
        if (want != s->flips[c].size) return 0;
        for (int q = 0; q < s->flips[c].size; q++) {
            int a = (int)(s->flips[c].pairs[q] >> 16), b = (int)(s->flips[c].pairs[q] & 0xFFFF);
            if (a == b || a >= s->m || b >= s->m || !fac_eq(s, c, a, b)) return 0;
            for (int r = q + 1; r < s->flips[c].size; r++) {
                int a2 = (int)(s->flips[c].pairs[r] >> 16), b2 = (int)(s->flips[c].pairs[r] & 0xFFFF);
                if ((a2 == a && b2 == b) || (a2 == b && b2 == a)) return 0;
            }
        }
    }
    return 1;
}

int fgo_check_flipset_walk(int ring, int n1, int n2, int n3, uint64_t seed, int iterations, int walkers,
                           int plus_iterations) {
    fgo_walker *w = (fgo_walker *)malloc(sizeof(fgo_walker));
    fgo_scheme *s0 = (fgo_scheme *)malloc(sizeof(fgo_scheme));
    if (!w || !s0) { free(w); free(s0); return -1; }
    fgo_init_naive(s0, ring, n1, n2, n3);
    fgo_walk_params p;
    memset(&p, 0, sizeof(p));
    p.iterations = 1;
    p.max_iterations = iterations;
    p.plus_iterations = plus_iterations;
    p.use_expand = 1;
    p.thr_expand = 429496730u; /* 10% per iteration: exercise expand often */
    p.thr_best_eq = 42949673u;
    p.select_mode = FGO_SELECT_REFERENCE;
    p.check_reduce = 1;
    p.zt_visit_cap = 16;
    p.mode = 0;
    int bad = 0;
    for (int q = 0; q < walkers; q++) {
        fgo_walker_init(w, s0, s0->m, seed, (uint32_t)q);
        for (int it = 0; it < iterations; it++) {
            fgo_walk(w, &p);
            if (!flipset_matches(&w->cur) || !fgo_validate(&w->cur)) bad++;
        }
    }
    free(w); free(s0);
    return bad;
}

/* ------------------------------------------------------------------ */
/* resize and meta operators (Z2): src/schemes/scheme_z2.cu:184-310,   */
/* 643-979,1165-1324 and resizeKernel (src/entities/flip_graph.cu:575-613) */
/* ------------------------------------------------------------------ */

/* Bit (i, j) of an a x b factor sits at i * b + j. */
static uint64_t bit_at(uint64_t x, int pos) { return (x >> pos) & 1u; }
static uint64_t embed(uint64_t x, int a, int b, int nb, int di, int dj);

/* Integer-ring factors carry signs in a second bitmask with the magnitude
 * layout, so every bit map below is applied to v and to s alike (s is 0 for
 * Z2). */
static uint64_t drop_row_bits(uint64_t x, int a, int b, int row) {
    uint64_t y = 0;
    int ii = 0;
    for (int i = 0; i < a; i++) {
        if (i == row) continue;
        for (int j = 0; j < b; j++) y |= bit_at(x, i * b + j) << (ii * b + j);
        ii++;
    }
    return y;
}

static uint64_t drop_col_bits(uint64_t x, int a, int b, int column) {
    uint64_t y = 0;
    for (int i = 0; i < a; i++) {
        int jj = 0;
        for (int j = 0; j < b; j++) {
            if (j == column) continue;
            y |= bit_at(x, i * b + j) << (i * (b - 1) + jj);
            jj++;
        }
    }
    return y;
}

/* excludeRow (scheme_z2.cu:205-224, scheme_integer.cu:294-313): drop row
 * `row` of every factor of component `matrix` (n[matrix] x n[matrix + 1]). */
static void exclude_row(fgo_scheme *s, int matrix, int row) {
    const int a = s->n[matrix], b = s->n[(matrix + 1) % 3];
    for (int t = 0; t < s->m; t++) {
        s->v[matrix][t] = drop_row_bits(s->v[matrix][t], a, b, row);
        s->s[matrix][t] = drop_row_bits(s->s[matrix][t], a, b, row);
    }
}

/* excludeColumn (scheme_z2.cu:184-203, scheme_integer.cu:273-292) */
static void exclude_column(fgo_scheme *s, int matrix, int column) {
    const int a = s->n[matrix], b = s->n[(matrix + 1) % 3];
    for (int t = 0; t < s->m; t++) {
        s->v[matrix][t] = drop_col_bits(s->v[matrix][t], a, b, column);
        s->s[matrix][t] = drop_col_bits(s->s[matrix][t], a, b, column);
    }
}

/* addColumn (scheme_z2.cu:226-239, scheme_integer.cu:315-328): an empty last
 * column; addRow (:241-254) leaves the bit positions unchanged. */
static void add_column(fgo_scheme *s, int matrix) {
    const int a = s->n[matrix], b = s->n[(matrix + 1) % 3];
    for (int t = 0; t < s->m; t++) {
        s->v[matrix][t] = embed(s->v[matrix][t], a, b, b + 1, 0, 0);
        s->s[matrix][t] = embed(s->s[matrix][t], a, b, b + 1, 0, 0);
    }
}

static void set_nn(fgo_scheme *s) {
    for (int c = 0; c < 3; c++) s->nn[c] = s->n[c] * s->n[(c + 1) % 3];
}

static void z2_triplet(fgo_scheme *s, uint64_t u, uint64_t v, uint64_t w) {
    tv a = {u, 0, 1}, b = {v, 0, 1}, c = {w, 0, 1};
    add_triplet(s, 0, 1, 2, a, b, c);
}

/* project (scheme_z2.cu:643-653, scheme_integer.cu:573-584; the integer
 * ring also fixes signs) */
void fgo_project(fgo_scheme *s, int p, int q) {
    exclude_row(s, p, q);
    exclude_column(s, (p + 2) % 3, q);
    s->n[p]--;
    set_nn(s);
    remove_zeroes(s);
    if (s->ring == FGO_RING_ZT) fgo_fix_signs(s);
    fgo_init_flips(s);
}

/* extend (scheme_z2.cu:655-687, scheme_integer.cu:586-618): the new terms
 * have single +1 coefficients */
void fgo_extend(fgo_scheme *s, int p) {
    const int n0 = s->n[0], n1 = s->n[1], n2 = s->n[2];
    if (p == 0) {
        add_column(s, 2);
        for (int i = 0; i < n2; i++)
            for (int j = 0; j < n1; j++)
                z2_triplet(s, 1ull << (n0 * n1 + j), 1ull << (j * n2 + i), 1ull << (i * (n0 + 1) + n0));
    } else if (p == 1) {
        add_column(s, 0);
        for (int i = 0; i < n0; i++)
            for (int j = 0; j < n2; j++)
                z2_triplet(s, 1ull << (i * (n1 + 1) + n1), 1ull << (n1 * n2 + j), 1ull << (j * n0 + i));
    } else {
        add_column(s, 1);
        for (int i = 0; i < n0; i++)
            for (int j = 0; j < n1; j++)
                z2_triplet(s, 1ull << (i * n1 + j), 1ull << (j * (n2 + 1) + n2), 1ull << (n2 * n0 + i));
    }
    s->n[p]++;
    set_nn(s);
    fgo_init_flips(s);
}

/* Factor x of an a x b block placed at offset (di, dj) of a rows x nb
 * matrix. */
static uint64_t embed(uint64_t x, int a, int b, int nb, int di, int dj) {
    uint64_t y = 0;
    for (int i = 0; i < a; i++)
        for (int j = 0; j < b; j++) y |= bit_at(x, i * b + j) << ((i + di) * nb + j + dj);
    return y;
}

/* product (scheme_z2.cu:689-747, scheme_integer.cu:620-678): dimension p
 * doubled, every term copied into the second diagonal block. */
void fgo_product_dim(fgo_scheme *s, int p) {
    int nn_[3] = {s->n[0], s->n[1], s->n[2]}, d[3];
    nn_[p] *= 2;
    for (int i = 0; i < 3; i++) d[i] = p == i ? s->n[i] : 0;
    const int m = s->m;
    for (int t = 0; t < m; t++) {
        for (int c = 0; c < 3; c++) {
            const int c1 = (c + 1) % 3;
            const uint64_t x = s->v[c][t], g = s->s[c][t];
            s->v[c][t] = embed(x, s->n[c], s->n[c1], nn_[c1], 0, 0);
            s->s[c][t] = embed(g, s->n[c], s->n[c1], nn_[c1], 0, 0);
            s->v[c][t + m] = embed(x, s->n[c], s->n[c1], nn_[c1], d[c], d[c1]);
            s->s[c][t + m] = embed(g, s->n[c], s->n[c1], nn_[c1], d[c], d[c1]);
            s->ok[c][t + m] = 1;
        }
    }
    for (int c = 0; c < 3; c++) s->n[c] = nn_[c];
    set_nn(s);
    s->m = 2 * m;
    fgo_init_flips(s);
}

/* merge (scheme_z2.cu:749-817, scheme_integer.cu:680-745): `o` joins as the
 * second block along p. */
void fgo_merge(fgo_scheme *s, const fgo_scheme *o, int p) {
    int nn_[3], d[3];
    for (int i = 0; i < 3; i++) {
        nn_[i] = i == p ? s->n[i] + o->n[i] : s->n[i];
        d[i] = i == p ? s->n[i] : 0;
    }
    for (int t = 0; t < s->m; t++)
        for (int c = 0; c < 3; c++) {
            const int c1 = (c + 1) % 3;
            s->v[c][t] = embed(s->v[c][t], s->n[c], s->n[c1], nn_[c1], 0, 0);
            s->s[c][t] = embed(s->s[c][t], s->n[c], s->n[c1], nn_[c1], 0, 0);
        }
    for (int t = 0; t < o->m; t++)
        for (int c = 0; c < 3; c++) {
            const int c1 = (c + 1) % 3;
            s->v[c][s->m + t] = embed(o->v[c][t], o->n[c], o->n[c1], nn_[c1], d[c], d[c1]);
            s->s[c][s->m + t] = embed(o->s[c][t], o->n[c], o->n[c1], nn_[c1], d[c], d[c1]);
            s->ok[c][s->m + t] = 1;
        }
    for (int c = 0; c < 3; c++) s->n[c] = nn_[c];
    set_nn(s);
    s->m += o->m;
    fgo_init_flips(s);
}

/* product (scheme_z2.cu:819-862, scheme_integer.cu:747-789): Kronecker
 * product of two schemes, term index1 * m2 + index2; an integer coefficient
 * is the product of the two (signs XOR). */
void fgo_product_scheme(fgo_scheme *s, const fgo_scheme *o) {
    fgo_scheme *a = (fgo_scheme *)malloc(sizeof(fgo_scheme));
    fgo_copy(a, s, 0);
    for (int c = 0; c < 3; c++) s->n[c] = a->n[c] * o->n[c];
    set_nn(s);
    s->m = a->m * o->m;
    for (int t1 = 0; t1 < a->m; t1++)
        for (int t2 = 0; t2 < o->m; t2++) {
            const int t = t1 * o->m + t2;
            for (int c = 0; c < 3; c++) {
                const int c1 = (c + 1) % 3;
                uint64_t y = 0, g = 0;
                for (int i = 0; i < a->nn[c]; i++) {
                    if (!bit_at(a->v[c][t1], i)) continue;
                    const int r1 = i / a->n[c1], q1 = i % a->n[c1];
                    for (int j = 0; j < o->nn[c]; j++) {
                        if (!bit_at(o->v[c][t2], j)) continue;
                        const int r2 = j / o->n[c1], q2 = j % o->n[c1];
                        const int pos = (r1 * o->n[c] + r2) * s->n[c1] + q1 * o->n[c1] + q2;
                        y |= 1ull << pos;
                        g |= (bit_at(a->s[c][t1], i) ^ bit_at(o->s[c][t2], j)) << pos;
                    }
                }
                s->v[c][t] = y;
                s->s[c][t] = g;
                s->ok[c][t] = 1;
            }
        }
    free(a);
    fgo_init_flips(s);
}

/* swapSize (scheme_z2.cu:916-979, scheme_integer.cu:843-905): the scheme of
 * the transposed problem obtained by exchanging dimensions p1 and p2; the
 * integer ring then fixes signs. */
void fgo_swap_size(fgo_scheme *s, int p1, int p2) {
    if (p1 == p2) return;
    if (p1 > p2) { int t = p1; p1 = p2; p2 = t; }
    int idx[3], nn_[3];
    if (p1 == 0 && p2 == 1) { idx[0] = 0; idx[1] = 2; idx[2] = 1; }
    else if (p1 == 0 && p2 == 2) { idx[0] = 1; idx[1] = 0; idx[2] = 2; }
    else { idx[0] = 2; idx[1] = 1; idx[2] = 0; }
    for (int i = 0; i < 3; i++) nn_[i] = s->n[(idx[i] + 1) % 3];
    for (int t = 0; t < s->m; t++) {
        uint64_t y[3], g[3];
        for (int c = 0; c < 3; c++) {
            const int a = nn_[c], b = nn_[(c + 1) % 3];
            const uint64_t x = s->v[idx[c]][t], sx = s->s[idx[c]][t];
            y[c] = g[c] = 0;
            for (int i = 0; i < a; i++)
                for (int j = 0; j < b; j++) {
                    y[c] |= bit_at(x, j * a + i) << (i * b + j);
                    g[c] |= bit_at(sx, j * a + i) << (i * b + j);
                }
        }
        for (int c = 0; c < 3; c++) {
            s->v[c][t] = y[c];
            s->s[c][t] = g[c];
        }
    }
    for (int c = 0; c < 3; c++) s->n[c] = nn_[c];
    set_nn(s);
    if (s->ring == FGO_RING_ZT) fgo_fix_signs(s);
    fgo_init_flips(s);
}

/* swapBasisRows / swapBasisColumns (scheme_z2.cu:864-914, scheme_integer.cu:791-841) */
void fgo_swap_basis_rows(fgo_scheme *s, int i1, int i2) {
    const int n0 = s->n[0], n1 = s->n[1], n2 = s->n[2];
    int rows[64];
    for (int r = 0; r < n0; r++) rows[r] = r;
    rows[i1] = i2;
    rows[i2] = i1;
    for (int t = 0; t < s->m; t++) {
        uint64_t u = 0, w = 0, su = 0, sw = 0;
        for (int i = 0; i < n0; i++)
            for (int j = 0; j < n1; j++) {
                u |= bit_at(s->v[0][t], rows[i] * n1 + j) << (i * n1 + j);
                su |= bit_at(s->s[0][t], rows[i] * n1 + j) << (i * n1 + j);
            }
        for (int i = 0; i < n2; i++)
            for (int j = 0; j < n0; j++) {
                w |= bit_at(s->v[2][t], i * n0 + rows[j]) << (i * n0 + j);
                sw |= bit_at(s->s[2][t], i * n0 + rows[j]) << (i * n0 + j);
            }
        s->v[0][t] = u;
        s->v[2][t] = w;
        s->s[0][t] = su;
        s->s[2][t] = sw;
    }
}

void fgo_swap_basis_columns(fgo_scheme *s, int j1, int j2) {
    const int n0 = s->n[0], n1 = s->n[1], n2 = s->n[2];
    int cols[64];
    for (int c = 0; c < n2; c++) cols[c] = c;
    cols[j1] = j2;
    cols[j2] = j1;
    for (int t = 0; t < s->m; t++) {
        uint64_t v = 0, w = 0, sv = 0, sw = 0;
        for (int i = 0; i < n1; i++)
            for (int j = 0; j < n2; j++) {
                v |= bit_at(s->v[1][t], i * n2 + cols[j]) << (i * n2 + j);
                sv |= bit_at(s->s[1][t], i * n2 + cols[j]) << (i * n2 + j);
            }
        for (int i = 0; i < n2; i++)
            for (int j = 0; j < n0; j++) {
                w |= bit_at(s->v[2][t], cols[i] * n0 + j) << (i * n0 + j);
                sw |= bit_at(s->s[2][t], cols[i] * n0 + j) << (i * n0 + j);
            }
        s->v[1][t] = v;
        s->v[2][t] = w;
        s->s[1][t] = sv;
        s->s[2][t] = sw;
    }
}

/* isValidProject / isValidExtension / isValidProduct (scheme_z2.cu:256-294),
 * with this build's rank capacity and factor width in place of MAX_RANK and
 * MAX_MATRIX_ELEMENTS. */
static int valid_project(const fgo_scheme *s, int i) {
    return s->n[i] - 1 >= FGO_MIN_PROJECT_N && s->n[(i + 1) % 3] >= FGO_MIN_PROJECT_N &&
           s->n[(i + 2) % 3] >= FGO_MIN_PROJECT_N;
}

static int valid_sizes(const int nn_[3], const fgo_resize_caps *k) {
    for (int p = 0; p < 3; p++) {
        if (nn_[p] * nn_[(p + 1) % 3] > k->max_coef) return 0;
        if (nn_[p] > FGO_MAX_EXTENSION_N) return 0;
    }
    return 1;
}

static int valid_extension(const fgo_scheme *s, int i, const fgo_resize_caps *k) {
    int nn_[3] = {s->n[0], s->n[1], s->n[2]};
    nn_[i] += 1;
    if (s->m + nn_[(i + 1) % 3] * nn_[(i + 2) % 3] > k->max_rank) return 0;
    return valid_sizes(nn_, k);
}

static int valid_product(const fgo_scheme *s, int i, const fgo_resize_caps *k) {
    int nn_[3] = {s->n[0], s->n[1], s->n[2]};
    nn_[i] *= 2;
    if (s->m * 2 > k->max_rank) return 0;
    return valid_sizes(nn_, k);
}

static uint32_t aux_below(fgo_rng *r, uint32_t n) { return fgo_rng_aux32(r) % n; }

/* tryProject / tryExtend / tryProduct (scheme_z2.cu:1165-1223) */
static int try_project(fgo_scheme *s, fgo_rng *r) {
    int idx[3], size = 0;
    for (int i = 0; i < 3; i++)
        if (valid_project(s, i)) idx[size++] = i;
    if (!size) return 0;
    const int p = idx[aux_below(r, (uint32_t)size)];
    const int q = (int)aux_below(r, (uint32_t)s->n[p]);
    fgo_project(s, p, q);
    while (try_reduce_mode(s, FGO_SELECT_CANONICAL)) {
    }
    return 1;
}

static int try_extend(fgo_scheme *s, fgo_rng *r, const fgo_resize_caps *k) {
    int idx[3], size = 0;
    for (int i = 0; i < 3; i++)
        if (valid_extension(s, i, k)) idx[size++] = i;
    if (!size) return 0;
    fgo_extend(s, idx[aux_below(r, (uint32_t)size)]);
    while (try_reduce_mode(s, FGO_SELECT_CANONICAL)) {
    }
    return 1;
}

static int try_product_dim(fgo_scheme *s, fgo_rng *r, const fgo_resize_caps *k) {
    int idx[3], size = 0;
    for (int i = 0; i < 3; i++)
        if (valid_product(s, i, k)) idx[size++] = i;
    if (!size) return 0;
    fgo_product_dim(s, idx[aux_below(r, (uint32_t)size)]);
    return 1;
}

/* tryMerge (scheme_z2.cu:1225-1254) */
static int try_merge(fgo_scheme *s, const fgo_scheme *o, fgo_rng *r, const fgo_resize_caps *k) {
    if (s->m + o->m > k->max_rank) return 0;
    const int e0 = s->n[0] == o->n[0], e1 = s->n[1] == o->n[1], e2 = s->n[2] == o->n[2];
    const int a0 = s->n[0] + o->n[0], a1 = s->n[1] + o->n[1], a2 = s->n[2] + o->n[2];
    int p[3], size = 0;
    if (a0 <= FGO_MAX_EXTENSION_N && a0 * s->n[1] <= k->max_coef && a0 * s->n[2] <= k->max_coef && e1 && e2) p[size++] = 0;
    if (e0 && a1 <= FGO_MAX_EXTENSION_N && s->n[0] * a1 <= k->max_coef && s->n[2] * a1 <= k->max_coef && e2) p[size++] = 1;
    if (e0 && e1 && a2 <= FGO_MAX_EXTENSION_N && s->n[0] * a2 <= k->max_coef && s->n[1] * a2 <= k->max_coef) p[size++] = 2;
    if (!size) return 0;
    fgo_merge(s, o, p[aux_below(r, (uint32_t)size)]);
    return 1;
}

/* tryProduct(scheme) (scheme_z2.cu:1256-1275) */
static int try_product_scheme(fgo_scheme *s, const fgo_scheme *o, const fgo_resize_caps *k) {
    if (s->m * o->m > k->max_rank) return 0;
    int sz[3];
    for (int i = 0; i < 3; i++) sz[i] = s->n[i] * o->n[i];
    for (int i = 0; i < 3; i++) {
        if (sz[i] > FGO_MAX_EXTENSION_N) return 0;
        if (sz[i] * sz[(i + 1) % 3] > k->max_coef) return 0;
    }
    fgo_product_scheme(s, o);
    return 1;
}

/* swapSize(state) (scheme_z2.cu:1315-1324); the redraw loop is bounded. */
static void swap_size_random(fgo_scheme *s, fgo_rng *r) {
    for (int tries = 0; tries < FGO_PLUS_MAX_TRIES; tries++) {
        const int p1 = (int)aux_below(r, 3u), p2 = (int)aux_below(r, 3u);
        if (p1 != p2) {
            fgo_swap_size(s, p1, p2);
            return;
        }
    }
}

/* resizeKernel body (src/entities/flip_graph.cu:575-613) for walker w:
 * every uniform compare u < q becomes aux32() < floor(q * 2^32) and every
 * curand() % n becomes aux32() % n, all from the walker's aux stream.
 * bests[i] is walker i's best scheme as the launch left it. */
void fgo_resize_walker(fgo_walker *w, const fgo_scheme *const *bests, int count, uint64_t thr_resize,
                       const fgo_resize_caps *k) {
    fgo_scheme *s = &w->cur;
    if ((uint64_t)fgo_rng_aux32(&w->rng) >= thr_resize) return; /* :579 */
    if (fgo_rng_aux32(&w->rng) < FGO_THR_HALF) swap_size_random(s, &w->rng); /* :585-586 */
    int merged = 0;
    for (int i = 0; i < 3 && !merged; i++) { /* :588-592 */
        const int index = (int)aux_below(&w->rng, (uint32_t)count);
        merged |= try_merge(s, bests[index], &w->rng, k);
    }
    if (!merged) { /* :594-609 */
        const uint32_t x = fgo_rng_aux32(&w->rng);
        if (x < FGO_THR_005) {
            try_project(s, &w->rng);
        } else if (x < FGO_THR_055) {
            const int index = (int)aux_below(&w->rng, (uint32_t)count);
            try_product_scheme(s, bests[index], k);
        } else if (x < FGO_THR_085) {
            try_product_dim(s, &w->rng, k);
        } else {
            try_extend(s, &w->rng, k);
        }
    }
}

/* One resize pass over records (walkers walker_base + q): every walker
 * reads the best schemes the records hold before the pass. */
int fgo_resize_records(int ring, fgo_walker_record *recs, int count, uint32_t walker_base, uint64_t seed,
                       uint64_t thr_resize, int max_rank, int max_coef) {
    fgo_resize_caps k = {max_rank, max_coef};
    fgo_walker *ws = (fgo_walker *)malloc(sizeof(fgo_walker) * (size_t)count);
    const fgo_scheme **bests = (const fgo_scheme **)malloc(sizeof(fgo_scheme *) * (size_t)count);
    if (!ws || !bests) { free(ws); free(bests); return -2; }
    for (int q = 0; q < count; q++) {
        if (fgo_walker_from_record(&ws[q], ring, &recs[q], seed, walker_base + (uint32_t)q)) {
            free(ws); free(bests);
            return -1;
        }
        bests[q] = &ws[q].best;
    }
    /* the bests are not touched by the pass, so walkers can go in order */
    fgo_scheme *snap = (fgo_scheme *)malloc(sizeof(fgo_scheme) * (size_t)count);
    if (!snap) { free(ws); free(bests); return -2; }
    for (int q = 0; q < count; q++) {
        fgo_copy(&snap[q], &ws[q].best, 0);
        bests[q] = &snap[q];
    }
    for (int q = 0; q < count; q++) {
        fgo_resize_walker(&ws[q], bests, count, thr_resize, &k);
        fgo_walker_to_record(&ws[q], &recs[q]);
    }
    free(snap);
    free(ws);
    free(bests);
    return 0;
}

/* Single operators on blobs (tests): op 0 project(p=a, q=b), 1 extend(a),
 * 2 product(a), 3 merge(other, a), 4 product(other), 5 swapSize(a, b),
 * 6 swapBasisRows(a, b), 7 swapBasisColumns(a, b).  Returns the output
 * blob length. */
int fgo_blob_resize_op(int ring, const int32_t *in, const int32_t *other, int op, int a, int b, int32_t *out) {
    fgo_scheme *s = (fgo_scheme *)malloc(sizeof(fgo_scheme)), *o = (fgo_scheme *)malloc(sizeof(fgo_scheme));
    int len = -1;
    if (!s || !o) goto done;
    if (fgo_from_blob(s, ring, in, 0) != 1) goto done;
    if (other && fgo_from_blob(o, ring, other, 0) != 1) goto done;
    switch (op) {
    case 0: fgo_project(s, a, b); break;
    case 1: fgo_extend(s, a); break;
    case 2: fgo_product_dim(s, a); break;
    case 3: fgo_merge(s, o, a); break;
    case 4: fgo_product_scheme(s, o); break;
    case 5: fgo_swap_size(s, a, b); break;
    case 6: fgo_swap_basis_rows(s, a, b); break;
    case 7: fgo_swap_basis_columns(s, a, b); break;
    default: goto done;
    }
    len = fgo_to_blob(s, out);
done:
    free(s);
    free(o);
    return len;
}

/* inverseMatrixZ2 (scheme_z2.cu:444-486) on 0/1 int matrices */
static int gf2_inverse(int n, const int *a, int *inv) {
    int aug[16 * 32];
    const int n2 = 2 * n;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            aug[i * n2 + j] = a[i * n + j];
            aug[i * n2 + n + j] = i == j;
        }
    for (int col = 0; col < n; col++) {
        int piv = -1;
        for (int row = col; row < n; row++)
            if (aug[row * n2 + col]) { piv = row; break; }
        if (piv < 0) return 0;
        if (piv != col)
            for (int i = 0; i < n2; i++) { int t = aug[col * n2 + i]; aug[col * n2 + i] = aug[piv * n2 + i]; aug[piv * n2 + i] = t; }
        for (int row = 0; row < n; row++)
            if (row != col && aug[row * n2 + col])
                for (int j = col; j < n2; j++) aug[row * n2 + j] ^= aug[col * n2 + j];
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) inv[i * n + j] = aug[i * n2 + n + j];
    return 1;
}

/* invertibleMatrixZ2 (scheme_z2.cu:488-492) with randomMatrixZ2
 * (src/utils/random.cuh:24-32): row i is the low n bits of one aux draw;
 * the redraws are bounded. */
static int random_invertible(fgo_rng *r, int n, int *a, int *inv) {
    for (int tries = 0; tries < FGO_PLUS_MAX_TRIES; tries++) {
        for (int i = 0; i < n; i++) {
            const uint32_t rnd = fgo_rng_aux32(r);
            for (int j = 0; j < n; j++) a[i * n + j] = (int)((rnd >> j) & 1u);
        }
        if (gf2_inverse(n, a, inv)) return 1;
    }
    return 0;
}

/* matmul (scheme_z2.cu:494-528): left (n1 x n1) . X (n1 x n2) . right (n2 x n2) */
static uint64_t gf2_matmul(uint64_t x, const int *left, const int *right, int n1, int n2) {
    int r1[64], r2[64];
    for (int i = 0; i < n1 * n2; i++) r1[i] = r2[i] = 0;
    for (int i = 0; i < n1; i++)
        for (int j = 0; j < n2; j++)
            for (int k = 0; k < n1; k++) r1[i * n2 + j] ^= left[i * n1 + k] && (int)((x >> (k * n2 + j)) & 1u);
    for (int i = 0; i < n1; i++)
        for (int j = 0; j < n2; j++)
            for (int k = 0; k < n2; k++) r2[i * n2 + j] ^= r1[i * n2 + k] && right[k * n2 + j];
    uint64_t y = 0;
    for (int i = 0; i < n1 * n2; i++)
        if (r2[i]) y |= 1ull << i;
    return y;
}

static void identity(int n, int *a) {
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) a[i * n + j] = i == j;
}

/* sandwiching (scheme_z2.cu:1277-1300): U -> A U B^-1, V -> B V C^-1,
 * W -> C W A^-1 for random invertible A, B, C drawn in that order; each is
 * applied once drawn (the products associate), and a matrix whose bounded
 * redraws all fail ends the edge there. */
static void sandwiching(fgo_scheme *s, fgo_rng *r) {
    int a[256], ai[256], i0[256], i1[256], i2[256];
    const int n0 = s->n[0], n1 = s->n[1], n2 = s->n[2];
    identity(n0, i0);
    identity(n1, i1);
    identity(n2, i2);
    if (!random_invertible(r, n0, a, ai)) return;
    for (int t = 0; t < s->m; t++) {
        s->v[0][t] = gf2_matmul(s->v[0][t], a, i1, n0, n1);
        s->v[2][t] = gf2_matmul(s->v[2][t], i2, ai, n2, n0);
    }
    if (random_invertible(r, n1, a, ai)) {
        for (int t = 0; t < s->m; t++) {
            s->v[0][t] = gf2_matmul(s->v[0][t], i0, ai, n0, n1);
            s->v[1][t] = gf2_matmul(s->v[1][t], a, i2, n1, n2);
        }
        if (random_invertible(r, n2, a, ai)) {
            for (int t = 0; t < s->m; t++) {
                s->v[1][t] = gf2_matmul(s->v[1][t], i1, ai, n1, n2);
                s->v[2][t] = gf2_matmul(s->v[2][t], a, i0, n2, n0);
            }
        }
    }
    fgo_init_flips(s);
}

/* swapBasis (scheme_z2.cu:1302-1313); the partner lists stay valid, as a
 * permutation keeps equal factors equal.  The integer ring then fixes signs
 * and rebuilds them (scheme_integer.cu:1219-1233). */
static void swap_basis(fgo_scheme *s, fgo_rng *r) {
    if (fgo_rng_aux32(r) % 2u) {
        const int i1 = (int)(fgo_rng_aux32(r) % (uint32_t)s->n[0]), i2 = (int)(fgo_rng_aux32(r) % (uint32_t)s->n[0]);
        fgo_swap_basis_rows(s, i1, i2);
    } else {
        const int j1 = (int)(fgo_rng_aux32(r) % (uint32_t)s->n[2]), j2 = (int)(fgo_rng_aux32(r) % (uint32_t)s->n[2]);
        fgo_swap_basis_columns(s, j1, j2);
    }
    if (s->ring == FGO_RING_ZT && fgo_fix_signs(s)) fgo_init_flips(s);
}

/* One integer tryFlip on a scheme (tests): the walker's main block `it` of
 * (seed, walker) supplies the candidate and coin words, its aux stream the
 * rest.  Returns 1 when a flip was applied (op gets i, j, k, index1, index2),
 * 0 when none, < 0 on error; out gets the scheme afterwards. */
int fgo_blob_try_flip_zt(const int32_t *in, uint64_t seed, uint32_t walker, uint64_t it, int select_mode,
                         int visit_cap, int check_reduce, int32_t *out, int32_t *op) {
    fgo_scheme *s = (fgo_scheme *)malloc(sizeof(fgo_scheme));
    if (!s) return -1;
    if (fgo_from_blob(s, FGO_RING_ZT, in, 0) < 0) { free(s); return -1; }
    fgo_rng r;
    fgo_rng_init(&r, seed, walker);
    r.it = it;
    uint32_t b[4];
    fgo_rng_main_block(&r, b);
    fgo_walk_params p;
    memset(&p, 0, sizeof p);
    p.select_mode = select_mode;
    p.zt_visit_cap = visit_cap;
    p.check_reduce = check_reduce;
    for (int q = 0; q < 5; q++) tl_zt_op[q] = -1;
    int ok = try_flip_zt(s, b, &r, &p);
    for (int q = 0; q < 5; q++) op[q] = ok ? tl_zt_op[q] : -1;
    fgo_to_blob(s, out);
    free(s);
    return ok;
}

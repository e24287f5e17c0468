/* CPU oracle of the additions reducer: a sequential restatement of
 * AdditionsReducer and PairsCounter (src/entities/additions_reducer.cuh,
 * src/entities/pairs_counter.cuh) and of one runReducersKernel pass
 * (src/entities/scheme_additions_reducer.cu:485-540).
 *
 * Test infrastructure only: the GPU reducer is checked against it.
 *
 * Random numbers come from Philox4x32-10 instead of cuRAND: the stream of
 * reducer r, component c (0 U, 1 V, 2 W) is the word sequence of blocks
 * philox(ctr = (block_lo, block_hi, r, 2 + c), key = seed); the stream
 * (r, 5) holds the decision draws shared by the three components.  A
 * uniform float in (0, 1] is ((x >> 8) + 1) * 2^-24.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fg_oracle.h"

enum {
    MODE_GREEDY = 0,
    MODE_GREEDY_ALT = 1,
    MODE_GREEDY_RANDOM = 2,
    MODE_GREEDY_INTERSECTIONS = 3,
    MODE_WEIGHTED_RANDOM = 4,
    MODE_MIX = 5,
    MODE_RANDOM = 6,
};

typedef struct {
    uint32_t key[2];
    uint32_t r, c;
    uint64_t w;      /* words consumed */
} rstream;

static uint32_t rs_next(rstream *s) {
    uint32_t ctr[4] = {(uint32_t)(s->w >> 2), (uint32_t)(s->w >> 34), s->r, 2 + s->c}, out[4];
    fgo_philox4x32_10(ctr, s->key, out);
    return out[s->w++ & 3];
}

static float rs_uniform(rstream *s) { return (float)((rs_next(s) >> 8) + 1) * (1.0f / 16777216.0f); }

/* ---- PairsCounter ---- */
typedef struct {
    int i, j, count;
} pair_t;

typedef struct {
    int max_pairs;
    int size, top, max_count;
    pair_t *pairs;
    /* lookup of (i, j) -> index; any map gives the same results since only
     * the insertion order of pairs matters */
    int hcap;
    int *slot;
} counter_t;

static void counter_clear(counter_t *p) {
    p->size = p->top = p->max_count = 0;
    for (int q = 0; q < p->hcap; q++) p->slot[q] = -1;
}

static unsigned pair_hash(int i, int j, int cap) {
    return ((unsigned)i * 73856093u ^ (unsigned)(j + 4096) * 19349663u) % (unsigned)cap;
}

/* insert (pairs_counter.cuh:55-66, 219-222, 230-259) */
static void counter_insert(counter_t *p, int i, int j) {
    if (abs(i) > abs(j)) {
        int t = i;
        i = j;
        j = t;
    }
    if (i < 0) {
        i = -i;
        j = -j;
    }
    unsigned h = pair_hash(i, j, p->hcap);
    for (;;) {
        int q = p->slot[h];
        if (q < 0) {
            if (p->size >= p->max_pairs) return; /* full: the pair is dropped */
            q = p->size++;
            p->pairs[q].i = i;
            p->pairs[q].j = j;
            p->pairs[q].count = 0;
            p->slot[h] = q;
        }
        if (p->pairs[q].i == i && p->pairs[q].j == j) {
            if (++p->pairs[q].count > p->max_count) p->max_count = p->pairs[q].count;
            return;
        }
        h = (h + 1) % (unsigned)p->hcap;
    }
}

/* sort (pairs_counter.cuh:68-101): pairs of the highest count move to the
 * front by swaps, then pairs seen once are dropped from the rest. */
static void counter_sort(counter_t *p) {
    if (p->max_count < 2) {
        p->top = p->size;
        return;
    }
    p->top = 0;
    for (int i = 0; i < p->size; i++) {
        if (p->pairs[i].count != p->max_count) continue;
        if (i != p->top) {
            pair_t t = p->pairs[i];
            p->pairs[i] = p->pairs[p->top];
            p->pairs[p->top] = t;
        }
        p->top++;
    }
    int j = p->top;
    for (int i = p->top; i < p->size; i++)
        if (p->pairs[i].count > 1) p->pairs[j++] = p->pairs[i];
    p->size = j;
}

static int intersects(const pair_t *a, const pair_t *b) {
    return a->i == b->i || a->i == abs(b->j) || abs(a->j) == b->i || abs(a->j) == abs(b->j);
}

/* getGreedyIntersections (pairs_counter.cuh:145-171).  The float / double
 * mix of the reference is kept: the intersecting term is a double added to
 * the float sum. */
static pair_t greedy_intersections(const counter_t *p, rstream *s, float scale) {
    int imax = 0;
    float best = 0;
    for (int i = 0; i < p->size; i++) {
        float inter = 0;
        for (int j = 0; j < p->size; j++) {
            if (i == j) continue;
            if (!intersects(&p->pairs[i], &p->pairs[j])) {
                inter += (float)(p->pairs[j].count - 1);
            } else {
                const uint32_t x = rs_next(s);
                inter = (float)((double)inter + ((x % 2) ? 0.0 : 0.69 * (double)p->pairs[j].count));
            }
        }
        const float score = (float)(p->pairs[i].count - 1) + scale * inter;
        if (score > best) {
            best = score;
            imax = i;
        }
    }
    return p->pairs[imax];
}

/* getWeightedRandom (pairs_counter.cuh:173-191) */
static pair_t weighted_random(const counter_t *p, rstream *s) {
    float total = 0;
    for (int i = 0; i < p->size; i++) total += (float)p->pairs[i].count;
    const float x = rs_uniform(s) * total;
    float sum = 0;
    for (int i = 0; i < p->size; i++) {
        sum += (float)p->pairs[i].count;
        if (x <= sum) return p->pairs[i];
    }
    return p->pairs[p->size - 1];
}

/* selectSubexpression (additions_reducer.cuh:264-282) with the counter
 * getters (pairs_counter.cuh:130-143, 193-196) */
static pair_t select_pair(const counter_t *p, int mode, rstream *s, float scale) {
    switch (mode) {
    case MODE_GREEDY:
        return p->pairs[0];
    case MODE_GREEDY_ALT:
        return p->pairs[rs_next(s) % (uint32_t)p->top];
    case MODE_GREEDY_RANDOM:
        if (rs_uniform(s) < scale) return p->pairs[rs_next(s) % (uint32_t)p->size];
        return p->pairs[rs_next(s) % (uint32_t)p->top];
    case MODE_GREEDY_INTERSECTIONS:
        return greedy_intersections(p, s, scale);
    case MODE_WEIGHTED_RANDOM:
        return weighted_random(p, s);
    default:
        return p->pairs[rs_next(s) % (uint32_t)p->size];
    }
}

/* ---- AdditionsReducer ---- */
struct fgo_reducer {
    int max_expr, max_len, max_fresh;
    int count, real, fresh, naive, mode;
    float scale;
    int *size;
    int *expr; /* [max_expr][max_len] */
    int *vars; /* [max_fresh][2] */
    counter_t sub;
};

fgo_reducer *fgo_reducer_new(int max_expr, int max_len, int max_fresh, int max_pairs) {
    fgo_reducer *r = calloc(1, sizeof *r);
    r->max_expr = max_expr;
    r->max_len = max_len;
    r->max_fresh = max_fresh;
    r->size = calloc((size_t)max_expr, sizeof(int));
    r->expr = calloc((size_t)max_expr * max_len, sizeof(int));
    r->vars = calloc((size_t)max_fresh * 2, sizeof(int));
    r->sub.max_pairs = max_pairs;
    r->sub.pairs = calloc((size_t)max_pairs, sizeof(pair_t));
    r->sub.hcap = 2 * max_pairs + 1;
    r->sub.slot = malloc(sizeof(int) * (size_t)r->sub.hcap);
    counter_clear(&r->sub);
    return r;
}

void fgo_reducer_free(fgo_reducer *r) {
    if (!r) return;
    free(r->size);
    free(r->expr);
    free(r->vars);
    free(r->sub.pairs);
    free(r->sub.slot);
    free(r);
}

void fgo_reducer_clear(fgo_reducer *r) { r->count = r->real = r->fresh = r->naive = 0; }

/* addExpression (additions_reducer.cuh:68-87) */
int fgo_reducer_add(fgo_reducer *r, const int *values, int n) {
    if (r->count >= r->max_expr) return 0;
    int size = 0;
    int *e = r->expr + (size_t)r->count * r->max_len;
    for (int i = 0; i < n; i++) {
        if (values[i] == 1 || values[i] == -1) {
            if (size >= r->max_len) return 0;
            e[size++] = values[i] * (i + 1);
        } else if (values[i] != 0) {
            return 0;
        }
    }
    if (n > r->real) r->real = n;
    r->size[r->count++] = size;
    r->naive += size - 1;
    return 1;
}

void fgo_reducer_set_mode(fgo_reducer *r, int mode) { /* setMode (:158-162) */
    r->mode = mode;
    r->scale = -1;
}

/* updateSubexpressions (:206-215) */
static void update_subexpressions(fgo_reducer *r) {
    counter_clear(&r->sub);
    for (int q = 0; q < r->count; q++) {
        const int *e = r->expr + (size_t)q * r->max_len;
        for (int i = 0; i < r->size[q]; i++)
            for (int j = i + 1; j < r->size[q]; j++) counter_insert(&r->sub, e[i], e[j]);
    }
    counter_sort(&r->sub);
}

/* binarySearch (:241-262) over |values| */
static int bsearch_abs(const int *e, int size, int value, int start) {
    int lo = start, hi = size - 1, v = abs(value);
    while (lo <= hi) {
        const int mid = lo + (hi - lo) / 2, a = abs(e[mid]);
        if (a < v) lo = mid + 1;
        else if (a > v) hi = mid - 1;
        else return mid;
    }
    return -1;
}

/* replaceExpression (:284-294) */
static void replace_expression(int *e, int *size, int i1, int i2, int var) {
    int j = i1;
    for (int i = i1 + 1; i < *size; i++)
        if (i != i2) e[j++] = e[i];
    (*size)--;
    e[j] = var;
}

/* replaceSubexpression (:217-239) */
static void replace_subexpression(fgo_reducer *r, int pi, int pj) {
    const int var = r->real + r->fresh + 1;
    for (int q = 0; q < r->count; q++) {
        if (r->size[q] < 2) continue;
        int *e = r->expr + (size_t)q * r->max_len;
        const int i = bsearch_abs(e, r->size[q], pi, 0);
        if (i == -1) continue;
        const int j = bsearch_abs(e, r->size[q], pj, i + 1);
        if (j == -1) continue;
        if (e[i] == pi && e[j] == pj) replace_expression(e, &r->size[q], i, j, var);
        else if (e[i] == -pi && e[j] == -pj) replace_expression(e, &r->size[q], i, j, -var);
    }
    r->vars[2 * r->fresh] = pi;
    r->vars[2 * r->fresh + 1] = pj;
    r->fresh++;
}

/* reduce (:89-105) */
static void reducer_reduce(fgo_reducer *r, rstream *s) {
    if (r->scale < 0) r->scale = rs_uniform(s) / 2;
    while (r->fresh < r->max_fresh) {
        update_subexpressions(r);
        if (!(r->sub.size > 0 && r->sub.max_count > 1)) break;
        const int step_mode = r->mode == MODE_MIX ? (int)(rs_next(s) % (uint32_t)MODE_MIX) : r->mode;
        const pair_t p = select_pair(&r->sub, step_mode, s, r->scale);
        replace_subexpression(r, p.i, p.j);
    }
}

/* partialInitialize (:134-142): replay the first fresh variables of best */
static void partial_initialize(fgo_reducer *r, const int *best_vars, int best_fresh, int count) {
    for (int q = 0; q < count && q < best_fresh; q++) replace_subexpression(r, best_vars[2 * q], best_vars[2 * q + 1]);
}

int fgo_reducer_additions(const fgo_reducer *r) { /* getAdditions (:164-172) */
    int a = r->fresh;
    for (int q = 0; q < r->count; q++) a += r->size[q] - 1;
    return a;
}

/* ---- one runReducersKernel pass ---- */

static const int kModes[15] = {3, 3, 3, 3, 3, 3, 3, 3, 1, 1, 1, 1, 2, 2, 5};
static const int kMaxFresh[3] = {250, 250, 500};
static const int kMaxPairs[3] = {64 * 63 / 2, 64 * 63 / 2, 350 * 349 / 2};

/* copySchemeToReducers (scheme_additions_reducer.cu:461-483): U and V rows
 * are expressions over the matrix entries, W columns expressions over the
 * products. */
static int load_component(fgo_reducer *r, const int32_t *scheme, int c) {
    const int n1 = scheme[0], n2 = scheme[1], n3 = scheme[2], m = scheme[3];
    const int nn[3] = {n1 * n2, n2 * n3, n3 * n1};
    const int32_t *U = scheme + 4, *V = U + (size_t)m * nn[0], *W = V + (size_t)m * nn[1];
    fgo_reducer_clear(r);
    if (c == 0) {
        for (int t = 0; t < m; t++)
            if (!fgo_reducer_add(r, U + (size_t)t * nn[0], nn[0])) return 0;
    } else if (c == 1) {
        for (int t = 0; t < m; t++)
            if (!fgo_reducer_add(r, V + (size_t)t * nn[1], nn[1])) return 0;
    } else {
        int *col = malloc(sizeof(int) * (size_t)(m > 0 ? m : 1));
        for (int k = 0; k < nn[2]; k++) {
            for (int t = 0; t < m; t++) col[t] = W[(size_t)t * nn[2] + k];
            if (!fgo_reducer_add(r, col, m)) {
                free(col);
                return 0;
            }
        }
        free(col);
    }
    return 1;
}

static const int32_t *scheme_at(const int32_t *blob, int64_t len, int index) {
    const int32_t *p = blob + 1;
    for (int q = 0; q < index; q++) {
        const int m = p[3], nn0 = p[0] * p[1], nn1 = p[1] * p[2], nn2 = p[2] * p[0];
        p += 4 + (size_t)m * (nn0 + nn1 + nn2);
        if (p - blob > len) return NULL;
    }
    return p;
}

/* One runReducersKernel pass for reducers 0..count-1.  blob: count-prefixed
 * schemes (reducer r takes scheme r % count_of_schemes, after fixSigns).
 * best_fresh[3] / best_vars[3][f_cap][2]: the best reducers' fresh
 * variables.  words[r * 4 + q] holds the words consumed from stream q of
 * reducer r (components 0-2, 3: shared decisions) and is advanced.
 * Outputs per (r, c), index rc = r * 3 + c:
 *   info[rc * 8 + ...] = expressions, real, fresh, naive, mode, scale bits, additions, 0
 *   size[rc * e_cap + q], expr[(rc * e_cap + q) * l_cap + x], vars[(rc * f_cap + v) * 2 + {0,1}]
 * Returns 0, or -1 when a capacity is exceeded. */
int fgo_reducer_iteration(const int32_t *blob, int64_t len, int count, uint64_t seed, int independent,
                          const int32_t *best_fresh, const int32_t *best_vars, uint64_t *words, int e_cap,
                          int l_cap, int f_cap, int32_t *info, int32_t *size, int32_t *expr, int32_t *vars) {
    const int nsch = blob[0];
    if (nsch < 1) return -1;
    int rc_err = 0;
    for (int r = 0; r < count && !rc_err; r++) {
        const int32_t *sch = scheme_at(blob, len, r % nsch);
        if (!sch) return -1;
        fgo_reducer *red[3];
        rstream st[4];
        for (int q = 0; q < 4; q++) {
            st[q].key[0] = (uint32_t)seed;
            st[q].key[1] = (uint32_t)(seed >> 32);
            st[q].r = (uint32_t)r;
            st[q].c = q;
            st[q].w = words[(size_t)r * 4 + q];
        }
        for (int c = 0; c < 3; c++) {
            red[c] = fgo_reducer_new(e_cap, l_cap, kMaxFresh[c] < f_cap ? kMaxFresh[c] : f_cap, kMaxPairs[c]);
            if (!load_component(red[c], sch, c)) rc_err = 1;
            /* setMode (scheme_additions_reducer.cu:505-507) */
            fgo_reducer_set_mode(red[c], r == 0 ? 0 : kModes[rs_next(&st[c]) % 15u]);
        }
        if (!rc_err) {
            if (independent) { /* :509-518 */
                for (int c = 0; c < 3; c++) {
                    const int take = (double)rs_uniform(&st[c]) < 0.3;
                    if (take && best_fresh[c] > 0)
                        partial_initialize(red[c], best_vars + (size_t)c * f_cap * 2, best_fresh[c],
                                           1 + (int)(rs_next(&st[c]) % (uint32_t)best_fresh[c]));
                }
            } else if ((double)rs_uniform(&st[3]) < 0.3) { /* :519-528 */
                for (int c = 0; c < 3; c++)
                    if (best_fresh[c] > 0)
                        partial_initialize(red[c], best_vars + (size_t)c * f_cap * 2, best_fresh[c],
                                           1 + (int)(rs_next(&st[c]) % (uint32_t)best_fresh[c]));
            }
            for (int c = 0; c < 3; c++) reducer_reduce(red[c], &st[c]); /* :537-539 */
        }
        for (int c = 0; c < 3; c++) {
            const size_t rcx = (size_t)r * 3 + c;
            fgo_reducer *x = red[c];
            int32_t *inf = info + rcx * 8;
            uint32_t sb;
            memcpy(&sb, &x->scale, 4);
            inf[0] = x->count;
            inf[1] = x->real;
            inf[2] = x->fresh;
            inf[3] = x->naive;
            inf[4] = x->mode;
            inf[5] = (int32_t)sb;
            inf[6] = fgo_reducer_additions(x);
            inf[7] = 0;
            for (int q = 0; q < x->count; q++) {
                size[rcx * e_cap + q] = x->size[q];
                for (int v = 0; v < x->size[q]; v++)
                    expr[(rcx * e_cap + q) * l_cap + v] = x->expr[(size_t)q * x->max_len + v];
            }
            for (int v = 0; v < x->fresh; v++) {
                vars[(rcx * f_cap + v) * 2] = x->vars[2 * v];
                vars[(rcx * f_cap + v) * 2 + 1] = x->vars[2 * v + 1];
            }
            fgo_reducer_free(x);
        }
        for (int q = 0; q < 4; q++) words[(size_t)r * 4 + q] = st[q].w;
    }
    return rc_err ? -1 : 0;
}

/* Single-component run for unit tests: expressions given as a dense
 * [n_expr][n_vars] matrix of {-1,0,1}; mode and stream (reducer r,
 * component c) chosen by the caller.  Writes the same outputs as
 * fgo_reducer_iteration for one (r, c). */
int fgo_reducer_single(const int32_t *matrix, int n_expr, int n_vars, int mode, uint64_t seed, int r, int c,
                       int max_fresh, int e_cap, int l_cap, int f_cap, int32_t *info, int32_t *size, int32_t *expr,
                       int32_t *vars) {
    fgo_reducer *x = fgo_reducer_new(e_cap, l_cap, max_fresh < f_cap ? max_fresh : f_cap, kMaxPairs[c]);
    for (int q = 0; q < n_expr; q++)
        if (!fgo_reducer_add(x, matrix + (size_t)q * n_vars, n_vars)) {
            fgo_reducer_free(x);
            return -1;
        }
    rstream st = {{(uint32_t)seed, (uint32_t)(seed >> 32)}, (uint32_t)r, (uint32_t)c, 0};
    fgo_reducer_set_mode(x, mode);
    reducer_reduce(x, &st);
    uint32_t sb;
    memcpy(&sb, &x->scale, 4);
    info[0] = x->count;
    info[1] = x->real;
    info[2] = x->fresh;
    info[3] = x->naive;
    info[4] = x->mode;
    info[5] = (int32_t)sb;
    info[6] = fgo_reducer_additions(x);
    info[7] = (int32_t)st.w;
    for (int q = 0; q < x->count; q++) {
        size[q] = x->size[q];
        for (int v = 0; v < x->size[q]; v++) expr[(size_t)q * l_cap + v] = x->expr[(size_t)q * x->max_len + v];
    }
    for (int v = 0; v < x->fresh; v++) {
        vars[2 * v] = x->vars[2 * v];
        vars[2 * v + 1] = x->vars[2 * v + 1];
    }
    fgo_reducer_free(x);
    return 0;
}

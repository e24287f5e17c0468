/*
 * fg_oracle.h -- CPU oracle for the flip-graph random walk.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker (or the timed CPU baseline).  The product path (libflipgraph.so,
 * the CLIs) never links, loads or calls anything under oracle/.
 *
 * This is a plain-C restatement of the reference's walk logic
 * (dronperminov/FlipGraphGPU).  Every function cites the reference
 * file:line it follows.  Two deliberate, documented differences:
 *
 *  - RNG: cuRAND is replaced by Philox4x32-10 (Salmon et al., SC'11),
 *    keyed per walker.  One Philox block feeds each walk iteration; rare
 *    operators draw from a second (aux) stream of the same generator.
 *  - Pair order: the reference samples `curand() % size` into a FlipSet
 *    whose order depends on its add/remove history
 *    (src/entities/flip_set.cu:7-31).  In CANONICAL mode the pair list is
 *    always taken in the order a fresh initFlips() builds
 *    (src/schemes/scheme_z2.cu:141-150: component, then index1, then
 *    index2).  The set of pairs, and therefore the distribution of flips,
 *    is the same.  REFERENCE mode keeps the history-ordered FlipSet
 *    exactly as the reference does (used for the CPU baseline).
 *
 * Pins: see tests/test_oracle_*.py (Philox known-answer vectors, closed
 * forms, Strassen, brute-force invariants, planted-mistake checks).
 */
#ifndef FG_ORACLE_H
#define FG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FGO_MAX_RANK 350  /* src/config.cuh:10 */
#define FGO_MAX_PAIRS 500 /* src/config.cuh:14 */
#define FGO_RING_Z2 0
#define FGO_RING_ZT 1
#define FGO_SELECT_CANONICAL 0
#define FGO_SELECT_REFERENCE 1
#define FGO_PLUS_MAX_TRIES 1000 /* bound on tryPlus rejection (reference loops forever) */

/* src/entities/flip_set.cuh:5-18 */
typedef struct {
    int size;
    uint32_t pairs[FGO_MAX_PAIRS];
} fgo_flipset;

/* src/schemes/scheme_z2.cuh:21-27 and src/schemes/scheme_integer.cuh:16-21.
 * Z2: v[c][t] is the factor bitmask.  ZT: (v, s) are Addition.values /
 * Addition.signs (src/entities/addition.cuh:8-12), ok is Addition.valid. */
typedef struct {
    int ring;
    int n[3];
    int nn[3];
    int m;
    uint64_t v[3][FGO_MAX_RANK];
    uint64_t s[3][FGO_MAX_RANK];
    uint8_t ok[3][FGO_MAX_RANK];
    fgo_flipset flips[3];
} fgo_scheme;

typedef struct {
    uint32_t key[2];
    uint32_t walker;
    uint64_t it;  /* main-stream block counter: one block per walk iteration */
    uint64_t aux; /* aux-stream block counter */
    uint32_t abuf[4];
    int aidx;     /* next unread aux word, 4 = empty */
} fgo_rng;

typedef struct {
    int iterations;        /* walk-loop iterations per launch */
    int max_iterations;    /* upper bound of the random iteration count */
    int random_iterations; /* reference: resize probability > 0 */
    int plus_iterations;
    int use_reduce;
    uint32_t thr_reduce;   /* P(reduce check) = thr / 2^32 */
    int use_expand;
    uint32_t thr_expand;
    uint32_t thr_best_eq;  /* equal-rank best copy, 0.01 * 2^32 */
    int select_mode;       /* FGO_SELECT_* */
    int check_reduce;      /* tryFlip(checkReduce) */
    int zt_visit_cap;      /* candidates one integer tryFlip visits one at a time before its
                              exhaustive pass (at most 16; <= 0: 16) */
    int mode;              /* 0: flip-graph walk; 1: complexity minimizer (best_rank holds the complexity);
                              3: the additions reducer's scheme flips (iterations = max flips) */
    int use_sandwich;      /* sandwiching edge (Z2): P = thr_sandwich / 2^32 per applied flip */
    uint32_t thr_sandwich;
    int use_basis;         /* basis-change edge (Z2) */
    uint32_t thr_basis;
    int min_window;        /* iterations whose ranks count toward min_m (0: the whole launch) */
    int look_period;       /* Z2 lookahead proposals: every look_period-th iteration of a launch tries */
    int look_candidates;   /* look_candidates extra pairs and takes the first that reduces (0: off) */
} fgo_walk_params;

typedef struct {
    fgo_scheme cur;
    fgo_scheme best;
    int best_rank;
    int flips_count;
    fgo_rng rng;
    uint64_t total_flips;
    uint64_t sum_m;
    int forced[3];  /* proposed first flip of the next launch (component -1: none) */
    int log[4];     /* first flip of the last launch: component (-1: none), index1, index2, rank */
    int min_m;      /* lowest rank during the last launch */
} fgo_walker;

/* Walker record exchanged with tests; same layout as fg_walker_record in
 * include/fg.h (written independently). */
#define FGO_REC_MAX_M 256
typedef struct {
    int32_t n[3];
    int32_t m;
    int32_t best_m;
    int32_t best_rank;
    int32_t flips_count;
    int32_t aux_idx;
    uint64_t it;
    uint64_t aux;
    uint64_t total_flips;
    uint64_t sum_m;
    uint64_t cur_v[3][FGO_REC_MAX_M];
    uint64_t cur_s[3][FGO_REC_MAX_M];
    uint64_t best_v[3][FGO_REC_MAX_M];
    uint64_t best_s[3][FGO_REC_MAX_M];
    int32_t best_n[3]; /* format of the best scheme (0: same as n) */
    int32_t reserved;
} fgo_walker_record;

/* ---- RNG ---- */
void fgo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void fgo_rng_init(fgo_rng *r, uint64_t seed, uint32_t walker);
void fgo_rng_main_block(fgo_rng *r, uint32_t out[4]);
uint32_t fgo_rng_aux32(fgo_rng *r);

/* ---- schemes ---- */
int fgo_scheme_size(void);
void fgo_init_naive(fgo_scheme *s, int ring, int n1, int n2, int n3);
int fgo_validate(const fgo_scheme *s);
int fgo_complexity(const fgo_scheme *s);
void fgo_init_flips(fgo_scheme *s);
int fgo_count_pairs(const fgo_scheme *s);
/* blob: int32 [n1 n2 n3 m | U m*nn0 | V m*nn1 | W m*nn2] (text-input order) */
int fgo_from_blob(fgo_scheme *s, int ring, const int32_t *blob, int check_validity);
int fgo_to_blob(const fgo_scheme *s, int32_t *blob);
int fgo_blob_len(const fgo_scheme *s);
void fgo_copy(fgo_scheme *dst, const fgo_scheme *src, int with_flips);

/* explicit operators (trace replay) */
void fgo_flip(fgo_scheme *s, int i, int j, int k, int i1, int i2, int check_reduce);
void fgo_plus(fgo_scheme *s, int i, int j, int k, int i1, int i2, int variant);
void fgo_split_z2(fgo_scheme *s, int i, int j, int k, int index, uint64_t a1);
void fgo_split_zt(fgo_scheme *s, int i, int j, int k, int index, uint64_t av, uint64_t as);
void fgo_reduce(fgo_scheme *s, int i, int i1, int i2, int sub);
int fgo_try_reduce(fgo_scheme *s);
int fgo_fix_signs(fgo_scheme *s);
int fgo_pick_canonical(const fgo_scheme *s, uint32_t r, int *c, int *t1, int *t2);

/* ---- walk ---- */
void fgo_walker_init(fgo_walker *w, const fgo_scheme *start, int best_rank, uint64_t seed, uint32_t walker);
void fgo_walk(fgo_walker *w, const fgo_walk_params *p);
void fgo_walker_to_record(const fgo_walker *w, fgo_walker_record *rec);
int fgo_walker_from_record(fgo_walker *w, int ring, const fgo_walker_record *rec, uint64_t seed, uint32_t walker);

/* Batch helpers for ctypes: walkers are passed as records.  Each runs
 * `launches` launches of fgo_walk.  nthreads > 1 partitions walkers over
 * pthreads (CPU baseline).  Returns 0 on success. */
int fgo_walk_records(int ring, fgo_walker_record *recs, int count, uint32_t walker_base,
                     uint64_t seed, const fgo_walk_params *p, int launches, int nthreads);
int fgo_naive_records(int ring, int n1, int n2, int n3, fgo_walker_record *recs, int count);
/* One launch of the Z2 walk with a proposed first flip per walker
 * (forced[count][4]: component (-1: none), index1, index2, unused); writes
 * the first flip taken (log[count][4]) and the launch's lowest rank
 * (minm[count]); either output may be null. */
int fgo_walk_records_forced(fgo_walker_record *recs, int count, uint32_t walker_base, uint64_t seed,
                            const fgo_walk_params *p, const int32_t *forced, int32_t *log, int32_t *minm);
int fgo_record_validate(int ring, const fgo_walker_record *rec, int best);

/* Blob-level helpers for ctypes. */
int fgo_blob_validate(int ring, const int32_t *blob);
int fgo_blob_complexity(int ring, const int32_t *blob);
int fgo_blob_naive(int ring, int n1, int n2, int n3, int32_t *out);
/* Replay an explicit operator trace.  ops: 8 int32 per op:
 * [code, a0..a6].  code 0 flip(i,j,k,i1,i2,checkReduce); 1 plus(i,j,k,i1,i2,variant);
 * 2 split(i,j,k,index, a_lo, a_hi[, s_lo]) (ZT: a = values, s = signs in a5/a6 low bits);
 * 3 reduce(i,i1,i2,sub); 4 tryReduce over the stored pair lists; 5 fixSigns;
 * 6 tryReduce over the fresh-list order (the walk's canonical order).
 * Returns new blob length or -1. */
int fgo_blob_replay(int ring, const int32_t *in, const int32_t *ops, int nops, int32_t *out);
int fgo_blob_pairs(int ring, const int32_t *blob, int mode, uint32_t *out, int cap);
int fgo_blob_try_flip_zt(const int32_t *in, uint64_t seed, uint32_t walker, uint64_t it, int select_mode,
                         int visit_cap, int check_reduce, int32_t *out, int32_t *op);
int fgo_check_flipset_walk(int ring, int n1, int n2, int n3, uint64_t seed, int iterations, int walkers,
                           int plus_iterations);

/* Resize and meta operators (Z2).  MAX_RANK and MAX_MATRIX_ELEMENTS of the
 * reference (src/config.cuh:10-11) are replaced by the pool's rank capacity
 * and factor width; the other limits are the reference's (:34-37). */
#define FGO_MIN_PROJECT_N 2
#define FGO_MAX_EXTENSION_N 16
#define FGO_THR_HALF 2147483648u   /* floor(0.5 * 2^32) */
#define FGO_THR_005 214748364u     /* floor(0.05 * 2^32) */
#define FGO_THR_055 2362232012u    /* floor(0.55 * 2^32) */
#define FGO_THR_085 3650722201u    /* floor(0.85 * 2^32) */
typedef struct {
    int max_rank;
    int max_coef;
} fgo_resize_caps;
void fgo_project(fgo_scheme *s, int p, int q);
void fgo_extend(fgo_scheme *s, int p);
void fgo_product_dim(fgo_scheme *s, int p);
void fgo_merge(fgo_scheme *s, const fgo_scheme *o, int p);
void fgo_product_scheme(fgo_scheme *s, const fgo_scheme *o);
void fgo_swap_size(fgo_scheme *s, int p1, int p2);
void fgo_swap_basis_rows(fgo_scheme *s, int i1, int i2);
void fgo_swap_basis_columns(fgo_scheme *s, int j1, int j2);
void fgo_resize_walker(fgo_walker *w, const fgo_scheme *const *bests, int count, uint64_t thr_resize,
                       const fgo_resize_caps *k);
int fgo_resize_records(int ring, fgo_walker_record *recs, int count, uint32_t walker_base, uint64_t seed,
                       uint64_t thr_resize, int max_rank, int max_coef);
int fgo_blob_resize_op(int ring, const int32_t *in, const int32_t *other, int op, int a, int b, int32_t *out);

/* Additions reducer (fg_reducer.c). */
typedef struct fgo_reducer fgo_reducer;
fgo_reducer *fgo_reducer_new(int max_expr, int max_len, int max_fresh, int max_pairs);
void fgo_reducer_free(fgo_reducer *r);
void fgo_reducer_clear(fgo_reducer *r);
int fgo_reducer_add(fgo_reducer *r, const int *values, int n);
void fgo_reducer_set_mode(fgo_reducer *r, int mode);
int fgo_reducer_additions(const fgo_reducer *r);
int fgo_reducer_iteration(const int32_t *blob, int64_t len, int count, uint64_t seed, int independent,
                          const int32_t *best_fresh, const int32_t *best_vars, uint64_t *words, int e_cap,
                          int l_cap, int f_cap, int32_t *info, int32_t *size, int32_t *expr, int32_t *vars);
int fgo_reducer_single(const int32_t *matrix, int n_expr, int n_vars, int mode, uint64_t seed, int r, int c,
                       int max_fresh, int e_cap, int l_cap, int f_cap, int32_t *info, int32_t *size, int32_t *expr,
                       int32_t *vars);

#ifdef __cplusplus
}
#endif
#endif

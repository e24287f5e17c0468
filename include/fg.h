/*
 * fg.h -- C ABI of the MI355X flip-graph search engine (libflipgraph.so).
 *
 * The reference (dronperminov/FlipGraphGPU) has no library API: its
 * boundary is the flip_graph / complexity_minimizer / additions_reducer
 * processes, whose host classes own managed walker pools and launch
 * one-thread-per-walker kernels.  These entry points replace those host
 * classes and kernels; the CLIs in flipgraphgpu_amd/csrc call only this
 * ABI.  Conventions: int return codes (0 = ok, < 0 = error, message via
 * fg_last_error()), never exit(); the context owns device memory, callers
 * own every buffer passed in; one host thread and one HIP stream per
 * context.
 */
#ifndef FG_H
#define FG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FG_API __attribute__((visibility("default")))

#define FG_RING_Z2 0 /* src/config.cuh:43-49, SchemeZ2 */
#define FG_RING_ZT 1 /* SchemeInteger, coefficients in {-1,0,1} */

#define FG_REC_MAX_M 256 /* term capacity of fg_walker_record */

typedef struct fg_ctx fg_ctx;

/* Flags of src/main_flip_graph.cu:10-25 that reach the walk. */
typedef struct {
    int32_t ring;               /* FG_RING_Z2 / FG_RING_ZT (compile-time in the reference, src/config.cuh:6) */
    int32_t walkers;            /* --schemes */
    int32_t max_iterations;     /* --max-iterations */
    int32_t plus_iterations;    /* --plus-iterations */
    int32_t random_iterations;  /* 1 when --resize-probability > 0 (src/entities/flip_graph.cu:530) */
    int32_t check_reduce;       /* tryFlip(checkReduce): 1 flip_graph, 0 complexity_minimizer */
    double expand_probability;  /* --expand-probability (divided by max iterations) */
    double reduce_probability;  /* --reduce-probability (divided by max iterations) */
    uint64_t seed;              /* --seed (0 is replaced by time(0) in the CLI, not here) */
    int32_t max_rank;           /* term capacity per walker; 0 = max(naive rank, input ranks) */
    int32_t zt_visit_cap;       /* integer ring: candidates one tryFlip visits one at a time before its exhaustive pass (0 = 16, at most 16) */
    int32_t use_graph;          /* 1: replay each launch (queue reset, walk, Brent check) as a captured HIP graph */
    int32_t grid;               /* persistent workgroups per launch; 0 = all that fit on the GPU at once */
    int32_t word_bits;          /* Z2 LDS factor words: 16 or 32; 0 = 16 only when it lets more walker blocks run at once */
    double sandwiching_probability; /* --sandwiching-probability (divided by max iterations; Z2 needs 32-bit words; empty for ZT as in the reference) */
    double basis_probability;       /* --basis-probability (divided by max iterations; Z2 needs 32-bit words) */
    int32_t partner_layout;     /* Z2 flip-partner bookkeeping: 0 = class masks when the scheme has at most 32
                                   terms and 16-coefficient factors (and word_bits is 0), else partner lists;
                                   1 = partner lists; 2 = class masks; 3 = class masks with one lane per
                                   walker (64 walkers per workgroup instead of lane pairs serving 32) */
    int32_t lookahead_period;   /* Z2 reduction-seeking proposals (no reference counterpart; 0 = off): every
                                   lookahead_period-th iteration of a launch first draws lookahead_candidates
                                   pairs with orientation coins from the aux stream and applies the first whose
                                   flip lowers the rank, else flips uniformly as usual */
    int32_t lookahead_candidates;
    int32_t reserved0;
} fg_config;

/* Per-walker state, for tests and checkpointing. */
typedef struct {
    int32_t n[3];
    int32_t m;
    int32_t best_m;
    int32_t best_rank;
    int32_t flips_count; /* flips since the last rank drop (the reference's flips[idx]) */
    int32_t aux_idx;
    uint64_t it;          /* main RNG stream position = walk iterations done */
    uint64_t aux;         /* aux RNG stream position */
    uint64_t total_flips; /* applied flips */
    uint64_t sum_m;       /* sum over applied flips of the rank at that flip */
    uint64_t cur_v[3][FG_REC_MAX_M];  /* Z2 masks / ZT coefficient magnitudes */
    uint64_t cur_s[3][FG_REC_MAX_M];  /* ZT signs (0 for Z2) */
    uint64_t best_v[3][FG_REC_MAX_M];
    uint64_t best_s[3][FG_REC_MAX_M];
    int32_t best_n[3];    /* format of the best scheme (0 on write: same as n) */
    int32_t reserved;
} fg_walker_record;

FG_API const char *fg_last_error(void);

/* FlipGraph::FlipGraph (src/entities/flip_graph.cu:4-188): allocate the
 * walker pool and seed the per-walker RNG streams. */
FG_API int fg_create(const fg_config *cfg, fg_ctx **out);
/* FlipGraph::~FlipGraph (src/entities/flip_graph.cu:417-442). */
FG_API int fg_destroy(fg_ctx *ctx);

/* FlipGraph::initializeNaive + initialize (src/entities/flip_graph.cu:226-233,
 * 190-199): every walker starts from the naive n1 x n2 x n3 scheme. */
FG_API int fg_init_naive(fg_ctx *ctx, int n1, int n2, int n3);

/* FlipGraph::initializeFromFile (src/entities/flip_graph.cu:201-224): blob is
 * int32 [count, then count x (n1 n2 n3 m, U m*n1n2, V m*n2n3, W m*n3n1)] --
 * the reference's --input-path text format as integers.  Each scheme is
 * reduced with tryReduce until it no longer shrinks, then walkers are filled
 * round-robin.  Returns -2 if a scheme fails the Brent check. */
FG_API int fg_init_from_schemes(fg_ctx *ctx, const int32_t *blob, int64_t len);

/* ComplexityMinimizer::read + initialize + initializeKernel
 * (src/entities/complexity_minimizer.cu:22-37, 67-77, 172-188): blob as for
 * fg_init_from_schemes (one format for all schemes, no reduction).  Walker w
 * starts from scheme w % count, its best scheme is its start and every
 * walker's best complexity starts at the lowest complexity among the inputs.
 * The pool then runs fg_minimize instead of fg_walk; fg_best picks the
 * walker of lowest best complexity.  Returns -2 if a scheme fails the Brent
 * check. */
FG_API int fg_init_minimizer(fg_ctx *ctx, const int32_t *blob, int64_t len);

/* ComplexityMinimizer::minimizeIteration / minimizeKernel
 * (src/entities/complexity_minimizer.cu:79-84, 190-214): every walker flips
 * without reductions (tryFlip(state, false)) up to max_iterations times,
 * stopping at the first flip that finds no pair, and keeps the scheme of
 * lowest complexity below its best complexity.  Runs the GPU Brent check of
 * every current scheme; returns the number of invalid ones or < 0. */
FG_API int fg_minimize(fg_ctx *ctx);

/* Scheme pool of the additions reducer's --max-flips mode
 * (SchemeAdditionsReducer::initialize + initializeKernel,
 * src/entities/scheme_additions_reducer.cu:144-175,433-446): blob holds one
 * integer-ring scheme (count 1, no reduction); every walker starts from it
 * and walker 0 keeps it.  Returns -2 if it fails the Brent check. */
FG_API int fg_init_flips(fg_ctx *ctx, const int32_t *blob, int64_t len);

/* flipSchemesKernel (src/entities/scheme_additions_reducer.cu:448-460): every
 * walker w >= 1 restarts from walker 0's scheme, draws a count in
 * [0, max_flips] from its own random stream and applies that many flips with
 * reductions (tryFlip), stopping at the first that finds no pair; walker 0
 * is left alone.  Runs the GPU Brent check of every scheme; returns the
 * number of invalid ones or < 0. */
FG_API int fg_flip_schemes(fg_ctx *ctx, int max_flips);

/* bestComplexities[] (minimizer pools only; best may be null otherwise) and
 * getComplexity() of every walker's current scheme after the last launch
 * (src/entities/complexity_minimizer.cu:108-128).  Each non-null output
 * holds `walkers` values. */
FG_API int fg_complexities(fg_ctx *ctx, int32_t *best, int32_t *cur);

/* schemesBest[src].copyTo(schemes[dst[q]]) for q < count
 * (src/entities/complexity_minimizer.cu:130-133): the best scheme of walker
 * src becomes the current scheme of every listed walker (src may be listed).
 * Random streams and best schemes are untouched. */
FG_API int fg_broadcast_best(fg_ctx *ctx, int src, const int32_t *dst, int count);

/* Several formats in one pool (no reference counterpart: the reference runs
 * one format per process; its resize edges mix formats only through
 * src/entities/flip_graph.cu:575-613).  Walkers are split between the
 * nformats formats (n1 n2 n3 triples) in proportion to weights: format f gets
 * floor(walkers * w_f / sum w) and the remaining walkers go one each to the
 * formats with the largest fractional parts (ties: lower f).  Every walker
 * starts from its format's naive scheme; each format's walkers are
 * contiguous, in the given order. */
FG_API int fg_init_batched(fg_ctx *ctx, int nformats, const int32_t *formats, const int32_t *weights);

/* Autotuner (no reference counterpart; the reference's --block-size, threads
 * per block, is fixed here at 32 walkers per workgroup -- one wave, two lanes
 * per walker).  Times one launch of `iterations` per candidate and keeps the
 * fastest in flips per second: persistent grids of 1, 2, ... times the CU
 * count up to the occupancy limit, and, for Z2 pools whose config leaves the
 * layout open (word_bits 0, partner_layout 0, no basis edges, no trajectory
 * log), every LDS layout the pool fits (class masks, 16-bit words with
 * partner lists, 32-bit words with partner lists).  The trials are ordinary
 * walk launches; every layout and grid gives the same trajectories.  *grid
 * gets the chosen grid. */
FG_API int fg_autotune(fg_ctx *ctx, int iterations, int32_t *grid);

/* The pool's walk layout (0: Z2, 32-bit words and partner lists; 1: integer
 * ring; 2: Z2, 16-bit words and partner lists; 3: Z2 class masks) and the
 * trials of the last fg_autotune, [n][3] = layout, grid, flips per second /
 * 1e6 (at most cap rows).  Returns the number of trials. */
FG_API int fg_layout_info(fg_ctx *ctx, int32_t *layout, int32_t *trials, int cap);

/* FlipGraph::randomWalk / randomWalkKernel (src/entities/flip_graph.cu:235-240,
 * 520-573): one launch.  iterations <= 0 uses max_iterations (random in
 * [1, max_iterations] per walker when random_iterations is set).  Runs the
 * GPU Brent check of every walker's current scheme afterwards; returns the
 * number of invalid schemes found (0 expected) or < 0 on error. */
FG_API int fg_walk(fg_ctx *ctx, int iterations);

/* fg_walk that also records, per walker, every operator the launch applies
 * to its scheme (the flip, plus, split and tryReduce calls of
 * randomWalkKernel, src/entities/flip_graph.cu:520-573, with the arguments
 * its random draws chose; src/schemes/scheme_z2.cu:531-619,1127-1154,
 * scheme_integer.cu:428-547,1062-1101).  The trajectory is the one fg_walk
 * takes.  ops[w][cap][8] gets walker w's first cap operators as
 * [code, a0..a6]: 0 flip(i, j, k, index1, index2, checkReduce),
 * 1 plus(i, j, k, index1, index2, variant), 2 split(i, j, k, index, a, s)
 * (Z2: a = the new factor mask, s = 0; integer ring: a = magnitudes,
 * s = signs), 6 tryReduce (over the fresh-list pair order, the walk's);
 * counts[w] gets the number of operators applied,
 * which may exceed cap (the rest are not stored).  Replaying a walker's
 * trace from its scheme before the launch gives its scheme after it.  Plain
 * walk pools only (no minimizer, flips, logging, lookahead or basis edges).
 * No reference counterpart: the reference records nothing.  Returns what
 * fg_walk returns. */
FG_API int fg_walk_trace(fg_ctx *ctx, int iterations, int cap, int32_t *ops, int32_t *counts);

/* resizeKernel (src/entities/flip_graph.cu:575-613) on a walk pool (Z2 with
 * 32-bit factor words, or the integer ring): each walker resizes with the
 * given probability -- half the
 * time it first swaps two dimensions (swapSize), then it tries to merge with
 * the best scheme of up to three random walkers (merge), else it projects
 * out a row (project, 5%), multiplies by a random walker's best scheme
 * (product of schemes, 50%), doubles a dimension (product, 30%) or extends
 * one (extend, 15%) (src/schemes/scheme_z2.cu:643-979,1165-1324,
 * scheme_integer.cu:573-905).  The reference's MAX_RANK and
 * MAX_MATRIX_ELEMENTS are the pool's term capacity and the width of a factor
 * word (32 coefficients for Z2, 16 for the integer ring).  Runs the GPU Brent
 * check of every scheme; returns the number of invalid ones or < 0. */
FG_API int fg_resize(fg_ctx *ctx, double probability);

/* schemes[i].copyTo(schemesBest[i]) of FlipGraph::updateRanks
 * (src/entities/flip_graph.cu:249-262) for the listed walkers: the current
 * scheme, rank and format become the walker's best. */
FG_API int fg_snapshot_best(fg_ctx *ctx, const int32_t *walkers, int count);

/* bestRanks[] := best_rank[walkers] (updateRanks, :264-265). */
FG_API int fg_set_best_ranks(fg_ctx *ctx, const int32_t *best_rank);

/* Best scheme over all walkers (smallest best rank; lowest walker index
 * among ties), as a single-scheme blob (no count prefix). */
FG_API int fg_best(fg_ctx *ctx, int32_t *out_blob, int64_t cap, int32_t *out_rank, int32_t *out_walker);

/* Blob of walker `walker`'s best (which=1) or current (which=0) scheme. */
FG_API int fg_walker_scheme(fg_ctx *ctx, int walker, int which, int32_t *out_blob, int64_t cap);

/* Scheme::validate (src/schemes/scheme_z2.cu:3-29, scheme_integer.cu:3-38) on
 * the GPU for every scheme of a count-prefixed blob; out_ok[i] = 1 if valid.
 * Returns the number of valid schemes or < 0. */
FG_API int fg_verify(int ring, const int32_t *blob, int64_t len, int32_t *out_ok);

/* The Brent check on the host (Scheme::validate, src/schemes/scheme_z2.cu:3-29,
 * scheme_integer.cu:3-38): the library's own re-verification of GPU results,
 * for a count-prefixed blob as fg_verify takes; needs no GPU.  Returns the
 * number of valid schemes or < 0. */
FG_API int fg_verify_host(int ring, const int32_t *blob, int64_t len, int32_t *out_ok);

/* getComplexity (src/schemes/scheme_z2.cu:127-139, scheme_integer.cu:139-147)
 * of every scheme of a count-prefixed blob, on the host.  Returns the count
 * or < 0. */
FG_API int fg_complexity_host(int ring, const int32_t *blob, int64_t len, int32_t *out);

/* fg_verify without any allocation (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The staging (a copy
 * of the blob, the packed factor planes and the flags) lives in a per-GPU
 * file-scope device array of 8 MiB, held until the process ends; a blob that
 * needs more is rejected (< 0, see fg_last_error).  Runs on the null stream
 * and returns when out_ok is written.  Calls that overlap on one GPU share its
 * scratch array and race.  Same arguments, results and errors as fg_verify. */
FG_API int fg_verify_noalloc(int ring, const int32_t *blob, int64_t len, int32_t *out_ok);

/* Aggregate counters since fg_init_*: applied flips, sum of ranks at those
 * flips, walk iterations. */
FG_API int fg_counters(fg_ctx *ctx, uint64_t *flips, uint64_t *sum_m, uint64_t *iterations);

/* Duration of the last fg_walk kernel launch, measured with HIP events on
 * the context's stream. */
FG_API int fg_last_walk_ms(fg_ctx *ctx, float *ms);

/* Per-walker summary for reports (FlipGraph::report, src/entities/flip_graph.cu:303-380):
 * current rank, best-scheme rank, bestRanks[], flips-since-last-drop and the
 * format (3 values per walker).  Any output pointer may be null; each
 * non-null one holds `walkers` values (3 * walkers for n). */
FG_API int fg_summary(fg_ctx *ctx, int32_t *m, int32_t *best_m, int32_t *best_rank, int32_t *flips_count, int32_t *n);

/* Number of walkers and the format of walker 0. */
FG_API int fg_shape(fg_ctx *ctx, int32_t *walkers, int32_t *n1, int32_t *n2, int32_t *n3);

/* Read / overwrite walker state (tests, checkpoints). */
FG_API int fg_read_walkers(fg_ctx *ctx, int begin, int count, fg_walker_record *out);
FG_API int fg_write_walkers(fg_ctx *ctx, int begin, int count, const fg_walker_record *in);

/* ---- Learned proposals (no reference counterpart: the walk's uniform pair
 * draw, src/schemes/scheme_z2.cu:982-1027, stays the fallback) ---- */

/* Trajectory log of a Z2 pool: from now on every fg_walk launch records, for
 * walkers w with w % every == 0, the scheme before the launch's first flip
 * and the pair flipped.  `horizon` launches later the row is labelled 1 if
 * the walker's rank fell below the rank at that flip in the meantime, else
 * 0.  At most max_batches launches are kept. */
FG_API int fg_log_enable(fg_ctx *ctx, int64_t max_batches, int every, int horizon);
/* Labels from a window of the logging launch itself (log horizon 1): a
 * logged flip is labelled 1 when the walker's rank falls below the rank at
 * the flip within the launch's first `iterations` iterations (the flip's
 * own included); 0 restores whole-launch labels. */
FG_API int fg_log_window(fg_ctx *ctx, int iterations);
/* Labelled rows, rows labelled 1, labelled launches. */
FG_API int fg_log_info(fg_ctx *ctx, int64_t *rows, int64_t *positives, int64_t *batches);
/* Labelled rows begin..begin+count-1: words [count][3][max_terms] (factor
 * masks, zero beyond the rank), m [count], pairs [count][3] (component,
 * index1 < index2), labels [count].  Any output may be null. */
FG_API int fg_log_read(fg_ctx *ctx, int64_t begin, int64_t count, uint32_t *words, int32_t *m, int32_t *pairs,
                       float *labels);
/* Proposed first flips per walker [walkers][4] = component (-1: none),
 * index1, index2, 1; fg_walk_forced runs one launch whose first iteration
 * flips the proposed pair when it still is a flip (uniform draw otherwise). */
FG_API int fg_set_forced(fg_ctx *ctx, const int32_t *forced);
FG_API int fg_get_forced(fg_ctx *ctx, int32_t *forced);
FG_API int fg_walk_forced(fg_ctx *ctx, int iterations);

typedef struct fg_scorer fg_scorer;
/* Term-set flip scorer for n1 x n2 x n3 Z2 schemes of up to max_terms terms
 * (BF16 weights on the MFMA units, FP32 accumulation). */
FG_API int fg_scorer_create(int n1, int n2, int n3, int max_terms, uint64_t seed, fg_scorer **out);
FG_API int fg_scorer_destroy(fg_scorer *s);
FG_API int64_t fg_scorer_num_params(fg_scorer *s);
/* 16 values: F D Z max_terms and the offsets of W0 b0 A1 B1 c1 A2 B2 c2 Wz bz wo bo. */
FG_API int fg_scorer_layout(fg_scorer *s, int32_t *out);
FG_API int fg_scorer_get_params(fg_scorer *s, float *out);
FG_API int fg_scorer_set_params(fg_scorer *s, const float *in);
/* Forward on a host batch: words [count][3][max_terms], m [count], pairs
 * [count][3]; scores [count].  With labels: mean BCE loss and the gradient
 * of every parameter (fg_scorer_layout order). */
FG_API int fg_scorer_eval(fg_scorer *s, int count, const uint32_t *words, const int32_t *m, const int32_t *pairs,
                          const float *labels, float *scores, float *loss, float *grads);
/* Adam on batches sampled from ctx's trajectory log; losses[steps] (may be null). */
FG_API int fg_scorer_train(fg_scorer *s, fg_ctx *ctx, int steps, int batch, float lr, uint64_t seed, float *losses);
/* For walkers w with w % every == 0: draw `candidates` equal-factor pairs
 * uniformly, score them and propose the best (fg_set_forced's layout). */
FG_API int fg_scorer_propose(fg_scorer *s, fg_ctx *ctx, int candidates, int every, uint32_t round);
/* Candidates and scores of the last proposal round: walkers [tuples], pairs
 * [tuples][candidates][3] (component -1 for a missing candidate), scores
 * [tuples][candidates].  A fused round (schemes of up to 64 terms and 96
 * factor bits, at most 16 candidates) reports every proposing walker (a
 * missing candidate scores -1e30); a chunked round (larger schemes) reports
 * its last chunk of at most 4096 walkers.  Returns the number of tuples (0
 * before any round); any output may be null. */
FG_API int fg_scorer_last_proposals(fg_scorer *s, int32_t *walkers, int32_t *pairs, float *scores, int64_t cap);
/* Duration in ms of the fused proposal kernel of the last round (HIP events on
 * the scorer's stream; 0 when the last round was not fused).  No counterpart in
 * the reference, which has no learned proposals. */
FG_API int fg_scorer_last_propose_ms(fg_scorer *s, float *ms);
/* fg_scorer_propose + fg_walk_forced. */
FG_API int fg_walk_learned(fg_ctx *ctx, fg_scorer *s, int iterations, int candidates, int every, uint32_t round);

/* ---- Additions reducer (src/entities/scheme_additions_reducer.cu,
 * src/entities/additions_reducer.cuh) ---- */
typedef struct fg_reducer fg_reducer;

/* SchemeAdditionsReducer constructor + read + initialize
 * (src/entities/scheme_additions_reducer.cu:3-27, 29-101, 144-175): blob is
 * one scheme [n1 n2 n3 m U V W] over the integers (fixSigns is applied, the
 * Brent check must pass; -2 otherwise).  count reducers run per pass;
 * schemes_count as --schemes-count; max_flips > 0 is not supported. */
FG_API int fg_reducer_create(const int32_t *blob, int64_t len, int count, int schemes_count, int max_flips,
                             uint64_t seed, fg_reducer **out);
FG_API int fg_reducer_destroy(fg_reducer *r);
/* 16 values: n1 n2 n3 m count schemes_count max_flips top e_cap l_cap f_cap
 * p_cap h_bits lds_bytes 0 0. */
FG_API int fg_reducer_info(fg_reducer *r, int32_t *out);
/* naive additions, max real variables and max subexpressions of U, V, W
 * (getMaxRealVariables / getMaxSubexpressions, src/schemes/scheme_integer.cu:149-211). */
FG_API int fg_reducer_scheme_stats(fg_reducer *r, int32_t *out9);
/* reduceIteration (src/entities/scheme_additions_reducer.cu:177-188):
 * one runReducersKernel pass of every reducer on the GPU, results read back. */
FG_API int fg_reducer_run(fg_reducer *r, int iteration);
/* updateBest (src/entities/scheme_additions_reducer.cu:287-307) without the
 * printing and saving: *improved = 0 (no), 1 (fewer additions) or 2 (same
 * additions, fewer fresh variables). */
FG_API int fg_reducer_update(fg_reducer *r, int32_t *improved);
/* 11 values: best additions U V W, best fresh variables U V W, naive
 * additions U V W, total additions, total fresh variables. */
FG_API int fg_reducer_best(fg_reducer *r, int32_t *out11);
/* Report rows (src/entities/scheme_additions_reducer.cu:309-363): for each
 * row and component 6 values: reducer, mode, scale (float bits), naive,
 * reduced, fresh.  Returns the number of rows. */
FG_API int fg_reducer_rows(fg_reducer *r, int rows, int32_t *out);
/* getMode (src/entities/additions_reducer.cuh:185-204). */
FG_API int fg_reducer_mode_name(int mode, float scale, char *buf, int cap);
/* One component of the best reducer (reducers[count] of the reference) or
 * of reducer `reducer` after the last pass, serialised as int32: count real
 * fresh naive mode scale_bits additions, 2 * fresh fresh-variable operands,
 * then per expression its size and entries (+-(index + 1)).  Returns the
 * number of values written. */
FG_API int64_t fg_reducer_best_part(fg_reducer *r, int component, int32_t *out, int64_t cap);
FG_API int64_t fg_reducer_part(fg_reducer *r, int reducer, int component, int32_t *out, int64_t cap);
/* Words consumed per random stream: [count][4] (components U V W, shared decisions). */
FG_API int fg_reducer_words(fg_reducer *r, uint64_t *out);

/* Device properties used for roofline reporting. */
FG_API int fg_device_info(int *cu_count, int *clock_khz, int *lds_per_block, char *name, int name_cap);
/* The device's 32-bit integer VALU rate, measured: `blocks` workgroups of
 * 256 threads whose lanes each issue trips x 256 independent v_xor_b32;
 * writes the lane operations per second of the fastest of three timed
 * launches (and its kernel time in ms when kernel_ms is not null).  No
 * reference counterpart: it checks on the device at hand the peak the walk's
 * roofline fraction is quoted against (CUs x clock x 128 lanes on gfx950,
 * whose SIMDs execute two wave64 VALU instructions per 4 cycles from two
 * waves). */
FG_API int fg_valu_probe(int blocks, int trips, double *lane_ops_per_s, double *kernel_ms);

#ifdef __cplusplus
}
#endif
#endif

// Host runtime behind include/fg.h: owns the device walker pool of one
// context, launches the walk and Brent-check kernels on the context's
// stream and moves walker state between host records and the
// structure-of-arrays device layout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "fg.h"
#include "kernels.h"
#include "runtime_internal.h"
#include "scheme_host.h"

namespace {

thread_local std::string g_err;

int fail(const std::string &msg, int code = -1) {
    g_err = msg;
    return code;
}

#define HIPCHK(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) return fail(std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
hipError_t dalloc(T **p, size_t n) {
    return hipMalloc(reinterpret_cast<void **>(p), std::max<size_t>(n, 1) * sizeof(T));
}

uint32_t threshold(double p, int max_iterations) {
    // P(u * maxIterations < p), u uniform in (0, 1]  ->  x < thr, x uniform 32-bit
    if (p <= 0 || max_iterations <= 0) return 0;
    const double x = p / (double)max_iterations * 4294967296.0;
    if (x >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

constexpr uint32_t kThrBestEq = 42949673u;  // 0.01 * 2^32 (flip_graph.cu:545)

// Device word of one factor: the Z2 mask, or for the integer ring the
// magnitudes in bits 0-15 and the signs in bits 16-31.
uint32_t pack_word(int ring, uint64_t v, uint64_t s) {
    return ring ? (uint32_t)((v & 0xFFFFu) | ((s & v & 0xFFFFu) << 16)) : (uint32_t)v;
}
uint64_t word_values(int ring, uint32_t w) { return ring ? (w & 0xFFFFu) : w; }
uint64_t word_signs(int ring, uint32_t w) { return ring ? (w >> 16) : 0; }
int max_coefficients(int ring) { return ring ? 16 : 32; }

}  // namespace

struct fg_ctx {
    fg_config cfg{};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int W = 0;
    int maxm = 0;
    bool ready = false;
    float last_ms = 0.f;
    size_t lds_limit = 0;
    size_t lds_per_cu = 0;

    uint32_t *cur = nullptr, *cur_s = nullptr, *best = nullptr, *best_s = nullptr;
    int32_t *m = nullptr, *best_m = nullptr, *best_rank = nullptr, *flips_count = nullptr, *aux_idx = nullptr;
    int32_t *fmt = nullptr, *bfmt = nullptr, *err = nullptr, *ok = nullptr;
    uint64_t *it = nullptr, *aux = nullptr, *total_flips = nullptr, *sum_m = nullptr;
    int32_t *item_begin = nullptr, *item_end = nullptr, *queue = nullptr;
    int32_t *cx = nullptr, *dst = nullptr;
    int32_t *forced = nullptr;  // [W][4] proposed first flips (learned walks)
    // set only for the duration of an fg_walk_trace launch
    int32_t *trace = nullptr, *trace_n = nullptr;
    int trace_cap = 0;
    int nitems = 0;
    bool minimize = false;  // pool set up by fg_init_minimizer
    bool flips = false;     // pool set up by fg_init_flips
    int variant = 0;        // walk kernel variant (kernels.h: kVarZ2 / kVarZT / kVarZ2h / kVarZ2m / kVarZ2s)
    bool bfmt_same = true;  // every walker's best scheme has the walker's current format
    int nn_max = 0;         // widest factor of the pool's formats
    bool masks_fit = false; // the class-mask layout can hold the pool
    bool meta = false;      // sandwiching / basis edges configured
    std::vector<int32_t> trials;  // last fg_autotune: (variant, grid, flips per second / 1e6) triples

    // trajectory log: ring of the last `horizon` launches, labelled rows
    struct Log {
        bool on = false;
        int every = 1, horizon = 10, nslot = 0;
        int window = 0;  // iterations of the logging launch whose ranks label its flip (0: whole launches)
        long long max_batches = 0, launches = 0, labelled = 0, positives = 0;
        uint32_t *ring_words = nullptr, *words = nullptr;
        int32_t *ring_meta = nullptr, *ring_minm = nullptr, *m = nullptr, *pairs = nullptr;
        float *y = nullptr;
        std::vector<int64_t> valid;   // labelled rows
        std::vector<int32_t> vpairs;  // their pairs (3 per row) and labels, on the host
        std::vector<float> vy;
        void release() {
            void *ptrs[] = {ring_words, words, ring_meta, ring_minm, m, pairs, y};
            for (void *p : ptrs)
                if (p) (void)hipFree(p);
            ring_words = words = nullptr;
            ring_meta = ring_minm = m = pairs = nullptr;
            y = nullptr;
            valid.clear();
            vpairs.clear();
            vy.clear();
            on = false;
            window = 0;
            launches = labelled = positives = 0;
        }
    } log;
    int cus = 0;
    int groups_per_cu = 0;
    int grid = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_iterations = -1;  // launch the captured graph replays: iterations, grid and kernel variant
    int graph_grid = -1;
    int graph_variant = -1;

    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        graph_iterations = -1;
        graph_grid = -1;
        graph_variant = -1;
    }
    void release() {
        drop_graph();
        void *ptrs[] = {cur, cur_s, best, best_s, m, best_m, best_rank, flips_count, aux_idx, fmt, bfmt,
                        err, ok, it, aux, total_flips, sum_m, item_begin, item_end, queue, cx, dst, forced};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
        cur = cur_s = best = best_s = nullptr;
        m = best_m = best_rank = flips_count = aux_idx = fmt = bfmt = err = ok = nullptr;
        it = aux = total_flips = sum_m = nullptr;
        item_begin = item_end = queue = cx = dst = forced = nullptr;
        log.release();
        nitems = 0;
        minimize = false;
        flips = false;
        ready = false;
    }
};

extern "C" const char *fg_last_error(void) { return g_err.c_str(); }

namespace fgrt {
// Error text for the other translation units of the library.
void set_error(const std::string &msg) { g_err = msg; }
}  // namespace fgrt

namespace {
// Walks a count-prefixed blob, calling fn(q, scheme) for each scheme.
template <class Fn>
int for_each_scheme(const char *what, int ring, const int32_t *blob, int64_t len, Fn fn) {
    if (!blob || len < 1) return fail(std::string(what) + ": bad arguments");
    if (ring != FG_RING_Z2 && ring != FG_RING_ZT) return fail(std::string(what) + ": unknown ring");
    const int count = blob[0];
    if (count < 0) return fail(std::string(what) + ": negative count");
    int64_t pos = 1;
    for (int q = 0; q < count; ++q) {
        fg::Scheme s;
        std::string err;
        if (!fg::scheme_from_blob(ring, blob, len, &pos, &s, &err)) return fail(std::string(what) + ": " + err);
        fn(q, s);
    }
    return count;
}
}  // namespace

extern "C" int fg_verify_host(int ring, const int32_t *blob, int64_t len, int32_t *out_ok) {
    if (!out_ok) return fail("fg_verify_host: null output");
    int good = 0;
    const int rc = for_each_scheme("fg_verify_host", ring, blob, len, [&](int q, const fg::Scheme &s) {
        out_ok[q] = fg::validate(s) ? 1 : 0;
        good += out_ok[q];
    });
    return rc < 0 ? rc : good;
}

extern "C" int fg_complexity_host(int ring, const int32_t *blob, int64_t len, int32_t *out) {
    if (!out) return fail("fg_complexity_host: null output");
    return for_each_scheme("fg_complexity_host", ring, blob, len,
                           [&](int q, const fg::Scheme &s) { out[q] = fg::complexity(s); });
}

extern "C" int fg_device_info(int *cu_count, int *clock_khz, int *lds_per_block, char *name, int name_cap) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (clock_khz) *clock_khz = prop.clockRate;
    if (lds_per_block) *lds_per_block = (int)prop.sharedMemPerBlock;
    if (name && name_cap > 0) {
        strncpy(name, prop.gcnArchName, (size_t)name_cap - 1);
        name[name_cap - 1] = 0;
    }
    return 0;
}

extern "C" int fg_create(const fg_config *cfg, fg_ctx **out) {
    if (!cfg || !out) return fail("fg_create: null argument");
    *out = nullptr;
    if (cfg->ring != FG_RING_Z2 && cfg->ring != FG_RING_ZT) return fail("fg_create: unknown ring");
    if (cfg->walkers <= 0) return fail("fg_create: walkers must be positive");
    if (cfg->max_iterations <= 0) return fail("fg_create: max_iterations must be positive");
    if (cfg->plus_iterations < 0) return fail("fg_create: plus_iterations must be >= 0");
    if (cfg->word_bits != 0 && cfg->word_bits != 16 && cfg->word_bits != 32) return fail("fg_create: word_bits must be 0, 16 or 32");
    if (cfg->lookahead_period < 0 || cfg->lookahead_candidates < 0 ||
        (cfg->lookahead_candidates > 0 && (cfg->lookahead_period < 1 || cfg->ring != FG_RING_Z2)))
        return fail("fg_create: lookahead proposals need a positive period and a Z2 pool");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(std::string("fg_create: this build targets gfx950 (MI355X), found ") + prop.gcnArchName);
    fg_ctx *c = new fg_ctx();
    c->cfg = *cfg;
    if (c->cfg.zt_visit_cap <= 0) c->cfg.zt_visit_cap = 64;
    c->lds_limit = prop.sharedMemPerBlock;
    c->lds_per_cu = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor : prop.sharedMemPerBlock;
    c->cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return fail("fg_create: stream/event creation failed");
    }
    *out = c;
    return 0;
}

extern "C" int fg_destroy(fg_ctx *ctx) {
    if (!ctx) return 0;
    ctx->release();
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return 0;
}

namespace {

// Fill the pool from start schemes (walker w gets starts[w % count]) and
// set bestRanks as FlipGraph::updateRanks does (flip_graph.cu:249-277): the
// smallest rank per sorted format key, seeded with the naive rank of the
// configured format.  mode: the walk kernel mode the pool runs (0 walk,
// 1 complexity minimizer, 3 the additions reducer's scheme flips).
int init_pool(fg_ctx *ctx, const std::vector<fg::Scheme> &starts, const std::vector<int> &assign, int n1, int n2,
              int n3, int mode = 0) {
    if (starts.empty()) return fail("init: no schemes");
    const int W = ctx->cfg.walkers;
    int maxm = ctx->cfg.max_rank;
    const int ring = ctx->cfg.ring;
    for (const auto &s : starts) {
        maxm = std::max(maxm, std::max(s.m, s.naive_rank()));
        if (s.ring != ring) return fail("init: ring mismatch");
        for (int c = 0; c < 3; ++c)
            if (s.nn[c] > max_coefficients(ring))
                return fail("init: factors with more than " + std::to_string(max_coefficients(ring)) +
                            " coefficients are not supported on the GPU path for this ring");
    }
    if (maxm > 255) return fail("init: more than 255 terms per scheme are not supported on the GPU path");
    // Z2 factors of up to 16 coefficients can live in 16-bit LDS words.  They
    // cost a few percent per iteration, so they are used only when they cut
    // the number of rounds in which the chip works through all walker blocks
    // (4x4x4 at 131072 walkers: 4 rounds -> 3, +11% flips/s; at 65536: 2 and
    // 2, so 32-bit words stay).
    int nn_max = 0;
    for (const auto &s : starts) nn_max = std::max(nn_max, std::max(s.nn[0], std::max(s.nn[1], s.nn[2])));
    int variant = ring ? fgk::kVarZT : fgk::kVarZ2;
    const bool meta = ctx->cfg.sandwiching_probability > 0 || ctx->cfg.basis_probability > 0;
    if (meta && !ring && ctx->cfg.word_bits == 16) return fail("init: sandwiching and basis-change edges need 32-bit factor words");
    // class masks: Z2 schemes of at most 32 terms with factors of up to 16
    // coefficients, unless 32-bit words, partner lists or the meta edges are asked for
    const bool masks_fit = !ring && nn_max <= 16 && maxm <= fgk::kMaskMaxTerms && !meta && ctx->cfg.word_bits != 32;
    if (ctx->cfg.partner_layout < 0 || ctx->cfg.partner_layout > 3) return fail("init: partner_layout must be 0, 1, 2 or 3");
    if (ctx->cfg.partner_layout >= 2 && !masks_fit)
        return fail("init: class masks need Z2 schemes of at most 32 terms, factors of up to 16 coefficients, "
                    "16-bit words and no sandwiching or basis edges");
    if (ctx->cfg.partner_layout == 3) {
        variant = fgk::kVarZ2s;
    } else if (ctx->cfg.partner_layout == 2 || (ctx->cfg.partner_layout == 0 && masks_fit && ctx->cfg.word_bits == 0)) {
        variant = fgk::kVarZ2m;
        if (ctx->cfg.partner_layout == 0) {
            // One lane per walker issues about two thirds of the instructions
            // per walker-iteration that a lane pair does, but a SIMD then
            // runs one wave where lane pairs run two.  It wins once the pool
            // gives every SIMD its wave (3x3x3 at 65536 walkers: 24.7 -> 21.8
            // ms per 10000-iteration launch); smaller pools keep lane pairs,
            // whose waves reach more SIMDs.
            int gs = 0;
            HIPCHK(fgk::prepare_walk(fgk::kVarZ2s, maxm, mode == 1 ? 1 : 0, &gs));
            const long solo_waves = (W + fgk::kSoloWalkersPerGroup - 1) / fgk::kSoloWalkersPerGroup;
            if (gs > 0 && solo_waves >= (long)gs * ctx->cus && (mode == 0 || mode == 1)) variant = fgk::kVarZ2s;
        }
    } else if (!ring && ctx->cfg.word_bits == 16) {
        if (nn_max > 16) return fail("init: 16-bit factor words hold factors of up to 16 coefficients");
        variant = fgk::kVarZ2h;
    } else if (!ring && nn_max <= 16 && ctx->cfg.word_bits != 32 && !meta) {
        // workgroups resident per CU as the occupancy query gives them
        // (LDS and registers), for either word size
        const long blocks = (W + fgk::kWalkersPerGroup - 1) / fgk::kWalkersPerGroup;
        int g32 = 0, g16 = 0;
        HIPCHK(fgk::prepare_walk(fgk::kVarZ2, maxm, 0, &g32));
        HIPCHK(fgk::prepare_walk(fgk::kVarZ2h, maxm, 0, &g16));
        const long fit32 = (long)g32 * ctx->cus, fit16 = (long)g16 * ctx->cus;
        const long rounds32 = fit32 > 0 ? (blocks + fit32 - 1) / fit32 : LONG_MAX;
        const long rounds16 = fit16 > 0 ? (blocks + fit16 - 1) / fit16 : LONG_MAX;
        if (rounds16 < rounds32) variant = fgk::kVarZ2h;
    }
    const size_t lds = fgk::walk_lds_bytes(variant, maxm);
    if (lds > ctx->lds_limit)
        return fail("init: " + std::to_string(maxm) + " terms per walker exceed the LDS of one workgroup");

    ctx->release();
    ctx->variant = variant;
    ctx->nn_max = nn_max;
    ctx->masks_fit = masks_fit;
    ctx->meta = meta;
    ctx->trials.clear();
    ctx->W = W;
    ctx->maxm = maxm;
    const size_t plane = (size_t)3 * maxm * W;
    std::vector<uint32_t> words(plane, 0), best_words(plane, 0);
    std::vector<int32_t> m(W), best_m(W), best_rank(W), zeros(W, 0), aidx(W, 4), fmt((size_t)4 * W, 0);
    std::vector<uint64_t> z64(W, 0);

    std::map<std::string, int> best_by_key;
    best_by_key[fg::format_key(n1, n2, n3)] = n1 * n2 * n3;
    for (int w = 0; w < W; ++w) {
        const fg::Scheme &s = starts[assign[w]];
        m[w] = s.m;
        best_m[w] = s.m;
        fmt[4 * w] = s.n[0];
        fmt[4 * w + 1] = s.n[1];
        fmt[4 * w + 2] = s.n[2];
        for (int c = 0; c < 3; ++c)
            for (int t = 0; t < s.m; ++t) {
                const uint32_t x = pack_word(ring, s.v[c][t], s.s[c][t]);
                words[((size_t)c * maxm + t) * W + w] = x;
                best_words[((size_t)w * 3 + c) * maxm + t] = x;
            }
        const std::string key = fg::format_key(s.n[0], s.n[1], s.n[2]);
        auto f = best_by_key.find(key);
        if (f == best_by_key.end() || s.m < f->second) best_by_key[key] = s.m;
    }
    for (int w = 0; w < W; ++w) {
        const fg::Scheme &s = starts[assign[w]];
        best_rank[w] = best_by_key[fg::format_key(s.n[0], s.n[1], s.n[2])];
    }

    // work items: blocks of up to kWalkersPerGroup walkers of one format,
    // queued by decreasing naive rank (largest formats first)
    struct Item {
        int begin, end, cost;
    };
    std::vector<Item> items;
    for (int w = 0; w < W;) {
        const fg::Scheme &s = starts[assign[w]];
        int e = w + 1;
        while (e < W && e - w < fgk::kWalkersPerGroup) {
            const fg::Scheme &t = starts[assign[e]];
            if (t.n[0] != s.n[0] || t.n[1] != s.n[1] || t.n[2] != s.n[2]) break;
            ++e;
        }
        items.push_back({w, e, s.naive_rank()});
        w = e;
    }
    std::stable_sort(items.begin(), items.end(), [](const Item &a, const Item &b) { return a.cost > b.cost; });
    std::vector<int32_t> ib(items.size()), ie(items.size());
    for (size_t q = 0; q < items.size(); ++q) {
        ib[q] = items[q].begin;
        ie[q] = items[q].end;
    }

    HIPCHK(dalloc(&ctx->cur, plane));
    HIPCHK(dalloc(&ctx->best, plane));
    HIPCHK(dalloc(&ctx->m, W));
    HIPCHK(dalloc(&ctx->best_m, W));
    HIPCHK(dalloc(&ctx->best_rank, W));
    HIPCHK(dalloc(&ctx->flips_count, W));
    HIPCHK(dalloc(&ctx->aux_idx, W));
    HIPCHK(dalloc(&ctx->fmt, (size_t)4 * W));
    HIPCHK(dalloc(&ctx->bfmt, (size_t)4 * W));
    HIPCHK(dalloc(&ctx->err, 1));
    HIPCHK(dalloc(&ctx->ok, W));
    HIPCHK(dalloc(&ctx->it, W));
    HIPCHK(dalloc(&ctx->aux, W));
    HIPCHK(dalloc(&ctx->total_flips, W));
    HIPCHK(dalloc(&ctx->sum_m, W));
    HIPCHK(dalloc(&ctx->item_begin, ib.size()));
    HIPCHK(dalloc(&ctx->item_end, ie.size()));
    HIPCHK(dalloc(&ctx->queue, 1));
    HIPCHK(dalloc(&ctx->cx, W));
    HIPCHK(dalloc(&ctx->dst, W));
    HIPCHK(dalloc(&ctx->forced, (size_t)4 * W));
    HIPCHK(fgk::launch_fill(ctx->forced, -1, 4 * W, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->minimize = mode == 1;
    ctx->flips = mode == 3;
    ctx->nitems = (int)ib.size();
    HIPCHK(hipMemcpy(ctx->item_begin, ib.data(), ib.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->item_end, ie.data(), ie.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(fgk::prepare_walk(variant, maxm, mode, &ctx->groups_per_cu));
    if (mode == 0 && ring == FG_RING_Z2) {
        int g2 = 0;
        HIPCHK(fgk::prepare_walk(variant, maxm, 2, &g2));
        HIPCHK(fgk::prepare_walk(variant, maxm, 6, &g2));
        if (variant == fgk::kVarZ2) HIPCHK(fgk::prepare_walk(variant, maxm, 4, &g2));
        if (variant == fgk::kVarZ2) HIPCHK(fgk::prepare_walk(variant, maxm, 5, &g2));
    }
    if (mode == 0 && ring == FG_RING_ZT) {
        int g2 = 0;
        HIPCHK(fgk::prepare_walk(variant, maxm, 4, &g2));
        HIPCHK(fgk::prepare_walk(variant, maxm, 5, &g2));
    }
    if (mode == 0 && ring == FG_RING_Z2 && ctx->cfg.lookahead_candidates > 0) {
        int g6 = 0;  // the lookahead kernel is the one the pool launches
        HIPCHK(fgk::prepare_walk(variant, maxm, 6, &g6));
        ctx->groups_per_cu = std::min(ctx->groups_per_cu, g6);
    }
    if (ctx->groups_per_cu < 1) return fail("init: the walk kernel does not fit on a CU with this scheme size");
    ctx->grid = ctx->cfg.grid > 0 ? ctx->cfg.grid : ctx->groups_per_cu * ctx->cus;
    ctx->grid = std::min(ctx->grid, ctx->nitems);
    HIPCHK(hipMemcpy(ctx->cur, words.data(), plane * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->best, best_words.data(), plane * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->m, m.data(), W * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->best_m, best_m.data(), W * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->best_rank, best_rank.data(), W * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->flips_count, zeros.data(), W * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->aux_idx, aidx.data(), W * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->fmt, fmt.data(), (size_t)16 * W, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->bfmt, fmt.data(), (size_t)16 * W, hipMemcpyHostToDevice));
    ctx->bfmt_same = true;
    HIPCHK(hipMemcpy(ctx->it, z64.data(), W * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->aux, z64.data(), W * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->total_flips, z64.data(), W * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->sum_m, z64.data(), W * 8, hipMemcpyHostToDevice));
    ctx->ready = true;
    return 0;
}

int verify_planes(fg_ctx *ctx, int ring, const uint32_t *words, const int32_t *m, const int32_t *fmt, int count,
                  int maxm, int32_t *ok_dev) {
    fgk::VerifyArgs v{};
    v.words = words;
    v.signs = nullptr;
    v.m = m;
    v.fmt = fmt;
    v.count = count;
    v.max_m = maxm;
    v.ok = ok_dev;
    HIPCHK(fgk::launch_fill(ok_dev, 1, count, ctx->stream));
    HIPCHK(fgk::launch_verify(ring, v, ctx->stream));
    return 0;
}

}  // namespace

extern "C" int fg_init_naive(fg_ctx *ctx, int n1, int n2, int n3) {
    if (!ctx) return fail("fg_init_naive: null context");
    if (n1 < 1 || n2 < 1 || n3 < 1) return fail("fg_init_naive: sizes must be positive");
    if (n1 * n2 > 64 || n2 * n3 > 64 || n3 * n1 > 64) return fail("fg_init_naive: a factor exceeds 64 coefficients");
    std::vector<fg::Scheme> starts{fg::naive_scheme(ctx->cfg.ring, n1, n2, n3)};
    return init_pool(ctx, starts, std::vector<int>(ctx->cfg.walkers, 0), n1, n2, n3);
}

extern "C" int fg_init_from_schemes(fg_ctx *ctx, const int32_t *blob, int64_t len) {
    if (!ctx || !blob || len < 1) return fail("fg_init_from_schemes: bad arguments");
    const int count = blob[0];
    if (count < 1) return fail("fg_init_from_schemes: scheme count must be positive");
    std::vector<fg::Scheme> starts;
    int64_t pos = 1;
    for (int q = 0; q < count && q < ctx->cfg.walkers; ++q) {
        fg::Scheme s;
        std::string err;
        if (!fg::scheme_from_blob(ctx->cfg.ring, blob, len, &pos, &s, &err))
            return fail("fg_init_from_schemes: scheme " + std::to_string(q) + ": " + err);
        while (fg::try_reduce(&s)) {
        }
        if (!fg::validate(s)) return fail("fg_init_from_schemes: scheme " + std::to_string(q) + " fails the Brent check", -2);
        starts.push_back(std::move(s));
    }
    const fg::Scheme &s0 = starts[0];
    std::vector<int> assign(ctx->cfg.walkers);
    for (int w = 0; w < ctx->cfg.walkers; ++w) assign[w] = w % (int)starts.size();  // initializeCopyKernel
    return init_pool(ctx, starts, assign, s0.n[0], s0.n[1], s0.n[2]);
}

extern "C" int fg_init_minimizer(fg_ctx *ctx, const int32_t *blob, int64_t len) {
    if (!ctx || !blob || len < 1) return fail("fg_init_minimizer: bad arguments");
    const int count = blob[0];
    if (count < 1) return fail("fg_init_minimizer: scheme count must be positive");
    std::vector<fg::Scheme> starts;
    int64_t pos = 1;
    for (int q = 0; q < count; ++q) {
        fg::Scheme s;
        std::string err;
        if (!fg::scheme_from_blob(ctx->cfg.ring, blob, len, &pos, &s, &err))
            return fail("fg_init_minimizer: scheme " + std::to_string(q) + ": " + err);
        if (!fg::validate(s)) return fail("fg_init_minimizer: scheme " + std::to_string(q) + " fails the Brent check", -2);
        if (!starts.empty() && (s.n[0] != starts[0].n[0] || s.n[1] != starts[0].n[1] || s.n[2] != starts[0].n[2]))
            return fail("fg_init_minimizer: schemes must share one format");
        starts.push_back(std::move(s));
    }
    // initialize() + initializeKernel (complexity_minimizer.cu:67-77, 172-188):
    // walker w starts from scheme w % count, every best complexity from the
    // lowest complexity among the inputs
    int best_c = INT_MAX;
    for (const auto &s : starts) best_c = std::min(best_c, fg::complexity(s));
    std::vector<int> assign(ctx->cfg.walkers);
    for (int w = 0; w < ctx->cfg.walkers; ++w) assign[w] = w % (int)starts.size();
    const fg::Scheme &s0 = starts[0];
    const int rc = init_pool(ctx, starts, assign, s0.n[0], s0.n[1], s0.n[2], 1);
    if (rc) return rc;
    std::vector<int32_t> bc(ctx->W, best_c);
    HIPCHK(hipMemcpy(ctx->best_rank, bc.data(), bc.size() * 4, hipMemcpyHostToDevice));
    return 0;
}

namespace {
int walk_launch(fg_ctx *ctx, int iterations, bool use_forced, double resize = 0.0);
}

extern "C" int fg_init_flips(fg_ctx *ctx, const int32_t *blob, int64_t len) {
    if (!ctx || !blob || len < 1) return fail("fg_init_flips: bad arguments");
    if (ctx->cfg.ring != FG_RING_ZT) return fail("fg_init_flips: scheme flips are built for the integer ring");
    if (blob[0] != 1) return fail("fg_init_flips: exactly one scheme expected");
    fg::Scheme s;
    std::string err;
    int64_t pos = 1;
    if (!fg::scheme_from_blob(ctx->cfg.ring, blob, len, &pos, &s, &err)) return fail("fg_init_flips: " + err);
    if (!fg::validate(s)) return fail("fg_init_flips: the scheme fails the Brent check", -2);
    std::vector<fg::Scheme> starts{s};
    return init_pool(ctx, starts, std::vector<int>(ctx->cfg.walkers, 0), s.n[0], s.n[1], s.n[2], 3);
}

extern "C" int fg_flip_schemes(fg_ctx *ctx, int max_flips) {
    if (!ctx || !ctx->ready) return fail("fg_flip_schemes: context not initialised");
    if (!ctx->flips) return fail("fg_flip_schemes: the pool was not set up by fg_init_flips");
    if (max_flips < 1) return fail("fg_flip_schemes: max_flips must be positive");
    // schemes[0].copyTo(schemes[idx]) (:456): walker 0's scheme never changes
    // and is its best (start) scheme, so the copy is a broadcast of that best
    if (ctx->W > 1) {
        std::vector<int32_t> dst(ctx->W - 1);
        for (int w = 1; w < ctx->W; ++w) dst[w - 1] = w;
        const int rc = fg_broadcast_best(ctx, 0, dst.data(), (int)dst.size());
        if (rc) return rc;
    }
    return walk_launch(ctx, max_flips, false);
}

extern "C" int fg_minimize(fg_ctx *ctx) {
    if (!ctx || !ctx->ready) return fail("fg_minimize: context not initialised");
    if (!ctx->minimize) return fail("fg_minimize: the pool was not set up by fg_init_minimizer");
    return fg_walk(ctx, 0);
}

extern "C" int fg_complexities(fg_ctx *ctx, int32_t *best, int32_t *cur) {
    if (!ctx || !ctx->ready) return fail("fg_complexities: context not initialised");
    if (best && !ctx->minimize) return fail("fg_complexities: best complexities are kept by minimizer pools only");
    if (best) HIPCHK(hipMemcpy(best, ctx->best_rank, (size_t)ctx->W * 4, hipMemcpyDeviceToHost));
    if (cur) HIPCHK(hipMemcpy(cur, ctx->cx, (size_t)ctx->W * 4, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int fg_broadcast_best(fg_ctx *ctx, int src, const int32_t *dst, int count) {
    if (!ctx || !ctx->ready) return fail("fg_broadcast_best: context not initialised");
    if (src < 0 || src >= ctx->W) return fail("fg_broadcast_best: source walker out of range");
    if (count < 0 || count > ctx->W || (count > 0 && !dst)) return fail("fg_broadcast_best: bad destination list");
    for (int q = 0; q < count; ++q)
        if (dst[q] < 0 || dst[q] >= ctx->W) return fail("fg_broadcast_best: destination walker out of range");
    if (count == 0) return 0;
    HIPCHK(hipMemcpyAsync(ctx->dst, dst, (size_t)count * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(fgk::launch_broadcast_best(ctx->cur, ctx->m, ctx->best, ctx->best_m, ctx->W, ctx->maxm, src, ctx->dst, count,
                                      ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

extern "C" int fg_init_batched(fg_ctx *ctx, int nformats, const int32_t *formats, const int32_t *weights) {
    if (!ctx || nformats < 1 || !formats || !weights) return fail("fg_init_batched: bad arguments");
    const int W = ctx->cfg.walkers;
    long long total = 0;
    for (int f = 0; f < nformats; ++f) {
        if (weights[f] < 0) return fail("fg_init_batched: negative weight");
        total += weights[f];
        const int a = formats[3 * f], b = formats[3 * f + 1], c = formats[3 * f + 2];
        if (a < 1 || b < 1 || c < 1 || a * b > 64 || b * c > 64 || c * a > 64)
            return fail("fg_init_batched: bad format " + std::to_string(f));
    }
    if (total <= 0) return fail("fg_init_batched: weights sum to zero");
    std::vector<long long> count(nformats);
    std::vector<std::pair<double, int>> frac;
    long long given = 0;
    for (int f = 0; f < nformats; ++f) {
        const double exact = (double)W * weights[f] / (double)total;
        count[f] = (long long)exact;
        given += count[f];
        frac.push_back({exact - (double)count[f], f});
    }
    std::stable_sort(frac.begin(), frac.end(), [](const std::pair<double, int> &a, const std::pair<double, int> &b) {
        return a.first > b.first;
    });
    for (size_t q = 0; given < W; ++q, ++given) ++count[frac[q % frac.size()].second];
    std::vector<fg::Scheme> starts;
    std::vector<int> assign;
    for (int f = 0; f < nformats; ++f) {
        starts.push_back(fg::naive_scheme(ctx->cfg.ring, formats[3 * f], formats[3 * f + 1], formats[3 * f + 2]));
        for (long long q = 0; q < count[f]; ++q) assign.push_back(f);
    }
    return init_pool(ctx, starts, assign, formats[0], formats[1], formats[2]);
}


extern "C" int fg_walk(fg_ctx *ctx, int iterations) {
    if (ctx && ctx->flips) return fail("fg_walk: a scheme-flips pool runs fg_flip_schemes");
    return walk_launch(ctx, iterations, false);
}

extern "C" int fg_walk_trace(fg_ctx *ctx, int iterations, int cap, int32_t *ops, int32_t *counts) {
    if (!ctx || !ctx->ready) return fail("fg_walk_trace: context not initialised");
    if (cap < 0 || (cap > 0 && !ops) || !counts) return fail("fg_walk_trace: bad arguments");
    if (ctx->minimize || ctx->flips) return fail("fg_walk_trace: traces are recorded for plain walks");
    const size_t slots = (size_t)ctx->W * (size_t)std::max(cap, 1) * 8;
    int g = 0;
    if (fgk::prepare_walk(ctx->variant, ctx->maxm, 7, &g) != hipSuccess || g < 1)
        return fail("fg_walk_trace: the trace kernel does not fit the pool's layout");
    int32_t *d_ops = nullptr, *d_n = nullptr;
    HIPCHK(dalloc(&d_ops, slots));
    if (dalloc(&d_n, (size_t)ctx->W) != hipSuccess) {
        (void)hipFree(d_ops);
        return fail("fg_walk_trace: out of device memory");
    }
    ctx->trace = d_ops;
    ctx->trace_n = d_n;
    ctx->trace_cap = cap;
    const int rc = walk_launch(ctx, iterations, false);
    ctx->trace = ctx->trace_n = nullptr;
    ctx->trace_cap = 0;
    int out = rc;
    if (rc >= 0) {
        if ((cap > 0 && hipMemcpy(ops, d_ops, (size_t)ctx->W * cap * 8 * 4, hipMemcpyDeviceToHost) != hipSuccess) ||
            hipMemcpy(counts, d_n, (size_t)ctx->W * 4, hipMemcpyDeviceToHost) != hipSuccess)
            out = fail("fg_walk_trace: copy of the trace failed");
    }
    (void)hipFree(d_ops);
    (void)hipFree(d_n);
    return out;
}

extern "C" int fg_resize(fg_ctx *ctx, double probability) {
    if (!ctx || !ctx->ready) return fail("fg_resize: context not initialised");
    if (ctx->variant == fgk::kVarZ2h || fgk::mask_variant(ctx->variant) || ctx->minimize || ctx->flips)
        return fail("fg_resize: resize edges run on walk pools with 32-bit factor words (word_bits 32 for Z2)");
    if (!(probability > 0)) return 0;
    return walk_launch(ctx, 1, false, probability);
}

extern "C" int fg_snapshot_best(fg_ctx *ctx, const int32_t *walkers, int count) {
    if (!ctx || !ctx->ready) return fail("fg_snapshot_best: context not initialised");
    if (count < 0 || count > ctx->W || (count > 0 && !walkers)) return fail("fg_snapshot_best: bad walker list");
    for (int q = 0; q < count; ++q)
        if (walkers[q] < 0 || walkers[q] >= ctx->W) return fail("fg_snapshot_best: walker out of range");
    if (count == 0) return 0;
    HIPCHK(hipMemcpyAsync(ctx->dst, walkers, (size_t)count * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(fgk::launch_snapshot_best(ctx->cur, ctx->best, ctx->m, ctx->best_m, ctx->fmt, ctx->bfmt, ctx->W, ctx->maxm,
                                     ctx->dst, count, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

extern "C" int fg_set_best_ranks(fg_ctx *ctx, const int32_t *best_rank) {
    if (!ctx || !ctx->ready || !best_rank) return fail("fg_set_best_ranks: bad arguments");
    HIPCHK(hipMemcpy(ctx->best_rank, best_rank, (size_t)ctx->W * 4, hipMemcpyHostToDevice));
    return 0;
}

namespace fgrt {
int walk_forced(fg_ctx *ctx, int iterations) { return walk_launch(ctx, iterations, true); }

bool pool_view(fg_ctx *ctx, PoolView *v) {
    if (!ctx || !ctx->ready) return false;
    v->ring = ctx->cfg.ring;
    v->W = ctx->W;
    v->maxm = ctx->maxm;
    v->cur = ctx->cur;
    v->m = ctx->m;
    v->fmt = ctx->fmt;
    v->forced = ctx->forced;
    v->stream = ctx->stream;
    v->key0 = (uint32_t)ctx->cfg.seed;
    v->key1 = (uint32_t)(ctx->cfg.seed >> 32);
    return true;
}

bool log_view(fg_ctx *ctx, LogView *v) {
    if (!ctx || !ctx->log.on) return false;
    v->words = ctx->log.words;
    v->m = ctx->log.m;
    v->pairs = ctx->log.pairs;
    v->y = ctx->log.y;
    v->valid = &ctx->log.valid;
    v->vpairs = &ctx->log.vpairs;
    v->vy = &ctx->log.vy;
    v->maxm = ctx->maxm;
    return true;
}
}  // namespace fgrt

namespace {
int walk_launch(fg_ctx *ctx, int iterations, bool use_forced, double resize) {
    if (!ctx || !ctx->ready) return fail("fg_walk: context not initialised");
    const bool learn = resize <= 0 && !ctx->minimize && !ctx->flips && (use_forced || ctx->log.on);
    if (learn && ctx->cfg.ring != FG_RING_Z2) return fail("fg_walk: proposals and trajectory logs are built for Z2 walks");
    fgk::WalkArgs a{};
    a.args_bytes = (int32_t)sizeof(fgk::WalkArgs);
    a.cur = ctx->cur;
    a.best = ctx->best;
    a.m = ctx->m;
    a.best_m = ctx->best_m;
    a.best_rank = ctx->best_rank;
    a.flips_count = ctx->flips_count;
    a.aux_idx = ctx->aux_idx;
    a.it = ctx->it;
    a.aux = ctx->aux;
    a.total_flips = ctx->total_flips;
    a.sum_m = ctx->sum_m;
    a.fmt = ctx->fmt;
    a.walkers = ctx->W;
    a.max_m = ctx->maxm;
    a.key0 = (uint32_t)ctx->cfg.seed;
    a.key1 = (uint32_t)(ctx->cfg.seed >> 32);
    a.iterations = iterations > 0 ? iterations : ctx->cfg.max_iterations;
    a.max_iterations = iterations > 0 ? iterations : ctx->cfg.max_iterations;
    a.random_iterations = ctx->cfg.random_iterations;
    a.plus_iterations = ctx->cfg.plus_iterations;
    a.use_reduce = ctx->cfg.reduce_probability > 0;
    a.thr_reduce = threshold(ctx->cfg.reduce_probability, ctx->cfg.max_iterations);
    a.use_expand = ctx->cfg.expand_probability > 0;
    a.thr_expand = threshold(ctx->cfg.expand_probability, ctx->cfg.max_iterations);
    a.thr_best_eq = kThrBestEq;
    a.check_reduce = ctx->cfg.check_reduce;
    a.zt_visit_cap = ctx->cfg.zt_visit_cap;
    a.err = ctx->err;
    a.item_begin = ctx->item_begin;
    a.item_end = ctx->item_end;
    a.nitems = ctx->nitems;
    a.queue = ctx->queue;
    a.cx = ctx->cx;
    const bool meta = ctx->cfg.sandwiching_probability > 0 || ctx->cfg.basis_probability > 0;
    if (meta && learn) return fail("fg_walk: proposals and trajectory logs do not combine with basis-change edges");
    const bool look = ctx->cfg.lookahead_candidates > 0 && ctx->cfg.ring == FG_RING_Z2;
    if (look && (learn || meta)) return fail("fg_walk: lookahead proposals do not combine with learned proposals or basis edges");
    const bool trace = ctx->trace != nullptr;
    if (trace && (resize > 0 || ctx->minimize || ctx->flips || learn || meta || look))
        return fail("fg_walk_trace: traces are recorded for plain walks");
    const int mode = trace ? 7
                           : (resize > 0 ? 4 : (ctx->minimize ? 1 : (ctx->flips ? 3 : (learn ? 2 : (meta ? 5 : (look ? 6 : 0))))));
    a.trace = ctx->trace;
    a.trace_n = ctx->trace_n;
    a.trace_cap = ctx->trace_cap;
    a.look_period = ctx->cfg.lookahead_period;
    a.look_candidates = ctx->cfg.lookahead_candidates;
    a.use_sandwich = ctx->cfg.sandwiching_probability > 0;
    a.thr_sandwich = threshold(ctx->cfg.sandwiching_probability, ctx->cfg.max_iterations);
    a.use_basis = ctx->cfg.basis_probability > 0;
    a.thr_basis = threshold(ctx->cfg.basis_probability, ctx->cfg.max_iterations);
    // best-scheme formats: walk launches that cannot change a format skip
    // their copies while every walker's best has its current format
    a.bfmt = (ctx->bfmt_same && !(resize > 0)) ? nullptr : ctx->bfmt;
    if (resize > 0 && ctx->bfmt_same) {
        // a resize changes walkers' formats: from here on the copies write them
        ctx->bfmt_same = false;
        ctx->drop_graph();
    }
    if (resize > 0) {
        // resizeKernel: MAX_RANK and MAX_MATRIX_ELEMENTS become the pool's
        // term capacity and the coefficients a factor word holds
        a.thr_resize = resize >= 1.0 ? (1ull << 32) : (uint64_t)std::floor(resize * 4294967296.0);
        a.resize_max_rank = std::min(ctx->maxm, 350);
        a.resize_max_coef = ctx->cfg.ring == FG_RING_ZT ? 16 : 32;
    }
    const int slot = ctx->log.on ? (int)(ctx->log.launches % ctx->log.horizon) : 0;
    if (learn) {
        a.forced = use_forced ? ctx->forced : nullptr;
        if (ctx->log.on) {
            a.log_every = ctx->log.every;
            a.log_words = ctx->log.ring_words + (size_t)slot * ctx->log.nslot * 3 * ctx->maxm;
            a.log_meta = ctx->log.ring_meta + (size_t)slot * ctx->log.nslot * 4;
            a.minm = ctx->log.ring_minm + (size_t)slot * ctx->W;
            a.min_window = ctx->log.window;
        }
    }
    if (ctx->minimize) {
        // minimizeKernel: maxIterations flips without reductions or expansions
        a.iterations = a.max_iterations = ctx->cfg.max_iterations;
        a.random_iterations = 0;
        a.use_reduce = a.use_expand = 0;
        a.check_reduce = 0;
    }
    if (ctx->flips) {
        // flipSchemesKernel: up to `iterations` flips with reductions, nothing else
        a.random_iterations = 0;
        a.use_reduce = a.use_expand = 0;
        a.check_reduce = 1;
    }
    // one launch: error flag reset, queue reset + persistent walk (timed),
    // then the Brent check of every current scheme
    // (in graph mode the events bracket the whole graph instead)
    auto enqueue = [&](bool timed) -> int {
        HIPCHK(fgk::launch_fill(ctx->err, 0, 1, ctx->stream));
        if (timed) HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
        HIPCHK(fgk::launch_walk(ctx->variant, mode, a, ctx->grid, ctx->stream));
        if (timed) HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
        return verify_planes(ctx, ctx->cfg.ring, ctx->cur, ctx->m, ctx->fmt, ctx->W, ctx->maxm, ctx->ok);
    };
    if (ctx->cfg.use_graph && mode != 2 && mode != 4 && mode != 7) {
        if (ctx->graph_iterations != a.iterations || ctx->graph_grid != ctx->grid || ctx->graph_variant != ctx->variant) {
            ctx->drop_graph();
            HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue(false);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
            if (rc) return rc;
            if (ec != hipSuccess) return fail(std::string("fg_walk: graph capture failed: ") + hipGetErrorString(ec));
            ctx->graph = g;
            HIPCHK(hipGraphInstantiate(&ctx->graph_exec, ctx->graph, nullptr, nullptr, 0));
            ctx->graph_iterations = a.iterations;
            ctx->graph_grid = ctx->grid;
            ctx->graph_variant = ctx->variant;
        }
        HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
        HIPCHK(hipGraphLaunch(ctx->graph_exec, ctx->stream));
        HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
    } else {
        const int rc = enqueue(true);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1));
    int32_t err = 0;
    HIPCHK(hipMemcpy(&err, ctx->err, sizeof err, hipMemcpyDeviceToHost));
    if (err & fgk::kErrArgs) return fail("fg_walk: the walk kernel's argument layout differs from the host's (a stale build)", -3);
    if (err) return fail("fg_walk: kernel invariant violation, flags " + std::to_string(err), -3);
    std::vector<int32_t> ok(ctx->W);
    HIPCHK(hipMemcpy(ok.data(), ctx->ok, ok.size() * 4, hipMemcpyDeviceToHost));
    int bad = 0;
    for (int x : ok) bad += x == 0;
    if (ctx->log.on) {
        // rows logged `horizon` launches ago now have their whole window
        auto &L = ctx->log;
        const long long s = L.launches - L.horizon + 1;
        if (s >= 0 && s < L.max_batches) {
            const size_t base = (size_t)s * L.nslot;
            HIPCHK(fgk::launch_log_label(L.ring_words + (size_t)(s % L.horizon) * L.nslot * 3 * ctx->maxm,
                                         L.ring_meta + (size_t)(s % L.horizon) * L.nslot * 4, L.ring_minm, L.horizon,
                                         ctx->W, L.every, L.nslot, ctx->maxm, L.words + base * 3 * ctx->maxm,
                                         L.m + base, L.pairs + base * 3, L.y + base, ctx->stream));
            std::vector<float> y(L.nslot);
            std::vector<int32_t> pr((size_t)L.nslot * 3);
            HIPCHK(hipMemcpyAsync(y.data(), L.y + base, (size_t)L.nslot * 4, hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipMemcpyAsync(pr.data(), L.pairs + base * 3, pr.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
            for (int i = 0; i < L.nslot; ++i)
                if (y[i] >= 0.f) {
                    L.valid.push_back((int64_t)(base + i));
                    L.vpairs.insert(L.vpairs.end(), pr.begin() + 3 * i, pr.begin() + 3 * i + 3);
                    L.vy.push_back(y[i]);
                    L.positives += y[i] > 0.5f;
                }
            ++L.labelled;
        }
        ++L.launches;
    }
    return bad;
}
}  // namespace

extern "C" int fg_log_enable(fg_ctx *ctx, int64_t max_batches, int every, int horizon) {
    if (!ctx || !ctx->ready) return fail("fg_log_enable: context not initialised");
    if (ctx->cfg.ring != FG_RING_Z2 || ctx->minimize) return fail("fg_log_enable: trajectory logs are built for Z2 walks");
    if (max_batches < 1 || every < 1 || horizon < 1) return fail("fg_log_enable: bad arguments");
    auto &L = ctx->log;
    L.release();
    L.every = every;
    L.horizon = horizon;
    L.nslot = (ctx->W + every - 1) / every;
    L.max_batches = max_batches;
    const size_t rows = (size_t)max_batches * L.nslot, wpr = (size_t)3 * ctx->maxm;
    HIPCHK(dalloc(&L.ring_words, (size_t)horizon * L.nslot * wpr));
    HIPCHK(dalloc(&L.ring_meta, (size_t)horizon * L.nslot * 4));
    HIPCHK(dalloc(&L.ring_minm, (size_t)horizon * ctx->W));
    HIPCHK(dalloc(&L.words, rows * wpr));
    HIPCHK(dalloc(&L.m, rows));
    HIPCHK(dalloc(&L.pairs, rows * 3));
    HIPCHK(dalloc(&L.y, rows));
    HIPCHK(fgk::launch_fill(L.ring_meta, -1, (int)std::min<size_t>((size_t)horizon * L.nslot * 4, INT_MAX), ctx->stream));
    HIPCHK(fgk::launch_fill(L.ring_minm, INT_MAX, (int)std::min<size_t>((size_t)horizon * ctx->W, INT_MAX), ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    L.on = true;
    return 0;
}

extern "C" int fg_log_window(fg_ctx *ctx, int iterations) {
    if (!ctx || !ctx->log.on) return fail("fg_log_window: no trajectory log");
    if (iterations < 0) return fail("fg_log_window: iterations must be >= 0");
    if (iterations > 0 && ctx->log.horizon != 1)
        return fail("fg_log_window: an in-launch window needs a log horizon of one launch");
    ctx->log.window = iterations;
    return 0;
}

extern "C" int fg_log_info(fg_ctx *ctx, int64_t *rows, int64_t *positives, int64_t *batches) {
    if (!ctx || !ctx->log.on) return fail("fg_log_info: no trajectory log");
    if (rows) *rows = (int64_t)ctx->log.valid.size();
    if (positives) *positives = ctx->log.positives;
    if (batches) *batches = ctx->log.labelled;
    return 0;
}

extern "C" int fg_log_read(fg_ctx *ctx, int64_t begin, int64_t count, uint32_t *words, int32_t *m, int32_t *pairs,
                           float *labels) {
    if (!ctx || !ctx->log.on) return fail("fg_log_read: no trajectory log");
    const std::vector<int64_t> &v = ctx->log.valid;
    if (begin < 0 || count < 0 || begin + count > (int64_t)v.size()) return fail("fg_log_read: range out of bounds");
    const size_t wpr = (size_t)3 * ctx->maxm;
    // one copy per array for each run of consecutive log rows
    for (int64_t q = 0; q < count;) {
        const size_t r = (size_t)v[begin + q];
        int64_t n = 1;
        while (q + n < count && (size_t)v[begin + q + n] == r + (size_t)n) ++n;
        if (words)
            HIPCHK(hipMemcpy(words + q * wpr, ctx->log.words + r * wpr, (size_t)n * wpr * 4, hipMemcpyDeviceToHost));
        if (m) HIPCHK(hipMemcpy(m + q, ctx->log.m + r, (size_t)n * 4, hipMemcpyDeviceToHost));
        if (pairs) HIPCHK(hipMemcpy(pairs + q * 3, ctx->log.pairs + r * 3, (size_t)n * 12, hipMemcpyDeviceToHost));
        if (labels) HIPCHK(hipMemcpy(labels + q, ctx->log.y + r, (size_t)n * 4, hipMemcpyDeviceToHost));
        q += n;
    }
    return 0;
}

extern "C" int fg_set_forced(fg_ctx *ctx, const int32_t *forced) {
    if (!ctx || !ctx->ready || !forced) return fail("fg_set_forced: bad arguments");
    HIPCHK(hipMemcpy(ctx->forced, forced, (size_t)ctx->W * 16, hipMemcpyHostToDevice));
    return 0;
}

extern "C" int fg_get_forced(fg_ctx *ctx, int32_t *forced) {
    if (!ctx || !ctx->ready || !forced) return fail("fg_get_forced: bad arguments");
    HIPCHK(hipMemcpy(forced, ctx->forced, (size_t)ctx->W * 16, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int fg_walk_forced(fg_ctx *ctx, int iterations) { return fgrt::walk_forced(ctx, iterations); }

namespace {
// Makes `variant` the pool's walk kernel: sets the LDS limits of every mode
// the pool may launch with it and returns its resident workgroups per CU.
int use_variant(fg_ctx *ctx, int variant, int *groups_per_cu) {
    const int mode = ctx->minimize ? 1 : (ctx->flips ? 3 : 0);
    int g = 0;
    if (fgk::prepare_walk(variant, ctx->maxm, mode, &g) != hipSuccess || g < 1) return -1;
    if (mode == 0 && ctx->cfg.ring == FG_RING_Z2) {
        int g2 = 0;
        if (fgk::prepare_walk(variant, ctx->maxm, 2, &g2) != hipSuccess) return -1;
        if (fgk::prepare_walk(variant, ctx->maxm, 6, &g2) != hipSuccess) return -1;
        if (ctx->cfg.lookahead_candidates > 0) g = std::min(g, g2);
    }
    ctx->variant = variant;
    ctx->groups_per_cu = g;
    *groups_per_cu = g;
    return 0;
}
}  // namespace

extern "C" int fg_autotune(fg_ctx *ctx, int iterations, int32_t *grid) {
    if (!ctx || !ctx->ready) return fail("fg_autotune: context not initialised");
    if (iterations <= 0) return fail("fg_autotune: iterations must be positive");
    // layouts worth timing: every Z2 layout the pool fits when the config
    // leaves the choice open (the integer ring has one layout)
    std::vector<int> variants;
    const bool open = ctx->cfg.ring == FG_RING_Z2 && ctx->cfg.word_bits == 0 && ctx->cfg.partner_layout == 0 &&
                      !ctx->meta && !ctx->log.on;
    if (open) {
        if (ctx->masks_fit) {
            variants.push_back(fgk::kVarZ2s);
            variants.push_back(fgk::kVarZ2m);
        }
        if (ctx->nn_max <= 16) variants.push_back(fgk::kVarZ2h);
        variants.push_back(fgk::kVarZ2);
    } else {
        variants.push_back(ctx->variant);
    }
    ctx->trials.clear();
    int best_v = ctx->variant, best_g = ctx->grid, best_gpc = ctx->groups_per_cu;
    double best_rate = -1;
    for (int v : variants) {
        int gpc = 0;
        if (use_variant(ctx, v, &gpc)) continue;
        const int maxg = std::max(1, std::min(gpc * ctx->cus, ctx->nitems));
        for (int mult = 1;; ++mult) {
            const int g = std::min(mult * ctx->cus, maxg);
            ctx->grid = g;
            ctx->drop_graph();
            uint64_t f0 = 0, f1 = 0;
            if (fg_counters(ctx, &f0, nullptr, nullptr) < 0) return -1;
            const int bad = fg_walk(ctx, iterations);
            if (bad < 0) return bad;
            if (fg_counters(ctx, &f1, nullptr, nullptr) < 0) return -1;
            const double rate = (double)(f1 - f0) / std::max(1e-3, (double)ctx->last_ms) * 1e3;
            ctx->trials.insert(ctx->trials.end(), {v, g, (int32_t)std::min(rate / 1e6, 2e9)});
            if (rate > best_rate) {
                best_rate = rate;
                best_v = v;
                best_g = g;
                best_gpc = gpc;
            }
            if (g >= maxg) break;
        }
    }
    int gpc = 0;
    if (use_variant(ctx, best_v, &gpc)) return fail("fg_autotune: could not restore the chosen layout");
    ctx->groups_per_cu = best_gpc;
    ctx->grid = best_g;
    ctx->drop_graph();
    if (grid) *grid = best_g;
    return 0;
}

extern "C" int fg_layout_info(fg_ctx *ctx, int32_t *variant, int32_t *trials, int cap) {
    if (!ctx || !ctx->ready) return fail("fg_layout_info: context not initialised");
    if (variant) *variant = ctx->variant;
    const int n = (int)ctx->trials.size() / 3;
    if (trials)
        for (int q = 0; q < n && q < cap; ++q)
            for (int e = 0; e < 3; ++e) trials[3 * q + e] = ctx->trials[3 * q + e];
    return n;
}

extern "C" int fg_last_walk_ms(fg_ctx *ctx, float *ms) {
    if (!ctx || !ms) return fail("fg_last_walk_ms: null argument");
    *ms = ctx->last_ms;
    return 0;
}

extern "C" int fg_counters(fg_ctx *ctx, uint64_t *flips, uint64_t *sum_m, uint64_t *iterations) {
    if (!ctx || !ctx->ready) return fail("fg_counters: context not initialised");
    std::vector<uint64_t> a(ctx->W), b(ctx->W), c(ctx->W);
    HIPCHK(hipMemcpy(a.data(), ctx->total_flips, ctx->W * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(b.data(), ctx->sum_m, ctx->W * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(c.data(), ctx->it, ctx->W * 8, hipMemcpyDeviceToHost));
    uint64_t sa = 0, sb = 0, sc = 0;
    for (int w = 0; w < ctx->W; ++w) {
        sa += a[w];
        sb += b[w];
        sc += c[w];
    }
    if (flips) *flips = sa;
    if (sum_m) *sum_m = sb;
    if (iterations) *iterations = sc;
    return 0;
}

extern "C" int fg_summary(fg_ctx *ctx, int32_t *m, int32_t *best_m, int32_t *best_rank, int32_t *flips_count,
                          int32_t *n) {
    if (!ctx || !ctx->ready) return fail("fg_summary: context not initialised");
    const size_t bytes = (size_t)ctx->W * 4;
    if (n) {
        std::vector<int32_t> fmt((size_t)4 * ctx->W);
        HIPCHK(hipMemcpy(fmt.data(), ctx->fmt, fmt.size() * 4, hipMemcpyDeviceToHost));
        for (int w = 0; w < ctx->W; ++w)
            for (int c = 0; c < 3; ++c) n[3 * w + c] = fmt[4 * w + c];
    }
    if (m) HIPCHK(hipMemcpy(m, ctx->m, bytes, hipMemcpyDeviceToHost));
    if (best_m) HIPCHK(hipMemcpy(best_m, ctx->best_m, bytes, hipMemcpyDeviceToHost));
    if (best_rank) HIPCHK(hipMemcpy(best_rank, ctx->best_rank, bytes, hipMemcpyDeviceToHost));
    if (flips_count) HIPCHK(hipMemcpy(flips_count, ctx->flips_count, bytes, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int fg_shape(fg_ctx *ctx, int32_t *walkers, int32_t *n1, int32_t *n2, int32_t *n3) {
    if (!ctx || !ctx->ready) return fail("fg_shape: context not initialised");
    int32_t fmt[4];
    HIPCHK(hipMemcpy(fmt, ctx->fmt, 16, hipMemcpyDeviceToHost));
    if (walkers) *walkers = ctx->W;
    if (n1) *n1 = fmt[0];
    if (n2) *n2 = fmt[1];
    if (n3) *n3 = fmt[2];
    return 0;
}

extern "C" int fg_read_walkers(fg_ctx *ctx, int begin, int count, fg_walker_record *out) {
    if (!ctx || !ctx->ready || !out) return fail("fg_read_walkers: bad arguments");
    if (begin < 0 || count < 0 || begin + count > ctx->W) return fail("fg_read_walkers: range out of bounds");
    if (ctx->maxm > FG_REC_MAX_M) return fail("fg_read_walkers: scheme capacity exceeds the record");
    const int W = ctx->W, maxm = ctx->maxm;
    const size_t plane = (size_t)3 * maxm * W;
    std::vector<uint32_t> cur(plane), best(plane);
    std::vector<int32_t> m(W), bm(W), br(W), fc(W), ai(W), fmt((size_t)4 * W);
    std::vector<uint64_t> it(W), aux(W), tf(W), sm(W);
    HIPCHK(hipMemcpy(cur.data(), ctx->cur, plane * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(best.data(), ctx->best, plane * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(m.data(), ctx->m, W * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(bm.data(), ctx->best_m, W * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(br.data(), ctx->best_rank, W * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(fc.data(), ctx->flips_count, W * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ai.data(), ctx->aux_idx, W * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(fmt.data(), ctx->fmt, (size_t)16 * W, hipMemcpyDeviceToHost));
    std::vector<int32_t> bfmt((size_t)4 * W);
    HIPCHK(hipMemcpy(bfmt.data(), ctx->bfmt, (size_t)16 * W, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(it.data(), ctx->it, W * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(aux.data(), ctx->aux, W * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(tf.data(), ctx->total_flips, W * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(sm.data(), ctx->sum_m, W * 8, hipMemcpyDeviceToHost));
    for (int q = 0; q < count; ++q) {
        const int w = begin + q;
        fg_walker_record &r = out[q];
        memset(&r, 0, sizeof r);
        for (int c = 0; c < 3; ++c) r.n[c] = fmt[4 * w + c];
        for (int c = 0; c < 3; ++c) r.best_n[c] = bfmt[4 * w + c];
        r.m = m[w];
        r.best_m = bm[w];
        r.best_rank = br[w];
        r.flips_count = fc[w];
        r.aux_idx = ai[w];
        r.it = it[w];
        r.aux = aux[w];
        r.total_flips = tf[w];
        r.sum_m = sm[w];
        const int ring = ctx->cfg.ring;
        for (int c = 0; c < 3; ++c) {
            for (int t = 0; t < r.m; ++t) {
                const uint32_t x = cur[((size_t)c * maxm + t) * W + w];
                r.cur_v[c][t] = word_values(ring, x);
                r.cur_s[c][t] = word_signs(ring, x);
            }
            for (int t = 0; t < r.best_m; ++t) {
                const uint32_t x = best[((size_t)w * 3 + c) * maxm + t];
                r.best_v[c][t] = word_values(ring, x);
                r.best_s[c][t] = word_signs(ring, x);
            }
        }
    }
    return 0;
}

extern "C" int fg_write_walkers(fg_ctx *ctx, int begin, int count, const fg_walker_record *in) {
    if (!ctx || !ctx->ready || !in) return fail("fg_write_walkers: bad arguments");
    if (begin < 0 || count < 0 || begin + count > ctx->W) return fail("fg_write_walkers: range out of bounds");
    const int W = ctx->W, maxm = ctx->maxm;
    for (int q = 0; q < count; ++q) {
        const fg_walker_record &r = in[q];
        if (r.m < 0 || r.m > maxm || r.best_m < 0 || r.best_m > maxm)
            return fail("fg_write_walkers: record rank exceeds the pool capacity");
        for (int c = 0; c < 3; ++c) {
            const int *bn = r.best_n[0] > 0 ? r.best_n : r.n;
            if (r.n[c] < 1 || r.n[c] * r.n[(c + 1) % 3] > max_coefficients(ctx->cfg.ring) ||
                bn[c] * bn[(c + 1) % 3] > max_coefficients(ctx->cfg.ring))
                return fail("fg_write_walkers: factor wider than the GPU word of this ring");
        }
    }
    const size_t plane = (size_t)3 * maxm * W;
    std::vector<uint32_t> cur(plane), best(plane);
    std::vector<int32_t> m(W), bm(W), br(W), fc(W), ai(W), fmt((size_t)4 * W);
    std::vector<uint64_t> it(W), aux(W), tf(W), sm(W);
    HIPCHK(hipMemcpy(cur.data(), ctx->cur, plane * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(best.data(), ctx->best, plane * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(m.data(), ctx->m, W * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(bm.data(), ctx->best_m, W * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(br.data(), ctx->best_rank, W * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(fc.data(), ctx->flips_count, W * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ai.data(), ctx->aux_idx, W * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(fmt.data(), ctx->fmt, (size_t)16 * W, hipMemcpyDeviceToHost));
    std::vector<int32_t> bfmt((size_t)4 * W);
    HIPCHK(hipMemcpy(bfmt.data(), ctx->bfmt, (size_t)16 * W, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(it.data(), ctx->it, W * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(aux.data(), ctx->aux, W * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(tf.data(), ctx->total_flips, W * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(sm.data(), ctx->sum_m, W * 8, hipMemcpyDeviceToHost));
    for (int q = 0; q < count; ++q) {
        const int w = begin + q;
        const fg_walker_record &r = in[q];
        for (int c = 0; c < 3; ++c) fmt[4 * w + c] = r.n[c];
        for (int c = 0; c < 3; ++c) bfmt[4 * w + c] = r.best_n[0] > 0 ? r.best_n[c] : r.n[c];
        m[w] = r.m;
        bm[w] = r.best_m;
        br[w] = r.best_rank;
        fc[w] = r.flips_count;
        ai[w] = r.aux_idx;
        it[w] = r.it;
        aux[w] = r.aux;
        tf[w] = r.total_flips;
        sm[w] = r.sum_m;
        const int ring = ctx->cfg.ring;
        for (int c = 0; c < 3; ++c) {
            for (int t = 0; t < r.m; ++t) cur[((size_t)c * maxm + t) * W + w] = pack_word(ring, r.cur_v[c][t], r.cur_s[c][t]);
            for (int t = 0; t < r.best_m; ++t)
                best[((size_t)w * 3 + c) * maxm + t] = pack_word(ring, r.best_v[c][t], r.best_s[c][t]);
        }
    }
    HIPCHK(hipMemcpy(ctx->cur, cur.data(), plane * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->best, best.data(), plane * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->m, m.data(), W * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->best_m, bm.data(), W * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->best_rank, br.data(), W * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->flips_count, fc.data(), W * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->aux_idx, ai.data(), W * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->fmt, fmt.data(), (size_t)16 * W, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->bfmt, bfmt.data(), (size_t)16 * W, hipMemcpyHostToDevice));
    {
        bool same = true;
        for (size_t q = 0; q < (size_t)4 * W && same; ++q) same = (q % 4 == 3) || fmt[q] == bfmt[q];
        if (same != ctx->bfmt_same) ctx->drop_graph();  // a captured launch holds the old argument
        ctx->bfmt_same = same;
    }
    HIPCHK(hipMemcpy(ctx->it, it.data(), W * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->aux, aux.data(), W * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->total_flips, tf.data(), W * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->sum_m, sm.data(), W * 8, hipMemcpyHostToDevice));
    return 0;
}

namespace {

int walker_scheme(fg_ctx *ctx, int w, int which, fg::Scheme *out) {
    const int W = ctx->W, maxm = ctx->maxm;
    int32_t m = 0, fmt[4];
    HIPCHK(hipMemcpy(&m, (which ? ctx->best_m : ctx->m) + w, 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(fmt, ctx->fmt + 4 * w, 16, hipMemcpyDeviceToHost));
    std::vector<uint32_t> col((size_t)3 * maxm);
    if (which) {  // best schemes are contiguous per walker
        HIPCHK(hipMemcpy(col.data(), ctx->best + (size_t)w * 3 * maxm, col.size() * 4, hipMemcpyDeviceToHost));
    } else {  // one strided column of the current-scheme plane: pitch W words
        HIPCHK(hipMemcpy2D(col.data(), 4, ctx->cur + w, (size_t)W * 4, 4, (size_t)3 * maxm, hipMemcpyDeviceToHost));
    }
    fg::Scheme s;
    s.ring = ctx->cfg.ring;
    for (int c = 0; c < 3; ++c) s.n[c] = fmt[c];
    for (int c = 0; c < 3; ++c) s.nn[c] = s.n[c] * s.n[(c + 1) % 3];
    s.m = m;
    s.resize(m);
    for (int c = 0; c < 3; ++c)
        for (int t = 0; t < m; ++t) {
            s.v[c][t] = word_values(s.ring, col[(size_t)c * maxm + t]);
            s.s[c][t] = word_signs(s.ring, col[(size_t)c * maxm + t]);
        }
    *out = std::move(s);
    return 0;
}

int emit_blob(const fg::Scheme &s, int32_t *out, int64_t cap) {
    std::vector<int32_t> b;
    fg::scheme_to_blob(s, &b);
    if ((int64_t)b.size() > cap) return fail("output blob capacity too small");
    std::copy(b.begin(), b.end(), out);
    return (int)b.size();
}

}  // namespace

extern "C" int fg_walker_scheme(fg_ctx *ctx, int walker, int which, int32_t *out_blob, int64_t cap) {
    if (!ctx || !ctx->ready || !out_blob) return fail("fg_walker_scheme: bad arguments");
    if (walker < 0 || walker >= ctx->W) return fail("fg_walker_scheme: walker out of range");
    fg::Scheme s;
    int rc = walker_scheme(ctx, walker, which, &s);
    if (rc) return rc;
    return emit_blob(s, out_blob, cap);
}

extern "C" int fg_best(fg_ctx *ctx, int32_t *out_blob, int64_t cap, int32_t *out_rank, int32_t *out_walker) {
    if (!ctx || !ctx->ready) return fail("fg_best: context not initialised");
    // a minimizer pool ranks walkers by best complexity, a walk pool by best rank
    std::vector<int32_t> bm(ctx->W);
    HIPCHK(hipMemcpy(bm.data(), ctx->minimize ? ctx->best_rank : ctx->best_m, ctx->W * 4, hipMemcpyDeviceToHost));
    int bw = 0;
    for (int w = 1; w < ctx->W; ++w)
        if (bm[w] < bm[bw]) bw = w;
    if (out_rank) *out_rank = bm[bw];
    if (out_walker) *out_walker = bw;
    if (!out_blob) return 0;
    fg::Scheme s;
    int rc = walker_scheme(ctx, bw, 1, &s);
    if (rc) return rc;
    return emit_blob(s, out_blob, cap);
}

extern "C" int fg_verify(int ring, const int32_t *blob, int64_t len, int32_t *out_ok) {
    if (!blob || len < 1 || !out_ok) return fail("fg_verify: bad arguments");
    if (ring != FG_RING_Z2 && ring != FG_RING_ZT) return fail("fg_verify: unknown ring");
    const int count = blob[0];
    if (count < 0) return fail("fg_verify: negative count");
    std::vector<fg::Scheme> ss(count);
    int64_t pos = 1;
    int maxm = 1;
    for (int q = 0; q < count; ++q) {
        std::string err;
        if (!fg::scheme_from_blob(ring, blob, len, &pos, &ss[q], &err)) return fail("fg_verify: " + err);
        for (int c = 0; c < 3; ++c)
            if (ss[q].nn[c] > max_coefficients(ring)) return fail("fg_verify: factor wider than the GPU word of this ring");
        maxm = std::max(maxm, ss[q].m);
    }
    if (count == 0) return 0;
    std::vector<uint32_t> words((size_t)3 * maxm * count, 0);
    std::vector<int32_t> m(count), fmt((size_t)4 * count, 0);
    for (int q = 0; q < count; ++q) {
        m[q] = ss[q].m;
        for (int c = 0; c < 3; ++c) fmt[4 * q + c] = ss[q].n[c];
        for (int c = 0; c < 3; ++c)
            for (int t = 0; t < ss[q].m; ++t)
                words[((size_t)c * maxm + t) * count + q] = pack_word(ring, ss[q].v[c][t], ss[q].s[c][t]);
    }
    uint32_t *dw = nullptr;
    int32_t *dm = nullptr, *df = nullptr, *dok = nullptr;
    hipStream_t st = nullptr;
    int rc = 0;
    auto cleanup = [&]() {
        if (dw) (void)hipFree(dw);
        if (dm) (void)hipFree(dm);
        if (df) (void)hipFree(df);
        if (dok) (void)hipFree(dok);
        if (st) (void)hipStreamDestroy(st);
    };
    if (dalloc(&dw, words.size()) != hipSuccess || dalloc(&dm, count) != hipSuccess ||
        dalloc(&df, (size_t)4 * count) != hipSuccess || dalloc(&dok, count) != hipSuccess ||
        hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        cleanup();
        return fail("fg_verify: device allocation failed");
    }
    fg_ctx tmp;
    tmp.stream = st;
    if (hipMemcpy(dw, words.data(), words.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dm, m.data(), (size_t)count * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(df, fmt.data(), fmt.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
        cleanup();
        return fail("fg_verify: copy failed");
    }
    rc = verify_planes(&tmp, ring, dw, dm, df, count, maxm, dok);
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = fail("fg_verify: kernel failed");
    if (!rc && hipMemcpy(out_ok, dok, (size_t)count * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = fail("fg_verify: copy back failed");
    tmp.stream = nullptr;
    cleanup();
    if (rc) return rc;
    int good = 0;
    for (int q = 0; q < count; ++q) good += out_ok[q] != 0;
    return good;
}

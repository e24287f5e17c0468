// Host side of the batched additions reducer behind include/fg.h: the
// SchemeAdditionsReducer object of the reference
// (src/entities/scheme_additions_reducer.cu) with its reducers on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fg.h"
#include "reducer_args.h"
#include "runtime_internal.h"
#include "scheme_host.h"

namespace fgk {
hipError_t launch_reduce(const ReduceArgs &a, hipStream_t stream);
}

namespace {

int rfail(const std::string &msg, int code = -1);

#define RCHK(call)                                                                                 \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) return rfail(std::string(#call) + ": " + hipGetErrorString(e_));   \
    } while (0)

// Result of one (reducer, component) after a pass, as read back.
struct Part {
    int count = 0, real = 0, fresh = 0, naive = 0, mode = 0, additions = 0;
    float scale = 0;
    std::vector<int16_t> size, expr, vars;
};

const char *mode_name(int mode, float scale, char *buf) {  // getMode (additions_reducer.cuh:185-204)
    switch (mode) {
    case 0: return "g";
    case 1: return "ga";
    case 2: snprintf(buf, 16, "gr%d", (int)(scale * 100)); return buf;
    case 3: snprintf(buf, 16, "gi%d", (int)(scale * 100)); return buf;
    case 4: return "wr";
    case 6: return "rnd";
    default: return "mix";
    }
}

}  // namespace

namespace fgrt {
void set_error(const std::string &msg);
}

namespace {
int rfail(const std::string &msg, int code) {
    fgrt::set_error(msg);
    return code;
}
}  // namespace

struct fg_reducer {
    int count = 0, schemes_count = 1, max_flips = 0, top = 10;
    uint64_t seed = 0;
    hipStream_t stream = nullptr;
    fg::Scheme scheme;
    int n[3] = {0, 0, 0}, nn[3] = {0, 0, 0}, m = 0;
    int e_cap = 0, l_cap = 0, f_cap = 500, p_cap = 0, h_bits = 6;
    size_t lds = 0, lds_limit = 0;
    int S = 1;               // schemes the reducers read (reducer r takes scheme r % S)
    int max_nn = 0;
    fg_ctx *pool = nullptr;  // --max-flips > 0: the flipped schemes (walker w = scheme w)

    int8_t *coef = nullptr;
    int32_t *sm = nullptr, *best_fresh_d = nullptr, *info = nullptr, *err = nullptr;
    int16_t *best_vars_d = nullptr, *size = nullptr, *expr = nullptr, *vars = nullptr;
    uint64_t *words = nullptr;
    uint64_t decision_word = 0;

    // host state of updateBest / report
    std::vector<Part> parts;  // [count * 3], read back after each pass
    Part best[3];
    int best_additions[3] = {0, 0, 0}, best_fresh[3] = {0, 0, 0};
    int reduced_additions = 0, reduced_fresh = 0;
    std::vector<int> indices[3];
    bool improved_additions = false;

    ~fg_reducer() {
        void *ptrs[] = {coef, sm, best_fresh_d, info, err, best_vars_d, size, expr, vars, words};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
        if (pool) fg_destroy(pool);
    }
};

namespace {

template <typename T>
hipError_t ralloc(T **p, size_t n) {
    return hipMalloc(reinterpret_cast<void **>(p), std::max<size_t>(n, 1) * sizeof(T));
}

int coefficient(const fg::Scheme &s, int c, int t, int k) {
    const uint64_t bit = 1ull << k;
    if (!(s.v[c][t] & bit)) return 0;
    return (s.s[c][t] & bit) ? -1 : 1;
}

// Number of (expression, a < b) occurrences of one component (the work of
// one updateSubexpressions call).
long long occurrences(const fg::Scheme &s, int c) {
    long long p = 0;
    if (c < 2) {
        for (int t = 0; t < s.m; ++t) {
            const long long z = __builtin_popcountll(s.v[c][t]);
            p += z * (z - 1) / 2;
        }
    } else {
        for (int k = 0; k < s.nn[2]; ++k) {
            long long z = 0;
            for (int t = 0; t < s.m; ++t) z += (s.v[2][t] >> k) & 1;
            p += z * (z - 1) / 2;
        }
    }
    return p;
}

// getMaxRealVariables / getMaxSubexpressions (scheme_integer.cu:149-211)
int max_real_variables(const fg::Scheme &s, int c) {
    int best = 0;
    if (c < 2) {
        for (int t = 0; t < s.m; ++t) best = std::max(best, __builtin_popcountll(s.v[c][t]));
    } else {
        for (int k = 0; k < s.nn[2]; ++k) {
            int z = 0;
            for (int t = 0; t < s.m; ++t) z += (s.v[2][t] >> k) & 1;
            best = std::max(best, z);
        }
    }
    return best;
}

int max_subexpressions(const fg::Scheme &s, int c) {
    int total = 0;
    auto check = [](int v1, int v2, bool &pos, bool &neg) {  // checkSubexpression (scheme_integer.cu:213-223)
        if (!v1 || !v2) return;
        if (v1 < 0) {
            v1 = -v1;
            v2 = -v2;
        }
        if (v2 > 0) pos = true;
        else neg = true;
        (void)v1;
    };
    if (c == 2) {
        for (int a = 0; a < s.m; ++a)
            for (int b = a + 1; b < s.m; ++b) {
                bool pos = false, neg = false;
                for (int k = 0; k < s.nn[2] && (!pos || !neg); ++k) check(coefficient(s, 2, a, k), coefficient(s, 2, b, k), pos, neg);
                total += (int)pos + (int)neg;
            }
    } else {
        for (int i1 = 0; i1 < s.nn[c]; ++i1)
            for (int i2 = i1 + 1; i2 < s.nn[c]; ++i2) {
                bool pos = false, neg = false;
                for (int t = 0; t < s.m && (!pos || !neg); ++t) check(coefficient(s, c, t, i1), coefficient(s, c, t, i2), pos, neg);
                total += (int)pos + (int)neg;
            }
    }
    return total;
}

// Pair occurrences of one component of scheme q of the coefficient array
// (the same count as occurrences() on a host scheme).
long long coef_occurrences(const fg_reducer *R, const std::vector<int8_t> &coef, int q, int m, int c) {
    const int8_t *base = coef.data() + (size_t)q * 3 * R->m * R->max_nn;
    long long p = 0;
    if (c < 2) {
        for (int t = 0; t < m; ++t) {
            long long z = 0;
            for (int k = 0; k < R->nn[c]; ++k) z += base[((size_t)c * R->m + t) * R->max_nn + k] != 0;
            p += z * (z - 1) / 2;
        }
    } else {
        for (int k = 0; k < R->nn[2]; ++k) {
            long long z = 0;
            for (int t = 0; t < m; ++t) z += base[((size_t)2 * R->m + t) * R->max_nn + k] != 0;
            p += z * (z - 1) / 2;
        }
    }
    return p;
}

// Pair-table capacity and LDS of one pass for the largest occurrence count.
int size_pairs(fg_reducer *R, long long occ) {
    R->p_cap = (int)std::max<long long>(occ, 64);
    R->h_bits = 6;
    while ((1LL << R->h_bits) < 2LL * R->p_cap) ++R->h_bits;
    R->lds = fgk::reducer_lds_bytes(R->e_cap, R->l_cap, R->f_cap, R->p_cap, R->h_bits);
    if (R->lds > R->lds_limit)
        return rfail("fg_reducer: a scheme needs " + std::to_string(R->lds) +
                     " bytes of LDS per reducer component, more than one workgroup has");
    return 0;
}

// flipSchemesKernel (scheme_additions_reducer.cu:448-460) on the flip pool,
// then the pool's schemes become the reducers' inputs.
int flip_and_load(fg_reducer *R) {
    const int bad = fg_flip_schemes(R->pool, R->max_flips);
    if (bad < 0) return bad;
    if (bad > 0) return rfail("fg_reducer_run: " + std::to_string(bad) + " flipped schemes fail the Brent check", -3);
    fgrt::PoolView v{};
    if (!fgrt::pool_view(R->pool, &v)) return rfail("fg_reducer_run: flip pool not initialised");
    std::vector<uint32_t> words((size_t)3 * v.maxm * v.W);
    std::vector<int32_t> m(v.W);
    RCHK(hipMemcpy(words.data(), v.cur, words.size() * 4, hipMemcpyDeviceToHost));
    RCHK(hipMemcpy(m.data(), v.m, m.size() * 4, hipMemcpyDeviceToHost));
    std::vector<int8_t> coef((size_t)R->S * 3 * R->m * R->max_nn, 0);
    long long occ = 0;
    for (int q = 0; q < R->S; ++q) {
        if (m[q] < 1 || m[q] > R->m) return rfail("fg_reducer_run: flipped scheme with a bad rank");
        for (int c = 0; c < 3; ++c)
            for (int t = 0; t < m[q]; ++t) {
                const uint32_t x = words[((size_t)c * v.maxm + t) * v.W + q];
                for (int k = 0; k < R->nn[c]; ++k)
                    if ((x >> k) & 1u)
                        coef[(((size_t)q * 3 + c) * R->m + t) * R->max_nn + k] = ((x >> (16 + k)) & 1u) ? -1 : 1;
            }
        for (int c = 0; c < 3; ++c) occ = std::max(occ, coef_occurrences(R, coef, q, m[q], c));
    }
    const int rc = size_pairs(R, occ);
    if (rc) return rc;
    RCHK(hipMemcpy(R->coef, coef.data(), coef.size(), hipMemcpyHostToDevice));
    RCHK(hipMemcpy(R->sm, m.data(), (size_t)R->S * 4, hipMemcpyHostToDevice));
    return 0;
}

int read_back(fg_reducer *R) {
    const size_t rc = (size_t)R->count * 3;
    std::vector<int32_t> info(rc * 8);
    std::vector<int16_t> size(rc * R->e_cap), expr(rc * R->e_cap * R->l_cap), vars(rc * R->f_cap * 2);
    RCHK(hipMemcpy(info.data(), R->info, info.size() * 4, hipMemcpyDeviceToHost));
    RCHK(hipMemcpy(size.data(), R->size, size.size() * 2, hipMemcpyDeviceToHost));
    RCHK(hipMemcpy(expr.data(), R->expr, expr.size() * 2, hipMemcpyDeviceToHost));
    RCHK(hipMemcpy(vars.data(), R->vars, vars.size() * 2, hipMemcpyDeviceToHost));
    R->parts.resize(rc);
    for (size_t q = 0; q < rc; ++q) {
        Part &p = R->parts[q];
        const int32_t *in = info.data() + q * 8;
        p.count = in[0];
        p.real = in[1];
        p.fresh = in[2];
        p.naive = in[3];
        p.mode = in[4];
        memcpy(&p.scale, &in[5], 4);
        p.additions = in[6];
        p.size.assign(size.begin() + q * R->e_cap, size.begin() + (q + 1) * R->e_cap);
        p.expr.assign(expr.begin() + q * R->e_cap * R->l_cap, expr.begin() + (q + 1) * R->e_cap * R->l_cap);
        p.vars.assign(vars.begin() + q * R->f_cap * 2, vars.begin() + q * R->f_cap * 2 + 2 * p.fresh);
    }
    return 0;
}

int upload_best(fg_reducer *R) {
    std::vector<int32_t> bf(3);
    std::vector<int16_t> bv((size_t)3 * R->f_cap * 2, 0);
    for (int c = 0; c < 3; ++c) {
        bf[c] = R->best[c].fresh;
        std::copy(R->best[c].vars.begin(), R->best[c].vars.end(), bv.begin() + (size_t)c * R->f_cap * 2);
    }
    RCHK(hipMemcpy(R->best_fresh_d, bf.data(), 12, hipMemcpyHostToDevice));
    RCHK(hipMemcpy(R->best_vars_d, bv.data(), bv.size() * 2, hipMemcpyHostToDevice));
    return 0;
}

bool better(int a1, int f1, int a2, int f2) { return a1 != a2 ? a1 < a2 : f1 < f2; }

// updateBestIndependent (scheme_additions_reducer.cu:190-251)
bool update_independent(fg_reducer *R) {
    bool updated = false;
    for (int c = 0; c < 3; ++c) {
        std::vector<int> &idx = R->indices[c];
        std::partial_sort(idx.begin(), idx.begin() + R->top, idx.end(), [&](int a, int b) {
            const Part &pa = R->parts[(size_t)a * 3 + c], &pb = R->parts[(size_t)b * 3 + c];
            return better(pa.additions, pa.fresh, pb.additions, pb.fresh);
        });
        const Part &p = R->parts[(size_t)idx[0] * 3 + c];
        if (p.additions < R->best_additions[c] || (p.additions == R->best_additions[c] && p.fresh < R->best_fresh[c])) {
            R->best_additions[c] = p.additions;
            R->best_fresh[c] = p.fresh;
            R->best[c] = p;
            updated = true;
        }
    }
    return updated;
}

// updateBestTogether (:253-285)
bool update_together(fg_reducer *R) {
    std::vector<int> &idx = R->indices[0];
    auto tot = [&](int r, bool fresh) {
        int s = 0;
        for (int c = 0; c < 3; ++c) s += fresh ? R->parts[(size_t)r * 3 + c].fresh : R->parts[(size_t)r * 3 + c].additions;
        return s;
    };
    std::partial_sort(idx.begin(), idx.begin() + R->top, idx.end(),
                      [&](int a, int b) { return better(tot(a, false), tot(a, true), tot(b, false), tot(b, true)); });
    const int t = idx[0];
    const int add = tot(t, false), fr = tot(t, true);
    if (add < R->reduced_additions || (add == R->reduced_additions && fr < R->reduced_fresh)) {
        for (int c = 0; c < 3; ++c) {
            R->best_additions[c] = R->parts[(size_t)t * 3 + c].additions;
            R->best_fresh[c] = R->parts[(size_t)t * 3 + c].fresh;
            R->best[c] = R->parts[(size_t)t * 3 + c];
        }
        return true;
    }
    return false;
}

}  // namespace

extern "C" int fg_reducer_create(const int32_t *blob, int64_t len, int count, int schemes_count, int max_flips,
                                 uint64_t seed, fg_reducer **out) {
    if (!blob || !out || len < 4) return rfail("fg_reducer_create: bad arguments");
    *out = nullptr;
    if (count < 1 || schemes_count < 1 || max_flips < 0) return rfail("fg_reducer_create: counts must be positive");
    const int m = blob[3], n1 = blob[0], n2 = blob[1], n3 = blob[2];
    if (n1 < 1 || n2 < 1 || n3 < 1 || m < 1) return rfail("fg_reducer_create: bad scheme header");
    // SchemeAdditionsReducer::read (scheme_additions_reducer.cu:29-101)
    if (m > 350)
        return rfail("Error: expressions count for U (" + std::to_string(m) + ") too big for compiled configuration (350)");
    if (n3 * n1 > 64)
        return rfail("Error: expressions count for W (" + std::to_string(n3 * n1) + ") too big for compiled configuration (64)");
    fg::Scheme s;
    int64_t pos = 0;
    std::string err;
    if (!fg::scheme_from_blob(FG_RING_ZT, blob, len, &pos, &s, &err) || !fg::validate(s))
        return rfail("Error: readed scheme is invalid", -2);
    const int max_real[3] = {32, 32, 175};
    const char *names = "UVW";
    for (int c = 0; c < 3; ++c) {
        const int rv = max_real_variables(s, c);
        if (rv > max_real[c])
            return rfail(std::string("Error: real variables for ") + names[c] + " (" + std::to_string(rv) +
                         ") too big for compiled configuration (" + std::to_string(max_real[c]) + ")");
    }
    for (int c = 0; c < 3; ++c) {
        const int sub = max_subexpressions(s, c);
        if (sub > fgk::kReducerMaxPairs[c])
            return rfail(std::string("Error: max subexpressions for ") + names[c] + " (" + std::to_string(sub) +
                         ") too big for compiled configuration (" + std::to_string(fgk::kReducerMaxPairs[c]) + ")");
    }

    fg_reducer *R = new fg_reducer();
    R->count = count;
    R->schemes_count = schemes_count;
    R->max_flips = max_flips;
    R->seed = seed;
    R->top = std::min(10, count);
    R->scheme = s;
    R->m = m;
    for (int c = 0; c < 3; ++c) {
        R->n[c] = s.n[c];
        R->nn[c] = s.nn[c];
    }
    R->e_cap = std::max(m, R->nn[2]);
    R->l_cap = std::max(std::max(R->nn[0], R->nn[1]), m);
    long long p = 0;
    for (int c = 0; c < 3; ++c) p = std::max(p, occurrences(s, c));
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        delete R;
        return rfail("fg_reducer_create: no HIP device");
    }
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        delete R;
        return rfail(std::string("fg_reducer_create: this build targets gfx950 (MI355X), found ") + prop.gcnArchName);
    }
    R->lds_limit = prop.sharedMemPerBlock;
    if (size_pairs(R, p)) {
        delete R;
        return -1;
    }
    // the reducers read scheme r % schemes; without flips every scheme is a
    // copy of the input (initializeKernel, :433-446), so one copy suffices
    const int S = max_flips > 0 ? schemes_count : 1;
    const int max_nn = std::max(std::max(R->nn[0], R->nn[1]), R->nn[2]);
    R->S = S;
    R->max_nn = max_nn;
    std::vector<int8_t> coef((size_t)S * 3 * m * max_nn, 0);
    for (int q = 0; q < S; ++q)
        for (int c = 0; c < 3; ++c)
            for (int t = 0; t < m; ++t)
                for (int k = 0; k < R->nn[c]; ++k)
                    coef[(((size_t)q * 3 + c) * m + t) * max_nn + k] = (int8_t)coefficient(s, c, t, k);
    std::vector<int32_t> sm(S, m);
    const size_t rc = (size_t)count * 3;
    if (hipStreamCreateWithFlags(&R->stream, hipStreamNonBlocking) != hipSuccess ||
        ralloc(&R->coef, coef.size()) != hipSuccess || ralloc(&R->sm, S) != hipSuccess ||
        ralloc(&R->best_fresh_d, 3) != hipSuccess || ralloc(&R->best_vars_d, (size_t)3 * R->f_cap * 2) != hipSuccess ||
        ralloc(&R->words, (size_t)count * 4) != hipSuccess || ralloc(&R->info, rc * 8) != hipSuccess ||
        ralloc(&R->size, rc * R->e_cap) != hipSuccess || ralloc(&R->expr, rc * R->e_cap * R->l_cap) != hipSuccess ||
        ralloc(&R->vars, rc * R->f_cap * 2) != hipSuccess || ralloc(&R->err, 1) != hipSuccess) {
        delete R;
        return rfail("fg_reducer_create: device allocation failed");
    }
    std::vector<uint64_t> w0((size_t)count * 4, 0);
    if (hipMemcpy(R->coef, coef.data(), coef.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(R->sm, sm.data(), sm.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(R->words, w0.data(), w0.size() * 8, hipMemcpyHostToDevice) != hipSuccess) {
        delete R;
        return rfail("fg_reducer_create: upload failed");
    }
    if (max_flips > 0) {
        // the flipped schemes live in a walker pool of the integer-ring walk
        fg_config pc{};
        pc.ring = FG_RING_ZT;
        pc.walkers = schemes_count;
        pc.max_iterations = max_flips;
        pc.plus_iterations = 1 << 30;
        pc.check_reduce = 1;
        pc.seed = seed;
        std::vector<int32_t> one(1 + (size_t)len);
        one[0] = 1;
        std::copy(blob, blob + len, one.begin() + 1);
        if (fg_create(&pc, &R->pool) < 0 || fg_init_flips(R->pool, one.data(), (int64_t)one.size()) < 0) {
            const std::string e = fg_last_error();
            delete R;
            return rfail("fg_reducer_create: flip pool: " + e);
        }
    }
    // initialize (:144-175): the best reducers start as the plain expressions
    for (int c = 0; c < 3; ++c) {
        Part &b = R->best[c];
        b.count = c < 2 ? m : R->nn[2];
        b.real = c < 2 ? R->nn[c] : m;
        b.size.assign(R->e_cap, 0);
        b.expr.assign((size_t)R->e_cap * R->l_cap, 0);
        int naive = 0;
        for (int e = 0; e < b.count; ++e) {
            int len = 0;
            for (int v = 0; v < b.real; ++v) {
                const int x = c < 2 ? coefficient(s, c, e, v) : coefficient(s, 2, v, e);
                if (x) b.expr[(size_t)e * R->l_cap + len++] = (int16_t)(x > 0 ? v + 1 : -(v + 1));
            }
            b.size[e] = (int16_t)len;
            naive += len - 1;
        }
        b.naive = naive;
        b.additions = naive;
        R->best_additions[c] = naive;
        R->indices[c].resize(count);
        for (int q = 0; q < count; ++q) R->indices[c][q] = q;
    }
    R->best_fresh[0] = R->nn[0] * (R->nn[0] - 1);
    R->best_fresh[1] = R->nn[1] * (R->nn[1] - 1);
    R->best_fresh[2] = m * (m - 1);
    R->reduced_additions = R->best_additions[0] + R->best_additions[1] + R->best_additions[2];
    R->reduced_fresh = R->best_fresh[0] + R->best_fresh[1] + R->best_fresh[2];
    if (upload_best(R)) {
        delete R;
        return -1;
    }
    *out = R;
    return 0;
}

extern "C" int fg_reducer_destroy(fg_reducer *r) {
    delete r;
    return 0;
}

extern "C" int fg_reducer_info(fg_reducer *r, int32_t *out) {
    if (!r || !out) return rfail("fg_reducer_info: bad arguments");
    const int32_t v[16] = {r->n[0], r->n[1], r->n[2], r->m, r->count, r->schemes_count, r->max_flips, r->top,
                           r->e_cap, r->l_cap, r->f_cap, r->p_cap, r->h_bits, (int32_t)r->lds, 0, 0};
    memcpy(out, v, sizeof v);
    return 0;
}

extern "C" int fg_reducer_scheme_stats(fg_reducer *r, int32_t *out) {
    if (!r || !out) return rfail("fg_reducer_scheme_stats: bad arguments");
    for (int c = 0; c < 3; ++c) {
        out[c] = r->best[c].naive;
        out[3 + c] = max_real_variables(r->scheme, c);
        out[6 + c] = max_subexpressions(r->scheme, c);
    }
    return 0;
}

extern "C" int fg_reducer_run(fg_reducer *r, int iteration) {
    if (!r) return rfail("fg_reducer_run: null reducer");
    if (r->max_flips > 0 && iteration > 1) {  // reduceIteration (:172-176)
        const int rc = flip_and_load(r);
        if (rc) return rc;
    }
    fgk::ReduceArgs a{};
    a.coef = r->coef;
    a.sm = r->sm;
    a.schemes = r->S;
    a.n1 = r->n[0];
    a.n2 = r->n[1];
    a.n3 = r->n[2];
    a.max_m = r->m;
    a.max_nn = std::max(std::max(r->nn[0], r->nn[1]), r->nn[2]);
    a.count = r->count;
    a.key0 = (uint32_t)r->seed;
    a.key1 = (uint32_t)(r->seed >> 32);
    a.independent = r->max_flips == 0;
    a.decision_word = r->decision_word;
    a.best_fresh = r->best_fresh_d;
    a.best_vars = r->best_vars_d;
    a.words = r->words;
    a.info = r->info;
    a.size = r->size;
    a.expr = r->expr;
    a.vars = r->vars;
    a.e_cap = r->e_cap;
    a.l_cap = r->l_cap;
    a.f_cap = r->f_cap;
    a.p_cap = r->p_cap;
    a.h_bits = r->h_bits;
    a.err = r->err;
    int32_t zero = 0;
    RCHK(hipMemcpyAsync(r->err, &zero, 4, hipMemcpyHostToDevice, r->stream));
    RCHK(fgk::launch_reduce(a, r->stream));
    RCHK(hipStreamSynchronize(r->stream));
    int32_t err = 0;
    RCHK(hipMemcpy(&err, r->err, 4, hipMemcpyDeviceToHost));
    if (!a.independent) ++r->decision_word;
    if (err) return rfail("fg_reducer_run: reducer capacity exceeded (flags " + std::to_string(err) + ")", -3);
    return read_back(r);
}

extern "C" int fg_reducer_update(fg_reducer *r, int32_t *improved) {
    if (!r || r->parts.empty()) return rfail("fg_reducer_update: no pass has run");
    // updateBest (:287-307)
    const bool updated = r->max_flips == 0 ? update_independent(r) : update_together(r);
    int add = 0, fr = 0;
    for (int c = 0; c < 3; ++c) {
        add += r->best_additions[c];
        fr += r->best_fresh[c];
    }
    int code = 0;
    if (updated) {
        code = add < r->reduced_additions ? 1 : 2;
        r->reduced_additions = add;
        r->reduced_fresh = fr;
        if (upload_best(r)) return -1;
    }
    if (improved) *improved = code;
    return 0;
}

extern "C" int fg_reducer_best(fg_reducer *r, int32_t *out) {
    if (!r || !out) return rfail("fg_reducer_best: bad arguments");
    for (int c = 0; c < 3; ++c) {
        out[c] = r->best_additions[c];
        out[3 + c] = r->best_fresh[c];
        out[6 + c] = r->best[c].naive;
    }
    out[9] = r->reduced_additions;
    out[10] = r->reduced_fresh;
    return 0;
}

extern "C" int fg_reducer_rows(fg_reducer *r, int rows, int32_t *out) {
    if (!r || !out || r->parts.empty()) return rfail("fg_reducer_rows: no pass has run");
    // report (:309-363): row i shows U of indices[0][i] and V / W of
    // indices[1][i] / indices[2][i] (all indices[0] when the schemes flip)
    int n = 0;
    for (int i = 0; i < rows && i < r->top && i < r->count; ++i, ++n)
        for (int c = 0; c < 3; ++c) {
            const int w = r->indices[r->max_flips > 0 ? 0 : c][i];
            const Part &p = r->parts[(size_t)w * 3 + c];
            int32_t *o = out + ((size_t)i * 3 + c) * 6;
            o[0] = w;
            o[1] = p.mode;
            memcpy(&o[2], &p.scale, 4);
            o[3] = p.naive;
            o[4] = p.additions;
            o[5] = p.fresh;
        }
    return n;
}

extern "C" int fg_reducer_mode_name(int mode, float scale, char *buf, int cap) {
    if (!buf || cap < 8) return rfail("fg_reducer_mode_name: bad buffer");
    char tmp[16];
    const char *s = mode_name(mode, scale, tmp);
    strncpy(buf, s, (size_t)cap - 1);
    buf[cap - 1] = 0;
    return 0;
}

// Serialises one part: count real fresh naive mode scale_bits additions,
// fresh pairs (2 * fresh), then per expression its size and entries.
static int64_t emit_part(const fg_reducer *r, const Part &p, int32_t *out, int64_t cap) {
    std::vector<int32_t> v{p.count, p.real, p.fresh, p.naive, p.mode, 0, p.additions};
    memcpy(&v[5], &p.scale, 4);
    for (int q = 0; q < 2 * p.fresh; ++q) v.push_back(p.vars[q]);
    for (int e = 0; e < p.count; ++e) {
        v.push_back(p.size[e]);
        for (int x = 0; x < p.size[e]; ++x) v.push_back(p.expr[(size_t)e * r->l_cap + x]);
    }
    if ((int64_t)v.size() > cap) return rfail("output buffer too small");
    std::copy(v.begin(), v.end(), out);
    return (int64_t)v.size();
}

extern "C" int64_t fg_reducer_best_part(fg_reducer *r, int component, int32_t *out, int64_t cap) {
    if (!r || !out || component < 0 || component > 2) return rfail("fg_reducer_best_part: bad arguments");
    return emit_part(r, r->best[component], out, cap);
}

extern "C" int64_t fg_reducer_part(fg_reducer *r, int reducer, int component, int32_t *out, int64_t cap) {
    if (!r || !out || component < 0 || component > 2 || reducer < 0 || reducer >= r->count || r->parts.empty())
        return rfail("fg_reducer_part: bad arguments");
    return emit_part(r, r->parts[(size_t)reducer * 3 + component], out, cap);
}

extern "C" int fg_reducer_words(fg_reducer *r, uint64_t *out) {
    if (!r || !out) return rfail("fg_reducer_words: bad arguments");
    RCHK(hipMemcpy(out, r->words, (size_t)r->count * 4 * 8, hipMemcpyDeviceToHost));
    for (int q = 0; q < r->count; ++q) out[(size_t)q * 4 + 3] = r->decision_word;
    return 0;
}

// Learned flip scorer behind include/fg.h: a term-set network that scores
// candidate flips of a Z2 scheme, trained on the GPU from the walks' own
// trajectory log, and the learned walk round that mixes its proposals with
// uniform flips.  No counterpart in the reference.
//
// Network (BF16 operands on the MFMA units, FP32 accumulation):
//   e_t = relu(x_t W0 + b0)                       x_t: the 3*n^2 factor bits of term t
//   h_t = relu(e_t A1 + mean(e) B1 + c1)          mean over the scheme's terms
//   g_t = relu(h_t A2 + mean(h) B2 + c2)
//   z   = relu([g_a + g_b, g_a * g_b] Wz + bz)    for the pair (component c, terms a, b)
//   s   = z . wo[c] + bo[c]                       score (logit of "rank falls within the horizon")
// Loss: binary cross-entropy with logits.  Optimiser: Adam.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "fg.h"
#include "runtime_internal.h"
#include "scorer_args.h"

namespace {

using fgk::sc::bf16;

int sfail(const std::string &msg, int code = -1) {
    fgrt::set_error(msg);
    return code;
}

#define SCHK(call)                                                                                \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) return sfail(std::string(#call) + ": " + hipGetErrorString(e_));  \
    } while (0)

template <typename T>
hipError_t salloc(T **p, size_t n) {
    return hipMalloc(reinterpret_cast<void **>(p), std::max<size_t>(n, 1) * sizeof(T));
}

uint64_t splitmix(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

struct fg_scorer {
    int n[3] = {0, 0, 0}, nn[3] = {0, 0, 0}, F = 0, D = 128, Z = 64, M = 0;
    uint64_t seed = 0;
    hipStream_t stream = nullptr;
    long np = 0, oW0 = 0, ob0 = 0, oA1 = 0, oB1 = 0, oc1 = 0, oA2 = 0, oB2 = 0, oc2 = 0, oWz = 0, obz = 0, owo = 0,
         obo = 0;
    float *w = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
    bf16 *wb = nullptr;
    // fused proposal rounds: transposed weights and the candidates of the last round
    bf16 *wt = nullptr;
    int32_t *fmap = nullptr, *fpairs = nullptr, *ftup = nullptr;
    float *fscores = nullptr;
    long capF = 0;  // candidate tuples the fused buffers hold
    long capL = 0;  // proposing walkers fmap holds
    int lastB = 0, lastK = 0;
    bool last_fused = false;  // the last round's candidates sit in the fused buffers (else: map / pairs / s)
    // HIP events around the fused kernel of the last round, and its duration
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.0f;
    int step = 0;

    // workspace (capacity in tuples and pairs; backward buffers on demand)
    int capB = 0, capP = 0;
    bool back = false;
    std::vector<void *> bufs;
    uint32_t *words = nullptr;
    int32_t *m = nullptr, *pairs = nullptr, *tup = nullptr, *map = nullptr;
    float *y = nullptr, *loss = nullptr;
    bf16 *X = nullptr, *Eb = nullptr, *MEb = nullptr, *Hb = nullptr, *MHb = nullptr, *Gb = nullptr, *Qp = nullptr;
    float *mask = nullptr, *inv_n = nullptr, *E = nullptr, *ME = nullptr, *P = nullptr, *Q = nullptr, *H = nullptr,
          *MH = nullptr, *G = nullptr, *Zf = nullptr, *s = nullptr;
    // backward
    float *ds = nullptr, *dZ = nullptr, *dQp = nullptr, *dG = nullptr, *dV = nullptr, *SDV = nullptr, *dH = nullptr,
          *dMH = nullptr, *dU = nullptr, *SDU = nullptr, *dE = nullptr, *dME = nullptr, *dP0 = nullptr;
    bf16 *dZb = nullptr, *dVb = nullptr, *SDVb = nullptr, *dUb = nullptr, *SDUb = nullptr, *dP0b = nullptr;

    void free_work() {
        for (void *p : bufs)
            if (p) (void)hipFree(p);
        bufs.clear();
        capB = capP = 0;
        back = false;
    }
    ~fg_scorer() {
        free_work();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        void *ps[] = {w, g, m1, m2, wb, wt, fmap, fpairs, ftup, fscores};
        for (void *p : ps)
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

template <typename T>
int take(fg_scorer *S, T **p, size_t n) {
    if (salloc(p, n) != hipSuccess) return sfail("fg_scorer: device allocation failed");
    S->bufs.push_back(*p);
    return 0;
}

// Workspace for B tuples and P pairs (rows = B * M).
int ensure(fg_scorer *S, int B, int P, bool back) {
    if (B <= S->capB && P <= S->capP && (!back || S->back)) return 0;
    const int nb = std::max(B, S->capB), npr = std::max(P, S->capP);
    const bool nbk = back || S->back;
    S->free_work();
    const size_t R = (size_t)nb * S->M, D = S->D;
    int rc = 0;
    rc |= take(S, &S->words, (size_t)nb * 3 * S->M);
    rc |= take(S, &S->m, nb);
    rc |= take(S, &S->map, nb);
    rc |= take(S, &S->pairs, (size_t)npr * 3);
    rc |= take(S, &S->tup, npr);
    rc |= take(S, &S->y, npr);
    rc |= take(S, &S->loss, 1);
    rc |= take(S, &S->X, R * S->F);
    rc |= take(S, &S->mask, R);
    rc |= take(S, &S->inv_n, nb);
    rc |= take(S, &S->E, R * D);
    rc |= take(S, &S->Eb, R * D);
    rc |= take(S, &S->ME, (size_t)nb * D);
    rc |= take(S, &S->MEb, (size_t)nb * D);
    rc |= take(S, &S->P, R * D);
    rc |= take(S, &S->Q, (size_t)nb * D);
    rc |= take(S, &S->H, R * D);
    rc |= take(S, &S->Hb, R * D);
    rc |= take(S, &S->MH, (size_t)nb * D);
    rc |= take(S, &S->MHb, (size_t)nb * D);
    rc |= take(S, &S->G, R * D);
    rc |= take(S, &S->Gb, R * D);
    rc |= take(S, &S->Qp, (size_t)npr * 2 * D);
    rc |= take(S, &S->Zf, (size_t)npr * S->Z);
    rc |= take(S, &S->s, npr);
    if (nbk) {
        rc |= take(S, &S->ds, npr);
        rc |= take(S, &S->dZ, (size_t)npr * S->Z);
        rc |= take(S, &S->dZb, (size_t)npr * S->Z);
        rc |= take(S, &S->dQp, (size_t)npr * 2 * D);
        rc |= take(S, &S->dG, R * D);
        rc |= take(S, &S->dV, R * D);
        rc |= take(S, &S->dVb, R * D);
        rc |= take(S, &S->SDV, (size_t)nb * D);
        rc |= take(S, &S->SDVb, (size_t)nb * D);
        rc |= take(S, &S->dH, R * D);
        rc |= take(S, &S->dMH, (size_t)nb * D);
        rc |= take(S, &S->dU, R * D);
        rc |= take(S, &S->dUb, R * D);
        rc |= take(S, &S->SDU, (size_t)nb * D);
        rc |= take(S, &S->SDUb, (size_t)nb * D);
        rc |= take(S, &S->dE, R * D);
        rc |= take(S, &S->dME, (size_t)nb * D);
        rc |= take(S, &S->dP0, R * D);
        rc |= take(S, &S->dP0b, R * D);
    }
    if (rc) return -1;
    S->capB = nb;
    S->capP = npr;
    S->back = nbk;
    return 0;
}

hipError_t gemm(fg_scorer *S, const bf16 *A, long sam, long sak, const bf16 *B, long sbk, long sbn, float *C, long ldc,
                bf16 *Cb, long ldcb, const float *bias, int M, int N, int K, int relu, int acc) {
    fgk::sc::GemmArgs a{};
    a.A = A;
    a.sa_m = sam;
    a.sa_k = sak;
    a.B = B;
    a.sb_k = sbk;
    a.sb_n = sbn;
    a.C = C;
    a.ldc = ldc;
    a.Cb = Cb;
    a.ldcb = ldcb;
    a.bias = bias;
    a.M = M;
    a.N = N;
    a.K = K;
    a.relu = relu;
    a.accumulate = acc;
    return fgk::sc::launch_gemm(a, S->stream);
}

// Forward over B featurised tuples and P pairs; scores to S->s.
int forward(fg_scorer *S, int B, int P) {
    const int R = B * S->M, D = S->D;
    const bf16 *wb = S->wb;
    const float *w = S->w;
    hipStream_t st = S->stream;
    SCHK(gemm(S, S->X, S->F, 1, wb + S->oW0, D, 1, S->E, D, S->Eb, D, w + S->ob0, R, D, S->F, 1, 0));
    SCHK(fgk::sc::launch_seg_mean(S->E, S->mask, S->inv_n, B, S->M, D, S->ME, S->MEb, st));
    SCHK(gemm(S, S->Eb, D, 1, wb + S->oA1, D, 1, S->P, D, nullptr, 0, nullptr, R, D, D, 0, 0));
    SCHK(gemm(S, S->MEb, D, 1, wb + S->oB1, D, 1, S->Q, D, nullptr, 0, nullptr, B, D, D, 0, 0));
    SCHK(fgk::sc::launch_set_combine(S->P, S->Q, w + S->oc1, S->mask, B, S->M, D, S->H, S->Hb, st));
    SCHK(fgk::sc::launch_seg_mean(S->H, S->mask, S->inv_n, B, S->M, D, S->MH, S->MHb, st));
    SCHK(gemm(S, S->Hb, D, 1, wb + S->oA2, D, 1, S->P, D, nullptr, 0, nullptr, R, D, D, 0, 0));
    SCHK(gemm(S, S->MHb, D, 1, wb + S->oB2, D, 1, S->Q, D, nullptr, 0, nullptr, B, D, D, 0, 0));
    SCHK(fgk::sc::launch_set_combine(S->P, S->Q, w + S->oc2, S->mask, B, S->M, D, S->G, S->Gb, st));
    SCHK(fgk::sc::launch_pair_gather(S->G, S->tup, S->pairs, P, S->M, D, S->Qp, st));
    SCHK(gemm(S, S->Qp, 2 * D, 1, wb + S->oWz, S->Z, 1, S->Zf, S->Z, nullptr, 0, w + S->obz, P, S->Z, 2 * D, 1, 0));
    SCHK(fgk::sc::launch_head_out(S->Zf, w + S->owo, w + S->obo, S->pairs, P, S->Z, S->s, st));
    return 0;
}

// Loss and gradients (into S->g) for the last forward, labels in S->y.
int backward(fg_scorer *S, int B, int P) {
    const int R = B * S->M, D = S->D, Zd = S->Z;
    const bf16 *wb = S->wb;
    float *g = S->g;
    hipStream_t st = S->stream;
    SCHK(fgk::sc::launch_fill_f32(S->loss, 0.f, 1, st));
    SCHK(fgk::sc::launch_fill_f32(g, 0.f, S->np, st));
    SCHK(fgk::sc::launch_bce(S->s, S->y, P, S->ds, S->loss, st));
    SCHK(fgk::sc::launch_head_back(S->Zf, S->ds, S->w + S->owo, S->pairs, P, Zd, S->dZ, S->dZb, g + S->owo, g + S->obo, st));
    SCHK(fgk::sc::launch_colsum(S->dZ, P, Zd, g + S->obz, st));
    SCHK(gemm(S, S->Qp, 1, 2 * D, S->dZb, Zd, 1, g + S->oWz, Zd, nullptr, 0, nullptr, 2 * D, Zd, P, 0, 0));
    SCHK(gemm(S, S->dZb, Zd, 1, wb + S->oWz, 1, Zd, S->dQp, 2 * D, nullptr, 0, nullptr, P, 2 * D, Zd, 0, 0));
    SCHK(fgk::sc::launch_fill_f32(S->dG, 0.f, (long)R * D, st));
    SCHK(fgk::sc::launch_pair_scatter(S->dQp, S->G, S->tup, S->pairs, P, S->M, D, S->dG, st));
    // layer 2
    SCHK(fgk::sc::launch_set_back(S->dG, nullptr, S->inv_n, S->G, S->mask, B, S->M, D, S->dV, S->dVb, st));
    SCHK(gemm(S, S->Hb, 1, D, S->dVb, D, 1, g + S->oA2, D, nullptr, 0, nullptr, D, D, R, 0, 0));
    SCHK(fgk::sc::launch_colsum(S->dV, R, D, g + S->oc2, st));
    SCHK(fgk::sc::launch_seg_sum(S->dV, S->mask, B, S->M, D, S->SDV, S->SDVb, st));
    SCHK(gemm(S, S->MHb, 1, D, S->SDVb, D, 1, g + S->oB2, D, nullptr, 0, nullptr, D, D, B, 0, 0));
    SCHK(gemm(S, S->dVb, D, 1, wb + S->oA2, 1, D, S->dH, D, nullptr, 0, nullptr, R, D, D, 0, 0));
    SCHK(gemm(S, S->SDVb, D, 1, wb + S->oB2, 1, D, S->dMH, D, nullptr, 0, nullptr, B, D, D, 0, 0));
    // layer 1
    SCHK(fgk::sc::launch_set_back(S->dH, S->dMH, S->inv_n, S->H, S->mask, B, S->M, D, S->dU, S->dUb, st));
    SCHK(gemm(S, S->Eb, 1, D, S->dUb, D, 1, g + S->oA1, D, nullptr, 0, nullptr, D, D, R, 0, 0));
    SCHK(fgk::sc::launch_colsum(S->dU, R, D, g + S->oc1, st));
    SCHK(fgk::sc::launch_seg_sum(S->dU, S->mask, B, S->M, D, S->SDU, S->SDUb, st));
    SCHK(gemm(S, S->MEb, 1, D, S->SDUb, D, 1, g + S->oB1, D, nullptr, 0, nullptr, D, D, B, 0, 0));
    SCHK(gemm(S, S->dUb, D, 1, wb + S->oA1, 1, D, S->dE, D, nullptr, 0, nullptr, R, D, D, 0, 0));
    SCHK(gemm(S, S->SDUb, D, 1, wb + S->oB1, 1, D, S->dME, D, nullptr, 0, nullptr, B, D, D, 0, 0));
    // embedding
    SCHK(fgk::sc::launch_set_back(S->dE, S->dME, S->inv_n, S->E, S->mask, B, S->M, D, S->dP0, S->dP0b, st));
    SCHK(gemm(S, S->X, 1, S->F, S->dP0b, D, 1, g + S->oW0, D, nullptr, 0, nullptr, S->F, D, R, 0, 0));
    SCHK(fgk::sc::launch_colsum(S->dP0, R, D, g + S->ob0, st));
    return 0;
}

int featurize(fg_scorer *S, const uint32_t *words, const int32_t *m, const int32_t *map, long pool_stride, int B) {
    fgk::sc::FeatArgs f{};
    f.words = words;
    f.m = m;
    f.map = map;
    f.pool_stride = pool_stride;
    f.B = B;
    f.M = S->M;
    f.nn0 = S->nn[0];
    f.nn1 = S->nn[1];
    f.nn2 = S->nn[2];
    f.X = S->X;
    f.mask = S->mask;
    f.inv_n = S->inv_n;
    SCHK(fgk::sc::launch_featurize(f, S->stream));
    return 0;
}

}  // namespace

extern "C" int fg_scorer_create(int n1, int n2, int n3, int max_terms, uint64_t seed, fg_scorer **out) {
    if (!out || n1 < 1 || n2 < 1 || n3 < 1 || max_terms < 2) return sfail("fg_scorer_create: bad arguments");
    *out = nullptr;
    if (n1 * n2 > 32 || n2 * n3 > 32 || n3 * n1 > 32) return sfail("fg_scorer_create: factors of up to 32 bits");
    fg_scorer *S = new fg_scorer();
    S->n[0] = n1;
    S->n[1] = n2;
    S->n[2] = n3;
    S->nn[0] = n1 * n2;
    S->nn[1] = n2 * n3;
    S->nn[2] = n3 * n1;
    S->F = S->nn[0] + S->nn[1] + S->nn[2];
    S->M = max_terms;
    S->seed = seed;
    const long D = S->D, Zd = S->Z;
    long o = 0;
    auto at = [&](long n) {
        const long r = o;
        o += n;
        return r;
    };
    S->oW0 = at((long)S->F * D);
    S->ob0 = at(D);
    S->oA1 = at(D * D);
    S->oB1 = at(D * D);
    S->oc1 = at(D);
    S->oA2 = at(D * D);
    S->oB2 = at(D * D);
    S->oc2 = at(D);
    S->oWz = at(2 * D * Zd);
    S->obz = at(Zd);
    S->owo = at(3 * Zd);
    S->obo = at(3);
    S->np = o;
    // uniform(-a, a), a = sqrt(6 / (fan_in + fan_out)); biases zero
    std::vector<float> w((size_t)S->np, 0.f);
    uint64_t st = seed ^ 0x5C0DE5EEDull;
    auto fill = [&](long off, long fan_in, long fan_out) {
        const double a = std::sqrt(6.0 / (double)(fan_in + fan_out));
        for (long q = 0; q < fan_in * fan_out; ++q) {
            const double u = (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
            w[off + q] = (float)((2 * u - 1) * a);
        }
    };
    fill(S->oW0, S->F, D);
    fill(S->oA1, D, D);
    fill(S->oB1, D, D);
    fill(S->oA2, D, D);
    fill(S->oB2, D, D);
    fill(S->oWz, 2 * D, Zd);
    fill(S->owo, Zd, 3);
    if (hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking) != hipSuccess || salloc(&S->w, S->np) != hipSuccess ||
        salloc(&S->g, S->np) != hipSuccess || salloc(&S->m1, S->np) != hipSuccess || salloc(&S->m2, S->np) != hipSuccess ||
        salloc(&S->wb, S->np) != hipSuccess) {
        delete S;
        return sfail("fg_scorer_create: device allocation failed");
    }
    std::vector<float> zero((size_t)S->np, 0.f);
    if (hipMemcpy(S->w, w.data(), w.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(S->m1, zero.data(), zero.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(S->m2, zero.data(), zero.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        fgk::sc::launch_to_bf16(S->w, S->wb, S->np, S->stream) != hipSuccess ||
        hipStreamSynchronize(S->stream) != hipSuccess) {
        delete S;
        return sfail("fg_scorer_create: upload failed");
    }
    *out = S;
    return 0;
}

extern "C" int fg_scorer_destroy(fg_scorer *s) {
    delete s;
    return 0;
}

extern "C" int64_t fg_scorer_num_params(fg_scorer *s) { return s ? s->np : sfail("fg_scorer_num_params: null scorer"); }

extern "C" int fg_scorer_layout(fg_scorer *s, int32_t *out) {
    if (!s || !out) return sfail("fg_scorer_layout: bad arguments");
    const int32_t v[16] = {s->F, s->D, s->Z, s->M, (int32_t)s->oW0, (int32_t)s->ob0, (int32_t)s->oA1, (int32_t)s->oB1,
                           (int32_t)s->oc1, (int32_t)s->oA2, (int32_t)s->oB2, (int32_t)s->oc2, (int32_t)s->oWz,
                           (int32_t)s->obz, (int32_t)s->owo, (int32_t)s->obo};
    memcpy(out, v, sizeof v);
    return 0;
}

extern "C" int fg_scorer_get_params(fg_scorer *s, float *out) {
    if (!s || !out) return sfail("fg_scorer_get_params: bad arguments");
    SCHK(hipMemcpy(out, s->w, (size_t)s->np * 4, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int fg_scorer_set_params(fg_scorer *s, const float *in) {
    if (!s || !in) return sfail("fg_scorer_set_params: bad arguments");
    SCHK(hipMemcpy(s->w, in, (size_t)s->np * 4, hipMemcpyHostToDevice));
    SCHK(fgk::sc::launch_to_bf16(s->w, s->wb, s->np, s->stream));
    SCHK(hipStreamSynchronize(s->stream));
    return 0;
}

extern "C" int fg_scorer_eval(fg_scorer *s, int count, const uint32_t *words, const int32_t *m, const int32_t *pairs,
                              const float *labels, float *scores, float *loss, float *grads) {
    if (!s || count < 1 || !words || !m || !pairs) return sfail("fg_scorer_eval: bad arguments");
    for (int q = 0; q < count; ++q) {
        if (m[q] < 1 || m[q] > s->M) return sfail("fg_scorer_eval: rank out of range");
        const int c = pairs[3 * q], a = pairs[3 * q + 1], b = pairs[3 * q + 2];
        if (c >= 0 && (c > 2 || a < 0 || b < 0 || a >= m[q] || b >= m[q] || a == b))
            return sfail("fg_scorer_eval: pair out of range");
    }
    const bool back = labels != nullptr && (loss || grads);
    if (ensure(s, count, count, back)) return -1;
    std::vector<int32_t> tup(count);
    for (int q = 0; q < count; ++q) tup[q] = q;
    SCHK(hipMemcpy(s->words, words, (size_t)count * 3 * s->M * 4, hipMemcpyHostToDevice));
    SCHK(hipMemcpy(s->m, m, (size_t)count * 4, hipMemcpyHostToDevice));
    SCHK(hipMemcpy(s->pairs, pairs, (size_t)count * 12, hipMemcpyHostToDevice));
    SCHK(hipMemcpy(s->tup, tup.data(), (size_t)count * 4, hipMemcpyHostToDevice));
    if (labels) SCHK(hipMemcpy(s->y, labels, (size_t)count * 4, hipMemcpyHostToDevice));
    if (featurize(s, s->words, s->m, nullptr, 0, count) || forward(s, count, count)) return -1;
    if (back && backward(s, count, count)) return -1;
    SCHK(hipStreamSynchronize(s->stream));
    if (scores) SCHK(hipMemcpy(scores, s->s, (size_t)count * 4, hipMemcpyDeviceToHost));
    if (back && loss) SCHK(hipMemcpy(loss, s->loss, 4, hipMemcpyDeviceToHost));
    if (back && grads) SCHK(hipMemcpy(grads, s->g, (size_t)s->np * 4, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int fg_scorer_train(fg_scorer *s, fg_ctx *ctx, int steps, int batch, float lr, uint64_t seed, float *losses) {
    if (!s || steps < 0 || batch < 1 || lr <= 0) return sfail("fg_scorer_train: bad arguments");
    fgrt::LogView L;
    if (!fgrt::log_view(ctx, &L)) return sfail("fg_scorer_train: the context has no trajectory log");
    if (L.maxm != s->M) return sfail("fg_scorer_train: the log's term capacity differs from the scorer's");
    const int64_t nrows = (int64_t)L.valid->size();
    if (nrows < 1) return sfail("fg_scorer_train: the trajectory log is empty");
    if (ensure(s, batch, batch, true)) return -1;
    std::vector<int32_t> map(batch), pr((size_t)batch * 3), tup(batch);
    std::vector<float> y(batch);
    for (int q = 0; q < batch; ++q) tup[q] = q;
    SCHK(hipMemcpy(s->tup, tup.data(), (size_t)batch * 4, hipMemcpyHostToDevice));
    uint64_t rng = seed;
    for (int it = 0; it < steps; ++it) {
        for (int q = 0; q < batch; ++q) {
            const int64_t k = (int64_t)(splitmix(rng) % (uint64_t)nrows);
            map[q] = (int32_t)(*L.valid)[k];
            for (int c = 0; c < 3; ++c) pr[3 * q + c] = (*L.vpairs)[3 * k + c];
            y[q] = (*L.vy)[k];
        }
        SCHK(hipMemcpyAsync(s->map, map.data(), (size_t)batch * 4, hipMemcpyHostToDevice, s->stream));
        SCHK(hipMemcpyAsync(s->pairs, pr.data(), (size_t)batch * 12, hipMemcpyHostToDevice, s->stream));
        SCHK(hipMemcpyAsync(s->y, y.data(), (size_t)batch * 4, hipMemcpyHostToDevice, s->stream));
        if (featurize(s, L.words, L.m, s->map, 0, batch) || forward(s, batch, batch) || backward(s, batch, batch))
            return -1;
        ++s->step;
        SCHK(fgk::sc::launch_adam(s->w, s->wb, s->g, s->m1, s->m2, s->np, lr, 0.9f, 0.999f, 1e-8f, s->step, s->stream));
        if (losses) {
            SCHK(hipMemcpyAsync(&losses[it], s->loss, 4, hipMemcpyDeviceToHost, s->stream));
        }
    }
    SCHK(hipStreamSynchronize(s->stream));
    return 0;
}

namespace {
// A proposal round with the fused kernel: candidates for every proposing
// walker in one launch, then one workgroup per walker runs the whole network
// and writes its best candidate.
int propose_fused(fg_scorer *s, const fgrt::PoolView &v, int candidates, int every, uint32_t round) {
    const int L = (v.W + every - 1) / every;
    const long need = (long)L * candidates;
    if (need > s->capF || L > s->capL || !s->fmap) {
        void *ps[] = {s->fmap, s->fpairs, s->ftup, s->fscores};
        for (void *p : ps)
            if (p) (void)hipFree(p);
        s->fmap = s->fpairs = s->ftup = nullptr;
        s->fscores = nullptr;
        s->capF = s->capL = 0;
        SCHK(salloc(&s->fmap, (size_t)std::max(L, 1)));
        SCHK(salloc(&s->fpairs, (size_t)need * 3));
        SCHK(salloc(&s->ftup, (size_t)need));
        SCHK(salloc(&s->fscores, (size_t)need));
        s->capF = need;
        s->capL = L;
    }
    if (!s->wt)
        SCHK(salloc(&s->wt, (size_t)fgk::sc::kPD_ * (fgk::sc::kProposeMaxF + 2 * fgk::sc::kPD_) +
                                (size_t)s->Z * 2 * s->D + (size_t)2 * fgk::sc::kPD_ * fgk::sc::kPD_));
    bf16 *W0t = s->wt, *A1t = W0t + (size_t)fgk::sc::kPD_ * fgk::sc::kProposeMaxF, *A2t = A1t + (size_t)fgk::sc::kPD_ * fgk::sc::kPD_;
    bf16 *Wzt = A2t + (size_t)fgk::sc::kPD_ * fgk::sc::kPD_;
    bf16 *B1t = Wzt + (size_t)s->Z * 2 * s->D, *B2t = B1t + (size_t)fgk::sc::kPD_ * fgk::sc::kPD_;
    SCHK(fgk::sc::launch_transpose_bf16(s->wb + s->oW0, s->F, s->D, fgk::sc::kProposeMaxF, W0t, s->stream));
    SCHK(fgk::sc::launch_transpose_bf16(s->wb + s->oA1, s->D, s->D, s->D, A1t, s->stream));
    SCHK(fgk::sc::launch_transpose_bf16(s->wb + s->oA2, s->D, s->D, s->D, A2t, s->stream));
    SCHK(fgk::sc::launch_transpose_bf16(s->wb + s->oWz, 2 * s->D, s->Z, 2 * s->D, Wzt, s->stream));
    SCHK(fgk::sc::launch_transpose_bf16(s->wb + s->oB1, s->D, s->D, s->D, B1t, s->stream));
    SCHK(fgk::sc::launch_transpose_bf16(s->wb + s->oB2, s->D, s->D, s->D, B2t, s->stream));
    std::vector<int32_t> map(L);
    for (int q = 0; q < L; ++q) map[q] = q * every;
    SCHK(hipMemcpyAsync(s->fmap, map.data(), (size_t)L * 4, hipMemcpyHostToDevice, s->stream));
    fgk::sc::CandArgs c{};
    c.words = v.cur;
    c.m = v.m;
    c.map = s->fmap;
    c.W = v.W;
    c.B = L;
    c.M = s->M;
    c.K = candidates;
    c.key0 = v.key0;
    c.key1 = v.key1;
    c.round = round;
    c.pairs = s->fpairs;
    c.tup = s->ftup;
    SCHK(fgk::sc::launch_candidates(c, s->stream));
    fgk::sc::ProposeArgs a{};
    a.words = v.cur;
    a.m = v.m;
    a.map = s->fmap;
    a.W = v.W;
    a.M = s->M;
    a.nn0 = s->nn[0];
    a.nn1 = s->nn[1];
    a.nn2 = s->nn[2];
    a.B = L;
    a.K = candidates;
    a.pairs = s->fpairs;
    a.W0t = W0t;
    a.A1t = A1t;
    a.A2t = A2t;
    a.B1t = B1t;
    a.B2t = B2t;
    a.Wz = s->wb + s->oWz;
    a.Wzt = Wzt;
    a.b0 = s->w + s->ob0;
    a.c1 = s->w + s->oc1;
    a.c2 = s->w + s->oc2;
    a.bz = s->w + s->obz;
    a.wo = s->w + s->owo;
    a.bo = s->w + s->obo;
    a.scores = s->fscores;
    a.forced = v.forced;
    if (!s->ev0) SCHK(hipEventCreate(&s->ev0));
    if (!s->ev1) SCHK(hipEventCreate(&s->ev1));
    SCHK(hipEventRecord(s->ev0, s->stream));
    SCHK(fgk::sc::launch_propose(a, s->stream));
    SCHK(hipEventRecord(s->ev1, s->stream));
    SCHK(hipStreamSynchronize(s->stream));
    SCHK(hipEventElapsedTime(&s->last_ms, s->ev0, s->ev1));
    s->lastB = L;
    s->lastK = candidates;
    s->last_fused = true;
    return 0;
}
}  // namespace

extern "C" int fg_scorer_last_propose_ms(fg_scorer *s, float *ms) {
    if (!s || !ms) return sfail("fg_scorer_last_propose_ms: bad arguments");
    *ms = s->lastB && s->last_fused ? s->last_ms : 0.0f;
    return 0;
}

extern "C" int fg_scorer_last_proposals(fg_scorer *s, int32_t *walkers, int32_t *pairs, float *scores, int64_t cap) {
    if (!s) return sfail("fg_scorer_last_proposals: null scorer");
    const long n = (long)s->lastB * s->lastK;
    if (n == 0) return 0;
    if (cap < n) return sfail("fg_scorer_last_proposals: buffer too small");
    const int32_t *wsrc = s->last_fused ? s->fmap : s->map;
    const int32_t *psrc = s->last_fused ? s->fpairs : s->pairs;
    const float *ssrc = s->last_fused ? s->fscores : s->s;
    if (walkers) SCHK(hipMemcpy(walkers, wsrc, (size_t)s->lastB * 4, hipMemcpyDeviceToHost));
    if (pairs) SCHK(hipMemcpy(pairs, psrc, (size_t)n * 12, hipMemcpyDeviceToHost));
    if (scores) SCHK(hipMemcpy(scores, ssrc, (size_t)n * 4, hipMemcpyDeviceToHost));
    return (int)s->lastB;
}

extern "C" int fg_scorer_propose(fg_scorer *s, fg_ctx *ctx, int candidates, int every, uint32_t round) {
    if (!s || candidates < 1 || every < 1) return sfail("fg_scorer_propose: bad arguments");
    fgrt::PoolView v;
    if (!fgrt::pool_view(ctx, &v)) return sfail("fg_scorer_propose: context not initialised");
    if (v.ring != FG_RING_Z2 || v.maxm != s->M) return sfail("fg_scorer_propose: pool does not match the scorer");
    if (s->M <= fgk::sc::kProposeMaxTerms && s->F <= fgk::sc::kProposeMaxF && candidates <= fgk::sc::kProposeMaxK)
        return propose_fused(s, v, candidates, every, round);
    s->lastB = s->lastK = 0;
    const int chunk = 4096;
    const int L = (v.W + every - 1) / every;
    if (ensure(s, std::min(chunk, L), std::min(chunk, L) * candidates, false)) return -1;
    std::vector<int32_t> map(chunk);
    for (int b0 = 0; b0 < L; b0 += chunk) {
        const int B = std::min(chunk, L - b0);
        for (int q = 0; q < B; ++q) map[q] = (b0 + q) * every;
        SCHK(hipMemcpyAsync(s->map, map.data(), (size_t)B * 4, hipMemcpyHostToDevice, s->stream));
        fgk::sc::CandArgs c{};
        c.words = v.cur;
        c.m = v.m;
        c.map = s->map;
        c.W = v.W;
        c.B = B;
        c.M = s->M;
        c.K = candidates;
        c.key0 = v.key0;
        c.key1 = v.key1;
        c.round = round;
        c.pairs = s->pairs;
        c.tup = s->tup;
        SCHK(fgk::sc::launch_candidates(c, s->stream));
        if (featurize(s, v.cur, v.m, s->map, v.W, B) || forward(s, B, B * candidates)) return -1;
        SCHK(fgk::sc::launch_choose(s->s, s->pairs, s->map, B, candidates, v.forced, s->stream));
        SCHK(hipStreamSynchronize(s->stream));
        // the last chunk's candidates and scores stay readable (fg_scorer_last_proposals)
        s->lastB = B;
        s->lastK = candidates;
        s->last_fused = false;
    }
    return 0;
}

extern "C" int fg_walk_learned(fg_ctx *ctx, fg_scorer *s, int iterations, int candidates, int every, uint32_t round) {
    const int rc = fg_scorer_propose(s, ctx, candidates, every, round);
    if (rc < 0) return rc;
    return fgrt::walk_forced(ctx, iterations);
}

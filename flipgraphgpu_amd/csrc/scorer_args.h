// Kernel interfaces of the learned flip scorer: a term-set network in BF16
// with FP32 accumulation on the MFMA units (scorer_kernels.hip), driven by
// scorer.cpp.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fgk {
namespace sc {

typedef __bf16 bf16;

// C(i, j) = epilogue(sum_k A(i, k) B(k, j)) with strided bf16 operands:
// A(i, k) = A[i * sa_m + k * sa_k], B(k, j) = B[k * sb_k + j * sb_n].
// Epilogue: + bias[j] (if bias), relu (if relu), + C(i, j) (if accumulate);
// the result goes to C (fp32, if non-null) and Cb (bf16, if non-null).
struct GemmArgs {
    const bf16 *A;
    long sa_m, sa_k;
    const bf16 *B;
    long sb_k, sb_n;
    float *C;
    long ldc;
    bf16 *Cb;
    long ldcb;
    const float *bias;
    int M, N, K;
    int relu;
    int accumulate;
};

hipError_t launch_gemm(const GemmArgs &g, hipStream_t s);

// Term features: row (b * M + t) holds the bits of factors U, V, W of term t
// of tuple b (nn0 + nn1 + nn2 columns, F = their sum), zero for t >= m_b.
// With w = map[b] (or b without a map) the words are tuple-major
// words[(w * 3 + c) * M + t], or pool planes words[(c * M + t) * W + w]
// when pool_stride = W > 0.
struct FeatArgs {
    const uint32_t *words;
    const int32_t *m;
    const int32_t *map;  // tuple -> walker (pool mode) or null
    long pool_stride;    // walkers in the pool (pool mode), 0 for tuple-major words
    int B, M, nn0, nn1, nn2;
    bf16 *X;             // [B * M][F]
    float *mask;         // [B * M] 1 for valid rows
    float *inv_n;        // [B] 1 / m_b
};
hipError_t launch_featurize(const FeatArgs &f, hipStream_t s);

// Per-tuple mean over valid rows of an fp32 [B * M][D] matrix: out[b][d]
// (fp32) and its bf16 copy.
hipError_t launch_seg_mean(const float *Hrows, const float *mask, const float *inv_n, int B, int M, int D,
                           float *out, bf16 *outb, hipStream_t s);
// Per-tuple sum over valid rows (backward of the mean, before 1/n).
hipError_t launch_seg_sum(const float *G, const float *mask, int B, int M, int D, float *out, bf16 *outb,
                          hipStream_t s);
// H = relu(P + Q[row / M] + bias) over rows, masked; writes fp32 and bf16.
hipError_t launch_set_combine(const float *P, const float *Q, const float *bias, const float *mask, int B, int M,
                              int D, float *H, bf16 *Hb, hipStream_t s);
// Pair inputs: Qp[p] = [g_t1 + g_t2, g_t1 * g_t2] for pair p of tuple tup[p].
hipError_t launch_pair_gather(const float *G, const int32_t *tup, const int32_t *pairs, int P, int M, int D,
                              bf16 *Qp, hipStream_t s);
// Scores s[p] = z[p] . wo[c_p] + bo[c_p].
hipError_t launch_head_out(const float *Z, const float *wo, const float *bo, const int32_t *pairs, int P, int Zd,
                           float *s, hipStream_t s_);
// BCE with logits: loss sum into *loss, ds = (sigmoid(s) - y) / P.
hipError_t launch_bce(const float *s, const float *y, int P, float *ds, float *loss, hipStream_t st);
// Backward of the output layer: dZ = ds * wo[c] masked by Z > 0 (fp32 + bf16),
// dwo[c] += ds * Z, dbo[c] += ds.
hipError_t launch_head_back(const float *Z, const float *ds, const float *wo, const int32_t *pairs, int P, int Zd,
                            float *dZ, bf16 *dZb, float *dwo, float *dbo, hipStream_t st);
// dG rows (fp32, accumulated) from dQp: dg_t1 += d1 + d2 * g_t2, dg_t2 += d1 + d2 * g_t1.
hipError_t launch_pair_scatter(const float *dQ, const float *G, const int32_t *tup, const int32_t *pairs, int P, int M,
                               int D, float *dG, hipStream_t st);
// dPre = (dH + dmean[row / M] * inv_n[b]) * (H > 0) * mask, fp32 + bf16.
hipError_t launch_set_back(const float *dH, const float *dmean, const float *inv_n, const float *H, const float *mask,
                           int B, int M, int D, float *dPre, bf16 *dPreb, hipStream_t st);
// Column sums of an fp32 matrix (bias gradients), accumulated into out.
hipError_t launch_colsum(const float *G, long rows, int D, float *out, hipStream_t st);
// Adam with bias correction on n parameters; refreshes the bf16 copy.
hipError_t launch_adam(float *w, bf16 *wb, const float *g, float *m1, float *m2, long n, float lr, float b1, float b2,
                       float eps, int step, hipStream_t st);
hipError_t launch_to_bf16(const float *x, bf16 *y, long n, hipStream_t st);
hipError_t launch_fill_f32(float *p, float v, long n, hipStream_t st);

// Candidate flips of pool walkers: for walker map[b], k random equal-factor
// pairs (c, t1 < t2) drawn uniformly from all such pairs (with
// replacement), Philox stream (draw, walker, 6) keyed by seed.  pairs[(b * K + k) * 3 + {c, t1, t2}]
// (c = -1 when the walker has no pair).
struct CandArgs {
    const uint32_t *words;  // pool planes words[(c * M + t) * W + w]
    const int32_t *m;
    const int32_t *map;
    long W;
    int B, M, K;
    uint32_t key0, key1;
    uint32_t round;
    int32_t *pairs;
    int32_t *tup;  // [B * K] tuple index of each pair (b)
};
hipError_t launch_candidates(const CandArgs &a, hipStream_t st);
// forced[map[b] * 4 + ...] = best-scoring candidate (c, t1, t2, 1), or c = -1.
hipError_t launch_choose(const float *scores, const int32_t *pairs, const int32_t *map, int B, int K, int32_t *forced,
                         hipStream_t st);

// Fused proposal round (inference only): one workgroup of 4 waves scores the
// K candidate pairs of one pool walker with the whole network kept on chip --
// term features, the three set layers on the MFMA units (each wave owns 32
// of the 128 columns for all terms), the per-scheme means, the pair head --
// and writes the best candidate into the proposal array.  The rounding
// points are those of the unfused forward pass: every MFMA operand is bf16,
// accumulation and the elementwise steps are FP32.  Schemes of up to
// kProposeMaxTerms terms and kProposeMaxF feature bits; K <= kProposeMaxK.
constexpr int kProposeMaxTerms = 64;
constexpr int kProposeMaxF = 96;
constexpr int kProposeMaxK = 16;
constexpr int kPD_ = 128;  // hidden width of the network
struct ProposeArgs {
    const uint32_t *words;  // pool planes words[(c * M + t) * W + w]
    const int32_t *m;
    const int32_t *map;     // tuple -> walker
    long W;
    int M;                  // pool term capacity (<= kProposeMaxTerms)
    int nn0, nn1, nn2;      // F = nn0 + nn1 + nn2 <= kProposeMaxF
    int B, K;
    const int32_t *pairs;   // [B * K][3] candidates (component -1: none)
    const bf16 *W0t;        // [128][kProposeMaxF]: W0 transposed, zero beyond F
    const bf16 *A1t, *A2t;  // [128][128]: A1, A2 transposed
    const bf16 *B1t, *B2t;  // [128][128]: B1, B2 transposed
    const bf16 *Wz;         // [256][64] as stored
    const bf16 *Wzt;        // [64][256]: Wz transposed
    const float *b0, *c1, *c2, *bz, *wo, *bo;
    float *scores;          // [B * K] (may be null)
    int32_t *forced;        // [W][4] proposals
};
hipError_t launch_propose(const ProposeArgs &a, hipStream_t st);
// Column-major bf16 copy of a [rows][cols] row-major bf16 matrix, rows padded
// with zeros to ld: out[c * ld + r].
hipError_t launch_transpose_bf16(const bf16 *in, int rows, int cols, int ld, bf16 *out, hipStream_t st);

}  // namespace sc
}  // namespace fgk

// Kernels of the learned flip scorer for CDNA4 (gfx950): a strided BF16
// GEMM on the MFMA units (v_mfma_f32_16x16x32_bf16, FP32 accumulation) and
// the elementwise / segmented kernels of the term-set network's forward and
// backward passes, plus candidate collection from the walker pool.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lane_common.h"
#include "scorer_args.h"

namespace fgk {
namespace sc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 64;   // output tile per workgroup (64 x 64)
constexpr int kTK = 32;     // K step (one MFMA)
constexpr int kPad = 8;     // LDS row padding (bf16 elements)

// 256 threads = 4 waves in a 2 x 2 arrangement, each wave owns a 32 x 32
// block made of 2 x 2 MFMA 16 x 16 tiles.  Operand tiles are staged in LDS:
// As[row][k] and Bs[col][k], so that each lane reads its 8 consecutive k
// values of one row (A) or one column (B) with a 16-byte load.
__global__ void __launch_bounds__(256) gemm_kernel(GemmArgs g) {
    __shared__ __attribute__((aligned(16))) bf16 As[kTile][kTK + kPad];
    __shared__ __attribute__((aligned(16))) bf16 Bs[kTile][kTK + kPad];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const long i0 = (long)blockIdx.y * kTile, j0 = (long)blockIdx.x * kTile;
    floatx4 acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
    const bf16 zero = (bf16)0.0f;
    for (int k0 = 0; k0 < g.K; k0 += kTK) {
        for (int q = 0; q < (kTile * kTK) / 256; ++q) {
            const int idx = tid + 256 * q;
            const int r = idx / kTK, k = idx % kTK;  // A: k fastest
            const long gi = i0 + r, gk = k0 + k;
            As[r][k] = (gi < g.M && gk < g.K) ? g.A[gi * g.sa_m + gk * g.sa_k] : zero;
            const int kb = idx / kTile, n = idx % kTile;  // B: n fastest
            const long gj = j0 + n, gkb = k0 + kb;
            Bs[n][kb] = (gj < g.N && gkb < g.K) ? g.B[gkb * g.sb_k + gj * g.sb_n] : zero;
        }
        __syncthreads();
        const int kk = 8 * (lane >> 4);
        bf16x8 af[2], bfv[2];
        for (int a = 0; a < 2; ++a)
            af[a] = *reinterpret_cast<const bf16x8 *>(&As[wm * 32 + a * 16 + (lane & 15)][kk]);
        for (int b = 0; b < 2; ++b)
            bfv[b] = *reinterpret_cast<const bf16x8 *>(&Bs[wn * 32 + b * 16 + (lane & 15)][kk]);
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfv[b], acc[a][b], 0, 0, 0);
        __syncthreads();
    }
    // D[i][j]: lane holds rows 4 * (lane / 16) + v of column lane % 16
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int v = 0; v < 4; ++v) {
                const long i = i0 + wm * 32 + a * 16 + 4 * (lane >> 4) + v;
                const long j = j0 + wn * 32 + b * 16 + (lane & 15);
                if (i >= g.M || j >= g.N) continue;
                float x = acc[a][b][v];
                if (g.bias) x += g.bias[j];
                if (g.relu) x = x > 0.f ? x : 0.f;
                if (g.accumulate) x += g.C[i * g.ldc + j];
                if (g.C) g.C[i * g.ldc + j] = x;
                if (g.Cb) g.Cb[i * g.ldcb + j] = (bf16)x;
            }
}

hipError_t launch_gemm(const GemmArgs &g, hipStream_t s) {
    if (g.M <= 0 || g.N <= 0) return hipSuccess;
    dim3 grid((unsigned)((g.N + kTile - 1) / kTile), (unsigned)((g.M + kTile - 1) / kTile));
    hipLaunchKernelGGL(gemm_kernel, grid, dim3(256), 0, s, g);
    return hipGetLastError();
}

static inline unsigned blocks(long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

__global__ void featurize_kernel(FeatArgs f) {
    const int F = f.nn0 + f.nn1 + f.nn2;
    const long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= (long)f.B * f.M) return;
    const int b = (int)(row / f.M), t = (int)(row % f.M);
    const long w = f.map ? f.map[b] : b;
    const int m = f.m[w];
    const bool valid = t < m;
    f.mask[row] = valid ? 1.f : 0.f;
    if (t == 0) f.inv_n[b] = m > 0 ? 1.f / (float)m : 0.f;
    const int nn[3] = {f.nn0, f.nn1, f.nn2};
    int col = 0;
    for (int c = 0; c < 3; ++c) {
        uint32_t x = 0;
        if (valid) x = f.pool_stride > 0 ? f.words[((long)c * f.M + t) * f.pool_stride + w] : f.words[((long)w * 3 + c) * f.M + t];
        for (int q = 0; q < nn[c]; ++q) f.X[row * F + col + q] = (bf16)(float)((x >> q) & 1u);
        col += nn[c];
    }
}

hipError_t launch_featurize(const FeatArgs &f, hipStream_t s) {
    hipLaunchKernelGGL(featurize_kernel, dim3(blocks((long)f.B * f.M)), dim3(256), 0, s, f);
    return hipGetLastError();
}

// one thread per (tuple, column)
__global__ void seg_mean_kernel(const float *H, const float *mask, const float *inv_n, int B, int M, int D, float *out,
                                bf16 *outb, int scale) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)B * D) return;
    const int b = (int)(g / D), d = (int)(g % D);
    float s = 0.f;
    for (int t = 0; t < M; ++t) {
        const long r = (long)b * M + t;
        if (mask[r] != 0.f) s += H[r * D + d];
    }
    if (scale) s *= inv_n[b];
    out[g] = s;
    if (outb) outb[g] = (bf16)s;
}

hipError_t launch_seg_mean(const float *H, const float *mask, const float *inv_n, int B, int M, int D, float *out,
                           bf16 *outb, hipStream_t s) {
    hipLaunchKernelGGL(seg_mean_kernel, dim3(blocks((long)B * D)), dim3(256), 0, s, H, mask, inv_n, B, M, D, out, outb, 1);
    return hipGetLastError();
}

hipError_t launch_seg_sum(const float *G, const float *mask, int B, int M, int D, float *out, bf16 *outb,
                          hipStream_t s) {
    hipLaunchKernelGGL(seg_mean_kernel, dim3(blocks((long)B * D)), dim3(256), 0, s, G, mask, (const float *)nullptr, B,
                       M, D, out, outb, 0);
    return hipGetLastError();
}

__global__ void set_combine_kernel(const float *P, const float *Q, const float *bias, const float *mask, int B, int M,
                                   int D, float *H, bf16 *Hb) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)B * M * D) return;
    const long row = g / D;
    const int d = (int)(g % D), b = (int)(row / M);
    float x = P[g] + Q[(long)b * D + d] + bias[d];
    x = (x > 0.f && mask[row] != 0.f) ? x : 0.f;
    H[g] = x;
    Hb[g] = (bf16)x;
}

hipError_t launch_set_combine(const float *P, const float *Q, const float *bias, const float *mask, int B, int M,
                              int D, float *H, bf16 *Hb, hipStream_t s) {
    hipLaunchKernelGGL(set_combine_kernel, dim3(blocks((long)B * M * D)), dim3(256), 0, s, P, Q, bias, mask, B, M, D, H, Hb);
    return hipGetLastError();
}

__global__ void pair_gather_kernel(const float *G, const int32_t *tup, const int32_t *pairs, int P, int M, int D,
                                   bf16 *Qp) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)P * D) return;
    const int p = (int)(g / D), d = (int)(g % D);
    const int c = pairs[3 * p];
    float s = 0.f, m = 0.f;
    if (c >= 0) {
        const long b = tup[p];
        const float x = G[(b * M + pairs[3 * p + 1]) * D + d], y = G[(b * M + pairs[3 * p + 2]) * D + d];
        s = x + y;
        m = x * y;
    }
    Qp[(long)p * 2 * D + d] = (bf16)s;
    Qp[(long)p * 2 * D + D + d] = (bf16)m;
}

hipError_t launch_pair_gather(const float *G, const int32_t *tup, const int32_t *pairs, int P, int M, int D, bf16 *Qp,
                              hipStream_t s) {
    hipLaunchKernelGGL(pair_gather_kernel, dim3(blocks((long)P * D)), dim3(256), 0, s, G, tup, pairs, P, M, D, Qp);
    return hipGetLastError();
}

__global__ void head_out_kernel(const float *Z, const float *wo, const float *bo, const int32_t *pairs, int P, int Zd,
                                float *s) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int c = pairs[3 * p];
    if (c < 0) {
        s[p] = -1e30f;
        return;
    }
    float x = bo[c];
    for (int q = 0; q < Zd; ++q) x += Z[(long)p * Zd + q] * wo[c * Zd + q];
    s[p] = x;
}

hipError_t launch_head_out(const float *Z, const float *wo, const float *bo, const int32_t *pairs, int P, int Zd,
                           float *s, hipStream_t st) {
    hipLaunchKernelGGL(head_out_kernel, dim3(blocks(P)), dim3(256), 0, st, Z, wo, bo, pairs, P, Zd, s);
    return hipGetLastError();
}

// One workgroup: the mean BCE loss summed in a fixed order (each thread a
// fixed stride of pairs, then a fixed tree), so training is reproducible.
__global__ void bce_kernel(const float *s, const float *y, int P, float *ds, float *loss) {
    float acc = 0.f;
    for (int p = threadIdx.x; p < P; p += blockDim.x) {
        const float x = s[p];
        const float sp = x > 0.f ? x + log1pf(expf(-x)) : log1pf(expf(x));  // softplus
        acc += (sp - y[p] * x) / (float)P;
        const float sig = 1.f / (1.f + expf(-x));
        ds[p] = (sig - y[p]) / (float)P;
    }
    __shared__ float part[256];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss += part[0];
}

hipError_t launch_bce(const float *s, const float *y, int P, float *ds, float *loss, hipStream_t st) {
    hipLaunchKernelGGL(bce_kernel, dim3(1), dim3(256), 0, st, s, y, P, ds, loss);
    return hipGetLastError();
}

__global__ void head_back_kernel(const float *Z, const float *ds, const float *wo, const int32_t *pairs, int P, int Zd,
                                 float *dZ, bf16 *dZb) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)P * Zd) return;
    const int p = (int)(g / Zd), q = (int)(g % Zd);
    const int c = pairs[3 * p];
    float d = 0.f;
    if (c >= 0) {
        const float z = Z[g];
        d = z > 0.f ? ds[p] * wo[c * Zd + q] : 0.f;
    }
    dZ[g] = d;
    dZb[g] = (bf16)d;
}

// Output-layer gradients: one thread per (component c, column q) sums over
// the pairs in order (q == Zd: the bias), so the sums are reproducible.
__global__ void head_grad_kernel(const float *Z, const float *ds, const int32_t *pairs, int P, int Zd, float *dwo,
                                 float *dbo) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= 3 * (Zd + 1)) return;
    const int c = g / (Zd + 1), q = g % (Zd + 1);
    float acc = 0.f;
    for (int p = 0; p < P; ++p)
        if (pairs[3 * p] == c) acc += q < Zd ? ds[p] * Z[(long)p * Zd + q] : ds[p];
    if (q < Zd) dwo[c * Zd + q] += acc;
    else dbo[c] += acc;
}

hipError_t launch_head_back(const float *Z, const float *ds, const float *wo, const int32_t *pairs, int P, int Zd,
                            float *dZ, bf16 *dZb, float *dwo, float *dbo, hipStream_t st) {
    hipLaunchKernelGGL(head_back_kernel, dim3(blocks((long)P * Zd)), dim3(256), 0, st, Z, ds, wo, pairs, P, Zd, dZ, dZb);
    hipLaunchKernelGGL(head_grad_kernel, dim3(blocks(3 * (Zd + 1))), dim3(256), 0, st, Z, ds, pairs, P, Zd, dwo, dbo);
    return hipGetLastError();
}

// Backward passes score one pair per tuple (index1 < index2), so the two
// rows a thread updates are written by no other thread: plain adds, no
// atomics, and the result does not depend on scheduling.
__global__ void pair_scatter_kernel(const float *dQ, const float *G, const int32_t *tup, const int32_t *pairs, int P,
                                    int M, int D, float *dG) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)P * D) return;
    const int p = (int)(g / D), d = (int)(g % D);
    if (pairs[3 * p] < 0) return;
    const long b = tup[p];
    const long r1 = b * M + pairs[3 * p + 1], r2 = b * M + pairs[3 * p + 2];
    const float d1 = dQ[(long)p * 2 * D + d], d2 = dQ[(long)p * 2 * D + D + d];
    dG[r1 * D + d] += d1 + d2 * G[r2 * D + d];
    dG[r2 * D + d] += d1 + d2 * G[r1 * D + d];
}

hipError_t launch_pair_scatter(const float *dQ, const float *G, const int32_t *tup, const int32_t *pairs, int P, int M,
                               int D, float *dG, hipStream_t st) {
    hipLaunchKernelGGL(pair_scatter_kernel, dim3(blocks((long)P * D)), dim3(256), 0, st, dQ, G, tup, pairs, P, M, D, dG);
    return hipGetLastError();
}

__global__ void set_back_kernel(const float *dH, const float *dmean, const float *inv_n, const float *H,
                                const float *mask, int B, int M, int D, float *dPre, bf16 *dPreb) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)B * M * D) return;
    const long row = g / D;
    const int d = (int)(g % D), b = (int)(row / M);
    float x = dH[g] + (dmean ? dmean[(long)b * D + d] * inv_n[b] : 0.f);
    x = (H[g] > 0.f && mask[row] != 0.f) ? x : 0.f;
    dPre[g] = x;
    dPreb[g] = (bf16)x;
}

hipError_t launch_set_back(const float *dH, const float *dmean, const float *inv_n, const float *H, const float *mask,
                           int B, int M, int D, float *dPre, bf16 *dPreb, hipStream_t st) {
    hipLaunchKernelGGL(set_back_kernel, dim3(blocks((long)B * M * D)), dim3(256), 0, st, dH, dmean, inv_n, H, mask, B, M,
                       D, dPre, dPreb);
    return hipGetLastError();
}

// one thread per column, strided over rows
__global__ void colsum_kernel(const float *G, long rows, int D, float *out) {
    const int d = blockIdx.x;
    float s = 0.f;
    for (long r = threadIdx.x; r < rows; r += blockDim.x) s += G[r * D + d];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    __shared__ float part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[d] += part[0] + part[1] + part[2] + part[3];
}

hipError_t launch_colsum(const float *G, long rows, int D, float *out, hipStream_t st) {
    hipLaunchKernelGGL(colsum_kernel, dim3(D), dim3(256), 0, st, G, rows, D, out);
    return hipGetLastError();
}

__global__ void adam_kernel(float *w, bf16 *wb, const float *g, float *m1, float *m2, long n, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi, b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    const float x = w[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
    w[i] = x;
    if (wb) wb[i] = (bf16)x;
}

hipError_t launch_adam(float *w, bf16 *wb, const float *g, float *m1, float *m2, long n, float lr, float b1, float b2,
                       float eps, int step, hipStream_t st) {
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(blocks(n)), dim3(256), 0, st, w, wb, g, m1, m2, n, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

__global__ void to_bf16_kernel(const float *x, bf16 *y, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (bf16)x[i];
}

hipError_t launch_to_bf16(const float *x, bf16 *y, long n, hipStream_t st) {
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks(n)), dim3(256), 0, st, x, y, n);
    return hipGetLastError();
}

__global__ void fill_f32_kernel(float *p, float v, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t launch_fill_f32(float *p, float v, long n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_f32_kernel, dim3(blocks(n)), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

// One wave per walker: equal-factor pairs per (component, t1) row are
// counted into LDS, then each of the K draws picks a row by prefix count
// and the partner by a scan of that row.
__global__ void __launch_bounds__(64) candidates_kernel(CandArgs a) {
    extern __shared__ uint32_t cl[];
    const int b = blockIdx.x, lane = threadIdx.x;
    if (b >= a.B) return;
    const long w = a.map ? a.map[b] : b;
    const int m = a.m[w];
    uint32_t *F = cl;                          // [3][M]
    int32_t *cnt = (int32_t *)(cl + 3 * a.M);  // [3 * M] inclusive prefix of pair counts
    for (int q = lane; q < 3 * m; q += 64) {
        const int c = q / m, t = q % m;
        F[c * a.M + t] = a.words[((long)c * a.M + t) * a.W + w];
    }
    __syncthreads();
    int carry = 0;
    for (int base = 0; base < 3 * m; base += 64) {
        const int q = base + lane;
        int k = 0;
        if (q < 3 * m) {
            const int c = q / m, t = q % m;
            const uint32_t v = F[c * a.M + t];
            for (int u = t + 1; u < m; ++u) k += F[c * a.M + u] == v;
        }
        int x = k;
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(x, d);
            if (lane >= d) x += y;
        }
        if (q < 3 * m) cnt[q] = carry + x;
        carry += __shfl(x, 63);
    }
    __syncthreads();
    const int total = carry;
    for (int k = lane; k < a.K; k += 64) {
        int32_t *out = a.pairs + ((long)b * a.K + k) * 3;
        a.tup[(long)b * a.K + k] = b;
        if (total == 0) {
            out[0] = -1;
            out[1] = out[2] = 0;
            continue;
        }
        uint32_t o0, o1, o2, o3;
        lane::philox10(a.round, (uint32_t)k, (uint32_t)w, 6u, a.key0, a.key1, o0, o1, o2, o3);
        const int r = (int)lane::draw_below(o0, (uint32_t)total);
        int lo = 0, hi = 3 * m - 1;  // first row with cnt > r
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (cnt[mid] > r) hi = mid;
            else lo = mid + 1;
        }
        const int c = lo / m, t = lo % m;
        int skip = r - (lo > 0 ? cnt[lo - 1] : 0);
        const uint32_t v = F[c * a.M + t];
        int t2 = t + 1;
        for (; t2 < m; ++t2)
            if (F[c * a.M + t2] == v && skip-- == 0) break;
        out[0] = c;
        out[1] = t;
        out[2] = t2;
    }
}

__global__ void transpose_bf16_kernel(const bf16 *in, int rows, int cols, int ld, bf16 *out) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)cols * ld) return;
    const int c = (int)(g / ld), r = (int)(g % ld);
    out[g] = r < rows ? in[(long)r * cols + c] : (bf16)0.0f;
}

hipError_t launch_transpose_bf16(const bf16 *in, int rows, int cols, int ld, bf16 *out, hipStream_t st) {
    hipLaunchKernelGGL(transpose_bf16_kernel, dim3(blocks((long)cols * ld)), dim3(256), 0, st, in, rows, cols, ld, out);
    return hipGetLastError();
}

namespace {
constexpr int kPD = 128;                         // hidden width
constexpr int kPZ = 64;                          // pair-head width
constexpr int kPXs = kProposeMaxF + 8;           // LDS row stride of the features (bf16)
constexpr int kPEs = kPD + 8;                    // LDS row stride of the activations (bf16)
constexpr int kPSmem = kProposeMaxTerms * kPD * 4;  // G in FP32; aliases the feature and activation tiles
constexpr int kPQs = 2 * kPD + 8;                // LDS row stride of the pair-head rows (bf16)
static_assert(kProposeMaxTerms * (kPXs + kPEs) * 2 <= kPSmem, "feature + activation tiles must fit the G tile");
static_assert(16 * kPQs * 2 <= kPSmem && kProposeMaxK <= 16, "pair-head rows must fit the G tile");

// acc[rt][ct] += A[64 x K] (LDS, row stride lda) * Bt^T, for this wave's 32
// columns; Bt is the transposed weight [128][ldb] in global memory.
__device__ __forceinline__ void propose_mma(floatx4 (&acc)[4][2], const bf16 *A, int lda, const bf16 *Bt, int ldb,
                                            int K, int wave, int lane) {
    for (int k0 = 0; k0 < K; k0 += 32) {
        bf16x8 bfv[2];
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
            bfv[ct] = *reinterpret_cast<const bf16x8 *>(Bt + (long)(32 * wave + 16 * ct + (lane & 15)) * ldb + k0 +
                                                        8 * (lane >> 4));
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            const bf16x8 af = *reinterpret_cast<const bf16x8 *>(A + (rt * 16 + (lane & 15)) * lda + k0 + 8 * (lane >> 4));
#pragma unroll
            for (int ct = 0; ct < 2; ++ct) acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bfv[ct], acc[rt][ct], 0, 0, 0);
        }
    }
}

// q[col] = sum_k mean[k] B[k][col] for this wave's 32 columns, on the MFMA
// units: the mean is row 0 of an otherwise zero 16-row A tile, Bt is B
// transposed [col][k]; row 0 of the result sits in lanes 0-15 (element 0).
__device__ __forceinline__ void propose_mean_product(const bf16 *Mv, const bf16 *Bt, float *Qv, int wave, int lane) {
    floatx4 q[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
    const bf16 zero = (bf16)0.0f;
#pragma unroll
    for (int k0 = 0; k0 < kPD; k0 += 32) {
        bf16x8 af;
        if ((lane & 15) == 0) {
            af = *reinterpret_cast<const bf16x8 *>(Mv + k0 + 8 * (lane >> 4));
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) af[e] = zero;
        }
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
            const bf16x8 bfv =
                *reinterpret_cast<const bf16x8 *>(Bt + (long)(32 * wave + 16 * ct + (lane & 15)) * kPD + k0 + 8 * (lane >> 4));
            q[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bfv, q[ct], 0, 0, 0);
        }
    }
    if (lane < 16) {
        Qv[32 * wave + lane] = q[0][0];
        Qv[32 * wave + 16 + lane] = q[1][0];
    }
}
}  // namespace

__global__ void __launch_bounds__(256) propose_kernel(ProposeArgs a) {
    __shared__ __attribute__((aligned(16))) unsigned char smem[kPSmem];
    __shared__ __attribute__((aligned(16))) bf16 Mv[kPD];
    __shared__ float Qv[kPD];
    __shared__ float sc[kProposeMaxK];
    __shared__ float scw[4][kProposeMaxK];  // head partial sums of each wave, added in wave order
    bf16 *Es = reinterpret_cast<bf16 *>(smem);               // [64][kPEs]
    bf16 *Xs = Es + kProposeMaxTerms * kPEs;                  // [64][kPXs]
    float *Gs = reinterpret_cast<float *>(smem);             // [64][128], after the last layer
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int b = blockIdx.x;
    const long w = a.map[b];
    const int m = a.m[w];
    const float inv_n = m > 0 ? 1.f / (float)m : 0.f;
    const bf16 one = (bf16)1.0f, zero = (bf16)0.0f;

    // term features: the factor bits of U, V, W (zero rows beyond the rank)
    for (int q = tid; q < kProposeMaxTerms * kPXs; q += 256) Xs[q] = zero;
    __syncthreads();
    if (tid < 3 * kProposeMaxTerms) {
        const int c = tid / kProposeMaxTerms, t = tid % kProposeMaxTerms;
        if (t < m) {
            const uint32_t x = a.words[((long)c * a.M + t) * a.W + w];
            const int off = c == 0 ? 0 : (c == 1 ? a.nn0 : a.nn0 + a.nn1);
            const int nnc = c == 0 ? a.nn0 : (c == 1 ? a.nn1 : a.nn2);
            for (int q = 0; q < nnc; ++q) Xs[t * kPXs + off + q] = ((x >> q) & 1u) ? one : zero;
        }
    }
    __syncthreads();

    floatx4 acc[4][2];
    auto clear = [&]() {
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int ct = 0; ct < 2; ++ct) acc[rt][ct] = floatx4{0.f, 0.f, 0.f, 0.f};
    };
    // column sums of the valid rows of this lane's tiles, reduced over the 4
    // row groups of the wave; then the bf16 mean goes to Mv
    auto means = [&](float (&v)[4][2][4]) {
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
            float sum = 0.f;
#pragma unroll
            for (int rt = 0; rt < 4; ++rt)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (rt * 16 + 4 * (lane >> 4) + q < m) sum += v[rt][ct][q];
            sum += __shfl_xor(sum, 16);
            sum += __shfl_xor(sum, 32);
            if (lane < 16) Mv[32 * wave + 16 * ct + lane] = (bf16)(sum * inv_n);
        }
    };
    float v[4][2][4];

    // e = relu(x W0 + b0); the K loop stops at the factor bits rounded up to
    // the MFMA's K step (the feature and weight columns beyond them are zero)
    clear();
    const int kF = (a.nn0 + a.nn1 + a.nn2 + 31) & ~31;
    propose_mma(acc, Xs, kPXs, a.W0t, kProposeMaxF, kF, wave, lane);
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int col = 32 * wave + 16 * ct + (lane & 15);
                const float x = acc[rt][ct][q] + a.b0[col];
                v[rt][ct][q] = x > 0.f ? x : 0.f;
            }
    means(v);
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                Es[(rt * 16 + 4 * (lane >> 4) + q) * kPEs + 32 * wave + 16 * ct + (lane & 15)] = (bf16)v[rt][ct][q];
    __syncthreads();

    // h = relu(e A1 + mean(e) B1 + c1), then g = relu(h A2 + mean(h) B2 + c2)
    for (int layer = 0; layer < 2; ++layer) {
        const bf16 *At = layer ? a.A2t : a.A1t;
        const bf16 *Bm = layer ? a.B2t : a.B1t;
        const float *cb = layer ? a.c2 : a.c1;
        propose_mean_product(Mv, Bm, Qv, wave, lane);
        clear();
        propose_mma(acc, Es, kPEs, At, kPD, kPD, wave, lane);
        __syncthreads();  // every wave is done with Es, Qv is complete
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int ct = 0; ct < 2; ++ct)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int col = 32 * wave + 16 * ct + (lane & 15), row = rt * 16 + 4 * (lane >> 4) + q;
                    const float x = acc[rt][ct][q] + Qv[col] + cb[col];
                    v[rt][ct][q] = (x > 0.f && row < m) ? x : 0.f;
                }
        if (layer == 0) {
            means(v);
#pragma unroll
            for (int rt = 0; rt < 4; ++rt)
#pragma unroll
                for (int ct = 0; ct < 2; ++ct)
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        Es[(rt * 16 + 4 * (lane >> 4) + q) * kPEs + 32 * wave + 16 * ct + (lane & 15)] = (bf16)v[rt][ct][q];
        } else {
#pragma unroll
            for (int rt = 0; rt < 4; ++rt)
#pragma unroll
                for (int ct = 0; ct < 2; ++ct)
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        Gs[(rt * 16 + 4 * (lane >> 4) + q) * kPD + 32 * wave + 16 * ct + (lane & 15)] = v[rt][ct][q];
        }
        __syncthreads();
    }

    // pair head: z = relu([g_a + g_b, g_a * g_b] Wz + bz), s = z . wo[c] + bo[c],
    // up to 16 candidates at once.  The candidate rows q = [bf16(g_a + g_b),
    // bf16(g_a * g_b)] are built from G (thread: candidate tid / 16, hidden
    // units 8 (tid % 16) .. + 7), then written as a [16][256] bf16 tile over
    // G; each wave multiplies it by its 16 columns of Wz on the MFMA units.
    bf16 *Qs = reinterpret_cast<bf16 *>(smem);  // [16][kPQs], after G is read
    {
        const int k = tid / 16, i0 = 8 * (tid % 16);
        const long p = (long)b * a.K + k;
        const int c = k < a.K ? a.pairs[3 * p] : -1;
        bf16x8 qs, qm;
        if (c >= 0) {
            const int t1 = a.pairs[3 * p + 1], t2 = a.pairs[3 * p + 2];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float x = Gs[t1 * kPD + i0 + e], y = Gs[t2 * kPD + i0 + e];
                qs[e] = (bf16)(x + y);
                qm[e] = (bf16)(x * y);
            }
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) qs[e] = qm[e] = zero;
        }
        __syncthreads();  // every G read is done
        *reinterpret_cast<bf16x8 *>(Qs + k * kPQs + i0) = qs;
        *reinterpret_cast<bf16x8 *>(Qs + k * kPQs + kPD + i0) = qm;
        __syncthreads();
    }
    {
        floatx4 z = {0.f, 0.f, 0.f, 0.f};
        const int col = 16 * wave + (lane & 15);
#pragma unroll
        for (int k0 = 0; k0 < 2 * kPD; k0 += 32) {
            const bf16x8 af = *reinterpret_cast<const bf16x8 *>(Qs + (lane & 15) * kPQs + k0 + 8 * (lane >> 4));
            const bf16x8 bfv = *reinterpret_cast<const bf16x8 *>(a.Wzt + (long)col * (2 * kPD) + k0 + 8 * (lane >> 4));
            z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bfv, z, 0, 0, 0);
        }
        // lane holds candidates 4 (lane / 16) + q of column col
        float part[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = 4 * (lane >> 4) + q;
            const int c = k < a.K ? a.pairs[3 * ((long)b * a.K + k)] : -1;
            float x = z[q] + a.bz[col];
            x = x > 0.f ? x : 0.f;
            part[q] = c >= 0 ? x * a.wo[c * kPZ + col] : 0.f;
#pragma unroll
            for (int d = 8; d > 0; d >>= 1) part[q] += __shfl_xor(part[q], d);
        }
        if ((lane & 15) == 0)
#pragma unroll
            for (int q = 0; q < 4; ++q) scw[wave][4 * (lane >> 4) + q] = part[q];
    }
    __syncthreads();
    if (tid < a.K) {
        const long p = (long)b * a.K + tid;
        const int c = a.pairs[3 * p];
        const float hs = ((scw[0][tid] + scw[1][tid]) + scw[2][tid]) + scw[3][tid];
        const float s = c >= 0 ? a.bo[c] + hs : -1e30f;
        sc[tid] = s;
        if (a.scores) a.scores[p] = s;
    }
    __syncthreads();
    if (tid == 0) {  // the first best valid candidate, as the choose kernel picks it
        int best = -1;
        float bs = 0.f;
        for (int k = 0; k < a.K; ++k) {
            if (a.pairs[3 * ((long)b * a.K + k)] < 0) continue;
            if (best < 0 || sc[k] > bs) {
                best = k;
                bs = sc[k];
            }
        }
        int32_t *f = a.forced + w * 4;
        if (best < 0) {
            f[0] = -1;
            f[1] = f[2] = f[3] = 0;
        } else {
            const long p = (long)b * a.K + best;
            f[0] = a.pairs[3 * p];
            f[1] = a.pairs[3 * p + 1];
            f[2] = a.pairs[3 * p + 2];
            f[3] = 1;
        }
    }
}

hipError_t launch_propose(const ProposeArgs &a, hipStream_t st) {
    if (a.B <= 0) return hipSuccess;
    hipLaunchKernelGGL(propose_kernel, dim3((unsigned)a.B), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_candidates(const CandArgs &a, hipStream_t st) {
    const size_t lds = (size_t)6 * a.M * 4;
    hipLaunchKernelGGL(candidates_kernel, dim3(a.B), dim3(64), lds, st, a);
    return hipGetLastError();
}

__global__ void choose_kernel(const float *scores, const int32_t *pairs, const int32_t *map, int B, int K,
                              int32_t *forced) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    int best = -1;
    float bs = 0.f;
    for (int k = 0; k < K; ++k) {
        const long p = (long)b * K + k;
        if (pairs[3 * p] < 0) continue;
        if (best < 0 || scores[p] > bs) {
            best = k;
            bs = scores[p];
        }
    }
    const long w = map ? map[b] : b;
    int32_t *f = forced + w * 4;
    if (best < 0) {
        f[0] = -1;
        f[1] = f[2] = f[3] = 0;
        return;
    }
    const long p = (long)b * K + best;
    f[0] = pairs[3 * p];
    f[1] = pairs[3 * p + 1];
    f[2] = pairs[3 * p + 2];
    f[3] = 1;
}

hipError_t launch_choose(const float *scores, const int32_t *pairs, const int32_t *map, int B, int K, int32_t *forced,
                         hipStream_t st) {
    hipLaunchKernelGGL(choose_kernel, dim3(blocks(B)), dim3(256), 0, st, scores, pairs, map, B, K, forced);
    return hipGetLastError();
}

}  // namespace sc
}  // namespace fgk

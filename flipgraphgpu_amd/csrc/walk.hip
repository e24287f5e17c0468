// Flip-graph random walk kernels for CDNA4 (gfx950): the Z2 and integer
// ring kernels around the per-lane logic of walk_z2_lane.h /
// walk_zt_lane.h, the bit-sliced Brent checks, and their host launchers.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"
#include "walk_args.h"
#include "walk_z2_lane.h"
#include "walk_zt_lane.h"
#include "resize_ops.h"
#include "walk_persistent.h"

namespace fgk {
#if defined(FGK_PHASE_CLOCKS)
__device__ unsigned long long g_phase[16];
#endif

FGK_WALK_KERNEL walk_z2_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::Z2Lane, lane::kModeWalk>(A, lds);
}

FGK_WALK_KERNEL walk_zt_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::ZTLane, lane::kModeWalk>(A, lds);
}

// Z2 with 16-bit factor words in LDS (factors of up to 16 coefficients).
FGK_WALK_KERNEL walk_z2h_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::Z2Lane16, lane::kModeWalk>(A, lds);
}

// Z2 with class masks (at most 32 terms, factors of up to 16 coefficients).
FGK_WALK_KERNEL walk_z2m_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::Z2LaneM, lane::kModeWalk>(A, lds);
}

// Complexity minimizer (minimizeKernel, src/entities/complexity_minimizer.cu:190-214).
FGK_WALK_KERNEL minimize_z2_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::Z2Lane, lane::kModeMinimize>(A, lds);
}

FGK_WALK_KERNEL minimize_zt_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::ZTLane, lane::kModeMinimize>(A, lds);
}

FGK_WALK_KERNEL minimize_z2h_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::Z2Lane16, lane::kModeMinimize>(A, lds);
}

FGK_WALK_KERNEL minimize_z2m_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::Z2LaneM, lane::kModeMinimize>(A, lds);
}

// Scheme flips of the additions reducer (flipSchemesKernel,
// src/entities/scheme_additions_reducer.cu:448-460), integer ring.
FGK_WALK_KERNEL flips_zt_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::ZTLane, lane::kModeFlips>(A, lds);
}

// resizeKernel (src/entities/flip_graph.cu:575-613): Z2 with 32-bit words, and the integer ring.
FGK_WALK_KERNEL resize_z2_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::Z2Lane, lane::kModeResize>(A, lds);
}

FGK_WALK_KERNEL resize_zt_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::ZTLane, lane::kModeResize>(A, lds);
}

// Walk with the sandwiching and basis-change edges: Z2 with 32-bit words, and the
// integer ring (whose sandwiching is empty in the reference).
FGK_WALK_KERNEL walk_meta_z2_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::Z2Lane, lane::kModeMeta>(A, lds);
}

FGK_WALK_KERNEL walk_meta_zt_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::ZTLane, lane::kModeMeta>(A, lds);
}

// Walk with reduction-seeking lookahead proposals (Z2, every layout).
FGK_WALK_KERNEL look_z2_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::Z2Lane, lane::kModeLook>(A, lds);
}

FGK_WALK_KERNEL look_z2h_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::Z2Lane16, lane::kModeLook>(A, lds);
}

FGK_WALK_KERNEL look_z2m_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::Z2LaneM, lane::kModeLook>(A, lds);
}

// Plain walk that records every operator it applies (flip-trace launches).
FGK_WALK_KERNEL trace_z2_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::Z2Lane, lane::kModeTrace>(A, lds);
}

FGK_WALK_KERNEL trace_zt_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::ZTLane, lane::kModeTrace>(A, lds);
}

FGK_WALK_KERNEL trace_z2h_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::Z2Lane16, lane::kModeTrace>(A, lds);
}

FGK_WALK_KERNEL trace_z2m_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::Z2LaneM, lane::kModeTrace>(A, lds);
}

// Walk with a proposed first flip and trajectory logging (Z2).
FGK_WALK_KERNEL learn_z2_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::Z2Lane, lane::kModeLearn>(A, lds);
}

FGK_WALK_KERNEL learn_z2h_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::Z2Lane16, lane::kModeLearn>(A, lds);
}

FGK_WALK_KERNEL learn_z2m_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    walk_persistent<lane::Z2LaneM, lane::kModeLearn>(A, lds);
}

// copyTo of the minimizer's re-broadcast (complexity_minimizer.cu:130-133):
// the best scheme of walker `src` becomes the current scheme of walkers
// dst[0..count).  One thread per (destination, component, term).
__global__ void broadcast_best_kernel(uint32_t *cur, int32_t *m, const uint32_t *best, const int32_t *best_m,
                                      int walkers, int max_m, int src, const int32_t *dst, int count) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long per = 3L * max_m;
    if (g >= per * count) return;
    const int q = (int)(g / per);
    const int c = (int)((g % per) / max_m);
    const int t = (int)(g % max_m);
    const int d = dst[q];
    if (d < 0 || d >= walkers) return;
    const int ms = best_m[src];
    if (t < ms) cur[((size_t)c * max_m + t) * walkers + d] = best[((size_t)src * 3 + c) * max_m + t];
    if (c == 0 && t == 0) m[d] = ms;
}

// schemes[i].copyTo(schemesBest[i]) of updateRanks (src/entities/flip_graph.cu:249-262)
// for walkers list[0..count): the current scheme, its rank and its format
// become the walker's best.  One thread per (walker, component, term).
__global__ void snapshot_best_kernel(const uint32_t *cur, uint32_t *best, const int32_t *m, int32_t *best_m,
                                     const int32_t *fmt, int32_t *bfmt, int walkers, int max_m, const int32_t *list,
                                     int count) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long per = 3L * max_m;
    if (g >= per * count) return;
    const int q = (int)(g / per);
    const int c = (int)((g % per) / max_m);
    const int t = (int)(g % max_m);
    const int w = list[q];
    if (w < 0 || w >= walkers) return;
    const int mw = m[w];
    if (t < mw) best[((size_t)w * 3 + c) * max_m + t] = cur[((size_t)c * max_m + t) * walkers + w];
    if (c == 0 && t == 0) {
        best_m[w] = mw;
        bfmt[4 * w] = fmt[4 * w];
        bfmt[4 * w + 1] = fmt[4 * w + 1];
        bfmt[4 * w + 2] = fmt[4 * w + 2];
    }
}

// Labels the flips logged `horizon` launches ago: one thread per log slot;
// label 1 when the walker's rank fell below the rank at the flip during the
// launches since (minm ring), 0 when it did not, -1 when no flip was logged.
__global__ void log_label_kernel(const uint32_t *rwords, const int32_t *rmeta, const int32_t *rminm, int horizon, int W,
                                 int every, int nslot, int maxm, uint32_t *words, int32_t *m, int32_t *pairs, float *y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nslot) return;
    const int32_t *meta = rmeta + (size_t)i * 4;
    const int w = i * every;
    if (meta[0] < 0 || w >= W) {
        y[i] = -1.f;
        pairs[3 * i] = -1;
        m[i] = 0;
        return;
    }
    int lo = 0x7FFFFFFF;
    for (int q = 0; q < horizon; ++q) {
        const int v = rminm[(size_t)q * W + w];
        lo = v < lo ? v : lo;
    }
    const int rank = meta[3];
    y[i] = lo < rank ? 1.f : 0.f;
    m[i] = rank;
    pairs[3 * i] = meta[0];
    pairs[3 * i + 1] = meta[1];
    pairs[3 * i + 2] = meta[2];
    for (int q = 0; q < 3 * maxm; ++q) {
        const int t = q % maxm;
        words[(size_t)i * 3 * maxm + q] = t < rank ? rwords[(size_t)i * 3 * maxm + q] : 0u;
    }
}

// Fills n int32 with v (used instead of a memset so launches can be captured
// in a HIP graph).
__global__ void fill_i32_kernel(int32_t *p, int32_t v, int n) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) p[g] = v;
}

// Brent equations (src/schemes/scheme_z2.cu:3-29), bit-sliced: one thread per
// (U index i, scheme s), schemes on the fast index so that a wave's loads of
// term t are consecutive words.  One pass over the terms XORs the W mask of
// each term with u_t[i] = 1 into acc[j] for every V index j with v_t[j] = 1;
// acc[j] must then equal the single W bit of c[i/n2][j%n3] (or be zero).
__global__ void verify_z2_kernel(VerifyArgs V) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= 32L * V.count) return;
    const int s = (int)(g % V.count);
    const int i = (int)(g / V.count);
    const int n0 = V.fmt[4 * s], n1 = V.fmt[4 * s + 1], n2 = V.fmt[4 * s + 2];
    const int nn0 = n0 * n1, nn1 = n1 * n2;
    if (i >= nn0) return;
    const int m = V.m[s];
    const size_t stride = (size_t)V.count;
    uint32_t acc[32];
#pragma unroll
    for (int j = 0; j < 32; ++j) acc[j] = 0;
    for (int t = 0; t < m; ++t) {
        const uint32_t u = V.words[((size_t)0 * V.max_m + t) * stride + s];
        if (!((u >> i) & 1u)) continue;
        const uint32_t v = V.words[((size_t)1 * V.max_m + t) * stride + s];
        const uint32_t w = V.words[((size_t)2 * V.max_m + t) * stride + s];
#pragma unroll
        for (int j = 0; j < 32; ++j) acc[j] ^= w & (0u - ((v >> j) & 1u));
    }
    const int i1 = i / n1, i2 = i % n1;
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 32; ++j) {
        if (j < nn1) {
            const int j1 = j / n2, j2 = j % n2;
            const uint32_t want = (i2 == j1) ? (1u << (j2 * n0 + i1)) : 0u;
            ok = ok && acc[j] == want;
        }
    }
    if (!ok) V.ok[s] = 0;
}

// Brent equations over the integers (src/schemes/scheme_integer.cu:3-38) for
// packed ternary words (magnitudes bits 0-15, signs bits 16-31): one thread
// per (U index i, V index j, scheme s), schemes on the fast index.  One pass
// over the terms adds the signed W row of every term with u_t[i] v_t[j] != 0
// into acc[k]; acc[k] must be 1 exactly at the product's entry and 0 elsewhere.
__global__ void verify_zt_kernel(VerifyArgs V) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= 256L * V.count) return;
    const int s = (int)(g % V.count);
    const int ij = (int)(g / V.count);
    const int i = ij / 16, j = ij % 16;
    const int n0 = V.fmt[4 * s], n1 = V.fmt[4 * s + 1], n2 = V.fmt[4 * s + 2];
    const int nn0 = n0 * n1, nn1 = n1 * n2, nn2 = n2 * n0;
    if (i >= nn0 || j >= nn1) return;
    const int m = V.m[s];
    const size_t stride = (size_t)V.count;
    int acc[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) acc[k] = 0;
    for (int t = 0; t < m; ++t) {
        const uint32_t u = V.words[((size_t)0 * V.max_m + t) * stride + s];
        if (!((u >> i) & 1u)) continue;
        const uint32_t v = V.words[((size_t)1 * V.max_m + t) * stride + s];
        if (!((v >> j) & 1u)) continue;
        const uint32_t w = V.words[((size_t)2 * V.max_m + t) * stride + s];
        const uint32_t sg = ((u >> (16 + i)) ^ (v >> (16 + j))) & 1u;  // sign of u_t[i] v_t[j]
        const uint32_t neg = (w >> 16) ^ (0u - sg);                     // bit k: the product's sign at k
#pragma unroll
        for (int k = 0; k < 16; ++k)
            if ((w >> k) & 1u) acc[k] += ((neg >> k) & 1u) ? -1 : 1;
    }
    const int i1 = i / n1, i2 = i % n1, j1 = j / n2, j2 = j % n2;
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        if (k < nn2) {
            const int k1 = k / n0, k2 = k % n0;
            const int target = (i2 == j1) && (i1 == k2) && (j2 == k1);
            ok = ok && acc[k] == target;
        }
    }
    if (!ok) V.ok[s] = 0;
}

size_t walk_lds_bytes(int variant, int max_m) {
    if (variant == kVarZ2m) return z2m_lds_bytes(max_m);
    if (variant == kVarZ2s) return z2s_lds_bytes(max_m);
    return variant == kVarZT ? zt_lds_bytes(max_m) : z2_lds_bytes(max_m, variant == kVarZ2h ? 2 : 4);
}

static const void *walk_kernel(int variant, int mode) {
    const void *k[4][8] = {
        {reinterpret_cast<const void *>(walk_z2_kernel), reinterpret_cast<const void *>(minimize_z2_kernel),
         reinterpret_cast<const void *>(learn_z2_kernel), nullptr, reinterpret_cast<const void *>(resize_z2_kernel),
         reinterpret_cast<const void *>(walk_meta_z2_kernel), reinterpret_cast<const void *>(look_z2_kernel),
         reinterpret_cast<const void *>(trace_z2_kernel)},
        {reinterpret_cast<const void *>(walk_zt_kernel), reinterpret_cast<const void *>(minimize_zt_kernel), nullptr,
         reinterpret_cast<const void *>(flips_zt_kernel), reinterpret_cast<const void *>(resize_zt_kernel),
         reinterpret_cast<const void *>(walk_meta_zt_kernel), nullptr, reinterpret_cast<const void *>(trace_zt_kernel)},
        {reinterpret_cast<const void *>(walk_z2h_kernel), reinterpret_cast<const void *>(minimize_z2h_kernel),
         reinterpret_cast<const void *>(learn_z2h_kernel), nullptr, nullptr, nullptr,
         reinterpret_cast<const void *>(look_z2h_kernel), reinterpret_cast<const void *>(trace_z2h_kernel)},
        {reinterpret_cast<const void *>(walk_z2m_kernel), reinterpret_cast<const void *>(minimize_z2m_kernel),
         reinterpret_cast<const void *>(learn_z2m_kernel), nullptr, nullptr, nullptr,
         reinterpret_cast<const void *>(look_z2m_kernel), reinterpret_cast<const void *>(trace_z2m_kernel)},
    };
    if (mode < 0 || mode > 7) return nullptr;
    if (variant == kVarZ2s) return solo_walk_kernel(mode);
    if (variant < 0 || variant > 3) return nullptr;
    return k[variant][mode];
}

hipError_t prepare_walk(int variant, int max_m, int mode, int *groups_per_cu) {
    const size_t lds = walk_lds_bytes(variant, max_m);
    const void *fn = walk_kernel(variant, mode);
    if (!fn) return hipErrorInvalidValue;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(groups_per_cu, fn, kThreadsPerGroup, lds);
}

hipError_t launch_walk(int variant, int mode, const WalkArgs &a, int grid, hipStream_t stream) {
    const size_t lds = walk_lds_bytes(variant, a.max_m);
    const void *fn = walk_kernel(variant, mode);
    if (!fn) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(64), 0, stream, a.queue, 0, 1);
    WalkArgs arg = a;
    void *params[] = {&arg};
    return hipLaunchKernel(fn, dim3(grid), dim3(kThreadsPerGroup), params, lds, stream);
}

hipError_t launch_broadcast_best(uint32_t *cur, int32_t *m, const uint32_t *best, const int32_t *best_m, int walkers,
                                 int max_m, int src, const int32_t *dst, int count, hipStream_t stream) {
    const long threads = 3L * max_m * count;
    if (threads == 0) return hipSuccess;
    const int block = 256;
    hipLaunchKernelGGL(broadcast_best_kernel, dim3((unsigned)((threads + block - 1) / block)), dim3(block), 0, stream,
                       cur, m, best, best_m, walkers, max_m, src, dst, count);
    return hipGetLastError();
}

hipError_t launch_snapshot_best(const uint32_t *cur, uint32_t *best, const int32_t *m, int32_t *best_m,
                                const int32_t *fmt, int32_t *bfmt, int walkers, int max_m, const int32_t *list,
                                int count, hipStream_t stream) {
    const long threads = 3L * max_m * count;
    if (threads == 0) return hipSuccess;
    const int block = 256;
    hipLaunchKernelGGL(snapshot_best_kernel, dim3((unsigned)((threads + block - 1) / block)), dim3(block), 0, stream,
                       cur, best, m, best_m, fmt, bfmt, walkers, max_m, list, count);
    return hipGetLastError();
}

hipError_t launch_log_label(const uint32_t *rwords, const int32_t *rmeta, const int32_t *rminm, int horizon, int W,
                            int every, int nslot, int maxm, uint32_t *words, int32_t *m, int32_t *pairs, float *y,
                            hipStream_t stream) {
    hipLaunchKernelGGL(log_label_kernel, dim3((nslot + 255) / 256), dim3(256), 0, stream, rwords, rmeta, rminm, horizon,
                       W, every, nslot, maxm, words, m, pairs, y);
    return hipGetLastError();
}

hipError_t launch_fill(int32_t *p, int32_t v, int n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_i32_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, p, v, n);
    return hipGetLastError();
}

hipError_t launch_verify(int ring, const VerifyArgs &v, hipStream_t stream) {
    // Z2: (i < 32, scheme); integer ring: (i < 16, j < 16, scheme)
    const long threads = (long)v.count * (ring ? 256 : 32);
    const int block = 256;
    const int grid = (int)((threads + block - 1) / block);
    if (grid == 0) return hipSuccess;
    if (ring) hipLaunchKernelGGL(verify_zt_kernel, dim3(grid), dim3(block), 0, stream, v);
    else hipLaunchKernelGGL(verify_z2_kernel, dim3(grid), dim3(block), 0, stream, v);
    return hipGetLastError();
}

}  // namespace fgk

#if defined(FGK_PHASE_CLOCKS)
// Phase profile build only (make phases): read and clear the per-phase cycle
// sums of the Z2 walk kernels.
extern "C" __attribute__((visibility("default"))) int fg_phase_read(unsigned long long *out16) {
    return hipMemcpyFromSymbol(out16, HIP_SYMBOL(fgk::g_phase), 16 * 8) == hipSuccess ? 0 : -1;
}
extern "C" __attribute__((visibility("default"))) int fg_phase_reset(void) {
    unsigned long long z[16] = {0};
    return hipMemcpyToSymbol(HIP_SYMBOL(fgk::g_phase), z, 16 * 8) == hipSuccess ? 0 : -1;
}
#endif

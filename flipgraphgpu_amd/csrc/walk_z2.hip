// Z2 flip-graph random walk for CDNA4 (gfx950): the kernel wrapper around
// the per-lane logic in walk_z2_lane.h, the bit-sliced Brent check, and
// their host launchers.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "walk_args.h"
#include "walk_z2_lane.h"

namespace fgk {

__global__ void __launch_bounds__(kWalkersPerGroup) walk_z2_kernel(WalkArgs A) {
    extern __shared__ uint8_t lds[];
    const int lane = threadIdx.x;
    z2::walk_lane(A, lds, lane, blockIdx.x * kWalkersPerGroup + lane);
}

// Brent equations (src/schemes/scheme_z2.cu:3-29), bit-sliced: one thread
// per (scheme, U index i); for each V index j the XOR over terms of the W
// masks with u_t[i] v_t[j] = 1 must equal the single W bit of c[i/n2][j%n3].
__global__ void verify_z2_kernel(VerifyArgs V) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int row_cap = 64;
    const int s = g / row_cap;
    const int i = g % row_cap;
    if (s >= V.count) return;
    const int n0 = V.fmt[4 * s], n1 = V.fmt[4 * s + 1], n2 = V.fmt[4 * s + 2];
    const int nn0 = n0 * n1, nn1 = n1 * n2;
    if (i >= nn0) return;
    const int m = V.m[s];
    const size_t stride = (size_t)V.count;
    bool ok = true;
    for (int j = 0; j < nn1 && ok; ++j) {
        uint32_t acc = 0;
        for (int t = 0; t < m; ++t) {
            const uint32_t u = V.words[((size_t)0 * V.max_m + t) * stride + s];
            const uint32_t v = V.words[((size_t)1 * V.max_m + t) * stride + s];
            if (((u >> i) & (v >> j) & 1u) != 0) acc ^= V.words[((size_t)2 * V.max_m + t) * stride + s];
        }
        const int i1 = i / n1, i2 = i % n1, j1 = j / n2, j2 = j % n2;
        const uint32_t want = (i2 == j1) ? (1u << (j2 * n0 + i1)) : 0u;
        ok = acc == want;
    }
    if (!ok) V.ok[s] = 0;
}

}  // namespace fgk

namespace fgk {

hipError_t launch_walk_z2(const WalkArgs &a, hipStream_t stream) {
    const size_t lds = z2_lds_bytes(a.max_m);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(walk_z2_kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const int grid = (a.walkers + kWalkersPerGroup - 1) / kWalkersPerGroup;
    hipLaunchKernelGGL(walk_z2_kernel, dim3(grid), dim3(kWalkersPerGroup), lds, stream, a);
    return hipGetLastError();
}

hipError_t launch_verify_z2(const VerifyArgs &v, hipStream_t stream) {
    const long threads = (long)v.count * 64;
    const int block = 256;
    const int grid = (int)((threads + block - 1) / block);
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(verify_z2_kernel, dim3(grid), dim3(block), 0, stream, v);
    return hipGetLastError();
}

}  // namespace fgk

// fg_valu_probe: the device's 32-bit integer VALU issue rate, measured.
// Every lane runs 16 independent chains of v_xor_b32 (with a literal
// operand) written as inline asm, so the compiler can neither fuse, fold nor
// drop them; blocks of 4 waves, several per SIMD, keep every SIMD issuing.
// On gfx950 a SIMD executes two wave64 VALU instructions every 4 cycles when
// two of its waves have one ready (a single wave issues at most one), so the
// peak is CUs x clock x 128 lanes, the rate behind the published FP32 vector
// figure (157.3 TFLOP/s = 256 CUs x 2.4 GHz x 128 lanes x 2 for an FMA).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "fg.h"
#include "runtime_internal.h"

namespace {

constexpr int kChains = 16;   // independent xor chains per lane
constexpr int kUnroll = 16;   // xors per chain per loop trip
constexpr int kThreads = 256;

__global__ void __launch_bounds__(kThreads) valu_probe_kernel(uint32_t *out, uint32_t seed, int trips) {
    uint32_t a[kChains];
    const uint32_t k = seed ^ (threadIdx.x * 0x9E3779B9u) ^ (blockIdx.x << 8);
#pragma unroll
    for (int e = 0; e < kChains; ++e) a[e] = k + (uint32_t)e * 0x85EBCA6Bu;
    for (int i = 0; i < trips; ++i) {
#pragma unroll
        for (int r = 0; r < kUnroll; ++r)
#pragma unroll
            for (int e = 0; e < kChains; ++e) asm volatile("v_xor_b32 %0, 0x5bd1e995, %0" : "+v"(a[e]));
    }
    uint32_t x = 0;
#pragma unroll
    for (int e = 0; e < kChains; ++e) x ^= a[e];
    out[(size_t)blockIdx.x * kThreads + threadIdx.x] = x;
}

int fail(const char *msg) {
    fgrt::set_error(msg);
    return -1;
}

}  // namespace

extern "C" int fg_valu_probe(int blocks, int trips, double *lane_ops_per_s, double *kernel_ms) {
    if (blocks < 1 || trips < 1 || !lane_ops_per_s) return fail("fg_valu_probe: bad arguments");
    uint32_t *out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t err = hipMalloc(reinterpret_cast<void **>(&out), (size_t)blocks * kThreads * sizeof(uint32_t));
    if (err == hipSuccess) err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    float best = 0.f;
    if (err == hipSuccess) {
        // one warm-up launch, then the fastest of three timed ones
        for (int rep = 0; rep < 4 && err == hipSuccess; ++rep) {
            err = hipEventRecord(e0, 0);
            if (err != hipSuccess) break;
            hipLaunchKernelGGL(valu_probe_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, 0, out, 0x1234567u + rep,
                               trips);
            err = hipGetLastError();
            if (err == hipSuccess) err = hipEventRecord(e1, 0);
            if (err == hipSuccess) err = hipEventSynchronize(e1);
            float ms = 0.f;
            if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
            if (err == hipSuccess && rep > 0) best = rep == 1 ? ms : std::min(best, ms);
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (out) (void)hipFree(out);
    if (err != hipSuccess) {
        fgrt::set_error(std::string("fg_valu_probe: ") + hipGetErrorString(err));
        return -1;
    }
    const double lane_ops = (double)blocks * kThreads * (double)trips * kUnroll * kChains;
    *lane_ops_per_s = best > 0.f ? lane_ops / (best * 1e-3) : 0.0;
    if (kernel_ms) *kernel_ms = best;
    return 0;
}

// fg_verify_noalloc: the Brent check of fg_verify for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).  All device staging lives in one
// file-scope array per GPU, `scratch`; the caller passes only the input blob
// and the output flags.
//
// Staging layout in scratch (each region 256-byte aligned):
//   blob   the caller's blob, copied as is (int32)
//   off    [count] position of each scheme's header in blob
//   words  [3][max_m][count] packed factor words (the Brent kernel's planes)
//   m      [count], fmt [count][4], ok [count]
// Every region the check reads is written by the same call, and ok is set
// to 1 before the check, so nothing depends on a previous call's contents.
// Calls that overlap on one GPU (several host threads) share its scratch
// and race.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>

#include "fg.h"
#include "kernels.h"
#include "runtime_internal.h"

namespace {

constexpr size_t kScratchBytes = size_t(8) << 20;  // 8 MiB per GPU, held until the process ends
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_mu;
void *g_scratch[kMaxDevices];  // device address of `scratch` on each GPU (null: not looked up yet)

int fail(const char *msg) {
    fgrt::set_error(msg);
    return -1;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// The scratch address on the current device, looked up once per device with
// that device current.
int scratch_ptr(uint8_t **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return fail("fg_verify_noalloc: no usable current device");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fail("fg_verify_noalloc: hipGetSymbolAddress of the scratch array failed");
        g_scratch[dev] = p;
    }
    *out = static_cast<uint8_t *>(g_scratch[dev]);
    return 0;
}

// One thread per (scheme q, component c, term t < max_m): packs the factor
// of term t from the blob into the Brent kernel's plane word -- the Z2 mask
// (odd coefficients), or for the integer ring the magnitudes in bits 0-15 and
// the signs in bits 16-31 -- and writes m and the format once per scheme.
// fg_verify also applies fixSigns to integer schemes; that negates two
// factors of a term and leaves every product, hence the check, unchanged.
__global__ void stage_blob_kernel(const int32_t *blob, const int32_t *off, int count, int max_m, int ring,
                                  uint32_t *words, int32_t *m, int32_t *fmt) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= 3L * max_m * count) return;
    const int q = (int)(g % count);
    const int ct = (int)(g / count);
    const int c = ct / max_m, t = ct % max_m;
    const int32_t *h = blob + off[q];
    const int n[3] = {h[0], h[1], h[2]};
    const int ms = h[3];
    const int nn[3] = {n[0] * n[1], n[1] * n[2], n[2] * n[0]};
    if (c == 0 && t == 0) {
        m[q] = ms;
        fmt[4 * q] = n[0];
        fmt[4 * q + 1] = n[1];
        fmt[4 * q + 2] = n[2];
        fmt[4 * q + 3] = 0;
    }
    if (t >= ms) return;
    int base = 4;
    for (int x = 0; x < c; ++x) base += ms * nn[x];
    const int32_t *row = h + base + t * nn[c];
    uint32_t v = 0, s = 0;
    for (int b = 0; b < nn[c]; ++b) {
        const int x = row[b];
        if (ring == FG_RING_Z2) {
            v |= (uint32_t)((x < 0 ? -x : x) & 1) << b;
        } else {
            v |= (uint32_t)(x != 0) << b;
            s |= (uint32_t)(x < 0) << b;
        }
    }
    words[((size_t)c * max_m + t) * count + q] = ring == FG_RING_Z2 ? v : ((v & 0xFFFFu) | ((s & v & 0xFFFFu) << 16));
}

}  // namespace

extern "C" int fg_verify_noalloc(int ring, const int32_t *blob, int64_t len, int32_t *out_ok) {
    if (!blob || len < 1 || !out_ok) return fail("fg_verify_noalloc: bad arguments");
    if (ring != FG_RING_Z2 && ring != FG_RING_ZT) return fail("fg_verify_noalloc: unknown ring");
    const int count = blob[0];
    if (count < 0) return fail("fg_verify_noalloc: negative count");
    if (count == 0) return 0;
    // the host walks the blob once: the checks of fg_verify's parser and the
    // largest rank, without building any host structure
    const int wmax = ring == FG_RING_Z2 ? 32 : 16;
    int64_t pos = 1;
    int maxm = 1;
    for (int q = 0; q < count; ++q) {
        if (pos + 4 > len) return fail("fg_verify_noalloc: truncated scheme header");
        const int n1 = blob[pos], n2 = blob[pos + 1], n3 = blob[pos + 2], ms = blob[pos + 3];
        if (n1 < 1 || n2 < 1 || n3 < 1 || ms < 0 || ms > 350 || n1 * n2 > 64 || n2 * n3 > 64 || n3 * n1 > 64)
            return fail("fg_verify_noalloc: invalid scheme sizes");
        if (n1 * n2 > wmax || n2 * n3 > wmax || n3 * n1 > wmax)
            return fail("fg_verify_noalloc: factor wider than the GPU word of this ring");
        const int64_t need = (int64_t)ms * (n1 * n2 + n2 * n3 + n3 * n1);
        if (pos + 4 + need > len) return fail("fg_verify_noalloc: truncated scheme coefficients");
        if (ring == FG_RING_ZT)
            for (int64_t x = pos + 4; x < pos + 4 + need; ++x)
                if (blob[x] < -1 || blob[x] > 1) return fail("fg_verify_noalloc: integer coefficient outside {-1, 0, 1}");
        maxm = ms > maxm ? ms : maxm;
        pos += 4 + need;
    }
    const size_t o_blob = 0;
    const size_t o_off = align256(o_blob + (size_t)pos * 4);
    const size_t o_words = align256(o_off + (size_t)count * 4);
    const size_t o_m = align256(o_words + (size_t)3 * maxm * count * 4);
    const size_t o_fmt = align256(o_m + (size_t)count * 4);
    const size_t o_ok = align256(o_fmt + (size_t)count * 16);
    const size_t total = o_ok + (size_t)count * 4;
    if (total > kScratchBytes)
        return fail("fg_verify_noalloc: the blob needs more staging than the 8 MiB scratch holds; split it or use fg_verify");
    uint8_t *S = nullptr;
    if (scratch_ptr(&S)) return -1;
    int32_t *d_blob = reinterpret_cast<int32_t *>(S + o_blob);
    int32_t *d_off = reinterpret_cast<int32_t *>(S + o_off);
    uint32_t *d_words = reinterpret_cast<uint32_t *>(S + o_words);
    int32_t *d_m = reinterpret_cast<int32_t *>(S + o_m);
    int32_t *d_fmt = reinterpret_cast<int32_t *>(S + o_fmt);
    int32_t *d_ok = reinterpret_cast<int32_t *>(S + o_ok);
    hipStream_t st = nullptr;  // the null stream: creating one would allocate
    if (hipMemcpy(d_blob, blob, (size_t)pos * 4, hipMemcpyHostToDevice) != hipSuccess)
        return fail("fg_verify_noalloc: copy of the blob failed");
    // header positions, sent in chunks from a fixed host buffer
    int32_t chunk[1024];
    int64_t p = 1;
    for (int q0 = 0; q0 < count; q0 += 1024) {
        const int nq = count - q0 < 1024 ? count - q0 : 1024;
        for (int q = 0; q < nq; ++q) {
            chunk[q] = (int32_t)p;
            p += 4 + (int64_t)blob[p + 3] * (blob[p] * blob[p + 1] + blob[p + 1] * blob[p + 2] + blob[p + 2] * blob[p]);
        }
        if (hipMemcpy(d_off + q0, chunk, (size_t)nq * 4, hipMemcpyHostToDevice) != hipSuccess)
            return fail("fg_verify_noalloc: copy of the scheme positions failed");
    }
    const long threads = 3L * maxm * count;
    hipLaunchKernelGGL(stage_blob_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, d_blob, d_off,
                       count, maxm, ring, d_words, d_m, d_fmt);
    if (hipGetLastError() != hipSuccess) return fail("fg_verify_noalloc: staging kernel launch failed");
    if (fgk::launch_fill(d_ok, 1, count, st) != hipSuccess) return fail("fg_verify_noalloc: fill launch failed");
    fgk::VerifyArgs v{};
    v.words = d_words;
    v.signs = nullptr;
    v.m = d_m;
    v.fmt = d_fmt;
    v.count = count;
    v.max_m = maxm;
    v.ok = d_ok;
    if (fgk::launch_verify(ring, v, st) != hipSuccess) return fail("fg_verify_noalloc: Brent check launch failed");
    if (hipMemcpy(out_ok, d_ok, (size_t)count * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return fail("fg_verify_noalloc: copy of the results failed");
    int good = 0;
    for (int q = 0; q < count; ++q) good += out_ok[q] != 0;
    return good;
}

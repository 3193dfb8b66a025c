// Issue-rate probe of the 32-bit integer multiplies the transforms are built
// from (v_mul_lo_u32, v_mul_hi_u32): every lane runs 8 independent chains,
// alternating low and high halves, long enough that launch overhead is
// negligible, on enough wavefronts to fill every SIMD. The measured rate is
// the peak the INT32-multiply roofline of the transform kernels is quoted
// against. Beside it, a plain device copy: the measured HBM bandwidth, and
// the dispatch of known size that scales the profiler's FETCH_SIZE and
// WRITE_SIZE counters to bytes (no counterpart in the reference, which has
// no GPU).
#include <stdio.h>

#include "../../include/fastfib_amd.h"
#include "ff_internal.h"
#include "ff_runtime.h"

namespace {

constexpr int kChains = 8;
constexpr int kUnroll = 16;

__global__ void __launch_bounds__(256) k_mul_probe(uint32_t *out, uint32_t k0, uint32_t k1, int iters) {
    uint32_t x[kChains];
#pragma unroll
    for (int c = 0; c < kChains; c++) x[c] = threadIdx.x * 2654435761u + c * 40503u + blockIdx.x;
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
#pragma unroll
            for (int c = 0; c < kChains; c++) {
                // written as instructions: the compiler would fold a chain of
                // products by one constant into one product by its power
                if (c & 1) asm volatile("v_mul_hi_u32 %0, %1, %2" : "=v"(x[c]) : "v"(x[c]), "s"(k1));
                else asm volatile("v_mul_lo_u32 %0, %1, %2" : "=v"(x[c]) : "v"(x[c]), "s"(k0));
            }
        }
    }
    uint32_t r = 0;
#pragma unroll
    for (int c = 0; c < kChains; c++) r ^= x[c];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = r;
}

// b = a, 16 bytes per lane and four of them in flight per loop trip.
__global__ void __launch_bounds__(256) k_copy_probe(const uint4 *__restrict__ a, uint4 *__restrict__ b, size_t n16) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n16; i += 4 * stride) {
        const uint4 x0 = a[i], x1 = a[i + stride], x2 = a[i + 2 * stride], x3 = a[i + 3 * stride];
        b[i] = x0;
        b[i + stride] = x1;
        b[i + 2 * stride] = x2;
        b[i + 3 * stride] = x3;
    }
    for (; i < n16; i += stride) b[i] = a[i];
}

}  // namespace

extern "C" int ff_probe_copy(size_t bytes, double *bytes_per_s, double *seconds) {
    if (!bytes_per_s || bytes < 16) return FF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ff_ensure_init();
    int cus = 0;
    FF_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ff_device()));
    const size_t n16 = bytes / 16;
    uint4 *a = (uint4 *)ff_dalloc(n16 * 16), *b = (uint4 *)ff_dalloc(n16 * 16);
    const unsigned blocks = (unsigned)cus * 8;
    hipEvent_t e0, e1;
    FF_CHECK(hipEventCreate(&e0));
    FF_CHECK(hipEventCreate(&e1));
    double best = 0, best_s = 0;
    for (int rep = 0; rep < 4; rep++) {  // the first run warms the clocks
        FF_CHECK(hipEventRecord(e0, g_st));
        hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, g_st, a, b, n16);
        FF_CHECK(hipGetLastError());
        FF_CHECK(hipEventRecord(e1, g_st));
        FF_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        FF_CHECK(hipEventElapsedTime(&ms, e0, e1));
        const double rate = 2.0 * (double)(n16 * 16) / (ms * 1e-3);  // read + written
        if (rep > 0 && rate > best) best = rate, best_s = ms * 1e-3;
    }
    FF_CHECK(hipEventDestroy(e0));
    FF_CHECK(hipEventDestroy(e1));
    ff_dfree(a);
    ff_dfree(b);
    *bytes_per_s = best;
    if (seconds) *seconds = best_s;
    return FF_OK;
}

extern "C" int ff_probe_int32_mul(double *mults_per_s, double *seconds) {
    if (!mults_per_s) return FF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ff_ensure_init();
    int cus = 0;
    FF_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ff_device()));
    const unsigned blocks = (unsigned)cus * 8;  // 8 workgroups of 4 waves per CU: 8 waves per SIMD
    const int threads = 256, iters = 4096;
    uint32_t *out = (uint32_t *)ff_dalloc((size_t)blocks * threads * 4);
    hipEvent_t a, b;
    FF_CHECK(hipEventCreate(&a));
    FF_CHECK(hipEventCreate(&b));
    double best = 0, best_s = 0;
    for (int rep = 0; rep < 4; rep++) {  // the first run warms the clocks
        FF_CHECK(hipEventRecord(a, g_st));
        hipLaunchKernelGGL(k_mul_probe, dim3(blocks), dim3(threads), 0, g_st, out, 0x9E3779B1u | 1u, 0x85EBCA77u,
                           iters);
        FF_CHECK(hipGetLastError());
        FF_CHECK(hipEventRecord(b, g_st));
        FF_CHECK(hipEventSynchronize(b));
        float ms = 0;
        FF_CHECK(hipEventElapsedTime(&ms, a, b));
        const double mults = (double)blocks * threads * iters * kUnroll * kChains;
        const double rate = mults / (ms * 1e-3);
        if (rep > 0 && rate > best) best = rate, best_s = ms * 1e-3;
    }
    FF_CHECK(hipEventDestroy(a));
    FF_CHECK(hipEventDestroy(b));
    ff_dfree(out);
    *mults_per_s = best;
    if (seconds) *seconds = best_s;
    return FF_OK;
}

// Linear limb passes on gfx950: the fast-doubling step, the last-step
// factors, small additions, and the chunked carry resolution they share.
//
// Each pass splits its limbs into chunks of FF_CH limbs; every chunk is
// computed from a zero carry-in (chunk 0 may start with a small constant),
// exactly as the reference's parallel chunks do (fib.c:19-62, 111-143), and
// records per output stream its signed carry-out e_k and a summary of its
// limbs. The carry entering chunk k is then c_k = f_{k-1}(c_{k-1}) with
//   f(c) = e + ovf(c),  ovf(c) = +1 if c > 0 and limb0 + c wraps and limbs
//   1.. are all ones, -1 if c < 0 and limb0 < -c and limbs 1.. are zero,
// a map on a small integer range; maps compose associatively, so one
// workgroup per stream scans them (instead of the reference's sequential
// add_small_at fix, fib.c:64-70,94-99), and an apply pass adds c_k into each
// chunk with the ripple confined to the chunk.
#include <stdint.h>

#include "ff_internal.h"
#include "ff_limbs.h"

typedef unsigned __int128 u128;

// Map domain: carries in [-8, 7], stored as index c + 8.
#define CDOM 16
#define COFF 8

__device__ __forceinline__ void summarize(const uint64_t *d, uint64_t i0, uint64_t i1, int e, FFCarrySum *out) {
    FFCarrySum s;
    s.limb0 = d[i0];
    uint32_t ones = 1, zero = 1;
    for (uint64_t i = i0 + 1; i < i1; i++) {
        const uint64_t v = d[i];
        ones &= (v == ~0ull);
        zero &= (v == 0);
    }
    s.ones = (uint8_t)ones;
    s.zero = (uint8_t)zero;
    s.e = (int8_t)e;
    s.pad_ = 0;
    *out = s;
}

__device__ __forceinline__ int sum_apply(const FFCarrySum &s, int c) {
    int ov = 0;
    if (c > 0 && s.ones && s.limb0 + (uint64_t)c < s.limb0) ov = 1;
    if (c < 0 && s.zero && s.limb0 < (uint64_t)(-c)) ov = -1;
    return s.e + ov;
}

// fib.c:36-62: A = s1 + s2, B = 4 s1 - s2 + cb0, C = B - A over one chunk.
__global__ __launch_bounds__(256) void k_step_local(FFStepArgs a) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nch = (a.len + FF_CH - 1) / FF_CH;
    if (k >= nch) return;
    const uint64_t i0 = k * FF_CH, i1 = i0 + FF_CH < a.len ? i0 + FF_CH : a.len;
    uint64_t cA = 0, prev = (i0 && i0 - 1 < a.n1) ? a.s1[i0 - 1] : 0;
    int64_t cB = k ? 0 : a.cb0, cC = 0;
    for (uint64_t i = i0; i < i1; i++) {
        const uint64_t x = i < a.n1 ? a.s1[i] : 0, y = i < a.n2 ? a.s2[i] : 0;
        const uint64_t x4 = (x << 2) | (prev >> 62);
        prev = x;
        const u128 av = (u128)x + y + cA;
        const uint64_t ai = (uint64_t)av;
        cA = (uint64_t)(av >> 64);
        const __int128 bv = (__int128)x4 - y + cB;
        const uint64_t bi = (uint64_t)bv;
        cB = (int64_t)(bv >> 64);
        const __int128 cv = (__int128)bi - ai + cC;
        cC = (int64_t)(cv >> 64);
        if (a.A) a.A[i] = ai;
        if (a.B) a.B[i] = bi;
        a.C[i] = (uint64_t)cv;
    }
    // stream 0: A or B (whichever is kept), stream 1: C
    if (a.A) summarize(a.A, i0, i1, (int)cA, &a.sums[k]);
    else summarize(a.B, i0, i1, (int)cB, &a.sums[k]);
    summarize(a.C, i0, i1, (int)(cC + cB - (int64_t)cA), &a.sums[nch + k]);
}

// fib.c:123-143: X = 2u + v, Y = 2u - v over one chunk.
__global__ __launch_bounds__(256) void k_last_local(FFLastArgs a) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nch = (a.len + FF_CH - 1) / FF_CH;
    if (k >= nch) return;
    const uint64_t i0 = k * FF_CH, i1 = i0 + FF_CH < a.len ? i0 + FF_CH : a.len;
    uint64_t cX = 0, prev = (i0 && i0 - 1 < a.nu) ? a.u[i0 - 1] : 0;
    int64_t cY = 0;
    for (uint64_t i = i0; i < i1; i++) {
        const uint64_t x = i < a.nu ? a.u[i] : 0, y = i < a.nv ? a.v[i] : 0;
        const uint64_t x2 = (x << 1) | (prev >> 63);
        prev = x;
        const u128 s = (u128)x2 + y + cX;
        a.X[i] = (uint64_t)s;
        cX = (uint64_t)(s >> 64);
        if (a.Y) {
            const __int128 d = (__int128)x2 - y + cY;
            a.Y[i] = (uint64_t)d;
            cY = (int64_t)(d >> 64);
        }
    }
    summarize(a.X, i0, i1, (int)cX, &a.sums[k]);
    if (a.Y) summarize(a.Y, i0, i1, (int)cY, &a.sums[nch + k]);
}

// Scan of the carry maps of one stream per workgroup: writes c_k for every
// chunk (c_0 = 0).
#define SC_THREADS 1024

__global__ __launch_bounds__(SC_THREADS) void k_carry_scan(FFCarryStreams cs) {
    const int st = blockIdx.x;
    const FFCarrySum *sums = cs.sums + (size_t)st * cs.nch;
    int8_t *cin = cs.cin + (size_t)st * cs.nch;
    const uint64_t nch = cs.nch;
    const uint64_t per = (nch + SC_THREADS - 1) / SC_THREADS;
    const uint64_t k0 = threadIdx.x * per, k1 = k0 + per < nch ? k0 + per : nch;
    // segment map: start from the identity
    int8_t m[CDOM];
#pragma unroll
    for (int c = 0; c < CDOM; c++) m[c] = (int8_t)(c - COFF);
    for (uint64_t k = k0; k < k1; k++) {
        const FFCarrySum s = sums[k];
#pragma unroll
        for (int c = 0; c < CDOM; c++) m[c] = (int8_t)sum_apply(s, m[c]);
    }
    __shared__ int8_t sm[2][SC_THREADS][CDOM];
#pragma unroll
    for (int c = 0; c < CDOM; c++) sm[0][threadIdx.x][c] = m[c];
    __syncthreads();
    int cur = 0;
    for (int off = 1; off < SC_THREADS; off <<= 1) {
        // inclusive: map_i = map_i o map_{i-off}
        if ((int)threadIdx.x >= off) {
#pragma unroll
            for (int c = 0; c < CDOM; c++) {
                const int v = sm[cur][threadIdx.x - off][c];
                sm[1 - cur][threadIdx.x][c] = sm[cur][threadIdx.x][v + COFF];
            }
        } else {
#pragma unroll
            for (int c = 0; c < CDOM; c++) sm[1 - cur][threadIdx.x][c] = sm[cur][threadIdx.x][c];
        }
        __syncthreads();
        cur = 1 - cur;
    }
    int c = threadIdx.x ? sm[cur][threadIdx.x - 1][COFF] : 0;
    for (uint64_t k = k0; k < k1; k++) {
        cin[k] = (int8_t)c;
        c = sum_apply(sums[k], c);
    }
}

// Adds c_k into chunk k of each stream (ripple confined to the chunk).
__global__ __launch_bounds__(256) void k_carry_apply(FFCarryStreams cs) {
    const int st = blockIdx.y;
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= cs.nch) return;
    const int c = cs.cin[(size_t)st * cs.nch + k];
    if (!c) return;
    uint64_t *d = cs.d[st];
    const uint64_t i0 = k * FF_CH, i1 = i0 + FF_CH < cs.len ? i0 + FF_CH : cs.len;
    if (c > 0) {
        uint64_t v = (uint64_t)c;
        for (uint64_t i = i0; i < i1 && v; i++) {
            const uint64_t old = d[i];
            d[i] = old + v;
            v = d[i] < old ? 1 : 0;
        }
    } else {
        uint64_t v = (uint64_t)(-c);
        for (uint64_t i = i0; i < i1 && v; i++) {
            const uint64_t old = d[i];
            d[i] = old - v;
            v = old < v ? 1 : 0;
        }
    }
}

// d[0..len) += v (v may be negative; the result stays >= 0): one thread, the
// ripple dies after a limb or two (big_add_u64_inplace, bigint.c:173-182).
__global__ void k_add_small(uint64_t *d, uint64_t len, int64_t v) {
    if (threadIdx.x || blockIdx.x) return;
    if (v > 0) {
        uint64_t c = (uint64_t)v;
        for (uint64_t i = 0; i < len && c; i++) {
            const uint64_t old = d[i];
            d[i] = old + c;
            c = d[i] < old ? 1 : 0;
        }
    } else if (v < 0) {
        uint64_t c = (uint64_t)(-v);
        for (uint64_t i = 0; i < len && c; i++) {
            const uint64_t old = d[i];
            d[i] = old - c;
            c = old < c ? 1 : 0;
        }
    }
}

// Normalized length and bit count of several arrays.
__global__ __launch_bounds__(256) void k_normlen_multi(const FFNormJob *jobs, uint64_t *res) {
    const FFNormJob J = jobs[blockIdx.x];
    __shared__ unsigned long long best;
    if (threadIdx.x == 0) best = 0;
    __syncthreads();
    uint64_t top = J.n;
    while (top > 0) {
        const uint64_t lo = top > 256 ? top - 256 : 0;
        const uint64_t i = lo + threadIdx.x;
        if (i < top && J.d[i] != 0) atomicMax(&best, (unsigned long long)(i + 1));
        __syncthreads();
        if (best) break;
        top = lo;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const uint64_t n = best ? best : 1;
        const uint64_t t = J.d[n - 1];
        res[2 * blockIdx.x] = n;
        res[2 * blockIdx.x + 1] = t ? 64 * (n - 1) + (64 - __clzll((long long)t)) : 0;
    }
}

extern "C" hipError_t ff_launch_step(const FFStepArgs *a, hipStream_t st) {
    const uint64_t nch = (a->len + FF_CH - 1) / FF_CH;
    hipLaunchKernelGGL(k_step_local, dim3((unsigned)((nch + 255) / 256)), dim3(256), 0, st, *a);
    return hipGetLastError();
}

extern "C" hipError_t ff_launch_last(const FFLastArgs *a, hipStream_t st) {
    const uint64_t nch = (a->len + FF_CH - 1) / FF_CH;
    hipLaunchKernelGGL(k_last_local, dim3((unsigned)((nch + 255) / 256)), dim3(256), 0, st, *a);
    return hipGetLastError();
}

extern "C" hipError_t ff_launch_carry(const FFCarryStreams *cs, hipStream_t st) {
    hipLaunchKernelGGL(k_carry_scan, dim3(cs->nstreams), dim3(SC_THREADS), 0, st, *cs);
    hipLaunchKernelGGL(k_carry_apply, dim3((unsigned)((cs->nch + 255) / 256), cs->nstreams), dim3(256), 0, st, *cs);
    return hipGetLastError();
}

extern "C" hipError_t ff_launch_add_small(uint64_t *d, uint64_t len, int64_t v, hipStream_t st) {
    hipLaunchKernelGGL(k_add_small, dim3(1), dim3(64), 0, st, d, len, v);
    return hipGetLastError();
}

extern "C" hipError_t ff_launch_normlen_multi(const FFNormJob *jobs_dev, int n, uint64_t *res_dev, hipStream_t st) {
    hipLaunchKernelGGL(k_normlen_multi, dim3(n), dim3(256), 0, st, jobs_dev, res_dev);
    return hipGetLastError();
}

// Linear limb passes on gfx950: the fast-doubling step, the last-step
// factors, small additions, and the chunked carry resolution they share.
//
// Each pass splits its limbs into chunks of FF_CH limbs; every chunk is
// computed from a zero carry-in (chunk 0 may start with a small constant),
// exactly as the reference's parallel chunks do (fib.c:19-62, 111-143), and
// records per output stream its signed carry-out e_k and a summary of its
// limbs. The carry entering chunk k is c_k = f_{k-1}(c_{k-1}) with
//   f(c) = e + ovf(c),  ovf(c) = +1 if c > 0 and limb0 + c wraps and limbs
//   1.. are all ones, -1 if c < 0 and limb0 < -c and limbs 1.. are zero.
// ovf can only be nonzero for a chunk whose limbs 1.. are all ones or all
// zero; when no chunk (but the last) is like that, c_k = e_{k-1} and no
// scan is needed. Otherwise the maps, on a small integer range, are composed
// by one workgroup per stream (instead of the reference's sequential
// add_small_at fix, fib.c:64-70,94-99). The apply pass adds c_k into each
// chunk with the ripple confined to the chunk and books the top limb.
#include <stdint.h>

#include "ff_internal.h"
#include "ff_limbs.h"

typedef unsigned __int128 u128;

// Map domain: carries in [-8, 7], stored as index c + 8.
#define CDOM 16
#define COFF 8

__device__ __forceinline__ FFCarrySum summarize(const uint64_t *v, int m, int e) {
    FFCarrySum s;
    s.limb0 = v[0];
    uint32_t ones = 1, zero = 1;
#pragma unroll
    for (int i = 1; i < FF_CH; i++) {
        if (i < m) {
            ones &= (v[i] == ~0ull);
            zero &= (v[i] == 0);
        }
    }
    s.ones = (uint8_t)ones;
    s.zero = (uint8_t)zero;
    s.e = (int8_t)e;
    s.pad_ = 0;
    s.pad2_ = 0;
    return s;
}

__device__ __forceinline__ int sum_apply(const FFCarrySum &s, int c) {
    int ov = 0;
    if (c > 0 && s.ones && s.limb0 + (uint64_t)c < s.limb0) ov = 1;
    if (c < 0 && s.zero && s.limb0 < (uint64_t)(-c)) ov = -1;
    return s.e + ov;
}

// fib.c:36-62: A = s1 + s2, B = 4 s1 - s2 + cb0, C = B - A over one chunk.
__global__ __launch_bounds__(FF_LIN_THREADS) void k_step_local(FFStepArgs a) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nch = (a.len + FF_CH - 1) / FF_CH;
    const bool live = k < nch;
    bool special = false;
    if (live) {
        const uint64_t *__restrict__ s1 = a.s1;
        const uint64_t *__restrict__ s2 = a.s2;
        const uint64_t i0 = k * FF_CH;
        const int m = (int)(a.len - i0 < FF_CH ? a.len - i0 : FF_CH);
        uint64_t x[FF_CH], y[FF_CH];
        const uint64_t prev0 = (i0 && i0 - 1 < a.n1) ? s1[i0 - 1] : 0;
#pragma unroll
        for (int j = 0; j < FF_CH; j++) {
            const uint64_t i = i0 + j;
            x[j] = (j < m && i < a.n1) ? s1[i] : 0;
            y[j] = (j < m && i < a.n2) ? s2[i] : 0;
        }
        uint64_t A[FF_CH], B[FF_CH], C[FF_CH];
        uint64_t cA = 0, prev = prev0;
        int64_t cB = k ? 0 : a.cb0, cC = 0;
#pragma unroll
        for (int j = 0; j < FF_CH; j++) {
            A[j] = B[j] = C[j] = 0;
            if (j < m) {  // a short last chunk stops at its m limbs
                const uint64_t x4 = (x[j] << 2) | (prev >> 62);
                prev = x[j];
                const u128 av = (u128)x[j] + y[j] + cA;
                A[j] = (uint64_t)av;
                cA = (uint64_t)(av >> 64);
                const __int128 bv = (__int128)x4 - y[j] + cB;
                B[j] = (uint64_t)bv;
                cB = (int64_t)(bv >> 64);
                const __int128 cv = (__int128)B[j] - A[j] + cC;
                C[j] = (uint64_t)cv;
                cC = (int64_t)(cv >> 64);
            }
        }
        uint64_t *__restrict__ dK = a.A ? a.A : a.B;
        uint64_t *__restrict__ dC = a.C;
#pragma unroll
        for (int j = 0; j < FF_CH; j++) {
            if (j < m) {
                dK[i0 + j] = a.A ? A[j] : B[j];
                dC[i0 + j] = C[j];
            }
        }
        const FFCarrySum sk = summarize(a.A ? A : B, m, a.A ? (int)cA : (int)cB);
        const FFCarrySum sc = summarize(C, m, (int)(cC + cB - (int64_t)cA));
        a.sums[k] = sk;
        a.sums[nch + k] = sc;
        special = (k + 1 < nch) && (sk.ones || sk.zero || sc.ones || sc.zero);
    }
    const int any = __syncthreads_or(special);
    if (threadIdx.x == 0) a.wgflag[blockIdx.x] = (uint32_t)any;
}

// fib.c:123-143: X = 2u + v, Y = 2u - v over one chunk.
__global__ __launch_bounds__(FF_LIN_THREADS) void k_last_local(FFLastArgs a) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nch = (a.len + FF_CH - 1) / FF_CH;
    const bool live = k < nch;
    bool special = false;
    if (live) {
        const uint64_t *__restrict__ u = a.u;
        const uint64_t *__restrict__ v = a.v;
        const uint64_t i0 = k * FF_CH;
        const int m = (int)(a.len - i0 < FF_CH ? a.len - i0 : FF_CH);
        uint64_t x[FF_CH], y[FF_CH];
        uint64_t prev = (i0 && i0 - 1 < a.nu) ? u[i0 - 1] : 0;
#pragma unroll
        for (int j = 0; j < FF_CH; j++) {
            const uint64_t i = i0 + j;
            x[j] = (j < m && i < a.nu) ? u[i] : 0;
            y[j] = (j < m && i < a.nv) ? v[i] : 0;
        }
        uint64_t X[FF_CH], Y[FF_CH];
        uint64_t cX = 0;
        int64_t cY = 0;
#pragma unroll
        for (int j = 0; j < FF_CH; j++) {
            X[j] = Y[j] = 0;
            if (j < m) {
                const uint64_t x2 = (x[j] << 1) | (prev >> 63);
                prev = x[j];
                const u128 s = (u128)x2 + y[j] + cX;
                X[j] = (uint64_t)s;
                cX = (uint64_t)(s >> 64);
                const __int128 d = (__int128)x2 - y[j] + cY;
                Y[j] = (uint64_t)d;
                cY = (int64_t)(d >> 64);
            }
        }
        uint64_t *__restrict__ dX = a.X;
#pragma unroll
        for (int j = 0; j < FF_CH; j++)
            if (j < m) dX[i0 + j] = X[j];
        const FFCarrySum sx = summarize(X, m, (int)cX);
        a.sums[k] = sx;
        special = (k + 1 < nch) && (sx.ones || sx.zero);
        if (a.Y) {
            uint64_t *__restrict__ dY = a.Y;
#pragma unroll
            for (int j = 0; j < FF_CH; j++)
                if (j < m) dY[i0 + j] = Y[j];
            const FFCarrySum sy = summarize(Y, m, (int)cY);
            a.sums[nch + k] = sy;
            special = special || ((k + 1 < nch) && (sy.ones || sy.zero));
        }
    }
    const int any = __syncthreads_or(special);
    if (threadIdx.x == 0) a.wgflag[blockIdx.x] = (uint32_t)any;
}

// Decides exact or fast carries, resets the length keys, and in the exact
// case scans the carry maps of one stream per workgroup: c_k for every chunk
// (c_0 = 0).
#define SC_THREADS 1024

__global__ __launch_bounds__(SC_THREADS) void k_carry_scan(FFCarryStreams cs) {
    const int st = blockIdx.x;
    const uint64_t nch = cs.nch;
    const uint64_t nwg = (nch + FF_LIN_THREADS - 1) / FF_LIN_THREADS;
    int mine = 0;
    for (uint64_t i = threadIdx.x; i < nwg; i += SC_THREADS) mine |= cs.wgflag[i] != 0;
    const int any = __syncthreads_or(mine);
    if (threadIdx.x == 0) {
        if (st == 0) *cs.mode = (uint32_t)any;
        *cs.lenkey[st] = 0;
    }
    if (!any) return;
    const FFCarrySum *sums = cs.sums + (size_t)st * nch;
    int8_t *cin = cs.cin + (size_t)st * nch;
    const uint64_t per = (nch + SC_THREADS - 1) / SC_THREADS;
    const uint64_t k0 = threadIdx.x * per, k1 = k0 + per < nch ? k0 + per : nch;
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

__device__ __forceinline__ void book_top(unsigned long long *key, const uint64_t *d, uint64_t i0, uint64_t i1) {
    for (uint64_t i = i1; i-- > i0;) {
        const uint64_t v = d[i];
        if (v) {
            atomicMax(key, ((unsigned long long)(i + 1) << 8) | (unsigned)(64 - __clzll((long long)v)));
            return;
        }
    }
}

// Adds c_k into chunk k of each stream (ripple confined to the chunk) and
// books the top nonzero limb.
__global__ __launch_bounds__(256) void k_carry_apply(FFCarryStreams cs) {
    const int st = blockIdx.y;
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= cs.nch) return;
    const int c = *cs.mode ? cs.cin[(size_t)st * cs.nch + k] : (k ? cs.sums[(size_t)st * cs.nch + k - 1].e : 0);
    uint64_t *d = cs.d[st];
    const uint64_t i0 = k * FF_CH, i1 = i0 + FF_CH < cs.len ? i0 + FF_CH : cs.len;
    if (c > 0) {
        uint64_t v = (uint64_t)c;
        for (uint64_t i = i0; i < i1 && v; i++) {
            const uint64_t old = d[i];
            d[i] = old + v;
            v = d[i] < old ? 1 : 0;
        }
    } else if (c < 0) {
        uint64_t v = (uint64_t)(-c);
        for (uint64_t i = i0; i < i1 && v; i++) {
            const uint64_t old = d[i];
            d[i] = old - v;
            v = old < v ? 1 : 0;
        }
    }
    if (i1 > cs.lo_top[st]) book_top(cs.lenkey[st], d, i0 > cs.lo_top[st] ? i0 : cs.lo_top[st], i1);
}

// d[0..len) += v (v may be negative; the result stays >= 0): one thread, the
// ripple dies after a limb or two (big_add_u64_inplace, bigint.c:173-182);
// then rebooks the top limb.
__global__ void k_add_small(uint64_t *d, uint64_t len, int64_t v, unsigned long long *lenkey, uint64_t lo_top) {
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
    *lenkey = 0;
    book_top(lenkey, d, lo_top, len);
}

extern "C" hipError_t ff_launch_step(const FFStepArgs *a, hipStream_t st) {
    const uint64_t nch = (a->len + FF_CH - 1) / FF_CH;
    ff_prof_begin(FK_STEP, st);
    hipLaunchKernelGGL(k_step_local, dim3((unsigned)((nch + FF_LIN_THREADS - 1) / FF_LIN_THREADS)),
                       dim3(FF_LIN_THREADS), 0, st, *a);
    ff_prof_end(FK_STEP, 8.0 * (a->n1 + a->n2 + 2 * a->len) + 2 * 16.0 * nch, st);
    return hipGetLastError();
}

extern "C" hipError_t ff_launch_last(const FFLastArgs *a, hipStream_t st) {
    const uint64_t nch = (a->len + FF_CH - 1) / FF_CH;
    ff_prof_begin(FK_STEP, st);
    hipLaunchKernelGGL(k_last_local, dim3((unsigned)((nch + FF_LIN_THREADS - 1) / FF_LIN_THREADS)),
                       dim3(FF_LIN_THREADS), 0, st, *a);
    ff_prof_end(FK_STEP, 8.0 * (a->nu + a->nv + (a->Y ? 2 : 1) * a->len) + 2 * 16.0 * nch, st);
    return hipGetLastError();
}

extern "C" hipError_t ff_launch_carry(const FFCarryStreams *cs, hipStream_t st) {
    ff_prof_begin(FK_CARRY, st);
    hipLaunchKernelGGL(k_carry_scan, dim3(cs->nstreams), dim3(SC_THREADS), 0, st, *cs);
    hipLaunchKernelGGL(k_carry_apply, dim3((unsigned)((cs->nch + 255) / 256), cs->nstreams), dim3(256), 0, st, *cs);
    ff_prof_end(FK_CARRY, cs->nstreams * 16.0 * cs->nch, st);
    return hipGetLastError();
}

extern "C" hipError_t ff_launch_add_small(uint64_t *d, uint64_t len, int64_t v, unsigned long long *lenkey,
                                          uint64_t lo_top, hipStream_t st) {
    hipLaunchKernelGGL(k_add_small, dim3(1), dim3(64), 0, st, d, len, v, lenkey, lo_top);
    return hipGetLastError();
}

// d[0] = v, d[1..n) = 0: a fill kernel rather than hipMemsetAsync, so the
// sequence stays safe to capture in a HIP graph.
__global__ void k_fill(uint64_t *d, uint64_t n, uint64_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        d[i] = i ? 0 : v;
}

extern "C" hipError_t ff_launch_fill(uint64_t *d, uint64_t n, uint64_t v, hipStream_t st) {
    const unsigned blocks = (unsigned)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024);
    hipLaunchKernelGGL(k_fill, dim3(blocks ? blocks : 1), dim3(256), 0, st, d, n, v);
    return hipGetLastError();
}

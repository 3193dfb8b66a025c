// Linear limb passes on gfx950: the last-step factors of F(n), small
// additions and fills.
//
// The reference computes its linear passes in parallel chunks, each from a
// zero carry-in, and fixes the chunk boundaries afterwards
// (fib.c:19-62, 111-143, add_small_at at fib.c:64-70, 94-99). Here a pass
// is one launch: every limb has its own thread and the carries are binary
// chains resolved by wavefront ballots and decoupled look-back across
// workgroups (ff_lookback.h).
#include <stdint.h>

#include "ff_internal.h"
#include "ff_limbs.h"
#include "ff_lookback.h"
#include "ff_runtime.h"

typedef unsigned __int128 u128;

// fib.c:123-143: X = 2u + v and Y = 2u - v (Y >= 0) over len limbs, both
// chains in one look-back round, four limbs per thread; 2u takes bit 63 of
// limb i - 1 of u.
#define LAST_LPT 4
#define LAST_LIMBS (LB_T * LAST_LPT)

__global__ __launch_bounds__(LB_T) void k_last(FFLastArgs a, uint32_t *fault, int resident) {
    const uint64_t nblk = (a.len + LAST_LIMBS - 1) / LAST_LIMBS;
    if (blockIdx.x >= nblk) return;
    __shared__ LbShared S;
    const uint32_t b = lb_ticket(S, a.lb, resident != 0, blockIdx.x);
    const uint64_t i0 = (uint64_t)b * LAST_LIMBS, ib = i0 + (uint64_t)threadIdx.x * LAST_LPT;
    u128 sx[LAST_LPT], sy[LAST_LPT];
    uint64_t prev = ib >= 1 && ib - 1 < a.nu ? a.u[ib - 1] : 0;
#pragma unroll
    for (int q = 0; q < LAST_LPT; q++) {
        const uint64_t i = ib + q;
        const bool live = i < a.len;
        const uint64_t x = live && i < a.nu ? a.u[i] : 0;
        const uint64_t y = live && i < a.nv ? a.v[i] : 0;
        const uint64_t x2 = (x << 1) | (prev >> 63);
        prev = x;
        sx[q] = live ? (u128)x2 + y : (u128)~0ull;
        sy[q] = live ? (u128)x2 + ~y : (u128)~0ull;  // Y = 2u + ~v + 1
    }
    bool g[2], p[2];
    lb_group_gp<LAST_LPT>(sx, &g[0], &p[0]);
    lb_group_gp<LAST_LPT>(sy, &g[1], &p[1]);
    const uint32_t c = lb_round<2>(S, a.lb + 2, b, g, p, 2u, fault);
    uint64_t out[2 * LAST_LPT];
    lb_group_apply<LAST_LPT>(sx, c & 1, out);
    lb_group_apply<LAST_LPT>(sy, (c >> 1) & 1, out + LAST_LPT);
#pragma unroll
    for (int q = 0; q < LAST_LPT; q++) {
        if (ib + q < a.len) {
            a.X[ib + q] = out[q];
            if (a.Y) a.Y[ib + q] = out[LAST_LPT + q];
        } else {
            out[q] = out[LAST_LPT + q] = 0;
        }
    }
    unsigned long long *const keys[2] = {a.keyX, a.keyY};
    if (a.Y) lb_keys<2, LAST_LPT>(a.lb + 2 + nblk, b, (uint32_t)nblk, i0, a.len, out, keys, fault);
    else lb_keys<1, LAST_LPT>(a.lb + 2 + nblk, b, (uint32_t)nblk, i0, a.len, out, keys, fault);
    if (lb_last(S, a.lb, a.lb + 3 + nblk, b, (uint32_t)nblk)) lb_clear(a.lb, 3 + nblk + lb_groups(nblk));
}

size_t ff_last_lb_words(uint64_t len) {
    const uint64_t nblk = (len + LAST_LIMBS - 1) / LAST_LIMBS;
    return 3 + nblk + lb_groups(nblk);
}

extern "C" hipError_t ff_launch_last(const FFLastArgs *a, hipStream_t st) {
    const uint64_t nblk = (a->len + LAST_LIMBS - 1) / LAST_LIMBS;
    ff_prof_begin(FK_STEP, st);
    const int res = nblk <= ff_resident_blocks((const void *)k_last, LB_T);
    hipLaunchKernelGGL(k_last, dim3((unsigned)nblk), dim3(LB_T), 0, st, *a, ff_lb_fault(), res);
    ff_prof_end(FK_STEP, 8.0 * (a->nu + a->nv + (a->Y ? 2 : 1) * a->len), st);
    return hipGetLastError();
}

// The top nonzero limb of d among the last three, as a length key.
__device__ __forceinline__ unsigned long long top_key(const uint64_t *d, uint64_t len) {
    const uint64_t lo = len >= 3 ? len - 3 : 0;
    for (uint64_t i = len; i-- > lo;)
        if (d[i]) return ((unsigned long long)(i + 1) << 8) | lb_blen(d[i]);
    return 0;
}

// d[0..len) += v (v may be negative; the result stays >= 0): one thread, the
// ripple dies after a limb or two (big_add_u64_inplace, bigint.c:173-182);
// then rebooks the top limb (among the last three).
__global__ void k_add_small(uint64_t *d, uint64_t len, int64_t v, unsigned long long *lenkey) {
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
    *lenkey = top_key(d, len);
}

extern "C" hipError_t ff_launch_add_small(uint64_t *d, uint64_t len, int64_t v, unsigned long long *lenkey,
                                          hipStream_t st) {
    hipLaunchKernelGGL(k_add_small, dim3(1), dim3(64), 0, st, d, len, v, lenkey);
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

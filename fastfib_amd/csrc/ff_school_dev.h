// Schoolbook products on the GPU, device side: the bodies of the
// one-workgroup and one-wavefront forms, shared by their list kernels and
// the batch work-queue kernel.
//
// The reference multiplies operands under 200 limbs without its NTT
// (mul_basic.c:64-106: schoolbook up to 64 limbs, bigint.c:257-320,
// Karatsuba above); so does this path when both operands are that short.
//
// Thread t owns columns k = 2t and 2t + 1 of the product: the column sum
// C_k = sum_{i+j=k} a_i b_j < nb 2^128 held in three words (lo, mid, hi,
// hi < 2^8). Writing the product as sum_k C_k 2^(64k), limb k gathers
// lo_k + mid_(k-1) + hi_(k-2) < 3 2^64, split as a 64-bit word w_k and a
// 2-bit overflow o_k; u_k = w_k + o_(k-1) <= 2^64 + 2, so the remaining
// carries are binary and resolve from wavefront ballots and the
// workgroup's wave summaries (the in-workgroup part of the look-back
// machinery, ff_lookback.h). A square forms each product a_i a_j (i < j)
// once and doubles it.
#pragma once
#include <stdint.h>

#include "ff_internal.h"
#include "ff_lookback.h"

#define FF_SCHOOL_T 256          // threads; two columns each: outputs up to 512 limbs
#define FF_SCHOOL_WAVE_LIMBS 128  // operand limbs per wave (na + nb + 1 <= 128)

struct FFSchoolShared {
    uint64_t A[FF_SCHOOL_LIMBS], B[FF_SCHOOL_LIMBS];
    uint64_t lo[2 * FF_SCHOOL_T + 2], mid[2 * FF_SCHOOL_T + 2];
    uint32_t hi[2 * FF_SCHOOL_T + 2];
    LbShared S;
    uint32_t st_word;
    unsigned long long best;
};

// four wavefronts' operands and column sums
struct FFSchoolWaveShared {
    uint64_t A[4][FF_SCHOOL_WAVE_LIMBS], B[4][FF_SCHOOL_WAVE_LIMBS];
    uint64_t lo[4][FF_SCHOOL_WAVE_LIMBS + 2], mid[4][FF_SCHOOL_WAVE_LIMBS + 2];
    uint32_t hi[4][FF_SCHOOL_WAVE_LIMBS + 2];
};

typedef unsigned __int128 ff_school_u128;

// (lo, mid, hi) += x * y
__device__ __forceinline__ void ff_school_mac(uint64_t x, uint64_t y, uint64_t &lo, uint64_t &mid, uint32_t &hi) {
    const uint64_t pl = x * y, ph = __umul64hi(x, y);  // ph <= 2^64 - 2
    lo += pl;
    const ff_school_u128 m = (ff_school_u128)mid + ph + (lo < pl ? 1u : 0u);
    mid = (uint64_t)m;
    hi += (uint32_t)(m >> 64);
}

__device__ __forceinline__ void school_body(const FFSchoolItem &it, FFSchoolShared &sh) {
    uint64_t *A = sh.A, *B = sh.B, *lo = sh.lo, *mid = sh.mid;
    uint32_t *hi = sh.hi;
    LbShared &S = sh.S;
    uint32_t &st_word = sh.st_word;
    unsigned long long &best = sh.best;
    const int t = threadIdx.x;
    const int na = (int)it.na, nb = (int)it.nb, nout = na + nb + 1;
    const bool sq = it.a == it.b;
    for (int i = t; i < na; i += FF_SCHOOL_T) A[i] = it.a[i];
    for (int i = t; i < nb; i += FF_SCHOOL_T) B[i] = it.b[i];
    if (t == 0) best = 0;
    __syncthreads();
    // column sums, stored two slots up (slots 0, 1 stay zero: the columns
    // below limb 0)
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const int k = 2 * t + q;
        uint64_t l = 0, m = 0;
        uint32_t h = 0;
        if (k < na + nb - 1) {
            const int i0 = k - nb + 1 > 0 ? k - nb + 1 : 0, i1 = k < na - 1 ? k : na - 1;
            if (!sq) {
                for (int i = i0; i <= i1; i++) ff_school_mac(A[i], B[k - i], l, m, h);
            } else {
                // pairs i < j once, doubled, then the square term
                const int half = k >= 1 ? (k - 1) / 2 : -1;  // the largest i with i < k - i
                const int h1 = half < i1 ? half : i1;
                for (int i = i0; i <= h1; i++) ff_school_mac(A[i], A[k - i], l, m, h);
                h = (h << 1) | (uint32_t)(m >> 63);
                m = (m << 1) | (l >> 63);
                l <<= 1;
                if (!(k & 1) && k / 2 <= i1 && k / 2 >= i0) ff_school_mac(A[k / 2], A[k / 2], l, m, h);
            }
        }
        lo[k + 2] = l, mid[k + 2] = m, hi[k + 2] = h;
    }
    if (t < 2) lo[t] = 0, mid[t] = 0, hi[t] = 0;
    __syncthreads();
    // limb k: lo_k + mid_(k-1) + hi_(k-2) as (w_k, o_k); u_k = w_k + o_(k-1)
    lb_u128 u[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const int k = 2 * t + q;
        const ff_school_u128 s = (ff_school_u128)lo[k + 2] + mid[k + 1] + hi[k];
        const ff_school_u128 sp = k >= 1 ? (ff_school_u128)lo[k + 1] + mid[k] + (k >= 2 ? hi[k - 1] : 0u) : 0;
        u[q] = (ff_school_u128)(uint64_t)s + (uint64_t)(sp >> 64);
        if (k >= nout) u[q] = (ff_school_u128)~0ull;  // past the product: a carry passes through
    }
    bool g, p;
    lb_group_gp<2>(u, &g, &p);
    const uint32_t c = lb_round<1>(S, &st_word, 0, &g, &p, 0, nullptr);
    uint64_t v[2];
    lb_group_apply<2>(u, c, v);
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const int k = 2 * t + q;
        if (k < nout) {
            it.out[k] = v[q];
            if (v[q]) atomicMax(&best, ((unsigned long long)(k + 1) << 8) | lb_blen(v[q]));
        }
    }
    __syncthreads();
    if (t == 0) *it.key = best;
}

// Products of at most 128 output limbs, one per wavefront (four per
// workgroup): two columns per lane, the carries from one ballot addition.

// live: the wave has an item (every wave reaches the workgroup barriers).
__device__ __forceinline__ void school_wave(const FFSchoolItem &it, bool live, uint64_t *A, uint64_t *B,
                                            uint64_t *lo, uint64_t *mid, uint32_t *hi) {
    const int lane = threadIdx.x & 63;
    const int na = live ? (int)it.na : 0, nb = live ? (int)it.nb : 0, nout = na + nb + 1;
    const bool sq = live && it.a == it.b;
    for (int i = lane; i < na; i += 64) A[i] = it.a[i];
    for (int i = lane; i < nb; i += 64) B[i] = it.b[i];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const int k = 2 * lane + q;
        uint64_t l = 0, m = 0;
        uint32_t h = 0;
        if (k < na + nb - 1) {
            const int i0 = k - nb + 1 > 0 ? k - nb + 1 : 0, i1 = k < na - 1 ? k : na - 1;
            if (!sq) {
                for (int i = i0; i <= i1; i++) ff_school_mac(A[i], B[k - i], l, m, h);
            } else {
                const int half = k >= 1 ? (k - 1) / 2 : -1;
                const int h1 = half < i1 ? half : i1;
                for (int i = i0; i <= h1; i++) ff_school_mac(A[i], A[k - i], l, m, h);
                h = (h << 1) | (uint32_t)(m >> 63);
                m = (m << 1) | (l >> 63);
                l <<= 1;
                if (!(k & 1) && k / 2 <= i1 && k / 2 >= i0) ff_school_mac(A[k / 2], A[k / 2], l, m, h);
            }
        }
        lo[k + 2] = l, mid[k + 2] = m, hi[k + 2] = h;
    }
    if (lane < 2) lo[lane] = 0, mid[lane] = 0, hi[lane] = 0;
    __syncthreads();
    if (!live) return;  // past the last barrier
    lb_u128 u[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const int k = 2 * lane + q;
        const ff_school_u128 s = (ff_school_u128)lo[k + 2] + mid[k + 1] + hi[k];
        const ff_school_u128 sp = k >= 1 ? (ff_school_u128)lo[k + 1] + mid[k] + (k >= 2 ? hi[k - 1] : 0u) : 0;
        u[q] = (ff_school_u128)(uint64_t)s + (uint64_t)(sp >> 64);
        if (k >= nout) u[q] = (ff_school_u128)~0ull;
    }
    bool g, p;
    lb_group_gp<2>(u, &g, &p);
    const uint64_t G = __ballot(g), Bp = G | __ballot(p);
    const uint32_t c = (uint32_t)((((G + Bp) ^ G ^ Bp) >> lane) & 1);
    uint64_t v[2];
    lb_group_apply<2>(u, c, v);
    unsigned long long best = 0;
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const int k = 2 * lane + q;
        if (k < nout) {
            it.out[k] = v[q];
            if (v[q]) best = ((unsigned long long)(k + 1) << 8) | lb_blen(v[q]);
        }
    }
    // the top nonzero limb: the highest lane with one (limb order follows lanes)
    const uint64_t has = __ballot(best != 0);
    if (has) {
        const int top = 63 - __clzll((long long)has);
        const unsigned long long kb = __shfl(best, top);
        if (lane == 0) *it.key = kb;
    } else if (lane == 0) {
        *it.key = 0;
    }
}

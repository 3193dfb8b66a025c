// CRT device code shared by the CRT kernels, the fused small doubling step
// and the chain of first doubling steps: Garner's recombination of the
// three residues, the limb sums of shifted coefficients, and the
// one-workgroup carry lookahead over up to 3 * SS_T limbs in registers.
#pragma once
#include <stdint.h>

#include "ff_internal.h"

typedef unsigned __int128 u128;

struct FFGarner {
    uint32_t p1, p2, p3;
    uint32_t pinv2, pinv3;
    uint32_t i12, i12p, i13, i13p, i23, i23p;  // pi^-1 mod pj, Montgomery form, and * pinv
    uint64_t p12;
};
FFGarner ff_garner_consts(void);  // ff_crt.hip

// One reduction of x < 2q below q: x - q wraps above x when x < q.
__device__ __forceinline__ uint32_t ff_sub_once(uint32_t x, uint32_t q) {
    const uint32_t d = x - q;
    return d < x ? d : x;
}

// X = x1 + x2 p1 + x3 p1 p2 from canonical residues (crt.c:33-64). The last
// product x3 p1 p2 (31 x 62 bits) is formed from p1 p2's 32-bit halves.
__device__ __forceinline__ u128 ff_garner(uint32_t r1, uint32_t r2, uint32_t r3, const FFGarner &G) {
    const uint32_t x1 = r1;
    const uint32_t x1m2 = ff_sub_once(x1, G.p2);  // p1 < 2 p2
    const uint32_t x2 = ff_mont_w(r2 - x1m2 + G.p2, G.i12, G.i12p, G.p2);
    const uint32_t x1m3 = ff_sub_once(x1, G.p3);  // p1 < 2 p3
    const uint32_t t3 = ff_mont_w(r3 - x1m3 + G.p3, G.i13, G.i13p, G.p3);
    const uint32_t x2m3 = ff_sub_once(x2, G.p3);  // p2 < 2 p3
    const uint32_t x3 = ff_mont_w(t3 - x2m3 + G.p3, G.i23, G.i23p, G.p3);
    const uint64_t lo = (uint64_t)x2 * G.p1 + x1;                       // < 2^62
    const uint64_t a = (uint64_t)x3 * (uint32_t)G.p12 + lo;             // < 2^63 + 2^62
    const uint64_t b = (uint64_t)x3 * (uint32_t)(G.p12 >> 32);          // < 2^61
    const uint64_t s = a + (b << 32);
    return ((u128)((b >> 32) + (s < a ? 1u : 0u)) << 64) | s;
}

// First coefficient touching limb i: coefficient j spans bits [j bw, j bw + 93).
// (A 32-bit division below 2^32 bits: a fraction of the 64-bit one's cost.)
__device__ __forceinline__ uint64_t first_coef(uint64_t i, int bw) {
    const uint64_t lo = 64 * i;
    if (lo < 93) return 0;
    if (lo - 93 < (1ull << 32)) return (uint32_t)(lo - 93) / (uint32_t)bw + 1;
    return (lo - 93) / bw + 1;
}

// T_i for the n consecutive limbs i0, i0 + 1, .. (n <= N) below i1, in one
// walk over the coefficients: the first one touching each limb follows the
// previous limb's by at most a few steps (one division per walk).
template <int N>
__device__ __forceinline__ void limb_sums(uint64_t i0, uint64_t i1, int bw, const uint64_t *xlo, const uint32_t *xhi,
                                          uint64_t jbase, uint64_t jend, u128 *T) {
    uint64_t j = first_coef(i0, bw);
    if (j < jbase) j = jbase;
#pragma unroll
    for (int q = 0; q < N; q++) {
        T[q] = 0;
        const uint64_t i = i0 + q;
        if (i >= i1) continue;
        const uint64_t bit = 64 * i;
        while (j < jend && j * (uint64_t)bw + 93 <= bit) j++;
        for (uint64_t jj = j; jj < jend; jj++) {
            const uint64_t jb = jj * (uint64_t)bw;
            if (jb >= bit + 64) break;
            const u128 X = ((u128)xhi[jj - jbase] << 64) | xlo[jj - jbase];
            T[q] += jb >= bit ? (uint64_t)(X << (unsigned)(jb - bit)) : (uint64_t)(X >> (unsigned)(bit - jb));
        }
    }
}

// T_i = sum over the coefficients touching limb i of digit i of X_j 2^(j bw).
__device__ __forceinline__ u128 limb_sum(uint64_t i, int bw, uint64_t j0, const uint64_t *xlo, const uint32_t *xhi,
                                         uint64_t jbase, uint64_t jend) {
    u128 T = 0;
    const uint64_t bit = 64 * i;
    for (uint64_t j = first_coef(i, bw) > j0 ? first_coef(i, bw) : j0; j < jend; j++) {
        const uint64_t jb = j * (uint64_t)bw;
        if (jb >= bit + 64) break;
        const u128 X = ((u128)xhi[j - jbase] << 64) | xlo[j - jbase];
        uint64_t d;
        if (jb >= bit) d = (uint64_t)(X << (unsigned)(jb - bit));
        else d = (uint64_t)(X >> (unsigned)(bit - jb));
        T += d;
    }
    return T;
}

// ------------------------------------------------------------------ one-workgroup steps
//
// Numbers of at most 3 SS_T limbs held in registers, thread t owning limbs
// t, t + SS_T, t + 2 SS_T: the CRT of a square's residues into them and
// additions by workgroup-wide carry lookahead (wavefront ballots, the
// SS_T / 64 wave summaries composed in LDS, chunks of SS_T limbs in order).

#define SS_CH 3  // chunks of SS_T limbs: len <= 3 SS_T

// Coefficients touching SS_T + 1 limbs at bw >= 32, with margin. The wave
// summaries are double-buffered by the lookahead's parity (SSRun.par), so
// consecutive lookaheads need one barrier each.
template <int SS_T>
struct SSShared {
    static constexpr int MAXC = ((SS_T + 1) * 64 + 93) / 32 + 64;
    uint64_t xlo[MAXC];
    uint32_t xhi[MAXC];
    uint64_t ex[SS_T];
    uint32_t wg[2][SS_T / 64], wp[2][SS_T / 64];
    uint32_t wg2[2][SS_T / 64], wp2[2][SS_T / 64];  // a second chain (ss_crt_pair)
    uint64_t hi_last, hi_last2;
};

// Per-thread state of a sequence of lookaheads (the same in every thread):
// the carry into the next chunk and the summary buffer to write next.
struct SSRun {
    uint32_t cin;
    int par;
};

// Carry into this thread's limb for a chunk: u (limb sum), g (generates),
// p (propagates), R.cin (carry into the chunk); returns u + carry and sets
// R.cin to the chunk's carry-out. Inside a wavefront the carries come from
// one 64-bit addition of the ballots (A = g, B = g | p: C = (A + B + c) ^ A
// ^ B); every wavefront then composes the SS_T / 64 wave summaries the same
// way, so one barrier suffices. Every thread of the block calls it.
template <int SS_T>
__device__ __forceinline__ uint64_t ss_lookahead(SSShared<SS_T> &S, SSRun &R, uint64_t u, bool g, bool p) {
    constexpr int NW = SS_T / 64;
    static_assert(NW < 64, "wave summaries fit one ballot");
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t G = __ballot(g), B = G | __ballot(p);
    if (lane == 0) {
        S.wg[R.par][w] = (G + B) < G;        // carry out with no carry in
        S.wp[R.par][w] = (B == ~0ull && G == 0);  // every lane propagates
    }
    __syncthreads();
    const bool wgl = lane < NW && S.wg[R.par][lane];
    const bool wpl = lane < NW && S.wp[R.par][lane];
    const uint64_t WG = __ballot(wgl), WB = WG | __ballot(wpl);
    const uint64_t WC = (WG + WB + (uint64_t)R.cin) ^ WG ^ WB;  // bit k: carry into wave k; bit NW: out
    R.cin = (uint32_t)((WC >> NW) & 1);
    R.par ^= 1;
    const uint64_t C = (G + B + ((WC >> w) & 1)) ^ G ^ B;
    return u + ((C >> lane) & 1);
}

// Limbs [0, len) of the product whose residues are y ([3][L]) into out[c]
// (limb c * SS_T + threadIdx.x).
template <int SS_T>
__device__ void ss_crt(SSShared<SS_T> &S, SSRun &R, const uint32_t *y, uint32_t L, int bw, uint32_t conv,
                       uint32_t len, const FFGarner &G, uint64_t out[SS_CH]) {
    const int t = threadIdx.x;
    R.cin = 0;
    for (int c = 0; c < SS_CH; c++) {
        out[c] = 0;
        const uint64_t i0 = (uint64_t)c * SS_T;
        if (i0 >= len || conv == 0) continue;  // uniform across the block
        const uint64_t i1 = i0 + SS_T < len ? i0 + SS_T : len;
        const uint64_t jbase = first_coef(i0 ? i0 - 1 : 0, bw);
        uint64_t jend = (64 * i1 + bw - 1) / bw;
        if (jend > conv) jend = conv;
        for (uint64_t j = jbase + t; j < jend; j += SS_T) {
            const u128 X = ff_garner(y[j], y[L + j], y[2 * (size_t)L + j], G);
            S.xlo[j - jbase] = (uint64_t)X;
            S.xhi[j - jbase] = (uint32_t)(X >> 64);
        }
        __syncthreads();
        const uint64_t i = i0 + t;
        const bool live = i < i1;
        const u128 T = live ? limb_sum(i, bw, jbase, S.xlo, S.xhi, jbase, jend) : 0;
        S.ex[t] = (uint64_t)(T >> 64);
        __syncthreads();
        // hi_last: the previous chunk's last high word (written below, read
        // here one chunk later, with two barriers between)
        const uint64_t hprev = t ? S.ex[t - 1] : (c ? S.hi_last : 0);
        const u128 u = (u128)(uint64_t)T + hprev;
        const bool g = live && (u >> 64) != 0;
        const bool p = !live || (!g && (uint64_t)u == ~0ull);
        __syncthreads();
        if (t == SS_T - 1) S.hi_last = S.ex[t];
        const uint64_t v = ss_lookahead(S, R, (uint64_t)u, g, p);
        out[c] = live ? v : 0;
    }
}

// ss_lookahead for two independent chains at once (u1, g1, p1 with carry
// R.cin; u2, g2, p2 with carry cin2): both chains' wave summaries behind one
// barrier.
template <int SS_T>
__device__ __forceinline__ void ss_lookahead2(SSShared<SS_T> &S, SSRun &R, uint32_t &cin2, uint64_t &u1, bool g1,
                                              bool p1, uint64_t &u2, bool g2, bool p2) {
    constexpr int NW = SS_T / 64;
    static_assert(NW < 64, "wave summaries fit one ballot");
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t Ga = __ballot(g1), Ba = Ga | __ballot(p1);
    const uint64_t Gb = __ballot(g2), Bb = Gb | __ballot(p2);
    if (lane == 0) {
        S.wg[R.par][w] = (Ga + Ba) < Ga;
        S.wp[R.par][w] = (Ba == ~0ull && Ga == 0);
        S.wg2[R.par][w] = (Gb + Bb) < Gb;
        S.wp2[R.par][w] = (Bb == ~0ull && Gb == 0);
    }
    __syncthreads();
    {
        const uint64_t WG = __ballot(lane < NW && S.wg[R.par][lane]);
        const uint64_t WB = WG | __ballot(lane < NW && S.wp[R.par][lane]);
        const uint64_t WC = (WG + WB + (uint64_t)R.cin) ^ WG ^ WB;
        R.cin = (uint32_t)((WC >> NW) & 1);
        const uint64_t C = (Ga + Ba + ((WC >> w) & 1)) ^ Ga ^ Ba;
        u1 += (C >> lane) & 1;
    }
    {
        const uint64_t WG = __ballot(lane < NW && S.wg2[R.par][lane]);
        const uint64_t WB = WG | __ballot(lane < NW && S.wp2[R.par][lane]);
        const uint64_t WC = (WG + WB + (uint64_t)cin2) ^ WG ^ WB;
        cin2 = (uint32_t)((WC >> NW) & 1);
        const uint64_t C = (Gb + Bb + ((WC >> w) & 1)) ^ Gb ^ Bb;
        u2 += (C >> lane) & 1;
    }
    R.par ^= 1;
}

// ss_crt for two products of conv coefficients each at once (residues y1,
// y2, [3][L]), limbs [0, len) into o1 and o2: one pass of barriers for both.
// The second product's coefficients and high words go to xlo2 / xhi2 (as
// many entries as S.xlo) and ex2 (SS_T entries), the caller's memory.
template <int SS_T>
__device__ void ss_crt_pair(SSShared<SS_T> &S, SSRun &R, uint64_t *xlo2, uint32_t *xhi2, uint64_t *ex2,
                            const uint32_t *y1, const uint32_t *y2, uint32_t L, int bw, uint32_t conv, uint32_t len,
                            const FFGarner &G, uint64_t o1[SS_CH], uint64_t o2[SS_CH]) {
    const int t = threadIdx.x;
    R.cin = 0;
    uint32_t cin2 = 0;
    for (int c = 0; c < SS_CH; c++) {
        o1[c] = o2[c] = 0;
        const uint64_t i0 = (uint64_t)c * SS_T;
        if (i0 >= len || conv == 0) continue;  // uniform across the block
        const uint64_t i1 = i0 + SS_T < len ? i0 + SS_T : len;
        const uint64_t jbase = first_coef(i0 ? i0 - 1 : 0, bw);
        uint64_t jend = (64 * i1 + bw - 1) / bw;
        if (jend > conv) jend = conv;
        for (uint64_t j = jbase + t; j < jend; j += SS_T) {
            const u128 X1 = ff_garner(y1[j], y1[L + j], y1[2 * (size_t)L + j], G);
            const u128 X2 = ff_garner(y2[j], y2[L + j], y2[2 * (size_t)L + j], G);
            S.xlo[j - jbase] = (uint64_t)X1;
            S.xhi[j - jbase] = (uint32_t)(X1 >> 64);
            xlo2[j - jbase] = (uint64_t)X2;
            xhi2[j - jbase] = (uint32_t)(X2 >> 64);
        }
        __syncthreads();
        const uint64_t i = i0 + t;
        const bool live = i < i1;
        const u128 T1 = live ? limb_sum(i, bw, jbase, S.xlo, S.xhi, jbase, jend) : 0;
        const u128 T2 = live ? limb_sum(i, bw, jbase, xlo2, xhi2, jbase, jend) : 0;
        S.ex[t] = (uint64_t)(T1 >> 64);
        ex2[t] = (uint64_t)(T2 >> 64);
        __syncthreads();
        // hi_last, hi_last2: the previous chunk's last high words (as in ss_crt)
        const uint64_t h1 = t ? S.ex[t - 1] : (c ? S.hi_last : 0);
        const uint64_t h2 = t ? ex2[t - 1] : (c ? S.hi_last2 : 0);
        const u128 v1 = (u128)(uint64_t)T1 + h1, v2 = (u128)(uint64_t)T2 + h2;
        const bool g1 = live && (v1 >> 64) != 0, g2 = live && (v2 >> 64) != 0;
        const bool p1 = !live || (!g1 && (uint64_t)v1 == ~0ull), p2 = !live || (!g2 && (uint64_t)v2 == ~0ull);
        __syncthreads();
        if (t == SS_T - 1) {
            S.hi_last = S.ex[t];
            S.hi_last2 = ex2[t];
        }
        uint64_t u1 = (uint64_t)v1, u2 = (uint64_t)v2;
        ss_lookahead2(S, R, cin2, u1, g1, p1, u2, g2, p2);
        o1[c] = live ? u1 : 0;
        o2[c] = live ? u2 : 0;
    }
}

// out = a + b' + cin0 over len limbs in registers (b' = b or ~b).
template <int SS_T>
__device__ void ss_add(SSShared<SS_T> &S, SSRun &R, const uint64_t a[SS_CH], const uint64_t b[SS_CH], bool neg,
                       uint32_t cin0, uint32_t len, uint64_t out[SS_CH]) {
    R.cin = cin0;
    for (int c = 0; c < SS_CH; c++) {
        const uint64_t i = (uint64_t)c * SS_T + threadIdx.x;
        if ((uint64_t)c * SS_T >= len) {
            out[c] = 0;
            continue;
        }
        const bool live = i < len;
        const uint64_t bv = neg ? ~b[c] : b[c];
        const uint64_t u = a[c] + bv;
        const bool g = live && u < a[c];
        const bool p = !live || u == ~0ull;
        const uint64_t v = ss_lookahead(S, R, u, g, p);
        out[c] = live ? v : 0;
    }
}

// The linear part of a doubling step from s1 = F(k)^2 and s2 = F(k-1)^2
// (fib.c:36-62): A = s1 + s2 = F(2k-1), B = 4 s1 - s2 + cb0 = F(2k+1),
// C = B - A = F(2k), cb0 = 2 (-1)^k; K = bit ? B : A.
template <int SS_T>
__device__ void ss_linear(SSShared<SS_T> &S, SSRun &R, const uint64_t s1[SS_CH], const uint64_t s2[SS_CH], int cb0,
                          uint32_t len, uint64_t K[SS_CH], uint64_t C[SS_CH], int bit) {
    const int t = threadIdx.x;
    uint64_t A[SS_CH], D[SS_CH], B[SS_CH];
    ss_add(S, R, s1, s2, false, 0, len, A);
    // 4 s1 (limb i takes the top two bits of limb i - 1), then + cb0
    for (int c = 0; c < SS_CH; c++) {
        D[c] = 0;
        if ((uint64_t)c * SS_T >= len) continue;  // uniform
        S.ex[t] = s1[c];
        __syncthreads();
        const uint64_t prev = t ? S.ex[t - 1] : (c ? S.hi_last : 0);
        __syncthreads();
        if (t == SS_T - 1) S.hi_last = s1[c];
        D[c] = (s1[c] << 2) | (prev >> 62);
    }
    {
        // cb0 = +2: add 2; cb0 = -2: add 2^(64 len) - 2 (limb 0 = ~1, the rest all ones)
        uint64_t k[SS_CH];
        for (int c = 0; c < SS_CH; c++) {
            const uint64_t i = (uint64_t)c * SS_T + t;
            k[c] = cb0 >= 0 ? (i == 0 ? (uint64_t)cb0 : 0) : (i == 0 ? ~(uint64_t)(-cb0 - 1) : ~0ull);
        }
        ss_add(S, R, D, k, false, 0, len, D);
    }
    ss_add(S, R, D, s2, true, 1, len, B);
    ss_add(S, R, B, A, true, 1, len, C);
    for (int c = 0; c < SS_CH; c++) K[c] = bit ? B[c] : A[c];
}

// Books limb i (value v) of a len-limb result into its length key: the top
// nonzero limb among the last three, as (limbs << 8) | bits of that limb.
__device__ __forceinline__ void ss_book(unsigned long long *key, uint64_t i, uint64_t v, uint64_t len) {
    const uint64_t lo_top = len >= 3 ? len - 3 : 0;
    if (i >= lo_top && i < len && v)
        atomicMax(key, ((unsigned long long)(i + 1) << 8) | (unsigned)(64 - __clzll((long long)v)));
}

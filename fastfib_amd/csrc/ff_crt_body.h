// CRT recombination and carry propagation, device side: the limb sums of
// one or two products' residues (crt_sums) and the carry-resolving body of a
// CRT job (crt_body), shared by the CRT kernels, the fused doubling step and
// the batch work-queue kernel. The method is described in ff_crt.hip.
#pragma once
#include <stdint.h>

#include "ff_crt_dev.h"
#include "ff_internal.h"
#include "ff_lookback.h"

#define CRT_LIMBS 256  // limbs per block, one per thread
// Coefficients a block can touch: limbs [i0 - 2, i0 + 256) span 258 * 64 bits
// and a coefficient 93 more, so at most (258 * 64 + 93) / bw + 2 of them;
// 2077 at the narrowest width, 8 (mul_split.c:151-152).
#define CRT_MAXC 2112


// One product's residues for the limb sums.
struct CrtSrc {
    const uint32_t *y;  // [3][L]
    uint32_t L;
    int bw;
    uint64_t conv;
};

// Limb sums of NP products for the workgroup's limbs: thread t holds limbs
// i0 + t LPT + q (q < LPT), u[P][q] = lo(T_i) + hi(T_{i-1}) of product P,
// and 2^64 - 1 at and past nout (a carry passes through); with prev,
// thread 0 also gets each product's u at limb i0 - 1 (uprev[P]). The products' Garner
// values are formed in one loop (their loads overlap). xlo / xhi: NP x MAXC
// LDS words, hx: NP x 256 + 4 NP. This form sums each limb's shifted 93-bit
// Garner values directly and serves every width (crt_sums picks it below
// 32 bits).
template <int LPT, int NP, int MAXC, bool COH = false>
__device__ __forceinline__ void crt_sums_x(const CrtSrc *src, uint64_t nout, uint64_t i0, const FFGarner &G,
                                           uint64_t *xlo, uint32_t *xhi, uint64_t *hx, bool prev, u128 (*u)[LPT],
                                           u128 *uprev) {
    const int t = threadIdx.x;
    const uint64_t i1 = i0 + CRT_LIMBS * LPT < nout ? i0 + CRT_LIMBS * LPT : nout;
    uint64_t jbase[NP], jend[NP];
#pragma unroll
    for (int P = 0; P < NP; P++) {
        // coefficients touching limbs [i0 - 2, i1)
        const uint64_t lim = src[P].conv < src[P].L ? src[P].conv : src[P].L;
        jbase[P] = first_coef(i0 >= 2 ? i0 - 2 : 0, src[P].bw);
        jend[P] = (64 * i1 + src[P].bw - 1) / src[P].bw;
        if (jend[P] > lim) jend[P] = lim;
        if (jend[P] < jbase[P]) jend[P] = jbase[P];
    }
    const uint64_t nj = NP == 1 ? jend[0] - jbase[0]
                                : (jend[0] - jbase[0] > jend[NP - 1] - jbase[NP - 1] ? jend[0] - jbase[0]
                                                                                   : jend[NP - 1] - jbase[NP - 1]);
    // every residue load issued before the first Garner step: one HBM
    // latency per workgroup instead of one per loop trip
    constexpr int ITERS = (MAXC + CRT_LIMBS - 1) / CRT_LIMBS;
    uint32_t r[NP][ITERS][3];
#pragma unroll
    for (int it = 0; it < ITERS; it++) {
        const uint64_t k = t + (uint64_t)it * CRT_LIMBS;
#pragma unroll
        for (int P = 0; P < NP; P++) {
            const uint64_t j = jbase[P] + k;
            if (k < nj && j < jend[P]) {
                const uint32_t *y = src[P].y;
                const size_t L = src[P].L;
                r[P][it][0] = ff_ld<COH>(y + j);
                r[P][it][1] = ff_ld<COH>(y + (L + j));
                r[P][it][2] = ff_ld<COH>(y + (2 * L + j));
            }
        }
    }
#pragma unroll
    for (int it = 0; it < ITERS; it++) {
        const uint64_t k = t + (uint64_t)it * CRT_LIMBS;
#pragma unroll
        for (int P = 0; P < NP; P++) {
            if (k < nj && jbase[P] + k < jend[P]) {
                const u128 X = ff_garner(r[P][it][0], r[P][it][1], r[P][it][2], G);
                xlo[P * MAXC + k] = (uint64_t)X;
                xhi[P * MAXC + k] = (uint32_t)(X >> 64);
            }
        }
    }
    __syncthreads();
    const uint64_t ib = i0 + (uint64_t)t * LPT;
#pragma unroll
    for (int P = 0; P < NP; P++) {
        const int bw = src[P].bw;
        const uint64_t *xl = xlo + P * MAXC;
        const uint32_t *xh = xhi + P * MAXC;
        u128 T[LPT];
        limb_sums<LPT>(ib, i1, bw, xl, xh, jbase[P], jend[P], T);
            hx[P * CRT_LIMBS + t] = (uint64_t)(T[LPT - 1] >> 64);
        // the sums below the workgroup's first limb, by lanes 2P and 2P + 1
        // of the first wavefront together: T at i0 - 1 and at i0 - 2
        if (t >> 1 == P && i0 > (uint64_t)(t & 1)) {
            const u128 Tb = limb_sum(i0 - 1 - (t & 1), bw, jbase[P], xl, xh, jbase[P], jend[P]);
            hx[NP * CRT_LIMBS + 4 * P + 2 * (t & 1)] = (uint64_t)Tb;
            hx[NP * CRT_LIMBS + 4 * P + 2 * (t & 1) + 1] = (uint64_t)(Tb >> 64);
        } else if (t >> 1 == P) {
            hx[NP * CRT_LIMBS + 4 * P + 2 * (t & 1)] = 0;
            hx[NP * CRT_LIMBS + 4 * P + 2 * (t & 1) + 1] = 0;
        }
#pragma unroll
        for (int q = 0; q < LPT; q++) {
            const u128 lo = (u128)(uint64_t)T[q];
            if (q) u[P][q] = lo + (uint64_t)(T[q - 1] >> 64);
            else u[P][0] = lo;
        }
    }
    __syncthreads();
#pragma unroll
    for (int P = 0; P < NP; P++) {
        // hx[NP 256 + 4P ..]: lo, hi of T at i0 - 1, then at i0 - 2
        const uint64_t *tb = hx + NP * CRT_LIMBS + 4 * P;
        const uint64_t hp = t ? hx[P * CRT_LIMBS + t - 1] : tb[1];
        if (prev && t == 0) uprev[P] = (u128)tb[0] + tb[3];
        u[P][0] += hp;
#pragma unroll
        for (int q = 0; q < LPT; q++)
            if (ib + q >= nout) u[P][q] = (u128)~0ull;
    }
    __syncthreads();  // the LDS is reused by the next call
}

// x / bw for the workgroup's bit positions: a 32-bit division while x fits.
__device__ __forceinline__ uint64_t crt_div_bw(uint64_t x, int bw) {
    if (x < (1ull << 32)) return (uint32_t)x / (uint32_t)bw;
    return x / (uint64_t)bw;
}

// Sum for limb ell (relative to the workgroup's first limb) of the digit
// sums Z_m (LDS z[0, nz)), Z_m at bit zrel + m bw relative to that limb's
// first bit: T = sum over m of the 64 bits [64 ell, 64 ell + 64) of
// Z_m 2^(zrel + m bw). Z_m < 2^40 and bw >= 32, so the m with a shift
// s = zrel + m bw - 64 ell in (-40, 64) are at most four consecutive ones;
// the first is found from a float reciprocal of bw and checked exactly.
__device__ __forceinline__ u128 crt_zsum(const uint64_t *z, int nz, int zrel, int bw, float rbw, int ell) {
    const int lo = 64 * ell - 40;  // touching: zrel + m bw > lo and < lo + 104
    int m0 = (int)floorf((float)(lo - zrel) * rbw) + 1;
    m0 -= (zrel + (m0 - 1) * bw > lo) ? 1 : 0;
    m0 += (zrel + m0 * bw <= lo) ? 1 : 0;
    uint64_t acc = 0;
    uint32_t hi = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int m = m0 + c;
        const int s = zrel + m * bw - 64 * ell;
        const bool in = m >= 0 && m < nz && s < 64;
        const uint64_t zv = z[in ? m : 0];
        uint64_t d = s >= 0 ? zv << (s & 63) : zv >> ((-s) & 63);
        d = in ? d : 0;
        acc += d;
        hi += acc < d ? 1u : 0u;
    }
    return ((u128)hi << 64) | acc;
}

// crt_sums for widths of at least 32 bits (every product whose coefficients
// fill a 32-bit word: the doubling steps, the last product, large
// products). Each Garner value X_j < 2^93 is first split into its bw-bit
// digits, which regroup as Z_j = d0(X_j) + d1(X_{j-1}) + d2(X_{j-2}) < 2^40
// with N = sum_j Z_j 2^(j bw) unchanged (3 bw >= 96 > 93: three digits); a
// limb is then touched by at most four Z, each one 64-bit shift. Z_j
// replaces X in the low LDS words once every thread has read its three X
// (two more barriers). Positions are kept relative to the workgroup's
// first limb in 32-bit integers.
template <int LPT, int NP, int MAXC, bool COH = false>
__device__ __forceinline__ void crt_sums_z(const CrtSrc *src, uint64_t nout, uint64_t i0, const FFGarner &G,
                                           uint64_t *xlo, uint32_t *xhi, uint64_t *hx, bool prev, u128 (*u)[LPT],
                                           u128 *uprev) {
    const int t = threadIdx.x;
    const uint64_t i1 = i0 + CRT_LIMBS * LPT < nout ? i0 + CRT_LIMBS * LPT : nout;
    uint64_t xb0[NP];
    int nx[NP], nz[NP], zoff[NP], zrel[NP];
#pragma unroll
    for (int P = 0; P < NP; P++) {
        const int bw = src[P].bw;
        const uint64_t lim = src[P].conv < src[P].L ? src[P].conv : src[P].L;
        // Z_j (bits [j bw, j bw + 40)) touching limbs [i0 - 2, i1)
        const uint64_t lob = i0 >= 2 ? 64 * (i0 - 2) : 0;
        const uint64_t zb0 = lob >= 40 ? crt_div_bw(lob - 40, bw) + 1 : 0;
        uint64_t zb1 = crt_div_bw(64 * i1 + bw - 1, bw);
        if (zb1 > lim + 2) zb1 = lim + 2;  // Z_j = 0 from lim + 2 on
        if (zb1 < zb0) zb1 = zb0;
        xb0[P] = zb0 >= 2 ? zb0 - 2 : 0;
        uint64_t xb1 = zb1 < lim ? zb1 : lim;
        if (xb1 < xb0[P]) xb1 = xb0[P];
        nx[P] = (int)(xb1 - xb0[P]);
        nz[P] = (int)(zb1 - zb0);
        zoff[P] = (int)(zb0 - xb0[P]);
        zrel[P] = (int)((int64_t)(zb0 * (uint64_t)bw) - (int64_t)(64 * i0));
    }
    // every residue load issued before the first Garner step
    constexpr int ITERS = (MAXC + CRT_LIMBS - 1) / CRT_LIMBS;
    uint32_t r[NP][ITERS][3];
#pragma unroll
    for (int it = 0; it < ITERS; it++) {
        const int k = t + it * CRT_LIMBS;
#pragma unroll
        for (int P = 0; P < NP; P++) {
            if (k < nx[P]) {
                const uint32_t *y = src[P].y + xb0[P];
                const size_t L = src[P].L;
                r[P][it][0] = ff_ld<COH>(y + k);
                r[P][it][1] = ff_ld<COH>(y + (L + k));
                r[P][it][2] = ff_ld<COH>(y + (2 * L + k));
            }
        }
    }
#pragma unroll
    for (int it = 0; it < ITERS; it++) {
        const int k = t + it * CRT_LIMBS;
#pragma unroll
        for (int P = 0; P < NP; P++) {
            if (k < nx[P]) {
                const u128 X = ff_garner(r[P][it][0], r[P][it][1], r[P][it][2], G);
                xlo[P * MAXC + k] = (uint64_t)X;
                xhi[P * MAXC + k] = (uint32_t)(X >> 64);
            }
        }
    }
    __syncthreads();
    uint64_t z[NP][ITERS];
#pragma unroll
    for (int P = 0; P < NP; P++) {
        const int bw = src[P].bw;
        const uint64_t mask = (1ull << bw) - 1;
        const uint64_t *xl = xlo + P * MAXC;
        const uint32_t *xh = xhi + P * MAXC;
#pragma unroll
        for (int it = 0; it < ITERS; it++) {
            const int m = t + it * CRT_LIMBS;
            const int k = m + zoff[P];  // X_j of Z_j; X_{j-1}, X_{j-2} below it
            uint64_t v = 0;
            if (m < nz[P]) {
                if (k < nx[P]) v = xl[k] & mask;
                if (k >= 1 && k - 1 < nx[P]) v += ((xl[k - 1] >> bw) | ((uint64_t)xh[k - 1] << (64 - bw))) & mask;
                if (k >= 2 && k - 2 < nx[P]) v += (uint64_t)(xh[k - 2] >> (2 * bw - 64));
            }
            z[P][it] = v;
        }
    }
    __syncthreads();
#pragma unroll
    for (int P = 0; P < NP; P++) {
#pragma unroll
        for (int it = 0; it < ITERS; it++) {
            const int m = t + it * CRT_LIMBS;
            if (m < nz[P]) xlo[P * MAXC + m] = z[P][it];
        }
    }
    __syncthreads();
#pragma unroll
    for (int P = 0; P < NP; P++) {
        const int bw = src[P].bw;
        const float rbw = 1.0f / (float)bw;
        const uint64_t *zl = xlo + P * MAXC;
        u128 T[LPT];
#pragma unroll
        for (int q = 0; q < LPT; q++) T[q] = crt_zsum(zl, nz[P], zrel[P], bw, rbw, t * LPT + q);
        hx[P * CRT_LIMBS + t] = (uint64_t)(T[LPT - 1] >> 64);
        // the sums below the workgroup's first limb, by lanes 2P and 2P + 1
        // of the first wavefront together: T at i0 - 1 and at i0 - 2
        if (t >> 1 == P) {
            const bool below = i0 > (uint64_t)(t & 1);
            const u128 Tb = below ? crt_zsum(zl, nz[P], zrel[P], bw, rbw, -1 - (t & 1)) : (u128)0;
            hx[NP * CRT_LIMBS + 4 * P + 2 * (t & 1)] = (uint64_t)Tb;
            hx[NP * CRT_LIMBS + 4 * P + 2 * (t & 1) + 1] = (uint64_t)(Tb >> 64);
        }
#pragma unroll
        for (int q = 0; q < LPT; q++) {
            const u128 lo = (u128)(uint64_t)T[q];
            if (q) u[P][q] = lo + (uint64_t)(T[q - 1] >> 64);
            else u[P][0] = lo;
        }
    }
    __syncthreads();
    const uint64_t ib = i0 + (uint64_t)t * LPT;
#pragma unroll
    for (int P = 0; P < NP; P++) {
        // hx[NP 256 + 4P ..]: lo, hi of T at i0 - 1, then at i0 - 2
        const uint64_t *tb = hx + NP * CRT_LIMBS + 4 * P;
        const uint64_t hp = t ? hx[P * CRT_LIMBS + t - 1] : tb[1];
        if (prev && t == 0) uprev[P] = (u128)tb[0] + tb[3];
        u[P][0] += hp;
#pragma unroll
        for (int q = 0; q < LPT; q++)
            if (ib + q >= nout) u[P][q] = (u128)~0ull;
    }
    __syncthreads();  // the LDS is reused by the next call
}

// Limb sums (see crt_sums_x): the digit form for widths of at least 32
// bits, the direct form below.
template <int LPT, int NP, int MAXC, bool COH = false>
__device__ __forceinline__ void crt_sums(const CrtSrc *src, uint64_t nout, uint64_t i0, const FFGarner &G,
                                         uint64_t *xlo, uint32_t *xhi, uint64_t *hx, bool prev, u128 (*u)[LPT],
                                         u128 *uprev) {
    bool wide = true;
#pragma unroll
    for (int P = 0; P < NP; P++) wide = wide && src[P].bw >= 32;
    if (wide) crt_sums_z<LPT, NP, MAXC, COH>(src, nout, i0, G, xlo, xhi, hx, prev, u, uprev);
    else crt_sums_x<LPT, NP, MAXC, COH>(src, nout, i0, G, xlo, xhi, hx, prev, u, uprev);
}

// Garner slots for the coefficients touching 256 LPT + 2 limbs at widths of
// at least 32 bits (the fused step's, and the CRT's at four limbs per thread).
#define CRT_MAXC_W32(LPT) (((CRT_LIMBS * (LPT) + 2) * 64 + 93) / 32 + 3)

struct CrtShared {
    uint64_t xlo[CRT_MAXC];
    uint32_t xhi[CRT_MAXC];
    uint64_t hx[CRT_LIMBS + 4];
    LbShared S;
};

template <int LPT, bool COH = false>
__device__ __forceinline__ void crt_body(const FFCrtJob &J, uint32_t bx, const FFGarner &G, uint32_t *fault,
                                         CrtShared &sh, bool resident) {
    const uint64_t nblk = (J.nout + CRT_LIMBS * LPT - 1) / (CRT_LIMBS * LPT);
    if (bx >= nblk) return;
    uint64_t *xlo = sh.xlo, *hx = sh.hx;
    uint32_t *xhi = sh.xhi;
    LbShared &S = sh.S;
    const uint32_t b = lb_ticket(S, J.lb, resident, bx);
    const uint64_t i0 = (uint64_t)b * CRT_LIMBS * LPT, ib = i0 + (uint64_t)threadIdx.x * LPT;
    u128 u[1][LPT], unused[1];
    const CrtSrc src{J.y, J.L, J.bw, J.conv};
    crt_sums<LPT, 1, CRT_MAXC, COH>(&src, J.nout, i0, G, xlo, xhi, hx, false, u, unused);
    bool g, p;
    lb_group_gp<LPT>(u[0], &g, &p);
    const uint32_t c = lb_round<1>(S, J.lb + 2, b, &g, &p, 0, fault);
    uint64_t v[LPT];
    lb_group_apply<LPT>(u[0], c, v);
#pragma unroll
    for (int q = 0; q < LPT; q++)
        if (ib + q < J.nout) J.out[ib + q] = v[q];
    unsigned long long *const keys[1] = {J.lenkey};
    lb_keys<1, LPT>(J.lb + 2 + nblk, b, (uint32_t)nblk, i0, J.nout, v, keys, fault);
    if (lb_last(S, J.lb, J.lb + 3 + nblk, b, (uint32_t)nblk)) lb_clear(J.lb, 3 + nblk + lb_groups(nblk));
}

// Limbs per thread of a job: one (shortest latency per workgroup) while a
// workgroup per 256 limbs fits one wave of residency (~1280 workgroups);
// above, four, where the coefficients are at least 32 bits wide (those
// touching 1026 limbs fit CRT_MAXC).
#define CRT_ONE_WAVE_LIMBS (1280u * CRT_LIMBS)
__host__ __device__ static inline int crt_lpt(int bw, uint64_t nout) { return bw >= 32 && nout > CRT_ONE_WAVE_LIMBS ? 4 : 1; }
// A work list (a batch) fills the GPU with many jobs: four limbs per thread
// wherever the width allows, for throughput, except in jobs of at most one
// block of limbs (the low levels of the decimal tree, small batch items):
// one workgroup holds such a job either way, and one limb per thread keeps
// its serial part short.
__host__ __device__ static inline int crt_lpt_list(int bw, uint64_t nout) {
    return bw >= 32 && nout > CRT_LIMBS ? 4 : 1;
}

// The Garner constants of the three primes (host side, for launches made
// outside ff_crt.hip).
void ff_garner_get(FFGarner *out);

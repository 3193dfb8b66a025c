// CRT recombination and carry propagation on gfx950.
//
// The three residue arrays y_p[i] (exact convolution coefficient i mod p,
// natural order, as the inverse top pass leaves them) are combined by
// Garner's method into X_j < p1 p2 p3 < 2^93 (crt.c:180-206, reference),
// and the product is N = sum_j X_j 2^(j bw) (crt.c:208-296). The reference
// packs coefficients into limbs with one running carry per chunk; here every
// output limb has its own thread:
//   N = sum_i T_i 2^(64 i),  T_i = sum_j digit_i(X_j 2^(j bw)),
// the base-2^64 digit i of each shifted coefficient; at most five
// coefficients touch a limb, so T_i < 5 * 2^64. With u_i = lo(T_i) +
// hi(T_{i-1}) <= 2^64 + 3 the remaining carries are 0/1: limb i =
// u_i + c_i mod 2^64 with c_{i+1} = g_i | (p_i & c_i), g_i = u_i >= 2^64,
// p_i = (u_i == 2^64 - 1). Inside a wavefront the carries come from one
// 64-bit addition of the ballots (A = g, B = g | p: C = (A + B) ^ A ^ B),
// across the block's waves from their carry-outs, across blocks from each
// block's generate/propagate pair:
//   pack   Garner into LDS, T_i, u_i, in-block carries; limbs written for a
//          zero carry into the block; block (G, P) recorded;
//   scan   only if some block propagates (P = 1: all its 256 limbs are all
//          ones, which needs adversarial operands), the exact carry into
//          every block by a lookahead scan; otherwise carry_in(w) = G(w-1);
//   apply  adds the block's carry (ripple bounded by the block) and books
//          the top nonzero limb.
// This replaces the reference's 8 sequential chunks plus carry fix
// (crt.c:326-336) with a parallel carry-lookahead over all limbs.
#include <stdint.h>

#include "ff_crt_dev.h"
#include "ff_internal.h"

static FFGarner g_garner;
static int g_garner_ready = 0;

static uint32_t h_pow(uint32_t a, uint64_t e, uint32_t p) {
    uint64_t r = 1, x = a % p;
    while (e) {
        if (e & 1) r = r * x % p;
        x = x * x % p;
        e >>= 1;
    }
    return (uint32_t)r;
}

static void garner_init(void) {
    if (g_garner_ready) return;
    FFPrime P[3];
    ff_prime_consts(P);
    FFGarner &G = g_garner;
    G.p1 = P[0].p, G.p2 = P[1].p, G.p3 = P[2].p;
    G.pinv2 = P[1].pinv, G.pinv3 = P[2].pinv;
    const uint64_t R2 = (1ull << 32) % G.p2, R3 = (1ull << 32) % G.p3;
    G.i12 = (uint32_t)(h_pow(G.p1 % G.p2, G.p2 - 2, G.p2) * R2 % G.p2);
    G.i13 = (uint32_t)(h_pow(G.p1 % G.p3, G.p3 - 2, G.p3) * R3 % G.p3);
    G.i23 = (uint32_t)(h_pow(G.p2 % G.p3, G.p3 - 2, G.p3) * R3 % G.p3);
    G.i12p = G.i12 * G.pinv2;
    G.i13p = G.i13 * G.pinv3;
    G.i23p = G.i23 * G.pinv3;
    G.p12 = (uint64_t)G.p1 * G.p2;
    g_garner_ready = 1;
}


#define CRT_LIMBS 256  // limbs per block, one per thread
// Coefficients a block can touch: limbs [i0 - 1, i0 + 256) span 257 * 64 bits
// and a coefficient 93 more, so at most (257 * 64 + 93) / bw + 2 of them;
// 2070 at the narrowest width, 8 (mul_split.c:151-152).
#define CRT_MAXC 2112


// One block per CRT_LIMBS output limbs (block bx of job J).
__device__ __forceinline__ void crt_pack_body(const FFCrtJob &J, uint32_t bx, const FFGarner &G) {
    const uint64_t nblk = (J.nout + CRT_LIMBS - 1) / CRT_LIMBS;
    if (bx >= nblk) return;
    const int bw = J.bw;
    const uint64_t i0 = (uint64_t)bx * CRT_LIMBS;
    const uint64_t i1 = i0 + CRT_LIMBS < J.nout ? i0 + CRT_LIMBS : J.nout;
    // coefficients touching limbs [i0 - 1, i1): limb i0 - 1 gives the block's first hi
    const uint64_t lim = J.conv < J.L ? J.conv : J.L;
    const uint64_t jbase = first_coef(i0 ? i0 - 1 : 0, bw);
    uint64_t jend = (64 * i1 + bw - 1) / bw;
    if (jend > lim) jend = lim;
    __shared__ uint64_t xlo[CRT_MAXC];
    __shared__ uint32_t xhi[CRT_MAXC];
    __shared__ uint64_t hi_s[CRT_LIMBS];
    __shared__ uint64_t wave_c[CRT_LIMBS / 64];
    const uint32_t *__restrict__ y1 = J.y;
    const uint32_t *__restrict__ y2 = J.y + J.L;
    const uint32_t *__restrict__ y3 = J.y + 2 * (size_t)J.L;
    for (uint64_t j = jbase + threadIdx.x; j < jend; j += CRT_LIMBS) {
        const u128 X = ff_garner(y1[j], y2[j], y3[j], G);
        xlo[j - jbase] = (uint64_t)X;
        xhi[j - jbase] = (uint32_t)(X >> 64);
    }
    __syncthreads();
    const uint64_t i = i0 + threadIdx.x;
    const bool live = i < i1;
    u128 T = live ? limb_sum(i, bw, jbase, xlo, xhi, jbase, jend) : 0;
    hi_s[threadIdx.x] = (uint64_t)(T >> 64);
    uint64_t hprev = 0;
    if (threadIdx.x == 0 && i0 > 0) hprev = (uint64_t)(limb_sum(i0 - 1, bw, jbase, xlo, xhi, jbase, jend) >> 64);
    __syncthreads();
    if (threadIdx.x > 0) hprev = hi_s[threadIdx.x - 1];
    const u128 u = (u128)(uint64_t)T + hprev;
    const uint64_t ulo = (uint64_t)u;
    const bool g = live && (u >> 64) != 0;
    const bool pp = live && !g && ulo == ~0ull;
    // carries inside each wavefront: lane k's (generate, propagate) are bit k
    // of A = g and B = g | p, so A + B + c_in carries exactly where the limbs
    // do, and C = (A + B + c_in) ^ A ^ B has bit k = carry into lane k
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t A = __ballot(g), B = A | __ballot(pp);
    const uint64_t wg = (uint64_t)(((u128)A + B) >> 64);  // carry out of the wave, zero carry in
    const uint64_t wp = (~B) == 0;                          // every lane propagates (then wg = 0)
    if (lane == 0) wave_c[wave] = wg | (wp << 1);
    __syncthreads();
    uint64_t cin = 0;  // carry into this wave from the earlier waves of the block
    for (int w = 0; w < wave; w++) {
        const uint64_t e = wave_c[w];
        cin = (e & 1) | ((e >> 1) & cin);
    }
    const uint64_t C = (A + B + cin) ^ A ^ B;
    const uint64_t cbit = (C >> lane) & 1;
    if (live) J.out[i] = ulo + cbit;
    // the block's generate/propagate pair
    if (threadIdx.x == 0) {
        uint64_t Gb = 0, Pb = 1;
        for (int w = 0; w < CRT_LIMBS / 64; w++) {
            const uint64_t e = wave_c[w];
            Gb = (e & 1) | ((e >> 1) & Gb);
            Pb = Pb & (e >> 1);
        }
        J.bgp[bx] = (uint32_t)(Gb | (Pb << 1));
    }
}

__global__ __launch_bounds__(CRT_LIMBS) void k_crt_pack(FFCrtSet js, FFGarner G) {
    crt_pack_body(js.j[blockIdx.y], blockIdx.x, G);
}

__device__ __forceinline__ int crt_find(const uint32_t *start, int n, uint32_t b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(CRT_LIMBS) void k_crt_pack_list(const FFCrtJob *jobs, const uint32_t *start, int n,
                                                             FFGarner G) {
    const int k = crt_find(start, n, blockIdx.x);
    crt_pack_body(jobs[k], blockIdx.x - start[k], G);
}

// Exact carry into every block, only when some block propagates.
#define SCAN_THREADS 1024

__device__ __forceinline__ void crt_scan_body(const FFCrtJob &J) {
    const uint64_t nblk = (J.nout + CRT_LIMBS - 1) / CRT_LIMBS;
    int mine = 0;
    for (uint64_t b = threadIdx.x; b < nblk; b += SCAN_THREADS) mine |= (J.bgp[b] >> 1) & 1;
    const int any = __syncthreads_or(mine);
    if (threadIdx.x == 0) {
        *J.mode = (uint32_t)any;
        *J.lenkey = 0;  // k_crt_apply books the top limb by atomicMax
    }
    if (!any) return;  // carry into block w is G(w-1)
    const uint64_t per = (nblk + SCAN_THREADS - 1) / SCAN_THREADS;
    const uint64_t b0 = threadIdx.x * per, b1 = b0 + per < nblk ? b0 + per : nblk;
    uint32_t Gs = 0, Ps = 1;
    for (uint64_t b = b0; b < b1; b++) {
        const uint32_t e = J.bgp[b];
        Gs = (e & 1) | ((e >> 1) & Gs);
        Ps = Ps & (e >> 1);
    }
    __shared__ uint32_t sg[SCAN_THREADS], sp[SCAN_THREADS];
    sg[threadIdx.x] = Gs;
    sp[threadIdx.x] = Ps;
    __syncthreads();
    for (int off = 1; off < SCAN_THREADS; off <<= 1) {
        const uint32_t g2 = sg[threadIdx.x], p2 = sp[threadIdx.x];
        uint32_t g1 = 0, p1 = 1;
        if ((int)threadIdx.x >= off) g1 = sg[threadIdx.x - off], p1 = sp[threadIdx.x - off];
        __syncthreads();
        sg[threadIdx.x] = g2 | (p2 & g1);
        sp[threadIdx.x] = p2 & p1;
        __syncthreads();
    }
    uint32_t c = threadIdx.x ? sg[threadIdx.x - 1] : 0;
    for (uint64_t b = b0; b < b1; b++) {
        J.bcin[b] = (uint8_t)c;
        const uint32_t e = J.bgp[b];
        c = (e & 1) | ((e >> 1) & c);
    }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_crt_scan(FFCrtSet js) { crt_scan_body(js.j[blockIdx.x]); }

__global__ __launch_bounds__(SCAN_THREADS) void k_crt_scan_list(const FFCrtJob *jobs) {
    crt_scan_body(jobs[blockIdx.x]);
}

// Top nonzero limb of d[i0..i1) booked into the product's length key.
__device__ __forceinline__ void book_top(const FFCrtJob &J, uint64_t i0, uint64_t i1) {
    for (uint64_t i = i1; i-- > i0;) {
        const uint64_t v = J.out[i];
        if (v) {
            const unsigned long long key = ((unsigned long long)(i + 1) << 8) | (unsigned)(64 - __clzll((long long)v));
            atomicMax(J.lenkey, key);
            return;
        }
    }
}

// Thread b of a job: adds block b's carry-in (ripple bounded by the block)
// and, for blocks holding limbs from nout - 3 up (where the top nonzero limb
// of an na x nb product lies), books the top limb.
__device__ __forceinline__ void crt_apply_body(const FFCrtJob &J, uint64_t b) {
    const uint64_t nblk = (J.nout + CRT_LIMBS - 1) / CRT_LIMBS;
    if (b >= nblk) return;
    const uint32_t c = *J.mode ? J.bcin[b] : (b ? J.bgp[b - 1] & 1 : 0);
    const uint64_t i0 = b * CRT_LIMBS, i1 = i0 + CRT_LIMBS < J.nout ? i0 + CRT_LIMBS : J.nout;
    if (c) {
        for (uint64_t i = i0; i < i1; i++) {
            if (++J.out[i] != 0) break;
        }
    }
    const uint64_t lo_top = J.nout >= 3 ? J.nout - 3 : 0;
    if (i1 > lo_top) book_top(J, i0 > lo_top ? i0 : lo_top, i1);
}

__global__ __launch_bounds__(256) void k_crt_apply(FFCrtSet js) {
    crt_apply_body(js.j[blockIdx.y], (uint64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ __launch_bounds__(256) void k_crt_apply_list(const FFCrtJob *jobs, const uint32_t *astart, int n) {
    const int k = crt_find(astart, n, blockIdx.x);
    crt_apply_body(jobs[k], (uint64_t)(blockIdx.x - astart[k]) * blockDim.x + threadIdx.x);
}

// Algorithmic bytes: pack reads 3 residues per coefficient and writes each
// output limb once; scan and apply touch one word per block of 256 limbs.
extern "C" hipError_t ff_launch_crt(const FFCrtSet *js, int njobs, uint64_t max_nout, double conv_total,
                                    double out_limbs_total, hipStream_t st) {
    garner_init();
    const uint64_t max_blk = (max_nout + CRT_LIMBS - 1) / CRT_LIMBS;
    dim3 g1((unsigned)max_blk, njobs);
    ff_prof_begin(FK_CRT_PACK, st);
    hipLaunchKernelGGL(k_crt_pack, g1, dim3(CRT_LIMBS), 0, st, *js, g_garner);
    ff_prof_end(FK_CRT_PACK, 12.0 * conv_total + 8.0 * out_limbs_total, st);
    ff_prof_begin(FK_CRT_SCAN, st);
    hipLaunchKernelGGL(k_crt_scan, dim3(njobs), dim3(SCAN_THREADS), 0, st, *js);
    ff_prof_end(FK_CRT_SCAN, 0.0, st);
    dim3 g3((unsigned)((max_blk + 255) / 256), njobs);
    ff_prof_begin(FK_CRT_APPLY, st);
    hipLaunchKernelGGL(k_crt_apply, g3, dim3(256), 0, st, *js);
    ff_prof_end(FK_CRT_APPLY, 4.0 * out_limbs_total / CRT_LIMBS, st);
    return hipGetLastError();
}

extern "C" hipError_t ff_launch_crt_list(const FFCrtJob *jobs, const uint32_t *start, const uint32_t *astart, int n,
                                         uint32_t blocks, uint32_t ablocks, double conv_total,
                                         double out_limbs_total, hipStream_t st) {
    if (!n) return hipSuccess;
    garner_init();
    ff_prof_begin(FK_CRT_PACK, st);
    hipLaunchKernelGGL(k_crt_pack_list, dim3(blocks), dim3(CRT_LIMBS), 0, st, jobs, start, n, g_garner);
    ff_prof_end(FK_CRT_PACK, 12.0 * conv_total + 8.0 * out_limbs_total, st);
    ff_prof_begin(FK_CRT_SCAN, st);
    hipLaunchKernelGGL(k_crt_scan_list, dim3(n), dim3(SCAN_THREADS), 0, st, jobs);
    ff_prof_end(FK_CRT_SCAN, 0.0, st);
    ff_prof_begin(FK_CRT_APPLY, st);
    hipLaunchKernelGGL(k_crt_apply_list, dim3(ablocks), dim3(256), 0, st, jobs, astart, n);
    ff_prof_end(FK_CRT_APPLY, 4.0 * out_limbs_total / CRT_LIMBS, st);
    return hipGetLastError();
}

// ------------------------------------------------------------------ fused small step
//
// A doubling step whose two squarings are one-block transforms (L <= 4096)
// is latency-bound: its CRT and linear pass would be six launches of a few
// microseconds each. Here one workgroup does all of it from the squares'
// residues: the CRT of F(k)^2 and F(k-1)^2 into registers (thread t owns
// limbs t, t + SS_T, t + 2 SS_T), the linear step by workgroup-wide carry
// lookahead, writing K (A or B) and C and booking their top limbs.
template <int SS_T>
__global__ __launch_bounds__(SS_T) void k_small_step(FFSmallStep a, FFGarner G) {
    __shared__ SSShared<SS_T> S;
    const int t = threadIdx.x;
    if (t == 0) {
        atomicExch(a.keyK, 0ull);
        atomicExch(a.keyC, 0ull);
        __threadfence();
    }
    uint64_t s1[SS_CH], s2[SS_CH], K[SS_CH], C[SS_CH];
    SSRun R{0, 0};
    ss_crt(S, R, a.y1, a.L1, a.bw1, a.conv1, a.len, G, s1);
    ss_crt(S, R, a.y2, a.L2, a.bw2, a.conv2, a.len, G, s2);
    ss_linear(S, R, s1, s2, a.cb0, a.len, K, C, a.bit);
    for (int c = 0; c < SS_CH; c++) {
        const uint64_t i = (uint64_t)c * SS_T + t;
        if (i >= a.len) break;
        a.K[i] = K[c];
        a.C[i] = C[c];
        ss_book(a.keyK, i, K[c], a.len);
        ss_book(a.keyC, i, C[c], a.len);
    }
}

FFGarner ff_garner_consts(void) {
    garner_init();
    return g_garner;
}

extern "C" hipError_t ff_launch_small_step(const FFSmallStep *a, hipStream_t st) {
    garner_init();
    ff_prof_begin(FK_STEP, st);
    // a quarter of the threads while every chunk fits 256 limbs (the early,
    // smallest steps), where barriers over 16 wavefronts would dominate
    if (a->len <= 3 * 256) hipLaunchKernelGGL(k_small_step<256>, dim3(1), dim3(256), 0, st, *a, g_garner);
    else hipLaunchKernelGGL(k_small_step<1024>, dim3(1), dim3(1024), 0, st, *a, g_garner);
    ff_prof_end(FK_STEP, 12.0 * (a->conv1 + a->conv2) + 16.0 * a->len, st);
    return hipGetLastError();
}

// CRT recombination and carry propagation on gfx950.
//
// The three residue arrays y_p[i] (exact convolution coefficient i mod p,
// natural order, as the inverse top pass leaves them) are combined by
// Garner's method into X_j < p1 p2 p3 < 2^93 (crt.c:33-64, reference),
// and the product is N = sum_j X_j 2^(j bw) (crt.c:66-154). The reference
// packs coefficients into limbs with one running carry per chunk; here every
// output limb has its own thread:
//   N = sum_i T_i 2^(64 i),  T_i = sum_j digit_i(X_j 2^(j bw)),
// the base-2^64 digit i of each shifted coefficient; at most five
// coefficients touch a limb, so T_i < 5 * 2^64. With u_i = lo(T_i) +
// hi(T_{i-1}) <= 2^64 + 3 the remaining carries are 0/1: limb i =
// u_i + c_i mod 2^64 with c_{i+1} = g_i | (p_i & c_i), g_i = u_i >= 2^64,
// p_i = (u_i == 2^64 - 1). One launch resolves them: inside a wavefront
// from one 64-bit addition of the ballots, across the workgroup's waves
// from their summaries, across workgroups by decoupled look-back
// (ff_lookback.h), and the job's last workgroup books the top limb. This
// replaces the reference's 8 sequential chunks plus carry fix
// (crt.c:184-194).
//
// The same machinery fuses a whole doubling step after its two squarings
// (k_step_fused): the CRT of both squares and the linear combination of
// fib.c:36-62 in one launch.
#include <stdint.h>

#include "ff_crt_body.h"
#include "ff_crt_dev.h"
#include "ff_internal.h"
#include "ff_lookback.h"
#include "ff_runtime.h"

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

void ff_garner_get(FFGarner *out) {
    garner_init();
    *out = g_garner;
}


template <int LPT>
__global__ __launch_bounds__(CRT_LIMBS) void k_crt(FFCrtSet js, FFGarner G, uint32_t *fault, int resident) {
    __shared__ CrtShared sh;
    crt_body<LPT>(js.j[blockIdx.y], blockIdx.x, G, fault, sh, resident != 0);
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

// Work-list form: job k owns blocks [start[k], start[k+1]) of the flat
// grid; a job's limbs per thread follow from its width.
__global__ __launch_bounds__(CRT_LIMBS) void k_crt_list(const FFCrtJob *jobs, const uint32_t *start, int n,
                                                        FFGarner G, uint32_t *fault) {
    __shared__ CrtShared sh;
    const int k = crt_find(start, n, blockIdx.x);
    if (crt_lpt_list(jobs[k].bw, jobs[k].nout) == 4) crt_body<4>(jobs[k], blockIdx.x - start[k], G, fault, sh, false);
    else crt_body<1>(jobs[k], blockIdx.x - start[k], G, fault, sh, false);
}

uint64_t ff_crt_blocks(uint64_t nout, int bw) {
    const uint64_t per = (uint64_t)CRT_LIMBS * crt_lpt_list(bw, nout);
    return (nout + per - 1) / per;
}

// Algorithmic bytes: 3 residues read per coefficient, each output limb
// written once.
extern "C" hipError_t ff_launch_crt(const FFCrtSet *js, int njobs, uint64_t max_nout, double conv_total,
                                    double out_limbs_total, hipStream_t st) {
    garner_init();
    // one launch: four limbs per thread when the jobs together pass one
    // wave of workgroups and every width allows it
    uint64_t total = 0;
    int wide = 1;
    for (int k = 0; k < njobs; k++) total += js->j[k].nout, wide &= js->j[k].bw >= 32;
    const int lpt = wide && total > (uint64_t)g_crt_wave_limbs ? 4 : 1;
    const FFCrtSet &set = *js;
    const uint64_t max_blk = (max_nout + (uint64_t)CRT_LIMBS * lpt - 1) / ((uint64_t)CRT_LIMBS * lpt);
    ff_prof_begin(FK_CRT_PACK, st);
    if (lpt == 4) {
        const int res = max_blk * njobs <= ff_resident_blocks((const void *)k_crt<4>, CRT_LIMBS);
        hipLaunchKernelGGL(k_crt<4>, dim3((unsigned)max_blk, njobs), dim3(CRT_LIMBS), 0, st, set, g_garner,
                           ff_lb_fault(), res);
    } else {
        const int res = max_blk * njobs <= ff_resident_blocks((const void *)k_crt<1>, CRT_LIMBS);
        hipLaunchKernelGGL(k_crt<1>, dim3((unsigned)max_blk, njobs), dim3(CRT_LIMBS), 0, st, set, g_garner,
                           ff_lb_fault(), res);
    }
    ff_prof_end(FK_CRT_PACK, 12.0 * conv_total + 8.0 * out_limbs_total, st);
    return hipGetLastError();
}

extern "C" hipError_t ff_launch_crt_list(const FFCrtJob *jobs, const uint32_t *start, int n, uint32_t blocks,
                                         double conv_total, double out_limbs_total, hipStream_t st) {
    if (!n) return hipSuccess;
    garner_init();
    ff_prof_begin(FK_CRT_PACK, st);
    hipLaunchKernelGGL(k_crt_list, dim3(blocks), dim3(CRT_LIMBS), 0, st, jobs, start, n, g_garner, ff_lb_fault());
    ff_prof_end(FK_CRT_PACK, 12.0 * conv_total + 8.0 * out_limbs_total, st);
    return hipGetLastError();
}

// ------------------------------------------------------------------ fused large step
//
// A doubling step from the residues of s1 = F(k)^2 and s2 = F(k-1)^2 in one
// launch (fib.c:36-62): per workgroup of 256 or 1024 limbs (one or four per
// thread), the CRT of both squares (their Garner values formed in one loop),
// then A = s1 + s2 = F(2k-1), E = 4 s1 + cb0, H = s1 + 2 s2,
// B = E - s2 = F(2k+1) and C = E - H = F(2k), every one a binary carry chain
// resolved by look-back in three rounds: {s1, s2}, {A, E, H}, {B, C}. The
// shifted limb i of 4 s1 (2 s2) takes the top two bits (top bit) of limb
// i - 1; for the workgroup's first limb they come from its own sums (limb
// i0 - 1 is summed too), unless the carry into that limb could reach them,
// in which case from the word the predecessor publishes with its last
// limbs' top bits. cb0 = -2 adds 2^(64 len) - 2 (the carry out of the top
// is dropped). The squares may instead be given as limbs (split products),
// skipping the first round.
template <int LPT>
__global__ __launch_bounds__(CRT_LIMBS) void k_step_fused(FFStepJob J, FFGarner G, uint32_t *fault, int resident) {
    constexpr uint64_t PER = (uint64_t)CRT_LIMBS * LPT;
    constexpr int MAXC = CRT_MAXC_W32(LPT);
    const uint64_t nblk = (J.len + PER - 1) / PER;
    if (blockIdx.x >= nblk) return;
    __shared__ uint64_t xlo[2 * MAXC];
    __shared__ uint32_t xhi[2 * MAXC];
    __shared__ uint64_t hx[2 * CRT_LIMBS + 8];
    __shared__ LbShared S;
    const int t = threadIdx.x;
    const uint32_t b = lb_ticket(S, J.lb, resident != 0, blockIdx.x);
    uint32_t *st = J.lb + 2;                // three rounds of nblk words
    uint32_t *lastw = J.lb + 2 + 3 * nblk;  // per workgroup: 8 | top bit of its last s2 limb << 2 | top two of s1's
    const uint64_t i0 = (uint64_t)b * PER, ib = i0 + (uint64_t)t * LPT;
    bool live[LPT];
#pragma unroll
    for (int q = 0; q < LPT; q++) live[q] = ib + q < J.len;
    bool g[3], p[3];
    uint32_t c;
    uint64_t s1[LPT], s2[LPT], prev1, prev2;  // top two bits of s1, top bit of s2, at limb ib - 1
    if (J.s1l) {
#pragma unroll
        for (int q = 0; q < LPT; q++) {
            s1[q] = live[q] && ib + q < J.n1 ? J.s1l[ib + q] : 0;
            s2[q] = live[q] && ib + q < J.n2 ? J.s2l[ib + q] : 0;
        }
        prev1 = ib >= 1 && ib - 1 < J.n1 ? J.s1l[ib - 1] >> 62 : 0;
        prev2 = ib >= 1 && ib - 1 < J.n2 ? J.s2l[ib - 1] >> 63 : 0;
    } else {
        u128 u[2][LPT], up[2] = {0, 0};
        const CrtSrc src[2] = {{J.y1, J.L1, J.bw1, J.conv1}, {J.y2, J.L2, J.bw2, J.conv2}};
        crt_sums<LPT, 2, MAXC>(src, J.len, i0, G, xlo, xhi, hx, true, u, up);
        lb_group_gp<LPT>(u[0], &g[0], &p[0]);
        lb_group_gp<LPT>(u[1], &g[1], &p[1]);
        c = lb_round<2>(S, st, b, g, p, 0, fault);
        lb_group_apply<LPT>(u[0], c & 1, s1);
        lb_group_apply<LPT>(u[1], (c >> 1) & 1, s2);
        if (t == CRT_LIMBS - 1)
            lb_store(lastw + b, 8u | (uint32_t)(s1[LPT - 1] >> 62) | ((uint32_t)(s2[LPT - 1] >> 63) << 2));
        hx[t] = s1[LPT - 1];
        hx[CRT_LIMBS + t] = s2[LPT - 1];
        uint64_t p1 = 0, p2 = 0;
        if (t == 0 && b > 0) {
            // a carry into limb i0 - 1 changes its top bits only through all-ones low bits
            const uint64_t lo1 = (uint64_t)up[0], lo2 = (uint64_t)up[1];
            const uint64_t low62 = (1ull << 62) - 1, low63 = (1ull << 63) - 1;
            if ((lo1 & low62) != low62 && (lo2 & low63) != low63) {
                p1 = lo1 >> 62, p2 = lo2 >> 63;
            } else {
                const uint32_t w = lb_wait(lastw + b - 1, fault);
                p1 = w & 3, p2 = (w >> 2) & 1;
            }
        }
        __syncthreads();
        prev1 = t ? hx[t - 1] >> 62 : p1;
        prev2 = t ? hx[CRT_LIMBS + t - 1] >> 63 : p2;
    }
    u128 ua[LPT], ue[LPT], uh[LPT];
#pragma unroll
    for (int q = 0; q < LPT; q++) {
        const uint64_t i = ib + q;
        const uint64_t d = (s1[q] << 2) | (q ? s1[q - 1] >> 62 : prev1);
        const uint64_t d2 = (s2[q] << 1) | (q ? s2[q - 1] >> 63 : prev2);
        const uint64_t kv = J.cb0 >= 0 ? (i == 0 ? (uint64_t)J.cb0 : 0) : (i == 0 ? ~(uint64_t)(-J.cb0 - 1) : ~0ull);
        ua[q] = live[q] ? (u128)s1[q] + s2[q] : (u128)~0ull;
        ue[q] = live[q] ? (u128)d + kv : (u128)~0ull;
        uh[q] = live[q] ? (u128)s1[q] + d2 : (u128)~0ull;
    }
    lb_group_gp<LPT>(ua, &g[0], &p[0]);
    lb_group_gp<LPT>(ue, &g[1], &p[1]);
    lb_group_gp<LPT>(uh, &g[2], &p[2]);
    c = lb_round<3>(S, st + nblk, b, g, p, 0, fault);
    uint64_t A[LPT], E[LPT], H[LPT], B[LPT], C[LPT];
    lb_group_apply<LPT>(ua, c & 1, A);
    lb_group_apply<LPT>(ue, (c >> 1) & 1, E);
    lb_group_apply<LPT>(uh, (c >> 2) & 1, H);
    u128 ub[LPT], uc[LPT];  // B = E + ~s2 + 1, C = E + ~H + 1
#pragma unroll
    for (int q = 0; q < LPT; q++) {
        ub[q] = live[q] ? (u128)E[q] + ~s2[q] : (u128)~0ull;
        uc[q] = live[q] ? (u128)E[q] + ~H[q] : (u128)~0ull;
    }
    lb_group_gp<LPT>(ub, &g[0], &p[0]);
    lb_group_gp<LPT>(uc, &g[1], &p[1]);
    c = lb_round<2>(S, st + 2 * nblk, b, g, p, 3, fault);
    lb_group_apply<LPT>(ub, c & 1, B);
    lb_group_apply<LPT>(uc, (c >> 1) & 1, C);
    uint64_t out[2 * LPT];
#pragma unroll
    for (int q = 0; q < LPT; q++) {
        out[q] = live[q] ? (J.bit ? B[q] : A[q]) : 0;
        out[LPT + q] = live[q] ? C[q] : 0;
        if (live[q]) {
            J.K[ib + q] = out[q];
            J.C[ib + q] = out[LPT + q];
        }
    }
    unsigned long long *const keys[2] = {J.keyK, J.keyC};
    lb_keys<2, LPT>(J.lb + 2 + 4 * nblk, b, (uint32_t)nblk, i0, J.len, out, keys, fault);
    if (lb_last(S, J.lb, J.lb + 3 + 4 * nblk, b, (uint32_t)nblk)) lb_clear(J.lb, 3 + 4 * nblk + lb_groups(nblk));
}

// Limbs per thread of a step: as for the CRT (one while a workgroup per
// 256 limbs fits one wave of residency).
static inline int step_lpt(uint64_t len) { return len > (uint64_t)g_crt_wave_limbs ? 4 : 1; }

size_t ff_step_lb_words(uint64_t len) {
    const uint64_t per = (uint64_t)CRT_LIMBS * step_lpt(len);
    const uint64_t nblk = (len + per - 1) / per;
    return 3 + 4 * nblk + lb_groups(nblk);
}

extern "C" hipError_t ff_launch_step_fused(const FFStepJob *J, double conv_total, hipStream_t st) {
    garner_init();
    if (!J->s1l && (J->bw1 < 32 || J->bw2 < 32)) return hipErrorInvalidValue;  // the LDS holds widths >= 32
    const int lpt = step_lpt(J->len);
    const uint64_t nblk = (J->len + (uint64_t)CRT_LIMBS * lpt - 1) / ((uint64_t)CRT_LIMBS * lpt);
    ff_prof_begin(FK_STEP_CRT, st);
    if (lpt == 4) {
        const int res = nblk <= ff_resident_blocks((const void *)k_step_fused<4>, CRT_LIMBS);
        hipLaunchKernelGGL(k_step_fused<4>, dim3((unsigned)nblk), dim3(CRT_LIMBS), 0, st, *J, g_garner, ff_lb_fault(),
                           res);
    } else {
        const int res = nblk <= ff_resident_blocks((const void *)k_step_fused<1>, CRT_LIMBS);
        hipLaunchKernelGGL(k_step_fused<1>, dim3((unsigned)nblk), dim3(CRT_LIMBS), 0, st, *J, g_garner, ff_lb_fault(),
                           res);
    }
    ff_prof_end(FK_STEP_CRT, 12.0 * conv_total + 16.0 * J->len, st);
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

// The first doubling steps of F(n), chained in one workgroup.
//
// While F(k) is small (its square a transform of at most 2^10 words), a
// doubling step is a few microseconds of work spread over dependent
// launches whose latency dominates. Here one workgroup of 1024 threads runs
// the whole chain of such steps with every value in LDS. Per step:
//   - the three primes' transforms of F(k) and F(k-1), stacked as the two
//     sub-transforms of one tile per prime (threads [320 pi, 320 pi + 320)
//     own prime pi, the last prime 384 threads: whole wavefronts), with the
//     pointwise squares between forward and inverse (fused rounds);
//   - the CRT of both squares and the linear step of the fused small step
//     (fib.c:36-62), every limb in a thread;
//   - the new pair (F(k'-1), F(k')) replaces the old one in LDS.
// The host starts the chain at F(k0 - 1), F(k0) for the top ten bits of n
// (k0 < 1024, at most 12 limbs) and passes each step's sizes, predicted
// from the closed form; every step books its new pair's length keys, which
// the host checks against the predictions at the end of F(n) like those of
// the other steps. The root tables' first 2^10 entries are copied into LDS
// once. The transforms are power-of-2 lengths only: a step
// never needs the 3 * 2^k lengths to fit in LDS, and the product does not
// depend on the length.
#include "ff_crt_dev.h"
#include "ff_internal.h"
#include "ff_ntt_dev.h"

#define EARLY_T 1024
#define EARLY_L (1 << FF_EARLY_LG)
#define EARLY_TILE (2 * EARLY_L + (2 * EARLY_L >> 4) + 16)

// The rounds read the root tables (their first 2^10 entries) from LDS: one
// workgroup cannot hide the latency of table reads from HBM.
struct FFLdsTables : FFNoTw {
    static constexpr bool lds_tables = true;
};

struct EarlyShared {
    uint32_t tw[3][4][EARLY_L];       // per prime: rt, rtp, irt, irtp
    uint64_t num[2][FF_EARLY_LIMBS];  // F(k - 1), F(k)
    uint32_t tile[3][EARLY_TILE];     // per prime: F(k), F(k - 1) stacked
    uint32_t y[2][3 * EARLY_L];       // residues of F(k)^2, F(k-1)^2: [prime][coefficient]
    SSShared<EARLY_T> ss;
};

__global__ __launch_bounds__(EARLY_T) void k_fib_early(FFEarlyArgs a, const FFTables *T, FFPrimes PR, FFGarner G) {
    __shared__ EarlyShared E;
    const int t = threadIdx.x;
    for (int i = t; i < FF_EARLY_LIMBS; i += EARLY_T) {
        E.num[0][i] = i < FF_EARLY_F0 ? a.f[0][i] : 0;
        E.num[1][i] = i < FF_EARLY_F0 ? a.f[1][i] : 0;
    }
    for (int i = t; i < 3 * EARLY_L; i += EARLY_T) {
        const int q = i / EARLY_L, j = i - q * EARLY_L;
        E.tw[q][0][j] = T->rt[q][j];
        E.tw[q][1][j] = T->rtp[q][j];
        E.tw[q][2][j] = T->irt[q][j];
        E.tw[q][3][j] = T->irtp[q][j];
    }
    if (t < 2 * a.nsteps) atomicExch(a.keys + t, 0ull);
    __threadfence();
    __syncthreads();
    const int pi = t < 640 ? t / 320 : 2;
    const int tid0 = 320 * pi, nthr = pi < 2 ? 320 : 384;
    const FFPrime P = PR.p[pi];
    const uint32_t p = P.p, pinv = P.pinv;
    SSRun R{0, 0};
    for (int s = 0; s < a.nsteps; s++) {
        const FFEarlyStep st = a.st[s];
        const int lg = st.lg;
        const int L = 1 << lg;
        const FFTile tl{E.tile[pi], 1, 1, 0, lg, 2, 1, tid0, nthr};
        // sub-transform 0: F(k), 1: F(k - 1), coefficients of bw bits
        const uint32_t *w1 = (const uint32_t *)E.num[1], *w2 = (const uint32_t *)E.num[0];
        for (int i = t - tid0; i < 2 * L; i += nthr) {
            const int pos = i & (L - 1);
            tl.s[ff_pos_addr(tl, i, 0)] =
                (i >> lg) ? ff_load_coef(w2, st.nw2, pos, st.bw, P) : ff_load_coef(w1, st.nw1, pos, st.bw, P);
        }
        __syncthreads();
        auto ld = [&](int pos, int c) -> uint32_t { return tl.s[ff_pos_addr(tl, pos, c)]; };
        auto sto = [&](int pos, int c, uint32_t v) { tl.s[ff_pos_addr(tl, pos, c)] = v; };
        ff_conv_io(tl, E.tw[pi][0], E.tw[pi][1], E.tw[pi][2], E.tw[pi][3], nullptr, p, pinv, false, false, ld, sto,
                   FFLdsTables{}, 2);
        __syncthreads();
        const uint32_t sc = a.sc[lg][pi];
        for (int i = t - tid0; i < 2 * L; i += nthr)
            E.y[i >> lg][pi * L + (i & (L - 1))] = ff_mont(tl.s[ff_pos_addr(tl, i, 0)], sc, p, pinv);
        __syncthreads();
        uint64_t s1[SS_CH], s2[SS_CH], K[SS_CH], C[SS_CH];
        ss_crt(E.ss, R, E.y[0], L, st.bw, st.conv1, st.len, G, s1);
        ss_crt(E.ss, R, E.y[1], L, st.bw, st.conv2, st.len, G, s2);
        ss_linear(E.ss, R, s1, s2, st.cb0, st.len, K, C, st.bit);
        // bit 0: (F(2k-1), F(2k)) = (A, C); bit 1: (F(2k), F(2k+1)) = (C, B)
        const uint64_t lo = st.bit ? C[0] : K[0], hi = st.bit ? K[0] : C[0];
        if (t < (int)st.len) {
            E.num[0][t] = lo;
            E.num[1][t] = hi;
            ss_book(a.keys + 2 * s, t, lo, st.len);
            ss_book(a.keys + 2 * s + 1, t, hi, st.len);
        }
        __syncthreads();
    }
    for (int i = t; i < (int)a.nlo; i += EARLY_T) a.lo[i] = E.num[0][i];
    for (int i = t; i < (int)a.nhi; i += EARLY_T) a.hi[i] = E.num[1][i];
}

hipError_t ff_launch_fib_early(const FFEarlyArgs *a, hipStream_t st) {
    if (a->nsteps < 1 || a->nsteps > FF_EARLY_MAXS || a->nlo > FF_EARLY_LIMBS || a->nhi > FF_EARLY_LIMBS)
        return hipErrorInvalidValue;
    for (int s = 0; s < a->nsteps; s++) {
        const FFEarlyStep &e = a->st[s];
        // every size within the LDS arrays and the one-chunk CRT
        if (e.lg < 4 || e.lg > FF_EARLY_LG || e.bw < 32 || e.bw > 35 || e.len > FF_EARLY_LIMBS ||
            e.conv1 > (1u << e.lg) || e.conv2 > e.conv1 || e.nw1 > 2 * FF_EARLY_LIMBS || e.nw2 > e.nw1)
            return hipErrorInvalidValue;
    }
    ff_prof_begin(FK_EARLY, st);
    hipLaunchKernelGGL(k_fib_early, dim3(1), dim3(EARLY_T), 0, st, *a, ff_tables_dev(), ff_primes(),
                       ff_garner_consts());
    ff_prof_end(FK_EARLY, 16.0 * a->st[a->nsteps - 1].len, st);
    return hipGetLastError();
}

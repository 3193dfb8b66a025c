// The first doubling steps of F(n), chained in one workgroup.
//
// While F(k) is small (its square a transform of at most 2^10 words), a
// doubling step is a few microseconds of work spread over dependent
// launches whose latency dominates. Here one workgroup of 1024 threads runs
// the whole chain of such steps with every value in LDS. Per step:
//   - the three primes' transforms of F(k) and F(k-1), stacked as the two
//     sub-transforms of one tile per prime (threads [320 pi, 320 pi + 320)
//     own prime pi, the last prime 384 threads: whole wavefronts), with the
//     new pair formed pointwise between forward and inverse (fused rounds):
//     F(2k - 1) = F(k)^2 + F(k-1)^2, F(2k) = F(k) (F(k) + 2 F(k-1)) and
//     F(2k + 1) = F(2k) + F(2k - 1), the identities behind the reference's
//     doubling step (fib.c:36-62) without its (-1)^k term;
//   - the CRT of both new numbers, every limb in a thread;
//   - the new pair (F(k'-1), F(k')) replaces the old one in LDS.
// The host starts the chain at F(k0 - 1), F(k0) for the top bits of n
// (k0 < 2^13, at most 89 limbs) and passes each step's sizes, predicted
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

// The middle round over both stacked sub-transforms at once: a = F(k) (sub
// 0) and b = F(k - 1) (sub 1) at one frequency in one thread. The pointwise
// step forms x = a (a + 2b) = F(2k) and y = a^2 + b^2 = F(2k - 1) and keeps
// (y, x), or (x, x + y) = (F(2k), F(2k + 1)) for bit 1: the new pair comes
// straight out of the inverse transforms, with no linear step after them.
// Coefficients of x + y are below 5 N 2^(2 bw) <= 5 * 2^9 * 2^70 < 2^82 for
// N <= 2^9 coefficients of F(k), well within p1 p2 p3 > 2^92.
template <int NL>
__device__ __forceinline__ void early_pair_mid(const FFTile &t, const uint32_t *const *tw, uint32_t p, uint32_t pinv,
                                               int bit) {
    constexpr int G = 1 << NL;
    const int ngroups = 1 << (t.rho - NL), off = 1 << t.rho;
    for (int g = ff_tile_tid(t); g < ngroups; g += ff_tile_nthr(t)) {
        uint32_t va[G], vb[G];
#pragma unroll
        for (int j = 0; j < G; j++) {
            va[j] = t.s[ff_pos_addr(t, (g << NL) | j, 0)];
            vb[j] = t.s[ff_pos_addr(t, off + ((g << NL) | j), 0)];
        }
        ff_dif_group<NL, true, true>(va, 0, 0, tw[0], tw[1], p, pinv);
        ff_dif_group<NL, true, true>(vb, 0, 0, tw[0], tw[1], p, pinv);
#pragma unroll
        for (int j = 0; j < G; j++) {
            const uint32_t a = va[j], b = vb[j];
            const uint32_t x = ff_mont(a, ff_add(a, ff_add(b, b, p), p), p, pinv);
            const uint32_t y = ff_add(ff_mont(a, a, p, pinv), ff_mont(b, b, p, pinv), p);
            va[j] = bit ? x : y;
            vb[j] = bit ? ff_add(x, y, p) : x;
        }
        ff_dit_group<NL, true, true>(va, 0, 0, tw[2], tw[3], p, pinv);
        ff_dit_group<NL, true, true>(vb, 0, 0, tw[2], tw[3], p, pinv);
#pragma unroll
        for (int j = 0; j < G; j++) {
            t.s[ff_pos_addr(t, (g << NL) | j, 0)] = va[j];
            t.s[ff_pos_addr(t, off + ((g << NL) | j), 0)] = vb[j];
        }
    }
}

// Forward transforms of both sub-transforms in radix-4 rounds, the pair
// middle round, inverse rounds; the tables (rt, rtp, irt, irtp) in LDS. The
// first round takes its values from ld(pos, c) (the coefficients), the last
// gives them to st(pos, c, v) (the scaled residues); t.rho >= 4, so those
// are a forward and an inverse round apart from the middle one.
template <class Ld, class St>
__device__ void early_pair_conv(const FFTile &t, const uint32_t *const *tw, uint32_t p, uint32_t pinv, int bit, Ld ld,
                                St st) {
    constexpr int rl = 2;
    const int R = (t.rho + rl - 1) / rl;
    int b = t.rho;
    for (int r = 0; r < R - 1; r++) {
        b -= rl;
        if (r) __syncthreads();
        ff_round_any<FF_R_DIF>(rl, r == 0, false, t, b, tw[0], tw[1], nullptr, nullptr, nullptr, p, pinv, ld, st,
                               FFLdsTables{});
    }
    const int nl0 = t.rho % rl == 0 ? rl : t.rho % rl;
    if (R > 1) __syncthreads();
    if (nl0 == 2) early_pair_mid<2>(t, tw, p, pinv, bit);
    else early_pair_mid<1>(t, tw, p, pinv, bit);
    b = nl0;
    for (int r = 1; r < R; r++) {
        __syncthreads();
        ff_round_any<FF_R_DIT>(rl, false, r == R - 1, t, b, nullptr, nullptr, tw[2], tw[3], nullptr, p, pinv, ld, st,
                               FFLdsTables{});
        b += rl;
    }
}

// early_pair_conv for 2^LG-point sub-transforms with every round's row
// bits a literal (as the tile's geometry): the LDS offsets of each round are
// compile-time constants.
template <int LG, int R, class Ld, class St>
__device__ __forceinline__ void early_fwd_fixed(const FFTile &t, const uint32_t *const *tw, uint32_t p, uint32_t pinv,
                                                Ld ld, St st) {
    constexpr int NR = (LG + 1) / 2;
    if constexpr (R < NR - 1) {
        constexpr int B = LG - 2 * (R + 1);
        if constexpr (R > 0) __syncthreads();
        ff_round_io<2, FF_R_DIF, R == 0, false>(t, B, tw[0], tw[1], nullptr, nullptr, nullptr, p, pinv, ld, st,
                                                FFLdsTables{});
        early_fwd_fixed<LG, R + 1>(t, tw, p, pinv, ld, st);
    }
}

template <int LG, int B, class Ld, class St>
__device__ __forceinline__ void early_inv_fixed(const FFTile &t, const uint32_t *const *tw, uint32_t p, uint32_t pinv,
                                                Ld ld, St st) {
    if constexpr (B < LG) {
        __syncthreads();
        ff_round_io<2, FF_R_DIT, false, B + 2 >= LG>(t, B, nullptr, nullptr, tw[2], tw[3], nullptr, p, pinv, ld, st,
                                                     FFLdsTables{});
        early_inv_fixed<LG, B + 2>(t, tw, p, pinv, ld, st);
    }
}

template <int LG, class Ld, class St>
__device__ __forceinline__ void early_pair_conv_fixed(const FFTile &t0, const uint32_t *const *tw, uint32_t p,
                                                      uint32_t pinv, int bit, Ld ld, St st) {
    static_assert(LG >= 4, "a forward and an inverse round apart from the middle one");
    const FFTile t{t0.s, 1, 1, 0, LG, 2, 1, t0.tid0, t0.nthr};
    constexpr int NL0 = LG % 2 ? 1 : 2;
    early_fwd_fixed<LG, 0>(t, tw, p, pinv, ld, st);
    __syncthreads();
    early_pair_mid<NL0>(t, tw, p, pinv, bit);
    early_inv_fixed<LG, NL0>(t, tw, p, pinv, ld, st);
}

// Coefficients of a product of len <= FF_EARLY_LIMBS limbs at bw >= 32, as
// SSShared::MAXC for that many limbs.
#define EARLY_MAXC (((FF_EARLY_LIMBS + 1) * 64 + 93) / 32 + 64)

// The second CRT chain's arrays (ss_crt_pair) share the tiles' memory: the
// tiles are dead from the last inverse round (which writes y) to the next
// step's first forward round.
struct EarlyCrt2 {
    uint64_t xlo[EARLY_MAXC];
    uint32_t xhi[EARLY_MAXC];
    uint64_t ex[EARLY_T];
};

struct EarlyShared {
    uint32_t tw[3][4][EARLY_L];       // per prime: rt, rtp, irt, irtp
    uint64_t num[2][FF_EARLY_LIMBS];  // F(k - 1), F(k)
    union {
        uint32_t tile[3][EARLY_TILE];  // per prime: F(k), F(k - 1) stacked
        EarlyCrt2 c2;
    };
    uint32_t y[2][3 * EARLY_L];       // residues of the new pair: [prime][coefficient]
    SSShared<EARLY_T> ss;
};
static_assert(sizeof(EarlyCrt2) <= sizeof(uint32_t) * 3 * EARLY_TILE, "second chain fits the tiles");

__global__ __launch_bounds__(EARLY_T) void k_fib_early(FFEarlyArgs a, const FFTables *T, FFPrimes PR, FFGarner G) {
    __shared__ EarlyShared E;
    static_assert(EARLY_T == EARLY_L, "thread t copies entry t of each table");
    const int t = threadIdx.x;
    const int pi = t < 640 ? t / 320 : 2;
    const int tid0 = 320 * pi, nthr = pi < 2 ? 320 : 384;
    // every load of the set-up (the twelve table entries of this thread,
    // the prime, the first step) issued before the first LDS store: the
    // tables are read through a generic pointer, which the compiler cannot
    // tell apart from LDS, so stores between them would order the loads
    const FFPrime P = PR.p[pi];
    const uint32_t p = P.p, pinv = P.pinv;
    FFEarlyStep st_next = a.st[0];
    uint32_t tv[3][4];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        tv[q][0] = T->rt[q][t];
        tv[q][1] = T->rtp[q][t];
        tv[q][2] = T->irt[q][t];
        tv[q][3] = T->irtp[q][t];
    }
    if (t < 2 * a.nsteps) atomicExch(a.keys + t, 0ull);
    for (int i = t; i < FF_EARLY_LIMBS; i += EARLY_T) {
        E.num[0][i] = i < FF_EARLY_F0 ? a.f[0][i] : 0;
        E.num[1][i] = i < FF_EARLY_F0 ? a.f[1][i] : 0;
    }
#pragma unroll
    for (int q = 0; q < 3; q++)
#pragma unroll
        for (int k = 0; k < 4; k++) E.tw[q][k][t] = tv[q][k];
    __threadfence();
    __syncthreads();
    SSRun R{0, 0};
    for (int s = 0; s < a.nsteps; s++) {
        // this step's sizes; the next step's load overlaps this step
        const FFEarlyStep st = st_next;
        if (s + 1 < a.nsteps) st_next = a.st[s + 1];
        const int lg = st.lg;
        const int L = 1 << lg;
        const FFTile tl{E.tile[pi], 1, 1, 0, lg, 2, 1, tid0, nthr};
        // sub-transform 0: F(k), 1: F(k - 1), coefficients of bw bits, read
        // from the pair by the first forward round; the last inverse round
        // writes the residues, scaled, to y
        const uint32_t *w1 = (const uint32_t *)E.num[1], *w2 = (const uint32_t *)E.num[0];
        auto ld = [&](int pos, int c) -> uint32_t {
            (void)c;
            const int j = pos & (L - 1);
            return (pos >> lg) ? ff_load_coef(w2, st.nw2, j, st.bw, P) : ff_load_coef(w1, st.nw1, j, st.bw, P);
        };
        const uint32_t sc = a.sc[lg][pi];
        auto sto = [&](int pos, int c, uint32_t v) {
            (void)c;
            E.y[pos >> lg][pi * L + (pos & (L - 1))] = ff_mont(v, sc, p, pinv);
        };
        const uint32_t *tw[4] = {E.tw[pi][0], E.tw[pi][1], E.tw[pi][2], E.tw[pi][3]};
        // the chain's longest transforms with literal round positions
        if (lg == 10) early_pair_conv_fixed<10>(tl, tw, p, pinv, st.bit, ld, sto);
        else if (lg == 9) early_pair_conv_fixed<9>(tl, tw, p, pinv, st.bit, ld, sto);
        else if (lg == 8) early_pair_conv_fixed<8>(tl, tw, p, pinv, st.bit, ld, sto);
        else early_pair_conv(tl, tw, p, pinv, st.bit, ld, sto);
        __syncthreads();
        // both new numbers, sums of products of F(k) and F(k - 1): conv1
        // coefficients each, at most (64 len + bw - 1) / bw <= EARLY_MAXC
        uint64_t s1[SS_CH], s2[SS_CH];
        ss_crt_pair(E.ss, R, E.c2.xlo, E.c2.xhi, E.c2.ex, E.y[0], E.y[1], L, st.bw, st.conv1, st.len, G, s1, s2);
        const uint64_t lo = s1[0], hi = s2[0];  // len <= FF_EARLY_LIMBS: one chunk
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
            e.conv1 > (1u << e.lg) || e.nw1 > 2 * FF_EARLY_LIMBS || e.nw2 > e.nw1)
            return hipErrorInvalidValue;
    }
    ff_prof_begin(FK_EARLY, st);
    hipLaunchKernelGGL(k_fib_early, dim3(1), dim3(EARLY_T), 0, st, *a, ff_tables_dev(), ff_primes(),
                       ff_garner_consts());
    ff_prof_end(FK_EARLY, 16.0 * a->st[a->nsteps - 1].len, st);
    return hipGetLastError();
}

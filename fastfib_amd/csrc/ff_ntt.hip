// Three-prime NTT passes for gfx950: tables, schedules and kernels.
//
// A product's transform of length L (2^k or 3 * 2^k, per prime) runs as
//   forward outer passes  -> inner blocks (forward, pointwise, inverse)
//   -> inverse outer passes,
// a four-step decomposition L = R * S done recursively: an outer pass does
// R-point transforms down the columns of a block (stride S, in an LDS tile
// of R rows x C columns), multiplies element (row frequency q, column c) by
// w_Lb^(c q), and leaves R independent blocks of S contiguous words; the
// innermost blocks (S <= 4096) are convolved entirely in LDS. The forward
// top pass reads the operand's b-bit coefficients straight from its limbs,
// and the inverse top pass folds in the CRT scale R^2 / L, so the residues
// come out in natural order as the exact convolution mod p. This mirrors
// the reference's top-pass + in-cache sub-block recursion (ntt_conv.c:8-16,
// 198-226) and its radix-3 outer layer (ntt_radix3.c:8-27), re-planned for
// LDS tiles and a grid of workgroups instead of CPU threads.
//
// Grid: blockIdx.x = tile, blockIdx.y = 3 * item + prime. All items of one
// launch share the schedule (same L, width and passes).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <vector>

#include "ff_crt_body.h"
#include "ff_internal.h"
#include "ff_ntt_dev.h"
#include "ff_runtime.h"
#include "ff_school_dev.h"

static const uint32_t kP[3] = {FF_P1, FF_P2, FF_P3};
static const uint32_t kG[3] = {5, 31, 13};

static FFTables *g_tables = nullptr;
static FFPrime g_prime[3];

static FFPrimes g_primes_arg;

static uint32_t h_pow(uint32_t a, uint64_t e, uint32_t p) {
    uint64_t r = 1, x = a % p;
    while (e) {
        if (e & 1) r = r * x % p;
        x = x * x % p;
        e >>= 1;
    }
    return (uint32_t)r;
}

static uint32_t h_mont_form(uint32_t x, uint32_t p) { return (uint32_t)(((uint64_t)x << 32) % p); }

extern "C" void ff_prime_consts(FFPrime out[3]) { memcpy(out, g_prime, sizeof g_prime); }
extern "C" const FFTables *ff_tables_dev(void) { return g_tables; }

// Builds the twiddle tables on the host (a few hundred KB) and uploads them.
extern "C" int ff_tables_init(void) {
    if (g_tables) return 0;
    FFTables *h = (FFTables *)malloc(sizeof(FFTables));
    if (!h) return -1;
    for (int i = 0; i < 3; i++) {
        const uint32_t p = kP[i];
        uint32_t inv = p;
        for (int it = 0; it < 5; it++) inv *= 2u - p * inv;  // p^-1 mod 2^32 by Newton
        FFPrime &P = g_prime[i];
        P.p = p;
        P.pinv = inv;
        const uint64_t R = (1ull << 32) % p;
        P.r2 = (uint32_t)(R * R % p);
        P.one = (uint32_t)R;
        const uint32_t w3 = h_pow(kG[i], (p - 1) / 3, p);
        P.w3 = h_mont_form(w3, p);
        P.w3p = P.w3 * inv;
        P.iw3 = h_mont_form((uint32_t)((uint64_t)w3 * w3 % p), p);
        P.iw3p = P.iw3 * inv;
        h->rt[i][0] = h->irt[i][0] = h->rtp[i][0] = h->irtp[i][0] = 0;
        for (uint32_t hh = 1; hh < FF_TW; hh <<= 1) {
            const uint32_t w = h_pow(kG[i], (p - 1) / (2 * hh), p);
            const uint32_t wi = h_pow(w, p - 2, p);
            uint64_t cur = 1, curi = 1;
            for (uint32_t j = 0; j < hh; j++) {
                h->rt[i][hh + j] = h_mont_form((uint32_t)cur, p);
                h->irt[i][hh + j] = h_mont_form((uint32_t)curi, p);
                h->rtp[i][hh + j] = h->rt[i][hh + j] * inv;
                h->irtp[i][hh + j] = h->irt[i][hh + j] * inv;
                cur = cur * w % p;
                curi = curi * wi % p;
            }
        }
        const uint32_t wb = h_pow(kG[i], (p - 1) / FF_NBASE, p);
        uint64_t cur = 1;
        for (uint32_t x = 0; x < FF_TL_SIZE; x++) {
            h->tl[i][x] = h_mont_form((uint32_t)cur, p);
            cur = cur * wb % p;
        }
        const uint32_t wh = h_pow(wb, FF_TL_SIZE, p);
        cur = 1;
        for (uint32_t x = 0; x < FF_TH_SIZE; x++) {
            h->th[i][x] = h_mont_form((uint32_t)cur, p);
            cur = cur * wh % p;
        }
    }
    memcpy(g_primes_arg.p, g_prime, sizeof g_prime);
    FFTables *d = nullptr;
    if (ff_hip_malloc((void **)&d, sizeof(FFTables)) != hipSuccess) {
        free(h);
        return -1;
    }
    if (hipMemcpy(d, h, sizeof(FFTables), hipMemcpyHostToDevice) != hipSuccess) {
        free(h);
        return -1;
    }
    free(h);
    g_tables = d;
    return 0;
}

// ------------------------------------------------------------------ kernels

#include "ff_ntt_body.h"

template <bool INV, bool TOP, bool R3, int SH = 0>
__global__ __launch_bounds__(FF_THREADS) void k_outer(FFPass ps, FFXformSet xs, const FFTables *T,
                                                      FFPrimes PR) {
    extern __shared__ uint32_t lds[];
    const int item = blockIdx.y / 3, pi = blockIdx.y - 3 * (blockIdx.y / 3);
    outer_body<INV, TOP, R3, false, SH>(ps, xs.x[item], pi, blockIdx.x, T, PR, lds);
}

// The fixed-geometry shape of a pass (ff_shape), 0 for none.
static int outer_shape(const FFPass &ps) {
    for (int sh = 1; sh <= FF_NSHAPES; sh++) {
        const FFShape s = ff_shape(sh);
        if (ps.rho == s.rho && ps.logC == s.logC && ps.ld == s.ld && ps.rl == s.rl && ps.r3 == s.r3) return sh;
    }
    return 0;
}

int ff_outer_shape(const FFPass *ps, bool top) {
    const int sh = outer_shape(*ps);
    return top || sh == 6 || sh == 7 ? sh : 0;
}

// An outer pass (INV: inverse; TOP: the top pass), by its shape: the
// fixed-geometry kernel where one is instantiated for it, else the general
// one.
template <bool INV, bool TOP>
static void launch_outer(const FFPass &ps, const FFXformSet *xs, dim3 grid, hipStream_t st, const FFTables *T,
                         FFPrimes PR) {
    const size_t lds = (size_t)ps.R * ps.ld * sizeof(uint32_t);
#define FF_SHCASE(R3, SH)                                                                                       \
    case SH: hipLaunchKernelGGL((k_outer<INV, TOP, R3, SH>), grid, dim3(FF_THREADS), lds, st, ps, *xs, T, PR); \
        return;
    const int sh = outer_shape(ps);
    if constexpr (TOP) {
        if (ps.r3) {
            switch (sh) {
                FF_SHCASE(true, 9)
                FF_SHCASE(true, 10)
                FF_SHCASE(true, 11)
                FF_SHCASE(true, 13)
                FF_SHCASE(true, 14)
            }
            hipLaunchKernelGGL((k_outer<INV, TOP, true>), grid, dim3(FF_THREADS), lds, st, ps, *xs, T, PR);
            return;
        }
        switch (sh) {
            FF_SHCASE(false, 1)
            FF_SHCASE(false, 2)
            FF_SHCASE(false, 3)
            FF_SHCASE(false, 4)
            FF_SHCASE(false, 5)
            FF_SHCASE(false, 6)
            FF_SHCASE(false, 7)
            FF_SHCASE(false, 8)
            FF_SHCASE(false, 12)
        }
        hipLaunchKernelGGL((k_outer<INV, TOP, false>), grid, dim3(FF_THREADS), lds, st, ps, *xs, T, PR);
    } else {
        switch (sh) {
            FF_SHCASE(false, 6)
            FF_SHCASE(false, 7)
        }
        hipLaunchKernelGGL((k_outer<INV, TOP, false>), grid, dim3(FF_THREADS), lds, st, ps, *xs, T, PR);
    }
#undef FF_SHCASE
}



template <int PM>
__global__ __launch_bounds__(2 * FF_THREADS, 2) void k_inner_pair(FFInner in, FFConvSet cs, const FFTables *T,
                                                                  FFPrimes PR) {
    extern __shared__ uint32_t lds[];
    const int item = blockIdx.y / 3, pi = blockIdx.y - 3 * (blockIdx.y / 3);
    inner_pair_body<PM>(in, cs.c[item], pi, blockIdx.x, T, PR, lds);
}

template <bool SMALL, bool R3, bool FIX12 = false>
__global__ __launch_bounds__(FF_THREADS, 4) void k_inner(FFInner in, FFConvSet cs, const FFTables *T, FFPrimes PR) {
    extern __shared__ uint32_t lds[];
    const int item = blockIdx.y / 3, pi = blockIdx.y - 3 * (blockIdx.y / 3);
    inner_body<SMALL, R3, false, FIX12>(in, cs.c[item], pi, blockIdx.x, T, PR, lds);
}



// ------------------------------------------------------------------ plans

static int ilog2(uint64_t x) { return 63 - __builtin_clzll(x); }

static uint32_t h_scale(uint32_t L, int i) {
    // R^2 / L mod p: Montgomery form of R / L, undoes the pointwise R^-1 and
    // the inverse transform's factor L (the reference's scale[] and scale3[],
    // ntt_tables.c:103-114).
    const uint32_t p = kP[i];
    const uint64_t R = (1ull << 32) % p;
    const uint64_t R2 = R * R % p;
    return (uint32_t)(R2 * h_pow(L % p, p - 2, p) % p);
}

uint32_t ff_inv_scale(uint32_t L, int i) { return h_scale(L, i); }
FFPrimes ff_primes(void) { return g_primes_arg; }

// Fills one outer pass for blocks of Lb words with R rows.
// Levels per round of an outer pass with 2^rho rows (times 3): rounds of
// four levels and one of the remainder, or of three where bit rho of
// g_outer_rl3 is set.
static int ff_outer_rl(int rho) { return (g_outer_rl3 >> rho) & 1 ? 3 : 4; }

static void fill_pass(FFPass *ps, uint32_t L, uint32_t Lb, uint32_t R, int top, int bw, int tile_words) {
    memset(ps, 0, sizeof *ps);
    ps->L = L;
    ps->Lb = Lb;
    ps->R = R;
    ps->Sb = Lb / R;
    ps->nblocks = L / Lb;
    ps->r3 = (R % 3 == 0);
    ps->rho = ilog2(ps->r3 ? R / 3 : R);
    uint32_t C = 16;
    while (C * 2 * R <= (uint32_t)tile_words && C * 2 <= ps->Sb) C *= 2;
    if (C > ps->Sb) C = ps->Sb;
    ps->C = (int)C;
    ps->logC = ilog2(C);
    // one spare word per row breaks LDS bank conflicts of narrow tiles, when
    // it still fits in 64 KB
    ps->ld = (int)C + ((C <= 32 && R * (C + 1) <= 16384) ? 1 : 0);
    ps->tpb = ps->Sb / C;
    ps->rootstep = FF_NBASE / Lb;
    ps->rstep3 = FF_NBASE / R;
    ps->top = top;
    ps->rl = g_outer_rl ? g_outer_rl : ff_outer_rl(ps->rho);
    ps->bw = bw;
    for (int i = 0; i < 3; i++) ps->sc[i] = h_scale(L, i);
}

// Schedule for a transform of length L with coefficient width bw.
// S_log: log2 of the innermost block (<= 12; 0 picks 12); tile_words: LDS
// words of an outer tile (0 picks 16384).
extern "C" int ff_plan_build(FFPlan *pl, uint32_t L, int bw, int lg, int r3, uint64_t conv, int S_log,
                             int tile_words) {
    memset(pl, 0, sizeof *pl);
    pl->L = L;
    pl->bw = bw;
    pl->lg = lg;
    pl->r3 = r3;
    pl->conv = conv;
    if (tile_words <= 0) tile_words = 16384;
    if (S_log <= 0 || S_log > 12) S_log = 12;
    FFInner &in = pl->inner;
    in.L = L;
    for (int i = 0; i < 3; i++) in.sc[i] = h_scale(L, i);
    in.bw = bw;
    if (L <= FF_TW) {
        in.small = 1;
        in.r3 = r3;
        in.s = r3 ? ilog2(L / 3) : ilog2(L);
        in.S = L;
        in.rstep3 = FF_NBASE / L;
        pl->npass = 0;
        return 0;
    }
    const int k = r3 ? ilog2(L / 3) : ilog2(L);
    const int s = k < S_log ? k : S_log;
    in.s = s;
    in.S = 1u << s;
    // blocks of at most 2^10 words (the small quadruplings' blocks) run
    // radix-4 rounds in the pair kernel: a radix-16 round of 2^10 words has
    // 64 groups for the half's 256 threads
    in.rl = g_pair_rl ? g_pair_rl : (s <= 10 ? 2 : 4);
    const int rbits = k - s;  // L / S = 3^r3 * 2^rbits
    // at most 2^10 rows per pass; the top pass carries the factor 3 and then
    // at most 2^8 power-of-2 rows; the top pass gets the smaller share
    int npass = 1;
    while (true) {
        const int top_bits = rbits / npass, other_bits = (rbits + npass - 1) / npass;
        if (top_bits <= (r3 ? 8 : 10) && other_bits <= 10) break;
        if (++npass > 3) return -1;
    }
    int bits[3] = {0, 0, 0};
    for (int i = 0; i < npass; i++) bits[i] = rbits / npass;
    for (int i = 0; i < rbits % npass; i++) bits[npass - 1 - i]++;
    pl->npass = npass;
    uint32_t Lb = L;
    for (int i = 0; i < npass; i++) {
        uint32_t R = (1u << bits[i]) * (i == 0 && r3 ? 3u : 1u);
        fill_pass(&pl->pass[i], L, Lb, R, i == 0, bw, tile_words);
        Lb /= R;
    }
    return 0;
}

static size_t outer_lds(const FFPass &ps) { return (size_t)ps.R * ps.ld * sizeof(uint32_t); }

static size_t inner_lds(const FFInner &in) {
    const uint32_t n = in.small ? in.L : in.S;
    return 2 * (n + (n >> 4) + 16) * sizeof(uint32_t);
}

extern "C" size_t ff_outer_lds_bytes(const FFPass *ps) { return outer_lds(*ps); }
extern "C" size_t ff_inner_lds_bytes(const FFInner *in) { return inner_lds(*in); }




size_t ff_queue_lds_crt(void) { return sizeof(CrtShared); }
size_t ff_queue_lds_school(void) { return sizeof(FFSchoolShared); }
size_t ff_queue_lds_school_wave(void) { return sizeof(FFSchoolWaveShared); }


// Algorithmic bytes (per prime, per array): a top forward pass reads the
// operand's limbs and writes 4L; every other outer pass reads and writes 4L.
extern "C" hipError_t ff_launch_forward(const FFPlan *pl, const FFXformSet *xs, int nx, double operand_bytes,
                                        hipStream_t st) {
    for (int i = 0; i < pl->npass; i++) {
        const FFPass &ps = pl->pass[i];
        dim3 grid(ps.nblocks * ps.tpb, 3 * nx);
        const int kind = i == 0 ? FK_FWD_TOP : FK_FWD;
        ff_prof_begin(kind, st);
        if (i == 0) launch_outer<false, true>(ps, xs, grid, st, g_tables, g_primes_arg);
        else launch_outer<false, false>(ps, xs, grid, st, g_tables, g_primes_arg);
        const double L4 = 4.0 * pl->L;
        ff_prof_end(kind, i == 0 ? 3.0 * (operand_bytes + nx * L4) : 3.0 * nx * 2 * L4, st);
    }
    return hipGetLastError();
}

// Algorithmic bytes per prime: the inner kernel reads a's (and b's)
// transform and writes the product's, 4L each; the one-block kernel reads
// the operands' limbs instead.
// Algorithmic 32-bit multiplies of one inner launch, three primes: per
// transform of an S-point block (S/2)(log2 S - 1) twiddled butterflies (the
// lowest level multiplies by 1) at 3 each (a Montgomery product by a root
// with its companion), and 4 per pointwise Montgomery product. A square
// transforms one array each way with one product per point; a general
// product two forward, one back; a weighted pair two forward, one or two
// back, with one or two products per point and output (x (w0 x + w1 y) +
// w2 y^2); at degree 4, three more for x, y and one for the scale.
static double conv_mults(const FFPlan *pl, const FFConvSet *cs, int nc) {
    const FFInner &in = pl->inner;
    const double S = in.small ? (double)in.L : (double)in.S, blocks = in.small ? 1.0 : (double)pl->L / in.S;
    const double bf = S / 2 * (in.s - 1) * 3;  // per block transform
    double m = 0;
    for (int i = 0; i < nc; i++) {
        const FFConv &c = cs->c[i];
        double fwd, inv, pw;
        if (!c.pair) {
            const bool sq = in.small ? c.bsrc == nullptr : c.fb == nullptr;
            fwd = sq ? 1 : 2, inv = 1, pw = 1;
        } else {
            const int outs = (c.pair & 2) ? 2 : 1;
            fwd = 2, inv = outs, pw = 0;
            for (int o = 0; o < outs; o++) {
                const uint8_t *w = o ? c.cv : c.cu;
                pw += 1 + (w[2] ? 1 : 0) + ((c.pair & 4) ? 4 : 0);
            }
        }
        m += blocks * ((fwd + inv) * bf + S * pw * 4);
    }
    return 3 * m;
}

extern "C" hipError_t ff_launch_conv(const FFPlan *pl, const FFConvSet *cs, int nc, int ngeneral,
                                     double operand_bytes, hipStream_t st) {
    const FFInner &in = pl->inner;
    if (cs->c[0].pair && in.small) return hipErrorInvalidValue;  // pairs take multi-pass plans
    const uint32_t n = in.small ? in.L : in.S;
    const size_t words = n + (n >> 4) + 16;
    dim3 grid(in.small ? 1 : pl->L / in.S, 3 * nc);
    const int kind = in.small ? FK_SMALL : FK_INNER;
    const double L4 = 4.0 * pl->L;
    ff_prof_begin(kind, st);
    // squares use tile A only: half the LDS, more workgroups per CU
    bool one_tile = true;
    for (int i = 0; i < nc; i++) {
        const FFConv &c = cs->c[i];
        one_tile = one_tile && !c.pair && !c.bspec && (in.small ? c.bsrc == nullptr : c.fb == nullptr);
    }
    const size_t lds = (one_tile ? 1 : 2) * words * sizeof(uint32_t);
    if (cs->c[0].pair & 6) {  // two outputs or degree 4: the 512-thread pair kernel
        const int pm = ff_pair_mode(in);
        if (pm == 1)
            hipLaunchKernelGGL(k_inner_pair<1>, grid, dim3(2 * FF_THREADS), lds, st, in, *cs, g_tables, g_primes_arg);
        else if (pm == 2)
            hipLaunchKernelGGL(k_inner_pair<2>, grid, dim3(2 * FF_THREADS), lds, st, in, *cs, g_tables, g_primes_arg);
        else
            hipLaunchKernelGGL(k_inner_pair<0>, grid, dim3(2 * FF_THREADS), lds, st, in, *cs, g_tables, g_primes_arg);
    } else if (!in.small && in.s == 12) {  // 2^12-word blocks: the fixed-path kernel
        hipLaunchKernelGGL((k_inner<false, false, true>), grid, dim3(FF_THREADS), lds, st, in, *cs, g_tables,
                           g_primes_arg);
    } else if (!in.small)
        hipLaunchKernelGGL((k_inner<false, false>), grid, dim3(FF_THREADS), lds, st, in, *cs, g_tables, g_primes_arg);
    else if (in.r3)
        hipLaunchKernelGGL((k_inner<true, true>), grid, dim3(FF_THREADS), lds, st, in, *cs, g_tables, g_primes_arg);
    else
        hipLaunchKernelGGL((k_inner<true, false>), grid, dim3(FF_THREADS), lds, st, in, *cs, g_tables, g_primes_arg);
    ff_prof_mults(conv_mults(pl, cs, nc));
    ff_prof_end(kind, in.small ? 3.0 * (operand_bytes + nc * L4) : 3.0 * (nc + ngeneral + nc) * L4, st);
    return hipGetLastError();
}

extern "C" hipError_t ff_launch_inverse(const FFPlan *pl, const FFXformSet *xs, int nx, hipStream_t st) {
    for (int i = pl->npass - 1; i >= 0; i--) {
        const FFPass &ps = pl->pass[i];
        dim3 grid(ps.nblocks * ps.tpb, 3 * nx);
        const int kind = i == 0 ? FK_INV_TOP : FK_INV;
        ff_prof_begin(kind, st);
        if (i == 0) launch_outer<true, true>(ps, xs, grid, st, g_tables, g_primes_arg);
        else launch_outer<true, false>(ps, xs, grid, st, g_tables, g_primes_arg);
        ff_prof_end(kind, 3.0 * nx * 8.0 * pl->L, st);
    }
    return hipGetLastError();
}

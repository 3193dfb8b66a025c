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

#include <vector>

#include "ff_internal.h"
#include "ff_ntt_dev.h"

static const uint32_t kP[3] = {FF_P1, FF_P2, FF_P3};
static const uint32_t kG[3] = {5, 31, 13};

static FFTables *g_tables = nullptr;
static FFPrime g_prime[3];

struct FFPrimes {
    FFPrime p[3];
};
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
    if (hipMalloc(&d, sizeof(FFTables)) != hipSuccess) {
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

#define FF_THREADS 256

// Outer pass over a tile of R rows x C columns of one block.
template <bool INV>
__global__ __launch_bounds__(FF_THREADS) void k_outer(FFPass ps, const FFXform *xs, const FFTables *T,
                                                      FFPrimes PR) {
    extern __shared__ uint32_t lds[];
    const int item = blockIdx.y / 3, pi = blockIdx.y - 3 * (blockIdx.y / 3);
    const FFXform X = xs[item];
    const FFPrime P = PR.p[pi];
    const uint32_t p = P.p;
    const uint32_t blk = blockIdx.x / ps.tpb, ct = blockIdx.x - blk * ps.tpb;
    const uint32_t c0 = ct << ps.logC;
    uint32_t *data = X.data + (size_t)pi * ps.L + (size_t)blk * ps.Lb;
    FFTile t{lds, ps.ld, ps.C, ps.logC, ps.rho, ps.r3 ? 3 : 1, 0};
    const int n = (int)ps.R << ps.logC;

    uint32_t scm = 0;
    if (INV && ps.top) scm = ps.sc[pi];
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int pos = i >> ps.logC, c = i & (ps.C - 1);
        const uint32_t col = c0 + c;
        const size_t g = (size_t)pos * ps.Sb + col;
        uint32_t v;
        if (!INV) {
            v = ps.top ? ff_load_coef(X.src, X.nw, g, ps.bw, P) : data[g];
        } else {
            uint32_t e = col * ff_row_freq(pos, ps.rho, ps.r3) * ps.rootstep;
            e = e ? FF_NBASE - e : 0;
            uint32_t tw = ff_root(T, pi, e, p, P.pinv);
            if (ps.top) tw = ff_mont(tw, scm, p, P.pinv);
            v = ff_mont(data[g], tw, p, P.pinv);
        }
        lds[ff_pos_addr(t, pos, c)] = v;
    }
    if (!INV) {
        if (ps.r3) ff_radix3_layer<false>(t, T, pi, P, ps.rstep3);
        ff_tile_transform<false>(t, T->rt[pi], T->rtp[pi], p);
    } else {
        ff_tile_transform<true>(t, T->irt[pi], T->irtp[pi], p);
        if (ps.r3) ff_radix3_layer<true>(t, T, pi, P, ps.rstep3);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int pos = i >> ps.logC, c = i & (ps.C - 1);
        const uint32_t col = c0 + c;
        const size_t g = (size_t)pos * ps.Sb + col;
        uint32_t v = lds[ff_pos_addr(t, pos, c)];
        if (!INV) {
            const uint32_t e = col * ff_row_freq(pos, ps.rho, ps.r3) * ps.rootstep;
            v = ff_mont(v, ff_root(T, pi, e, p, P.pinv), p, P.pinv);
        }
        data[g] = v;
    }
}

// Inner blocks: forward transform of a (and b), pointwise product, inverse,
// all in LDS. small = 1: the whole transform, read from the limbs and
// scaled for the CRT at the end.
__global__ __launch_bounds__(FF_THREADS) void k_inner(FFInner in, const FFXform *xa, const FFXform *xb,
                                                      const FFConv *cv, const FFTables *T, FFPrimes PR) {
    extern __shared__ uint32_t lds[];
    const int item = blockIdx.y / 3, pi = blockIdx.y - 3 * (blockIdx.y / 3);
    const FFPrime P = PR.p[pi];
    const uint32_t p = P.p;
    const FFConv C = cv[item];
    const bool square = in.small ? (xb[item].src == nullptr) : (C.fb == nullptr);
    const uint32_t n = in.small ? in.L : in.S;
    const int words = (int)(n + (n >> 4) + 16);
    uint32_t *A = lds, *B = lds + words;
    FFTile ta{A, 1, 1, 0, in.small && in.r3 ? in.s : in.s, in.small && in.r3 ? 3 : 1, 1};
    FFTile tb = ta;
    tb.s = B;
    const size_t off = (size_t)pi * in.L + (in.small ? 0 : (size_t)blockIdx.x * in.S);

    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        uint32_t va, vb = 0;
        if (in.small) {
            va = ff_load_coef(xa[item].src, xa[item].nw, i, in.bw, P);
            if (!square) vb = ff_load_coef(xb[item].src, xb[item].nw, i, in.bw, P);
        } else {
            va = C.fa[off + i];
            if (!square) vb = C.fb[off + i];
        }
        A[ff_pos_addr(ta, (int)i, 0)] = va;
        if (!square) B[ff_pos_addr(tb, (int)i, 0)] = vb;
    }
    if (in.small && in.r3) {
        ff_radix3_layer<false>(ta, T, pi, P, in.rstep3);
        if (!square) ff_radix3_layer<false>(tb, T, pi, P, in.rstep3);
    }
    ff_tile_transform<false>(ta, T->rt[pi], T->rtp[pi], p);
    if (!square) ff_tile_transform<false>(tb, T->rt[pi], T->rtp[pi], p);
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const int a = ff_pos_addr(ta, (int)i, 0);
        const uint32_t x = A[a];
        A[a] = ff_mont(x, square ? x : B[a], p, P.pinv);
    }
    __syncthreads();
    ff_tile_transform<true>(ta, T->irt[pi], T->irtp[pi], p);
    if (in.small && in.r3) ff_radix3_layer<true>(ta, T, pi, P, in.rstep3);
    uint32_t *dst = C.fa + off;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        uint32_t v = A[ff_pos_addr(ta, (int)i, 0)];
        if (in.small) v = ff_mont(v, in.sc[pi], p, P.pinv);
        dst[i] = v;
    }
}

// ------------------------------------------------------------------ plans

static int ilog2(uint64_t x) { return 63 - __builtin_clzll(x); }

static uint32_t h_scale(uint32_t L, int i) {
    // R^2 / L mod p: Montgomery form of R / L, undoes the pointwise R^-1 and
    // the inverse transform's factor L (the reference's scale[] and scale3[],
    // ntt_tables.c:299-309).
    const uint32_t p = kP[i];
    const uint64_t R = (1ull << 32) % p;
    const uint64_t R2 = R * R % p;
    return (uint32_t)(R2 * h_pow(L % p, p - 2, p) % p);
}

// Fills one outer pass for blocks of Lb words with R rows.
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

extern "C" hipError_t ff_launch_forward(const FFPlan *pl, const FFXform *x_dev, int nx, hipStream_t st) {
    for (int i = 0; i < pl->npass; i++) {
        const FFPass &ps = pl->pass[i];
        dim3 grid(ps.nblocks * ps.tpb, 3 * nx);
        hipLaunchKernelGGL(k_outer<false>, grid, dim3(FF_THREADS), outer_lds(ps), st, ps, x_dev, g_tables,
                           g_primes_arg);
    }
    return hipGetLastError();
}

extern "C" hipError_t ff_launch_conv(const FFPlan *pl, const FFXform *a_dev, const FFXform *b_dev,
                                     const FFConv *c_dev, int nc, hipStream_t st) {
    const FFInner &in = pl->inner;
    const uint32_t n = in.small ? in.L : in.S;
    const size_t words = n + (n >> 4) + 16;
    dim3 grid(in.small ? 1 : pl->L / in.S, 3 * nc);
    hipLaunchKernelGGL(k_inner, grid, dim3(FF_THREADS), 2 * words * sizeof(uint32_t), st, in, a_dev, b_dev, c_dev,
                       g_tables, g_primes_arg);
    return hipGetLastError();
}

extern "C" hipError_t ff_launch_inverse(const FFPlan *pl, const FFXform *x_dev, int nx, hipStream_t st) {
    for (int i = pl->npass - 1; i >= 0; i--) {
        const FFPass &ps = pl->pass[i];
        dim3 grid(ps.nblocks * ps.tpb, 3 * nx);
        hipLaunchKernelGGL(k_outer<true>, grid, dim3(FF_THREADS), outer_lds(ps), st, ps, x_dev, g_tables,
                           g_primes_arg);
    }
    return hipGetLastError();
}

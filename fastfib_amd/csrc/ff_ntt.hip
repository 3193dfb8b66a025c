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

#define FF_THREADS 256

// Four-step twiddles of an outer pass: element (row position pos, tile
// column c), column col = c0 + c, is scaled by w^(col * f(pos) * rootstep),
// f the row's frequency (ff_row_freq), w the 3 * 2^25-th root; INV: by the
// inverse root, times the CRT scale on the top pass. group() serves the 2^NL
// consecutive rows of one butterfly group: their frequencies are
// f0 + step * rev(j), so the twiddles are base * g^rev(j) from two table
// roots (base = w^(col f0 rootstep), g = w^(col step rootstep)).
template <bool INV>
struct FFOuterTw {
    static constexpr bool on = true;
    static constexpr bool lds_tables = false;
    // outer passes have no middle round
    __device__ uint32_t mid(uint32_t x, uint32_t y, uint32_t p, uint32_t pinv) const { return ff_mont(x, y, p, pinv); }
    const FFTables *T;
    int pi, rho, r3;
    bool top;
    uint32_t c0, rootstep, scm, p, pinv;

    __device__ __forceinline__ uint32_t root(uint32_t e) const {
        if (INV) e = e ? FF_NBASE - e : 0;
        return ff_root(T, pi, e, p, pinv);
    }
    __device__ __forceinline__ uint32_t one(int pos, int c) const {
        uint32_t tw = root((c0 + c) * ff_row_freq(pos, rho, r3) * rootstep);
        if (INV && top) tw = ff_mont(tw, scm, p, pinv);
        return tw;
    }
    template <int NL>
    __device__ __forceinline__ void group(int pos, int c, uint32_t *tw) const {
        const uint32_t cs = (c0 + c) * rootstep;
        const uint32_t step = (r3 ? 3u : 1u) << (rho - NL);
        uint32_t base = root(cs * ff_row_freq(pos, rho, r3));
        if (INV && top) base = ff_mont(base, scm, p, pinv);
        uint32_t pw[1 << NL];
        ff_geo<NL>(base, root(cs * step), pw, p, pinv);
#pragma unroll
        for (int j = 0; j < (1 << NL); j++) tw[j] = pw[ff_rev_c(j, NL)];
    }
};

// The tile of a forward top pass from the operand's limbs: element (row
// pos, column c) is the bw-bit coefficient pos * Sb + c0 + c of the nw
// 32-bit words at src, reduced mod p (src_load8, ntt_kernels.c:40-74,
// reference). Coefficient positions fit 32 bits (L <= 3 * 2^25 coefficients
// of at most 35 bits). The three words holding a coefficient are read by
// buffer loads, which return zero past the operand's last word, so
// coefficients past the operand need no bounds test; each thread issues
// the loads of FF_UNPACK_B elements before it uses the first of them. The
// high bits of a coefficient wider than 32 bits add hi * 2^32 mod p, a
// Montgomery product by R^2.
#define FF_UNPACK_B 8
__device__ __forceinline__ void ff_unpack_tile(const FFTile &t, const uint32_t *src, uint32_t nw, uint32_t Sb,
                                               uint32_t c0, int bw, const FFPrime &P) {
    const int n = (t.nsub << t.rho) << t.logC;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void *)src, (short)0, (int)(nw * 4u), 0x00020000);
    const uint32_t p = P.p;
    const uint32_t lomask = bw >= 32 ? ~0u : ((1u << bw) - 1);
    const uint32_t himask = bw > 32 ? ((1u << (bw - 32)) - 1) : 0u;
    const uint32_t r2p = P.r2 * P.pinv;
    const int tid = ff_tile_tid(t), nthr = ff_tile_nthr(t);
    for (int k0 = 0; k0 * nthr < n; k0 += FF_UNPACK_B) {
        uint32_t w[FF_UNPACK_B][3], sh[FF_UNPACK_B];
#pragma unroll
        for (int e = 0; e < FF_UNPACK_B; e++) {
            const int i = tid + (k0 + e) * nthr;
            const uint32_t g = (uint32_t)(i >> t.logC) * Sb + c0 + (uint32_t)(i & (t.C - 1));
            const uint32_t bit = g * (uint32_t)bw;
            const uint32_t off = (bit >> 5) << 2;
            sh[e] = bit & 31;
            // past the tile (i >= n) the loads are harmless and their values unused
            w[e][0] = __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0);
            w[e][1] = __builtin_amdgcn_raw_buffer_load_b32(rs, off + 4, 0, 0);
            w[e][2] = __builtin_amdgcn_raw_buffer_load_b32(rs, off + 8, 0, 0);
        }
#pragma unroll
        for (int e = 0; e < FF_UNPACK_B; e++) {
            const int i = tid + (k0 + e) * nthr;
            const uint32_t lo = __builtin_amdgcn_alignbit(w[e][1], w[e][0], sh[e]) & lomask;
            const uint32_t hi = __builtin_amdgcn_alignbit(w[e][2], w[e][1], sh[e]) & himask;
            uint32_t r = lo, q = r - p;
            r = q < r ? q : r;
            q = r - p;
            r = q < r ? q : r;  // lo < 2^32 < 3p
            if (bw > 32) r = ff_add(r, ff_mont_w(hi, P.r2, r2p, p), p);
            if (i < n) t.s[ff_pos_addr(t, i >> t.logC, i & (t.C - 1))] = r;
        }
    }
}

// Outer pass over one tile of R rows x C columns (tile index `tile` among
// the pass's nblocks * tpb) of array X, prime pi. COH: the array is read and
// written by coherent accesses (ff_ld / ff_st).
template <bool INV, bool TOP, bool R3, bool COH = false>
__device__ __forceinline__ void outer_body(const FFPass &ps, const FFXform &X, int pi, uint32_t tile,
                                           const FFTables *T, const FFPrimes &PR, uint32_t *lds) {
    const FFPrime P = PR.p[pi];
    const uint32_t p = P.p;
    const uint32_t blk = tile / ps.tpb, ct = tile - blk * ps.tpb;
    const uint32_t c0 = ct << ps.logC;
    const ff_gptr<uint32_t> data = ff_g(X.data) + (size_t)pi * ps.L + (size_t)blk * ps.Lb;
    FFTile t{lds, ps.ld, ps.C, ps.logC, ps.rho, R3 ? 3 : 1, 0};
    const int n = (int)ps.R << ps.logC;

    const uint32_t pinv = P.pinv;
    const uint32_t scm = (INV && TOP) ? ps.sc[pi] : 0;
    const bool top = TOP;
    const int bw = ps.bw;
    const uint32_t Sb = ps.Sb, rootstep = ps.rootstep;
    const int rho = ps.rho, r3 = R3;
    const ff_gptr<const uint32_t> src = ff_g(X.src);
    const uint32_t nw = X.nw;
    // element (row position pos, tile column c) <-> data[pos * Sb + c0 + c]
    auto load_fwd = [&](int pos, int c) -> uint32_t {
        const size_t g = (size_t)pos * Sb + c0 + c;
        return top ? ff_load_coef(src, nw, g, bw, P) : ff_ld<COH>(data + g);
    };
    // the twiddles are applied by the rounds (tw), not by these accessors
    auto store_fwd = [&](int pos, int c, uint32_t v) { ff_st<COH>(data + ((size_t)pos * Sb + c0 + c), v); };
    auto load_inv = [&](int pos, int c) -> uint32_t { return ff_ld<COH>(data + ((size_t)pos * Sb + c0 + c)); };
    const FFOuterTw<INV> tw{T, pi, rho, r3, top, c0, rootstep, scm, p, pinv};
    auto store_inv = [&](int pos, int c, uint32_t v) { ff_st<COH>(data + ((size_t)pos * Sb + c0 + c), v); };
    auto lds_load = [&](int pos, int c) -> uint32_t { return lds[ff_pos_addr(t, pos, c)]; };
    auto lds_store = [&](int pos, int c, uint32_t v) { lds[ff_pos_addr(t, pos, c)] = v; };

    if (!INV) {
        if (!R3 && !TOP) {
            ff_forward_io(t, T->rt[pi], T->rtp[pi], p, pinv, true, true, load_fwd, store_fwd, tw);
        } else {
            // the coefficients are unpacked into the LDS tile, and the
            // radix-3 layer works on the tile
            ff_unpack_tile(t, X.src, nw, Sb, c0, bw, P);
            if (R3) ff_radix3_layer<false>(t, T, pi, P, ps.rstep3);
            else __syncthreads();
            ff_forward_io(t, T->rt[pi], T->rtp[pi], p, pinv, false, true, lds_load, store_fwd, tw);
        }
    } else {
        if (!R3) {
            ff_inverse_io(t, T->irt[pi], T->irtp[pi], p, pinv, true, true, load_inv, store_inv, tw);
        } else {
            ff_inverse_io(t, T->irt[pi], T->irtp[pi], p, pinv, true, false, load_inv, lds_store, tw);
            ff_radix3_layer<true>(t, T, pi, P, ps.rstep3);
            for (int i = threadIdx.x; i < n; i += blockDim.x) store_inv(i >> ps.logC, i & (ps.C - 1),
                                                                        lds_load(i >> ps.logC, i & (ps.C - 1)));
        }
    }
}

template <bool INV, bool TOP, bool R3>
__global__ __launch_bounds__(FF_THREADS) void k_outer(FFPass ps, FFXformSet xs, const FFTables *T,
                                                      FFPrimes PR) {
    extern __shared__ uint32_t lds[];
    const int item = blockIdx.y / 3, pi = blockIdx.y - 3 * (blockIdx.y / 3);
    outer_body<INV, TOP, R3>(ps, xs.x[item], pi, blockIdx.x, T, PR, lds);
}

// Work-list form: item k owns blocks [start[k], start[k+1]) of the flat
// grid, 3 * nblocks * tpb of them (the prime is the major index).
__device__ __forceinline__ int ff_find_item(const uint32_t *start, int n, uint32_t b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <bool INV, bool TOP, bool R3>
__global__ __launch_bounds__(FF_THREADS) void k_outer_list(const FFOuterItem *items, const uint32_t *start, int n,
                                                           const FFTables *T, FFPrimes PR) {
    extern __shared__ uint32_t lds[];
    const int k = ff_find_item(start, n, blockIdx.x);
    const FFOuterItem &it = items[k];
    const uint32_t local = blockIdx.x - start[k];
    const uint32_t ntile = it.ps.nblocks * it.ps.tpb;
    const int pi = (int)(local / ntile);
    outer_body<INV, TOP, R3>(it.ps, it.x, pi, local - (uint32_t)pi * ntile, T, PR, lds);
}


// Inner blocks: forward transform of a (and b), pointwise product, inverse,
// all in LDS. small = 1: the whole transform, read from the limbs and
// scaled for the CRT at the end.
template <bool SMALL, bool R3, bool COH = false>
__device__ __forceinline__ void inner_body(const FFInner &in, const FFConv &C, int pi, uint32_t blkx,
                                           const FFTables *T, const FFPrimes &PR, uint32_t *lds) {
    const FFPrime P = PR.p[pi];
    const uint32_t p = P.p;
    const bool square = SMALL ? (C.bsrc == nullptr) : (C.fb == nullptr);
    const uint32_t n = SMALL ? in.L : in.S;
    const int words = (int)(n + (n >> 4) + 16);
    uint32_t *A = lds, *B = lds + words;
    FFTile ta{A, 1, 1, 0, in.s, R3 ? 3 : 1, 1};
    FFTile tb = ta;
    tb.s = B;
    const size_t off = (size_t)pi * in.L + (SMALL ? 0 : (size_t)blkx * in.S);

    const uint32_t pinv = P.pinv;
    const bool small = SMALL;
    const int bw = in.bw;
    const uint32_t sc = in.sc[pi];
    const ff_gptr<uint32_t> fa = ff_g(C.fa), dst = fa + off;
    const ff_gptr<const uint32_t> fb = ff_g(C.fb), asrc = ff_g(C.asrc), bsrc = ff_g(C.bsrc);
    // element i of the block <-> limbs' coefficient i (one-block form) or
    // the forward transform's word off + i
    auto load_a = [&](int pos, int c) -> uint32_t {
        (void)c;
        return small ? ff_load_coef(asrc, C.anw, (uint64_t)pos, bw, P) : ff_ld<COH>(fa + (off + pos));
    };
    auto load_b = [&](int pos, int c) -> uint32_t {
        (void)c;
        return small ? ff_load_coef(bsrc, C.bnw, (uint64_t)pos, bw, P) : ff_ld<COH>(fb + (off + pos));
    };
    auto store_a = [&](int pos, int c, uint32_t v) {
        (void)c;
        ff_st<COH>(dst + pos, small ? ff_mont(v, sc, p, pinv) : v);
    };
    auto lds_a = [&](int pos, int c) -> uint32_t { return A[ff_pos_addr(ta, pos, c)]; };
    auto lds_a_st = [&](int pos, int c, uint32_t v) { A[ff_pos_addr(ta, pos, c)] = v; };
    auto lds_b_st = [&](int pos, int c, uint32_t v) { B[ff_pos_addr(tb, pos, c)] = v; };

    if (!SMALL) {
        if (!square) {
            ff_forward_io(tb, T->rt[pi], T->rtp[pi], p, pinv, true, false, load_b, lds_b_st);
            __syncthreads();
        }
        if (C.pair == 1) {
            // a weighted pair with one output: the middle round forms
            // cu0 a^2 + cu1 ab + cu2 b^2 from a's group and B
            FFPairMid pm;
            pm.w0 = C.cu[0], pm.w1 = C.cu[1], pm.w2 = C.cu[2];
            ff_conv_io(ta, T->rt[pi], T->rtp[pi], T->irt[pi], T->irtp[pi], B, p, pinv, true, true, load_a, store_a,
                       pm);
        } else {
            ff_conv_io(ta, T->rt[pi], T->rtp[pi], T->irt[pi], T->irtp[pi], square ? nullptr : B, p, pinv, true,
                       true, load_a, store_a);
        }
        return;
    }
    const bool bgiven = C.bspec == 1;  // B's spectrum read, not computed
    // rounds of two levels up to 2^10 points: a radix-16 round of a short
    // transform keeps only n / 16 of the threads busy (16 at n = 256)
    const int small_rl = ta.rho <= 10 ? 2 : 4;
    if (C.bspec == 2) {
        // a's spectrum alone, stored by tile position for the products sharing it
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) lds_a_st((int)i, 0, load_a((int)i, 0));
        if (R3) ff_radix3_layer<false>(ta, T, pi, P, in.rstep3);
        else __syncthreads();
        ff_forward_io(ta, T->rt[pi], T->rtp[pi], p, pinv, false, false, lds_a, lds_a_st);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) ff_st<COH>(dst + i, lds_a((int)i, 0));
        return;
    }
    if (!R3) {
        // one block read from limbs: unpack element by element into LDS
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
            lds_a_st((int)i, 0, load_a((int)i, 0));
            if (bgiven) lds_b_st((int)i, 0, ff_ld<COH>(fb + (off + i)));
            else if (!square) lds_b_st((int)i, 0, load_b((int)i, 0));
        }
        __syncthreads();
        if (!square && !bgiven) {
            ff_forward_io(tb, T->rt[pi], T->rtp[pi], p, pinv, false, false, lds_a, lds_b_st);
            __syncthreads();
        }
        ff_conv_io(ta, T->rt[pi], T->rtp[pi], T->irt[pi], T->irtp[pi], square ? nullptr : B, p, pinv, false, true,
                   lds_a, store_a, FFNoTw{}, small_rl);
        return;
    }
    // 3 * 2^k in one block: the radix-3 layers work on the LDS tiles
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        lds_a_st((int)i, 0, load_a((int)i, 0));
        if (bgiven) lds_b_st((int)i, 0, ff_ld<COH>(fb + (off + i)));
        else if (!square) lds_b_st((int)i, 0, load_b((int)i, 0));
    }
    ff_radix3_layer<false>(ta, T, pi, P, in.rstep3);
    if (!square && !bgiven) {
        ff_radix3_layer<false>(tb, T, pi, P, in.rstep3);
        ff_forward_io(tb, T->rt[pi], T->rtp[pi], p, pinv, false, false, lds_a, lds_b_st);
        __syncthreads();
    }
    ff_conv_io(ta, T->rt[pi], T->rtp[pi], T->irt[pi], T->irtp[pi], square ? nullptr : B, p, pinv, false, false,
               lds_a, lds_a_st, FFNoTw{}, small_rl);
    ff_radix3_layer<true>(ta, T, pi, P, in.rstep3);
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) store_a((int)i, 0, lds_a((int)i, 0));
}

// Pair mode (multi-pass plans): 2 * FF_THREADS threads, the first half on
// the block of a, the second on the block of b, each through the forward
// levels into its LDS tile (tile A, tile B). The middle round then fuses the
// pointwise combination with the first inverse levels: each thread of the
// first half forms u = cu0 X + cu1 Y + cu2 Z of its group (X = a^2, Y = ab,
// Z = b^2 pointwise), each of the second half v likewise, both from tiles A
// and B; a barrier between those reads and the writes back keeps the two
// halves off each other's operands. The remaining inverse rounds end in
// HBM. With u alone the second half runs the rounds on an empty tile
// (nsub = 0), keeping the barrier counts equal.
//
// Degree 4 (pair & 4): the weights apply to X = x^2, Y = xy, Z = y^2 for
// x = a^2 + 2ab, y = a^2 + b^2 (F(2k), F(2k-1) from F(k), F(k-1)), and one
// product by R^3 restores the single R^-1 of a plain pointwise product that
// the inverse pass's CRT scale expects.

// w x mod p for a weight w < 8.
__device__ __forceinline__ uint32_t ff_wmul(uint32_t x, int w, uint32_t p) {
    const uint32_t x2 = ff_add(x, x, p), x4 = ff_add(x2, x2, p);
    uint32_t r = (w & 1) ? x : 0u;
    if (w & 2) r = ff_add(r, x2, p);
    if (w & 4) r = ff_add(r, x4, p);
    return r;
}

struct FFPairPoint {
    int w0, w1, w2;
    int deg4;
    uint32_t r3;  // R^3 mod p
    __device__ uint32_t operator()(uint32_t x, uint32_t y, uint32_t p, uint32_t pinv) const {
        if (deg4) {
            // x' = a (a + 2b), y' = a^2 + b^2
            const uint32_t aa = ff_mont(x, x, p, pinv), bb = ff_mont(y, y, p, pinv);
            const uint32_t y2 = ff_add(y, y, p);
            x = ff_mont(x, ff_add(x, y2, p), p, pinv);
            y = ff_add(aa, bb, p);
        }
        // w0 x^2 + w1 xy + w2 y^2 = x (w0 x + w1 y) + w2 y^2
        uint32_t u = ff_mont(x, ff_add(ff_wmul(x, w0, p), ff_wmul(y, w1, p), p), p, pinv);
        if (w2) u = ff_add(u, ff_wmul(ff_mont(y, y, p, pinv), w2, p), p);
        if (deg4) u = ff_mont(u, r3, p, pinv);
        return u;
    }
};

// The middle round on row bits [0, NL) of a one-column tile: this thread's
// output from tiles A and B, the first NL inverse levels, then (after a
// barrier) its own tile, or HBM through st when it is the last round. Every
// thread of the workgroup runs the same number of iterations.
template <int NL, class St>
__device__ __forceinline__ void pair_mid_round(const FFTile &t, bool active, const uint32_t *A, const uint32_t *B,
                                               const uint32_t *irt, const uint32_t *irtp, uint32_t p, uint32_t pinv,
                                               const FFPairPoint &pt, bool to_g, St st) {
    const int G = 1 << NL;
    const int ngroups = 1 << (t.rho - NL);
    const int nthr = ff_tile_nthr(t);
    const int iters = (ngroups + nthr - 1) / nthr;
    for (int it = 0; it < iters; it++) {
        const int g = ff_tile_tid(t) + it * nthr;
        const bool act = active && g < ngroups;
        uint32_t v[G];
        if (act) {
#pragma unroll
            for (int j = 0; j < G; j++) {
                const int ad = ff_pos_addr(t, (g << NL) | j, 0);
                v[j] = pt(A[ad], B[ad], p, pinv);
            }
            ff_dit_group<NL, false, true>(v, 0, 0, irt, irtp, p, pinv);
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int j = 0; j < G; j++) {
                const int pos = (g << NL) | j;
                if (to_g) st(pos, 0, v[j]);
                else t.s[ff_pos_addr(t, pos, 0)] = v[j];
            }
        }
    }
}

__device__ __forceinline__ void inner_pair_body(const FFInner &in, const FFConv &C, int pi, uint32_t blkx,
                                                const FFTables *T, const FFPrimes &PR, uint32_t *lds) {
    const FFPrime P = PR.p[pi];
    const uint32_t p = P.p, pinv = P.pinv;
    const uint32_t n = in.S;
    const int words = (int)(n + (n >> 4) + 16);
    const int half = threadIdx.x >= FF_THREADS;
    FFTile th{lds + half * words, 1, 1, 0, in.s, 1, 1, half * FF_THREADS, FF_THREADS};
    const size_t off = (size_t)pi * in.L + (size_t)blkx * in.S;
    const uint32_t *src = (half ? C.fb : C.fa) + off;
    uint32_t *dst = (half ? const_cast<uint32_t *>(C.fb) : C.fa) + off;
    auto load_g = [&](int pos, int c) -> uint32_t {
        (void)c;
        return src[pos];
    };
    auto lds_ld = [&](int pos, int c) -> uint32_t { return th.s[ff_pos_addr(th, pos, c)]; };
    auto lds_st = [&](int pos, int c, uint32_t v) { th.s[ff_pos_addr(th, pos, c)] = v; };
    auto store_g = [&](int pos, int c, uint32_t v) {
        (void)c;
        dst[pos] = v;
    };
    // blocks of at most 2^10 words (the small quadruplings' blocks) run
    // radix-4 rounds: a radix-16 round of 2^10 words has 64 groups for the
    // half's 256 threads
    const int rl = in.s <= 10 ? 2 : 4;
    ff_forward_io(th, T->rt[pi], T->rtp[pi], p, pinv, true, false, load_g, lds_st, FFNoTw{}, rl);
    __syncthreads();
    const int two = (C.pair & 3) == 2;
    const FFPairPoint pt{half ? C.cv[0] : C.cu[0], half ? C.cv[1] : C.cu[1], half ? C.cv[2] : C.cu[2],
                         (C.pair & 4) != 0, ff_mont(P.r2, P.r2, p, pinv)};
    const bool active = !half || two;
    const int R = (in.s + rl - 1) / rl;
    const int nl0 = in.s % rl == 0 ? rl : in.s % rl;
    const uint32_t *A = lds, *B = lds + words;
    switch (nl0) {
    case 4: pair_mid_round<4>(th, active, A, B, T->irt[pi], T->irtp[pi], p, pinv, pt, R == 1, store_g); break;
    case 3: pair_mid_round<3>(th, active, A, B, T->irt[pi], T->irtp[pi], p, pinv, pt, R == 1, store_g); break;
    case 2: pair_mid_round<2>(th, active, A, B, T->irt[pi], T->irtp[pi], p, pinv, pt, R == 1, store_g); break;
    default: pair_mid_round<1>(th, active, A, B, T->irt[pi], T->irtp[pi], p, pinv, pt, R == 1, store_g); break;
    }
    if (!active) th.nsub = 0;
    int b = nl0;
    for (int r = 1; r < R; r++) {
        __syncthreads();
        ff_round_any<FF_R_DIT>(rl, false, r == R - 1, th, b, nullptr, nullptr, T->irt[pi], T->irtp[pi], nullptr, p,
                               pinv, lds_ld, store_g);
        b += rl;
    }
}

__global__ __launch_bounds__(2 * FF_THREADS, 2) void k_inner_pair(FFInner in, FFConvSet cs, const FFTables *T,
                                                                  FFPrimes PR) {
    extern __shared__ uint32_t lds[];
    const int item = blockIdx.y / 3, pi = blockIdx.y - 3 * (blockIdx.y / 3);
    inner_pair_body(in, cs.c[item], pi, blockIdx.x, T, PR, lds);
}

template <bool SMALL, bool R3>
__global__ __launch_bounds__(FF_THREADS, 4) void k_inner(FFInner in, FFConvSet cs, const FFTables *T, FFPrimes PR) {
    extern __shared__ uint32_t lds[];
    const int item = blockIdx.y / 3, pi = blockIdx.y - 3 * (blockIdx.y / 3);
    inner_body<SMALL, R3>(in, cs.c[item], pi, blockIdx.x, T, PR, lds);
}

// Work-list form: item k owns blocks [start[k], start[k+1]), 3 * nblk of them.
template <bool SMALL, bool R3>
__global__ __launch_bounds__(FF_THREADS, 4) void k_inner_list(const FFInnerItem *items, const uint32_t *start, int n,
                                                              const FFTables *T, FFPrimes PR) {
    extern __shared__ uint32_t lds[];
    const int k = ff_find_item(start, n, blockIdx.x);
    const FFInnerItem &it = items[k];
    const uint32_t local = blockIdx.x - start[k];
    const uint32_t nblk = SMALL ? 1 : it.in.L / it.in.S;
    const int pi = (int)(local / nblk);
    inner_body<SMALL, R3>(it.in, it.c, pi, local - (uint32_t)pi * nblk, T, PR, lds);
}

// Inner blocks of a transform on their own, in place: the forward levels
// (INV false) or the inverse levels (INV true), no pointwise product (the
// split multiply forms its pointwise sums between the two).
template <bool INV>
__global__ __launch_bounds__(FF_THREADS, 4) void k_inner_part_list(const FFInnerItem *items, const uint32_t *start,
                                                                   int n, const FFTables *T, FFPrimes PR) {
    extern __shared__ uint32_t lds[];
    const int k = ff_find_item(start, n, blockIdx.x);
    const FFInnerItem &it = items[k];
    const uint32_t local = blockIdx.x - start[k];
    const uint32_t nblk = it.in.L / it.in.S;
    const int pi = (int)(local / nblk);
    const uint32_t blkx = local - (uint32_t)pi * nblk;
    const FFPrime P = PR.p[pi];
    FFTile ta{lds, 1, 1, 0, it.in.s, 1, 1};
    uint32_t *d = it.c.fa + (size_t)pi * it.in.L + (size_t)blkx * it.in.S;
    auto ld = [&](int pos, int c) -> uint32_t {
        (void)c;
        return d[pos];
    };
    auto st = [&](int pos, int c, uint32_t v) {
        (void)c;
        d[pos] = v;
    };
    if (!INV) ff_forward_io(ta, T->rt[pi], T->rtp[pi], P.p, P.pinv, true, true, ld, st);
    else ff_inverse_io(ta, T->irt[pi], T->irtp[pi], P.p, P.pinv, true, true, ld, st);
}

// ------------------------------------------------------------------ batch work queue
//
// A whole batch in one launch: workgroups (as many as the GPU holds at
// once) claim the tasks of the table one at a time, in table order, with one
// atomic. The table lists every block of every list launch of the grouped
// form (forward passes, shared one-block spectra, inner blocks, inverse
// passes, CRT jobs, schoolbook products, zero products) in that order, so
// each task's dependencies - the earlier passes of its own product, and the
// transform of a b operand it shares - are earlier tasks, already claimed
// by running workgroups: waits always end, whatever the residency. A
// product's next pass starts as soon as its own previous pass ends, not the
// whole batch's, and the launches' drain and fill between pass kinds are
// gone. CRT jobs take look-back tickets in claim order (never by grid index).
//
// Visibility: the transform arrays that pass between tasks are written and
// read only by agent-scope (coherent) accesses (ff_ld / ff_st; the L2s of
// the XCDs are not coherent with each other), every wave's stores complete
// before a finished task's counter is bumped, and a waiting task's loads are
// issued after its counters are seen. Agent-scope fences would do the same
// for ordinary accesses by writing back and invalidating whole L2s at every
// task, which measured several times slower than the grouped launches.
#define FQ_SPIN_MAX (1u << 20)

__device__ __noinline__ void fq_wait(const uint32_t *c, uint32_t need, uint32_t *fault) {
    for (uint32_t it = 0;; it++) {
        if (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= need) return;
        // a wait that runs out marks the fault word (the host then fails
        // loudly); once it is marked, later waits give up at once
        if ((it & 255) == 255 && lb_load(fault)) return;
        if (it >= FQ_SPIN_MAX) {
            atomicOr(fault, 1u);
            return;
        }
        __builtin_amdgcn_s_sleep(2);
    }
}

// Claims: the table is dealt to nq parts (task t to part t mod nq), each
// with its own claim counter (on its own cache line: one contended address
// serializes every claim of the launch); workgroup w claims from part
// w mod nq, chunk tasks of its part per atomic, and runs them in order.
// Every part keeps at least one workgroup (the grid is at least nq), and a
// part's claims follow the table order, so the lowest unfinished task is
// always held by a workgroup whose earlier tasks are done: waits end.
template <int MINW>
__global__ __launch_bounds__(FF_THREADS, MINW) void k_queue(const FFQTask *tasks, uint32_t ntasks, const char *desc,
                                                      uint32_t *ctr, uint32_t nctr, uint32_t nq, uint32_t chunk,
                                                      const FFTables *T, FFPrimes PR, FFGarner G, uint32_t *fault) {
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t s_task, s_last;
    uint32_t *const cnt = ctr + FQ_HEADER;
    const uint32_t part = blockIdx.x % nq;
    uint32_t next = 0, left = 0;  // the next task of the held chunk, tasks left in it
    for (;;) {
        if (!left) {
            if (threadIdx.x == 0) s_task = atomicAdd(ctr + FQ_LINE * part, 1u);
            __syncthreads();
            // uniform across the workgroup: the task's fields and its item's
            // load as scalars, as in the list kernels
            const uint32_t c = __builtin_amdgcn_readfirstlane(s_task);
            next = part + nq * c * chunk;
            left = chunk;
        }
        const uint32_t t = next;
        if (t >= ntasks) break;
        next += nq;
        left--;
        const FFQTask task = tasks[t];
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < 2; k++)
                if (task.dep[k] != FQ_NONE) fq_wait(cnt + task.dep[k], task.need[k], fault);
        }
        // the data other tasks wrote is read by coherent loads, issued after
        // this barrier (so after the counters were seen)
        __syncthreads();
        const char *item = desc + task.off;
        const uint32_t sub = task.sub;
        if (task.kind < FQ_INNER) {
            const FFOuterItem &it = *(const FFOuterItem *)item;
            const uint32_t ntile = it.ps.nblocks * it.ps.tpb;
            const int pi = (int)(sub / ntile);
            const uint32_t tile = sub - (uint32_t)pi * ntile;
            switch (task.kind) {
            case FQ_OUTER | 2 | 4: outer_body<false, true, true, true>(it.ps, it.x, pi, tile, T, PR, lds); break;
            case FQ_OUTER | 2: outer_body<false, true, false, true>(it.ps, it.x, pi, tile, T, PR, lds); break;
            case FQ_OUTER: outer_body<false, false, false, true>(it.ps, it.x, pi, tile, T, PR, lds); break;
            case FQ_OUTER | 1 | 2 | 4: outer_body<true, true, true, true>(it.ps, it.x, pi, tile, T, PR, lds); break;
            case FQ_OUTER | 1 | 2: outer_body<true, true, false, true>(it.ps, it.x, pi, tile, T, PR, lds); break;
            default: outer_body<true, false, false, true>(it.ps, it.x, pi, tile, T, PR, lds); break;
            }
        } else if (task.kind <= FQ_INNER_SMALL_R3) {
            const FFInnerItem &it = *(const FFInnerItem *)item;
            if (task.kind == FQ_INNER) {
                const uint32_t nblk = it.in.L / it.in.S;
                const int pi = (int)(sub / nblk);
                inner_body<false, false, true>(it.in, it.c, pi, sub - (uint32_t)pi * nblk, T, PR, lds);
            } else if (task.kind == FQ_INNER_SMALL) {
                inner_body<true, false, true>(it.in, it.c, (int)sub, 0, T, PR, lds);
            } else {
                inner_body<true, true, true>(it.in, it.c, (int)sub, 0, T, PR, lds);
            }
        } else if (task.kind == FQ_CRT) {
            const FFCrtJob &J = *(const FFCrtJob *)item;
            CrtShared &sh = *(CrtShared *)lds;
            if (crt_lpt_list(J.bw, J.nout) == 4) crt_body<4, true>(J, sub, G, fault, sh, false);
            else crt_body<1, true>(J, sub, G, fault, sh, false);
        } else if (task.kind == FQ_SCHOOL) {
            school_body(*(const FFSchoolItem *)item, *(FFSchoolShared *)lds);
        } else if (task.kind == FQ_SCHOOL_WAVE) {
            FFSchoolWaveShared &sh = *(FFSchoolWaveShared *)lds;
            const int w = threadIdx.x >> 6;
            const bool live = (uint32_t)w < sub;
            const FFSchoolItem *its = (const FFSchoolItem *)item;
            school_wave(its[live ? w : 0], live, sh.A[w], sh.B[w], sh.lo[w], sh.mid[w], sh.hi[w]);
        } else {
            if (threadIdx.x == 0) **(uint64_t *const *)item = 0;
        }
        // this task's coherent stores complete (every wave) before its counter moves
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        if (threadIdx.x == 0 && task.done != FQ_NONE) atomicAdd(cnt + task.done, 1u);
    }
    // every task is claimed; the last workgroup to get here (all others have
    // finished their last task) clears the counters for the next launch
    if (threadIdx.x == 0) s_last = atomicAdd(ctr + FQ_LINE * FQ_MAXQ, 1u) == gridDim.x - 1;
    __syncthreads();
    if (s_last) {
        for (uint32_t i = threadIdx.x; i < FQ_HEADER + nctr; i += blockDim.x)
            __hip_atomic_store(ctr + i, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
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

static size_t inner_lds(const FFInner &in) {
    const uint32_t n = in.small ? in.L : in.S;
    return 2 * (n + (n >> 4) + 16) * sizeof(uint32_t);
}

extern "C" size_t ff_outer_lds_bytes(const FFPass *ps) { return outer_lds(*ps); }
extern "C" size_t ff_inner_lds_bytes(const FFInner *in) { return inner_lds(*in); }

extern "C" hipError_t ff_launch_outer_list(bool inv, bool top, bool r3, const FFOuterItem *items,
                                           const uint32_t *start, int n, uint32_t blocks, size_t lds, double bytes,
                                           hipStream_t st) {
    if (!blocks) return hipSuccess;
    const int kind = inv ? (top ? FK_INV_TOP : FK_INV) : (top ? FK_FWD_TOP : FK_FWD);
    const dim3 g(blocks), b(FF_THREADS);
    ff_prof_begin(kind, st);
    if (!inv && top && r3) hipLaunchKernelGGL((k_outer_list<false, true, true>), g, b, lds, st, items, start, n, g_tables, g_primes_arg);
    else if (!inv && top) hipLaunchKernelGGL((k_outer_list<false, true, false>), g, b, lds, st, items, start, n, g_tables, g_primes_arg);
    else if (!inv) hipLaunchKernelGGL((k_outer_list<false, false, false>), g, b, lds, st, items, start, n, g_tables, g_primes_arg);
    else if (top && r3) hipLaunchKernelGGL((k_outer_list<true, true, true>), g, b, lds, st, items, start, n, g_tables, g_primes_arg);
    else if (top) hipLaunchKernelGGL((k_outer_list<true, true, false>), g, b, lds, st, items, start, n, g_tables, g_primes_arg);
    else hipLaunchKernelGGL((k_outer_list<true, false, false>), g, b, lds, st, items, start, n, g_tables, g_primes_arg);
    ff_prof_end(kind, bytes, st);
    return hipGetLastError();
}

extern "C" hipError_t ff_launch_inner_list(bool small, bool r3, const FFInnerItem *items, const uint32_t *start,
                                           int n, uint32_t blocks, size_t lds, double bytes, hipStream_t st) {
    if (!blocks) return hipSuccess;
    const int kind = small ? FK_SMALL : FK_INNER;
    // one-block lists of short transforms: n / 4 threads (what radix-4
    // rounds keep busy), whole wavefronts, so that more items run per CU.
    // The list's LDS size bounds its longest item: lds = 8 (n + n/16 + 16).
    uint32_t threads = FF_THREADS;
    if (small) {
        const size_t words = lds / 8;
        const size_t nmax = words > 16 ? (words - 16) * 16 / 17 + 1 : 1;
        const size_t want = (nmax / 4 + 63) / 64 * 64;
        threads = want < 64 ? 64 : (want < FF_THREADS ? (uint32_t)want : FF_THREADS);
    }
    const dim3 g(blocks), b(threads);
    ff_prof_begin(kind, st);
    if (!small) hipLaunchKernelGGL((k_inner_list<false, false>), g, b, lds, st, items, start, n, g_tables, g_primes_arg);
    else if (r3) hipLaunchKernelGGL((k_inner_list<true, true>), g, b, lds, st, items, start, n, g_tables, g_primes_arg);
    else hipLaunchKernelGGL((k_inner_list<true, false>), g, b, lds, st, items, start, n, g_tables, g_primes_arg);
    ff_prof_end(kind, bytes, st);
    return hipGetLastError();
}

extern "C" hipError_t ff_launch_inner_part_list(bool inv, const FFInnerItem *items, const uint32_t *start, int n,
                                                uint32_t blocks, size_t lds, double bytes, hipStream_t st) {
    if (!blocks) return hipSuccess;
    const int kind = inv ? FK_INV : FK_FWD;
    ff_prof_begin(kind, st);
    if (!inv) hipLaunchKernelGGL((k_inner_part_list<false>), dim3(blocks), dim3(FF_THREADS), lds, st, items, start, n,
                                 g_tables, g_primes_arg);
    else hipLaunchKernelGGL((k_inner_part_list<true>), dim3(blocks), dim3(FF_THREADS), lds, st, items, start, n,
                            g_tables, g_primes_arg);
    ff_prof_end(kind, bytes, st);
    return hipGetLastError();
}

size_t ff_queue_lds_crt(void) { return sizeof(CrtShared); }
size_t ff_queue_lds_school(void) { return sizeof(FFSchoolShared); }
size_t ff_queue_lds_school_wave(void) { return sizeof(FFSchoolWaveShared); }

// One workgroup per task up to what the GPU holds at once with this LDS
// size (more would only queue behind them).
hipError_t ff_launch_queue(const FFQTask *tasks, uint32_t ntasks, const char *desc, uint32_t *ctr, uint32_t nctr,
                           size_t lds, double bytes, hipStream_t st) {
    if (!ntasks) return hipSuccess;
    // register budget: waves per SIMD the compiler must allow (1: its own
    // choice, no spills; 3 or 4: more workgroups resident, some spills)
    const int w = g_queue_waves;
    const void *k = w >= 4 ? (const void *)k_queue<4> : w == 3 ? (const void *)k_queue<3> : (const void *)k_queue<1>;
    static std::map<std::pair<const void *, size_t>, uint32_t> resident;
    auto r = resident.find({k, lds});
    if (r == resident.end()) {
        int per_cu = 0, cus = 0;
        FF_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, FF_THREADS, lds));
        FF_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ff_device()));
        r = resident.emplace(std::make_pair(k, lds), (uint32_t)(per_cu > 0 ? per_cu : 1) * (uint32_t)cus).first;
    }
    const uint32_t chunk = (uint32_t)g_queue_chunk;
    const uint32_t chunks = (ntasks + chunk - 1) / chunk;
    const uint32_t grid = chunks < r->second ? chunks : r->second;
    const uint32_t nq = (uint32_t)g_queue_parts < grid ? (uint32_t)g_queue_parts : grid;
    FFGarner G;
    ff_garner_get(&G);
    ff_prof_begin(FK_QUEUE, st);
    if (w >= 4)
        hipLaunchKernelGGL(k_queue<4>, dim3(grid), dim3(FF_THREADS), lds, st, tasks, ntasks, desc, ctr, nctr, nq, chunk,
                           g_tables,
                           g_primes_arg, G, ff_lb_fault());
    else if (w == 3)
        hipLaunchKernelGGL(k_queue<3>, dim3(grid), dim3(FF_THREADS), lds, st, tasks, ntasks, desc, ctr, nctr, nq, chunk,
                           g_tables,
                           g_primes_arg, G, ff_lb_fault());
    else
        hipLaunchKernelGGL(k_queue<1>, dim3(grid), dim3(FF_THREADS), lds, st, tasks, ntasks, desc, ctr, nctr, nq, chunk,
                           g_tables,
                           g_primes_arg, G, ff_lb_fault());
    ff_prof_end(FK_QUEUE, bytes, st);
    return hipGetLastError();
}

// Algorithmic bytes (per prime, per array): a top forward pass reads the
// operand's limbs and writes 4L; every other outer pass reads and writes 4L.
extern "C" hipError_t ff_launch_forward(const FFPlan *pl, const FFXformSet *xs, int nx, double operand_bytes,
                                        hipStream_t st) {
    for (int i = 0; i < pl->npass; i++) {
        const FFPass &ps = pl->pass[i];
        dim3 grid(ps.nblocks * ps.tpb, 3 * nx);
        const int kind = i == 0 ? FK_FWD_TOP : FK_FWD;
        ff_prof_begin(kind, st);
        if (i == 0 && ps.r3)
            hipLaunchKernelGGL((k_outer<false, true, true>), grid, dim3(FF_THREADS), outer_lds(ps), st, ps, *xs,
                               g_tables, g_primes_arg);
        else if (i == 0)
            hipLaunchKernelGGL((k_outer<false, true, false>), grid, dim3(FF_THREADS), outer_lds(ps), st, ps, *xs,
                               g_tables, g_primes_arg);
        else
            hipLaunchKernelGGL((k_outer<false, false, false>), grid, dim3(FF_THREADS), outer_lds(ps), st, ps, *xs,
                               g_tables, g_primes_arg);
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
        hipLaunchKernelGGL(k_inner_pair, grid, dim3(2 * FF_THREADS), lds, st, in, *cs, g_tables, g_primes_arg);
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
        if (i == 0 && ps.r3)
            hipLaunchKernelGGL((k_outer<true, true, true>), grid, dim3(FF_THREADS), outer_lds(ps), st, ps, *xs,
                               g_tables, g_primes_arg);
        else if (i == 0)
            hipLaunchKernelGGL((k_outer<true, true, false>), grid, dim3(FF_THREADS), outer_lds(ps), st, ps, *xs,
                               g_tables, g_primes_arg);
        else
            hipLaunchKernelGGL((k_outer<true, false, false>), grid, dim3(FF_THREADS), outer_lds(ps), st, ps, *xs,
                               g_tables, g_primes_arg);
        ff_prof_end(kind, 3.0 * nx * 8.0 * pl->L, st);
    }
    return hipGetLastError();
}

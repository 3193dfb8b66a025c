/*
 * ORACLE — test infrastructure, not product code (see oracle.h).
 *
 * Scalar restatement of the reference path in plain C with u64 `%`
 * arithmetic: no Montgomery form, no AVX2, no threads. It computes the same
 * exact results by the same schedule rules (width, length, split pieces,
 * doubling identities, Barrett tree), in natural coefficient order.
 * Each function names the reference file:line it follows.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

const uint32_t ffo_prime[3] = {2113929217u, 2013265921u, 1811939329u}; /* ntt_arith.c:48-52 */
const uint32_t ffo_gen[3] = {5, 31, 13};

static int g_max_lg = 25; /* NTT_MAX_LG, common.h:46-48 */

void ffo_set_max_lg(int max_lg) { g_max_lg = max_lg; }
int ffo_get_max_lg(void) { return g_max_lg; }
void ffo_free(void *p) { free(p); }

static void *xalloc(size_t bytes) {
    void *p = malloc(bytes ? bytes : 1);
    if (!p) abort();
    return p;
}

static void *xzalloc(size_t bytes) {
    void *p = calloc(bytes ? bytes : 1, 1);
    if (!p) abort();
    return p;
}

/* ---------------------------------------------------------------- modular */

static uint32_t mulm(uint32_t a, uint32_t b, uint32_t p) { return (uint32_t)((uint64_t)a * b % p); }
static uint32_t addm(uint32_t a, uint32_t b, uint32_t p) { uint32_t s = a + b; return s >= p ? s - p : s; }
static uint32_t subm(uint32_t a, uint32_t b, uint32_t p) { return a >= b ? a - b : a + p - b; }

/* a^e mod p (ntt_arith.c:57-66). */
static uint32_t powm(uint32_t a, uint64_t e, uint32_t p) {
    uint64_t r = 1, x = a % p;
    while (e) {
        if (e & 1) r = r * x % p;
        x = x * x % p;
        e >>= 1;
    }
    return (uint32_t)r;
}

static uint32_t invm(uint32_t a, uint32_t p) { return powm(a, p - 2, p); }

/* ------------------------------------------------------------- transforms */

/* In-place DFT of length M = 2^k with root w (primitive M-th): bit reversal
   then iterative Cooley-Tukey; natural order in and out. */
static void dft_pow2(uint32_t *x, size_t M, uint32_t w, uint32_t p) {
    for (size_t i = 1, j = 0; i < M; i++) {
        size_t bit = M >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { uint32_t t = x[i]; x[i] = x[j]; x[j] = t; }
    }
    for (size_t len = 2; len <= M; len <<= 1) {
        uint32_t wl = powm(w, M / len, p);
        for (size_t s = 0; s < M; s += len) {
            uint32_t t = 1;
            for (size_t j = 0; j < len / 2; j++) {
                uint32_t u = x[s + j], v = mulm(x[s + j + len / 2], t, p);
                x[s + j] = addm(u, v, p);
                x[s + j + len / 2] = subm(u, v, p);
                t = mulm(t, wl, p);
            }
        }
    }
}

/* DFT of length L = 2^k or 3*2^k with root w. For L = 3M the radix-3 DIF
   layer of ntt_radix3.c:8-22 comes first: for n1 < M with a_r = x[n1+rM],
   y_r[n1] = w^(r n1) * sum_r' a_r' w3^(r r'); then X[3k1+r] = DFT_M(y_r)[k1]. */
static void dft_any(uint32_t *x, size_t L, uint32_t w, uint32_t p) {
    if ((L & (L - 1)) == 0) {
        dft_pow2(x, L, w, p);
        return;
    }
    size_t M = L / 3;
    uint32_t w3 = powm(w, M, p), w3s = mulm(w3, w3, p), wm = powm(w, 3, p);
    uint32_t *y = xalloc(L * sizeof(uint32_t));
    uint32_t t1 = 1, step = w;
    for (size_t n1 = 0; n1 < M; n1++) {
        uint32_t a0 = x[n1], a1 = x[n1 + M], a2 = x[n1 + 2 * M];
        uint32_t t2 = mulm(t1, t1, p);
        y[n1] = addm(addm(a0, a1, p), a2, p);
        y[M + n1] = mulm(addm(addm(a0, mulm(a1, w3, p), p), mulm(a2, w3s, p), p), t1, p);
        y[2 * M + n1] = mulm(addm(addm(a0, mulm(a1, w3s, p), p), mulm(a2, w3, p), p), t2, p);
        t1 = mulm(t1, step, p);
    }
    for (int r = 0; r < 3; r++) dft_pow2(y + r * M, M, wm, p);
    for (size_t k1 = 0; k1 < M; k1++)
        for (int r = 0; r < 3; r++) x[3 * k1 + r] = y[r * M + k1];
    free(y);
}

static uint32_t root_of(int pi, size_t L) {
    uint32_t p = ffo_prime[pi];
    return powm(ffo_gen[pi], (p - 1) / L, p);
}

void ffo_ntt_forward(int pi, uint32_t *x, size_t L) { dft_any(x, L, root_of(pi, L), ffo_prime[pi]); }

/* The contract of ntt_conv / ntt_conv_any (ntt_conv.c:198-226,
   ntt_radix3.c:196-219) up to the reference's index reversal and factor L,
   which its CRT undoes (crt.c:38-43); here both are undone directly. */
void ffo_cyclic_conv(int pi, const uint32_t *a, const uint32_t *b, size_t L, uint32_t *out) {
    uint32_t p = ffo_prime[pi], w = root_of(pi, L), wi = invm(w, p), li = invm((uint32_t)(L % p), p);
    uint32_t *fa = xalloc(L * sizeof(uint32_t)), *fb = xalloc(L * sizeof(uint32_t));
    memcpy(fa, a, L * sizeof(uint32_t));
    memcpy(fb, b, L * sizeof(uint32_t));
    dft_any(fa, L, w, p);
    if (b == a) memcpy(fb, fa, L * sizeof(uint32_t));
    else dft_any(fb, L, w, p);
    for (size_t i = 0; i < L; i++) fa[i] = mulm(fa[i], fb[i], p);
    dft_any(fa, L, wi, p);
    for (size_t i = 0; i < L; i++) out[i] = mulm(fa[i], li, p);
    free(fa);
    free(fb);
}

/* ------------------------------------------------------------- schedule */

/* ntt_width_ok (mul_ntt.c:55-59): min(ca, cb) * 2^(2w) < p1 p2 p3. */
int ffo_width_ok(size_t bits_a, size_t bits_b, unsigned w) {
    const u128 M = (u128)ffo_prime[0] * ffo_prime[1] * ffo_prime[2];
    size_t ca = (bits_a + w - 1) / w, cb = (bits_b + w - 1) / w;
    return ((u128)(ca < cb ? ca : cb) << (2 * w)) < M;
}

/* ntt_length (mul_ntt.c:78-93): smallest 2^lg >= conv (>= 16), or 3*2^(lg-2)
   when that is enough and 2^(lg-2) >= 16. */
size_t ffo_ntt_length(size_t conv, int *lg_out, int *r3_out) {
    size_t L = 16;
    int lg = 4, r3 = 0;
    while (L < conv) {
        L <<= 1;
        lg++;
    }
    if (L / 4 * 3 >= conv && L / 4 >= 16) {
        L = L / 4 * 3;
        lg -= 2;
        r3 = 1;
    }
    *lg_out = lg;
    *r3_out = r3;
    return L;
}

/* big_mul_ntt's width/length choice (mul_ntt.c:103-118) and, past the
   longest transform, big_mul_ntt_split's width and piece length
   (mul_split.c:151-177). */
void ffo_plan_product(size_t bits_a, size_t bits_b, int square, ffo_plan *pl) {
    memset(pl, 0, sizeof *pl);
    size_t L = 0;
    for (unsigned w = 32; w <= 35; w++) {
        if (!ffo_width_ok(bits_a, bits_b, w)) break;
        size_t ca = (bits_a + w - 1) / w, cb = (bits_b + w - 1) / w;
        int lg_w, r3_w;
        size_t L_w = ffo_ntt_length(ca + cb - 1, &lg_w, &r3_w);
        if (L == 0 || L_w < L) {
            L = L_w;
            pl->L = L_w, pl->lg = lg_w, pl->r3 = r3_w, pl->bw = w, pl->conv = ca + cb - 1;
        }
    }
    size_t maxL = (size_t)1 << g_max_lg;
    if (L != 0 && L <= maxL) return;

    pl->split = 1;
    unsigned bw = 35;
    while (bw > 8 && !ffo_width_ok(bits_a, bits_b, bw)) bw--;
    int lgmax = g_max_lg, lgp = lgmax;
    double best = 0;
    for (int lg = lgmax; lg >= lgmax - 4 && lg >= 10; lg--) {
        size_t h = (size_t)1 << (lg - 1);
        size_t ka = ((bits_a + bw - 1) / bw + h - 1) / h, kb = ((bits_b + bw - 1) / bw + h - 1) / h;
        double ca = (double)ka, cb = (double)kb;
        double tr = square ? 3 * ca - 1 : 2 * (ca + cb) - 1, pairs = square ? ca * (ca + 1) / 2 : ca * cb;
        double cost = (tr * lg + 5 * pairs) * (double)((size_t)1 << lg);
        if (lg == lgmax || cost < best) best = cost, lgp = lg;
    }
    size_t H = (size_t)1 << (lgp - 1);
    size_t ca = (bits_a + bw - 1) / bw, cb = (bits_b + bw - 1) / bw;
    pl->bw = bw;
    pl->lgp = lgp;
    pl->ka = (int)((ca + H - 1) / H);
    pl->kb = square ? pl->ka : (int)((cb + H - 1) / H);
    pl->conv = ca + cb - 1;
    pl->L = (size_t)(pl->ka + pl->kb) * H; /* Lbig: the residue array length */
    pl->lg = lgp;
    pl->r3 = 0;
}

/* ------------------------------------------------------------ operands */

static size_t norm_len(const uint64_t *d, size_t n) {
    while (n > 1 && d[n - 1] == 0) n--;
    return n ? n : 1;
}

static size_t bit_len(const uint64_t *d, size_t n) {
    n = norm_len(d, n);
    if (d[n - 1] == 0) return 0;
    return 64 * (n - 1) + (64 - (size_t)__builtin_clzll(d[n - 1]));
}

/* Field k of bw bits (src_load8, ntt_kernels.c:40-74), zero past the limbs. */
uint64_t ffo_field(const uint64_t *d, size_t n, size_t k, unsigned bw) {
    size_t bit = k * bw, wi = bit >> 6;
    unsigned sh = bit & 63;
    if (wi >= n) return 0;
    uint64_t v = d[wi] >> sh;
    if (sh + bw > 64 && wi + 1 < n) v |= d[wi + 1] << (64 - sh);
    return v & ((1ULL << bw) - 1);
}

/* --------------------------------------------------------------- CRT */

/* Garner (crt.c:33-64) and packing at bit i*bw with one running carry
   (crt.c:66-154), on natural-order normalized residues. */
void ffo_crt_pack(const uint32_t *const y[3], size_t conv, unsigned bw, uint64_t *out, size_t nout) {
    const uint32_t p1 = ffo_prime[0], p2 = ffo_prime[1], p3 = ffo_prime[2];
    const uint32_t i12 = invm(p1 % p2, p2), i13 = invm(p1 % p3, p3), i23 = invm(p2 % p3, p3);
    const u128 P12 = (u128)p1 * p2;
    const uint64_t mask = (1ULL << bw) - 1;
    u128 buf = 0, c = 0;
    unsigned nbits = 0;
    size_t w = 0;
    for (size_t i = 0; i < conv; i++) {
        uint32_t r1 = y[0][i], r2 = y[1][i], r3 = y[2][i];
        uint32_t x1 = r1;
        uint32_t x2 = mulm(subm(r2, x1 % p2, p2), i12, p2);
        uint32_t t3 = mulm(subm(r3, x1 % p3, p3), i13, p3);
        uint32_t x3 = mulm(subm(t3, x2 % p3, p3), i23, p3);
        u128 X = (u128)x1 + (u128)x2 * p1 + (u128)x3 * P12;
        u128 acc = X + c;
        buf |= (u128)((uint64_t)acc & mask) << nbits;
        c = acc >> bw;
        nbits += bw;
        if (nbits >= 64) {
            if (w < nout) out[w] = (uint64_t)buf;
            w++;
            buf >>= 64;
            nbits -= 64;
        }
    }
    buf |= c << nbits;
    for (; w < nout; w++, buf >>= 64) out[w] = (uint64_t)buf;
}

/* ------------------------------------------------------------ multiply */

static uint64_t *from_u64(uint64_t v, size_t *n) {
    uint64_t *d = xalloc(8);
    d[0] = v;
    *n = 1;
    return d;
}

uint64_t *ffo_mul_school(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, size_t *nout) {
    uint64_t *d = xzalloc((na + nb) * 8);
    for (size_t i = 0; i < na; i++) {
        uint64_t c = 0;
        for (size_t j = 0; j < nb; j++) {
            u128 t = (u128)a[i] * b[j] + d[i + j] + c;
            d[i + j] = (uint64_t)t;
            c = (uint64_t)(t >> 64);
        }
        d[i + nb] = c;
    }
    *nout = norm_len(d, na + nb);
    return d;
}

/* One prime's residues of a whole split product: pieces of H = Lp/2
   coefficients, each transformed once; output piece m = sum over i+j = m of
   A_i B_j (squares: i <= j, the others doubled), inverse, added at m*H
   (mul_split.c:98-129). Natural order, normalized. */
static void split_prime(int pi, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, int square,
                        const ffo_plan *pl, size_t ca, size_t cb, uint32_t *Y) {
    const uint32_t p = ffo_prime[pi];
    const size_t Lp = (size_t)1 << pl->lgp, H = Lp / 2;
    const int ka = pl->ka, kb = pl->kb;
    const uint32_t w = root_of(pi, Lp), wi = invm(w, p), li = invm((uint32_t)(Lp % p), p);
    uint32_t *A = xalloc((size_t)ka * Lp * 4), *B = A;
    for (int i = 0; i < ka; i++) {
        uint32_t *f = A + (size_t)i * Lp;
        for (size_t t = 0; t < Lp; t++) {
            size_t k = (size_t)i * H + t;
            f[t] = t < H && k < ca ? (uint32_t)(ffo_field(a, na, k, pl->bw) % p) : 0;
        }
        dft_any(f, Lp, w, p);
    }
    if (!square) {
        B = xalloc((size_t)kb * Lp * 4);
        for (int j = 0; j < kb; j++) {
            uint32_t *f = B + (size_t)j * Lp;
            for (size_t t = 0; t < Lp; t++) {
                size_t k = (size_t)j * H + t;
                f[t] = t < H && k < cb ? (uint32_t)(ffo_field(b, nb, k, pl->bw) % p) : 0;
            }
            dft_any(f, Lp, w, p);
        }
    }
    uint32_t *C = xalloc(Lp * 4);
    for (int m = 0; m < ka + kb - 1; m++) {
        int i0 = m - kb + 1 > 0 ? m - kb + 1 : 0, i1 = m < ka - 1 ? m : ka - 1;
        if (square && i1 > m / 2) i1 = m / 2;
        for (size_t x = 0; x < Lp; x++) {
            uint32_t acc = 0;
            for (int i = i0; i <= i1; i++) {
                int j = m - i;
                uint32_t t = mulm(A[(size_t)i * Lp + x], B[(size_t)j * Lp + x], p);
                acc = addm(acc, t, p);
                if (square && i < j) acc = addm(acc, t, p);
            }
            C[x] = acc;
        }
        dft_any(C, Lp, wi, p);
        for (size_t t = 0; t < Lp; t++) {
            size_t g = (size_t)m * H + t;
            Y[g] = addm(Y[g], mulm(C[t], li, p), p);
        }
    }
    free(C);
    if (B != A) free(B);
    free(A);
}

/* big_mul_ntt (mul_ntt.c:97-142) and its split branch (mul_split.c:149-187). */
uint64_t *ffo_mul_ntt(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, int square, size_t *nout) {
    if (square) b = a, nb = na;
    na = norm_len(a, na);
    nb = norm_len(b, nb);
    if ((na == 1 && a[0] == 0) || (nb == 1 && b[0] == 0)) return from_u64(0, nout);
    size_t bits_a = bit_len(a, na), bits_b = bit_len(b, nb);
    ffo_plan pl;
    ffo_plan_product(bits_a, bits_b, square, &pl);
    const size_t ca = (bits_a + pl.bw - 1) / pl.bw, cb = (bits_b + pl.bw - 1) / pl.bw;
    const size_t L = pl.L;
    uint32_t *y[3];
    for (int pi = 0; pi < 3; pi++) {
        const uint32_t p = ffo_prime[pi];
        y[pi] = xzalloc(L * 4);
        if (pl.split) {
            split_prime(pi, a, na, b, nb, square, &pl, ca, cb, y[pi]);
            continue;
        }
        uint32_t *fa = xalloc(L * 4), *fb = xalloc(L * 4);
        for (size_t k = 0; k < L; k++) {
            fa[k] = k < ca ? (uint32_t)(ffo_field(a, na, k, pl.bw) % p) : 0;
            fb[k] = k < cb ? (uint32_t)(ffo_field(b, nb, k, pl.bw) % p) : 0;
        }
        ffo_cyclic_conv(pi, fa, square ? fa : fb, L, y[pi]);
        free(fa);
        free(fb);
    }
    size_t limbs = na + nb + 1;
    uint64_t *d = xzalloc(limbs * 8);
    ffo_crt_pack((const uint32_t *const *)y, pl.conv, pl.bw, d, limbs);
    for (int i = 0; i < 3; i++) free(y[i]);
    *nout = norm_len(d, limbs);
    return d;
}

/* big_mul / big_sqr thresholds (mul_basic.c:64-77,109-112): the NTT from
   NTT_THRESHOLD_LIMBS = 200 limbs (the oracle's tables are always full, so
   ntt_fits holds); below it any exact method gives the same product. */
uint64_t *ffo_mul(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, int square, size_t *nout) {
    if (square) b = a, nb = na;
    na = norm_len(a, na);
    nb = norm_len(b, nb);
    size_t minlen = na < nb ? na : nb;
    if (minlen >= 200) return ffo_mul_ntt(a, na, b, nb, square, nout);
    return ffo_mul_school(a, na, b, nb, nout);
}

/* ------------------------------------------------------------ F(n) */

/* fib_step_linear / step_proc on one chunk (fib.c:36-62, 81-109). */
void ffo_fib_step(const uint64_t *s1, size_t n1, const uint64_t *s2, size_t n2, int k_odd, int bit,
                  uint64_t **lo, size_t *nlo, uint64_t **hi, size_t *nhi) {
    size_t len = n1 + 1;
    uint64_t *A = xalloc(len * 8), *B = xalloc(len * 8), *C = xalloc(len * 8);
    uint64_t cA = 0, prev = 0;
    int64_t cB = k_odd ? -2 : 2, cC = 0;
    for (size_t i = 0; i < len; i++) {
        uint64_t x = i < n1 ? s1[i] : 0, y = i < n2 ? s2[i] : 0;
        uint64_t x4 = (x << 2) | (prev >> 62);
        prev = x;
        u128 av = (u128)x + y + cA;
        A[i] = (uint64_t)av;
        cA = (uint64_t)(av >> 64);
        __int128 bv = (__int128)x4 - y + cB;
        B[i] = (uint64_t)bv;
        cB = (int64_t)(bv >> 64);
        __int128 cv = (__int128)B[i] - A[i] + cC;
        C[i] = (uint64_t)cv;
        cC = (int64_t)(cv >> 64);
    }
    if (bit) {
        *lo = C, *nlo = norm_len(C, len);
        *hi = B, *nhi = norm_len(B, len);
        free(A);
    } else {
        *lo = A, *nlo = norm_len(A, len);
        *hi = C, *nhi = norm_len(C, len);
        free(B);
    }
}

/* last_factors (fib.c:123-161): X = 2u + v and, if Y, Y = 2u - v. */
static void last_factors(const uint64_t *u, size_t nu, const uint64_t *v, size_t nv, uint64_t **X, size_t *nX,
                         uint64_t **Y, size_t *nY) {
    size_t len = (nu > nv ? nu : nv) + 1;
    uint64_t *x = xalloc(len * 8), *y = Y ? xalloc(len * 8) : NULL;
    uint64_t cX = 0, prev = 0;
    int64_t cY = 0;
    for (size_t i = 0; i < len; i++) {
        uint64_t xu = i < nu ? u[i] : 0, yv = i < nv ? v[i] : 0;
        uint64_t x2 = (xu << 1) | (prev >> 63);
        prev = xu;
        u128 s = (u128)x2 + yv + cX;
        x[i] = (uint64_t)s;
        cX = (uint64_t)(s >> 64);
        if (y) {
            __int128 d = (__int128)x2 - yv + cY;
            y[i] = (uint64_t)d;
            cY = (int64_t)(d >> 64);
        }
    }
    *X = x, *nX = norm_len(x, len);
    if (Y) *Y = y, *nY = norm_len(y, len);
}

static void add_small(uint64_t **d, size_t *n, int64_t v) {
    if (v >= 0) {
        uint64_t c = (uint64_t)v;
        for (size_t i = 0; c && i < *n; i++) {
            (*d)[i] += c;
            c = (*d)[i] < c;
        }
        if (c) {
            *d = realloc(*d, (*n + 1) * 8);
            (*d)[(*n)++] = c;
        }
    } else {
        uint64_t c = (uint64_t)(-v);
        for (size_t i = 0; c && i < *n; i++) {
            uint64_t old = (*d)[i];
            (*d)[i] = old - c;
            c = old < c;
        }
        *n = norm_len(*d, *n);
    }
}

/* fibonacci (fib.c:211-261): pair (F(k-1), F(k)) from the top bit down,
   two squarings and one linear step per bit, one multiply for the last. */
uint64_t *ffo_fibonacci(uint64_t n, size_t *nout) {
    if (n == 0) return from_u64(0, nout);
    size_t nkm1, nk;
    uint64_t *fkm1 = from_u64(0, &nkm1), *fk = from_u64(1, &nk);
    int k_odd = 1, top = 63 - __builtin_clzll(n);
    for (int i = top - 1; i >= 1; i--) {
        size_t n1, n2;
        uint64_t *s1 = ffo_mul(fk, nk, fk, nk, 1, &n1);
        uint64_t *s2 = ffo_mul(fkm1, nkm1, fkm1, nkm1, 1, &n2);
        free(fk);
        free(fkm1);
        int bit = (int)((n >> i) & 1);
        ffo_fib_step(s1, n1, s2, n2, k_odd, bit, &fkm1, &nkm1, &fk, &nk);
        free(s1);
        free(s2);
        k_odd = bit;
    }
    if (top == 0) {
        free(fkm1);
        *nout = nk;
        return fk;
    }
    uint64_t *r;
    size_t nr;
    if (n & 1) {
        uint64_t *x, *y;
        size_t nx, ny;
        last_factors(fk, nk, fkm1, nkm1, &x, &nx, &y, &ny);
        r = ffo_mul(x, nx, y, ny, 0, &nr);
        free(x);
        free(y);
        add_small(&r, &nr, k_odd ? -2 : 2);
    } else {
        uint64_t *x;
        size_t nx;
        last_factors(fkm1, nkm1, fk, nk, &x, &nx, NULL, NULL);
        r = ffo_mul(fk, nk, x, nx, 0, &nr);
        free(x);
    }
    free(fk);
    free(fkm1);
    *nout = norm_len(r, nr);
    return r;
}

/* ------------------------------------------------------------ decimal */

typedef struct {
    uint64_t *d;
    size_t n;
} Num;

static Num num_u64(uint64_t v) { Num r; r.d = from_u64(v, &r.n); return r; }
static int num_zero(const Num *a) { return a->n == 1 && a->d[0] == 0; }
static size_t num_bits(const Num *a) { return bit_len(a->d, a->n); }

static Num num_copy(const Num *a) {
    Num r = {xalloc(a->n * 8), a->n};
    memcpy(r.d, a->d, a->n * 8);
    return r;
}

static int num_cmp(const Num *a, const Num *b) {
    if (a->n != b->n) return a->n < b->n ? -1 : 1;
    for (size_t i = a->n; i-- > 0;)
        if (a->d[i] != b->d[i]) return a->d[i] < b->d[i] ? -1 : 1;
    return 0;
}

static Num num_add(const Num *a, const Num *b) {
    if (a->n < b->n) { const Num *t = a; a = b; b = t; }
    Num r = {xalloc((a->n + 1) * 8), a->n + 1};
    uint64_t c = 0;
    for (size_t i = 0; i < a->n; i++) {
        u128 s = (u128)a->d[i] + (i < b->n ? b->d[i] : 0) + c;
        r.d[i] = (uint64_t)s;
        c = (uint64_t)(s >> 64);
    }
    r.d[a->n] = c;
    r.n = norm_len(r.d, r.n);
    return r;
}

/* a - b, a >= b. */
static Num num_sub(const Num *a, const Num *b) {
    Num r = {xalloc(a->n * 8), a->n};
    uint64_t br = 0;
    for (size_t i = 0; i < a->n; i++) {
        uint64_t y = i < b->n ? b->d[i] : 0;
        u128 t = (u128)a->d[i] - y - br;
        r.d[i] = (uint64_t)t;
        br = (uint64_t)(t >> 64) ? 1 : 0;
    }
    r.n = norm_len(r.d, r.n);
    return r;
}

static Num num_shr(const Num *a, size_t k) {
    size_t s = k / 64, b = k % 64;
    if (s >= a->n) return num_u64(0);
    Num r = {xalloc((a->n - s) * 8), a->n - s};
    for (size_t i = 0; i < r.n; i++) {
        uint64_t lo = a->d[i + s], hi = i + s + 1 < a->n ? a->d[i + s + 1] : 0;
        r.d[i] = b ? (lo >> b) | (hi << (64 - b)) : lo;
    }
    r.n = norm_len(r.d, r.n);
    return r;
}

static Num num_shl(const Num *a, size_t k) {
    if (num_zero(a)) return num_u64(0);
    size_t s = k / 64, b = k % 64, n = a->n + s + 1;
    Num r = {xzalloc(n * 8), n};
    for (size_t i = 0; i < a->n; i++) {
        r.d[i + s] |= a->d[i] << b;
        if (b) r.d[i + s + 1] |= a->d[i] >> (64 - b);
    }
    r.n = norm_len(r.d, n);
    return r;
}

static Num num_pow2(size_t k) {
    Num r = {xzalloc((k / 64 + 1) * 8), k / 64 + 1};
    r.d[k / 64] = 1ULL << (k % 64);
    return r;
}

static Num num_mul(const Num *a, const Num *b) {
    Num r;
    r.d = ffo_mul(a->d, a->n, b->d, b->n, a == b, &r.n);
    return r;
}

static void num_addsmall(Num *a, uint64_t v, int sub) {
    add_small(&a->d, &a->n, sub ? -(int64_t)v : (int64_t)v);
}

/* big_recip (decimal.c:39-71). */
static Num recip(const Num *P, size_t s) {
    if (s <= 62) return num_u64((uint64_t)(((u128)1 << (2 * s)) / P->d[0]));
    size_t h = (s + 1) / 2 + 16;
    Num Ph = num_shr(P, s - h);
    Num mh = recip(&Ph, h);
    Num y = num_shl(&mh, s - h);
    free(Ph.d);
    free(mh.d);
    Num Z = num_pow2(2 * s);
    Num d = num_mul(P, &y);
    int below = num_cmp(&d, &Z) <= 0;
    Num e = below ? num_sub(&Z, &d) : num_sub(&d, &Z);
    Num ye = num_mul(&y, &e);
    Num corr = num_shr(&ye, 2 * s);
    Num y2 = below ? num_add(&y, &corr) : num_sub(&y, &corr);
    free(d.d), free(e.d), free(ye.d), free(corr.d), free(Z.d), free(y.d);
    num_addsmall(&y2, 2, 1);
    return y2;
}

uint64_t *ffo_recip(const uint64_t *P, size_t np, size_t s, size_t *nout) {
    Num p = {(uint64_t *)P, norm_len(P, np)};
    Num r = recip(&p, s);
    *nout = r.n;
    return r.d;
}

/* big_divmod_barrett (decimal.c:76-89). */
static void barrett(const Num *x, const Num *P, const Num *m, size_t s, Num *q, Num *r) {
    Num q1 = num_shr(x, s - 1);
    Num t = num_mul(&q1, m);
    *q = num_shr(&t, s + 1);
    Num qp = num_mul(q, P);
    *r = num_sub(x, &qp);
    free(q1.d), free(t.d), free(qp.d);
    while (num_cmp(r, P) >= 0) {
        Num t2 = num_sub(r, P);
        free(r->d);
        *r = t2;
        num_addsmall(q, 1, 0);
    }
}

/* big_divmod_any (decimal.c:95-125). */
static void divmod_any(const Num *x, const Num *P, size_t s, Num *q, Num *r) {
    size_t bx = num_zero(x) ? 0 : num_bits(x);
    if (bx < s) {
        *q = num_u64(0);
        *r = num_copy(x);
        return;
    }
    size_t t = bx - s + 1, k = t + 64 < s ? t + 64 : s;
    Num Pk = num_shr(P, s - k), Xk = num_shr(x, s - k);
    Num mk = recip(&Pk, k), rk;
    barrett(&Xk, &Pk, &mk, k, q, &rk);
    free(Pk.d), free(Xk.d), free(mk.d), free(rk.d);
    Num qp = num_mul(q, P);
    while (num_cmp(&qp, x) > 0) {
        num_addsmall(q, 1, 1);
        Num t2 = num_sub(&qp, P);
        free(qp.d);
        qp = t2;
    }
    *r = num_sub(x, &qp);
    free(qp.d);
    while (num_cmp(r, P) >= 0) {
        Num t2 = num_sub(r, P);
        free(r->d);
        *r = t2;
        num_addsmall(q, 1, 0);
    }
}

#define DEC_LEAF_LG 4                  /* decimal.c:24-27 */
#define DEC_P0 10000000000000000000ULL /* 10^19 */

/* dec_leaf (decimal.c:139-150): exactly 19*chunks digits by repeated
   division by 10^19. */
static void dec_leaf(const Num *x, size_t chunks, char *out) {
    uint64_t *t = xalloc(x->n * 8);
    memcpy(t, x->d, x->n * 8);
    size_t n = x->n;
    for (size_t c = chunks; c-- > 0;) {
        uint64_t rem = 0;
        for (size_t i = n; i-- > 0;) {
            u128 cur = ((u128)rem << 64) | t[i];
            t[i] = (uint64_t)(cur / DEC_P0);
            rem = (uint64_t)(cur % DEC_P0);
        }
        while (n > 1 && t[n - 1] == 0) n--;
        char *o = out + 19 * c;
        for (int j = 18; j >= 0; j--, rem /= 10) o[j] = (char)('0' + rem % 10);
    }
    free(t);
}

typedef struct {
    Num pw[48], inv[48];
    size_t bits[48];
} DecCtx;

/* dec_emit (decimal.c:192-203). */
static void dec_emit(const DecCtx *c, Num x, int lvl, char *out) {
    if (lvl <= DEC_LEAF_LG) {
        dec_leaf(&x, (size_t)2 << lvl, out);
        free(x.d);
        return;
    }
    Num q, r;
    barrett(&x, &c->pw[lvl], &c->inv[lvl], c->bits[lvl], &q, &r);
    free(x.d);
    dec_emit(c, q, lvl - 1, out);
    dec_emit(c, r, lvl - 1, out + ((size_t)19 << lvl));
}

/* big_to_string (decimal.c:229-281). */
char *ffo_to_string(const uint64_t *a, size_t na) {
    Num A = {(uint64_t *)a, norm_len(a, na)};
    if (num_zero(&A)) {
        char *s = xalloc(2);
        s[0] = '0', s[1] = 0;
        return s;
    }
    DecCtx *c = xzalloc(sizeof *c);
    size_t ba = num_bits(&A);
    int T = 0;
    c->pw[0] = num_u64(DEC_P0);
    c->bits[0] = num_bits(&c->pw[0]);
    while (2 * (c->bits[T] - 1) < ba) {
        c->pw[T + 1] = num_mul(&c->pw[T], &c->pw[T]);
        c->bits[T + 1] = num_bits(&c->pw[T + 1]);
        T++;
    }
    for (int i = DEC_LEAF_LG + 1; i < T; i++) c->inv[i] = recip(&c->pw[i], c->bits[i]);
    size_t half = (size_t)19 << T, total = 2 * half;
    char *buf = xalloc(total + 1);
    if (T <= DEC_LEAF_LG) {
        dec_leaf(&A, (size_t)2 << T, buf);
    } else {
        Num q, r;
        divmod_any(&A, &c->pw[T], c->bits[T], &q, &r);
        dec_emit(c, q, T - 1, buf);
        dec_emit(c, r, T - 1, buf + half);
    }
    for (int i = 0; i <= T; i++) free(c->pw[i].d);
    for (int i = DEC_LEAF_LG + 1; i < T; i++) free(c->inv[i].d);
    free(c);
    size_t z = 0;
    while (z + 1 < total && buf[z] == '0') z++;
    memmove(buf, buf + z, total - z);
    buf[total - z] = 0;
    return buf;
}

/* ------------------------------------------------------------ fingerprint */

/* tests/memrun.c:20-27 / tests/fingerprint.py:22. */
void ffo_fingerprint(const uint64_t *d, size_t n, uint64_t *bits, uint64_t *low64, uint64_t out[4]) {
    static const uint64_t ps[4] = {18446744073709551557ULL, 18446744073709551533ULL, 9223372036854775783ULL,
                                   4611686018427387847ULL};
    n = norm_len(d, n);
    *bits = bit_len(d, n);
    *low64 = d[0];
    for (int i = 0; i < 4; i++) {
        u128 m = 0;
        for (size_t k = n; k-- > 0;) m = ((m << 64) | d[k]) % ps[i];
        out[i] = (uint64_t)m;
    }
}

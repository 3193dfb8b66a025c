// Products past the longest transform (mul_split.c:1-187): every operand
// cut into pieces of H = Lp/2 coefficients, each piece transformed once,
// output piece m = sum over i + j = m of the pointwise products A_i B_j,
// one inverse per output piece, the pieces overlap-added (they overlap by
// H) into one residue array laid out like a single transform's output,
// which the ordinary CRT reads. The sums are the full product's
// coefficients, so the CRT bound is the one of a single big transform.
//
// On the GPU each stage is one launch for all pieces and primes: the
// forward passes and inner blocks of the ka + kb pieces as work lists, one
// kernel forming every output piece's pointwise sum, the inverse inner
// blocks and passes of the ka + kb - 1 output pieces as work lists, one
// overlap-add kernel, one CRT. Every piece stays in HBM (A and B take
// 12 Lp bytes per piece, the output pieces as much again).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../../include/fastfib_amd.h"
#include "ff_arith.h"
#include "ff_engine.h"
#include "ff_internal.h"
#include "ff_limbs.h"
#include "ff_ntt_dev.h"
#include "ff_runtime.h"

// C_m[x] = sum_{i + j = m} A_i[x] B_j[x] (mod p, Montgomery), every output
// piece m, prime and word x; squares (B == A) take each pair i < j twice
// (split_acc_range, mul_split.c:62-82).
struct SplitPrimes {
    FFPrime p[3];
};

__global__ void k_split_acc(const uint32_t *A, const uint32_t *B, uint32_t *C, int ka, int kb, uint64_t Lp,
                            SplitPrimes PR) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t per_m = 3 * Lp;
    const int nm = ka + kb - 1;
    if (t >= (uint64_t)nm * per_m) return;
    const int m = (int)(t / per_m);
    const uint64_t r = t - (uint64_t)m * per_m;
    const int pi = (int)(r / Lp);
    const uint64_t x = r - (uint64_t)pi * Lp;
    const FFPrime P = PR.p[pi];
    const bool sq = A == B;
    int i0 = m - kb + 1 > 0 ? m - kb + 1 : 0, i1 = m < ka - 1 ? m : ka - 1;
    if (sq && i1 > m / 2) i1 = m / 2;
    uint32_t acc = 0;
    for (int i = i0; i <= i1; i++) {
        const int j = m - i;
        const uint32_t v = ff_mont(A[((uint64_t)i * 3 + pi) * Lp + x], B[((uint64_t)j * 3 + pi) * Lp + x], P.p,
                                   P.pinv);
        acc = ff_add(acc, v, P.p);
        if (sq && i < j) acc = ff_add(acc, v, P.p);
    }
    C[((uint64_t)m * 3 + pi) * Lp + x] = acc;
}

// Y[g] = C_m[t] + C_{m-1}[t + H] (mod p) for g = m H + t, every prime; the
// output pieces' coefficients in natural order (coefficient t at word t).
// Pieces of at most 4096 coefficients have no outer pass to read the limbs:
// coefficient x of piece k into its array (all primes).
struct SplitSrc {
    const uint32_t *src;
    uint32_t nw;
    uint32_t *data;
};

__global__ void k_split_unpack(const SplitSrc *pieces, uint64_t Lp, int bw, SplitPrimes PR, uint64_t total) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const uint64_t per = 3 * Lp;
    const uint64_t k = t / per, r = t - k * per;
    const int pi = (int)(r / Lp);
    const uint64_t x = r - (uint64_t)pi * Lp;
    const SplitSrc s = pieces[k];
    s.data[(uint64_t)pi * Lp + x] = ff_load_coef(s.src, s.nw, x, bw, PR.p[pi]);
}

// sc: per-prime scale R^2 / Lp to apply here (pieces without an outer
// pass), or zeros when the inverse top pass applied it.
__global__ void k_split_add(const uint32_t *C, int nm, uint64_t Lp, uint64_t Lbig, uint32_t *Y, SplitPrimes PR,
                            uint32_t sc0, uint32_t sc1, uint32_t sc2) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * Lbig) return;
    const int pi = (int)(t / Lbig);
    const uint64_t g = t - (uint64_t)pi * Lbig;
    const uint64_t H = Lp / 2;
    const int m = (int)(g / H);
    const uint64_t u = g - (uint64_t)m * H;
    const uint32_t p = PR.p[pi].p;
    uint32_t v = 0;
    if (m < nm) v = C[((uint64_t)m * 3 + pi) * Lp + u];
    if (m >= 1 && m - 1 < nm) v = ff_add(v, C[((uint64_t)(m - 1) * 3 + pi) * Lp + u + H], p);
    const uint32_t sc = pi == 0 ? sc0 : (pi == 1 ? sc1 : sc2);
    if (sc) v = ff_mont(v, sc, p, PR.p[pi].pinv);
    Y[(uint64_t)pi * Lbig + g] = v;
}

namespace {

// Uploads a work list (items then start offsets) into one device block.
template <class Item>
void *upload_list(const std::vector<Item> &items, const std::vector<uint32_t> &start, const Item **di,
                  const uint32_t **ds) {
    const size_t bi = (items.size() * sizeof(Item) + 255) / 256 * 256;
    char *d = (char *)ff_dalloc(bi + start.size() * 4 + 256);
    FF_CHECK(hipStreamSynchronize(g_st));
    ff_copy(d, items.data(), items.size() * sizeof(Item), hipMemcpyHostToDevice);
    ff_copy(d + bi, start.data(), start.size() * 4, hipMemcpyHostToDevice);
    *di = (const Item *)d;
    *ds = (const uint32_t *)(d + bi);
    return d;
}

}  // namespace

// The piece transform length by the reference's cost model
// (mul_split.c:162-174): lg in [lgmax - 4, lgmax], lg >= 10, minimizing
// (transforms * lg + 5 * pairs) * 2^lg.
static int split_lg(uint64_t bits_a, uint64_t bits_b, unsigned bw, bool sq) {
    const int lgmax = g_max_lg;
    int lgp = lgmax;
    double best = 0;
    for (int lg = lgmax; lg >= lgmax - 4 && lg >= 10; lg--) {
        const uint64_t h = 1ull << (lg - 1);
        const double ca = (double)(((bits_a + bw - 1) / bw + h - 1) / h);
        const double cb = (double)(((bits_b + bw - 1) / bw + h - 1) / h);
        const double tr = sq ? 3 * ca - 1 : 2 * (ca + cb) - 1, pairs = sq ? ca * (ca + 1) / 2 : ca * cb;
        const double cost = (tr * lg + 5 * pairs) * (double)(1ull << lg);
        if (lg == lgmax || cost < best) best = cost, lgp = lg;
    }
    return lgp;
}

void ff_split_plan(uint64_t bits_a, uint64_t bits_b, bool sq, unsigned *bw_out, int *lgp_out, int force_lg) {
    // the widest width the CRT bound allows (mul_split.c:152-154)
    unsigned bw = 35;
    while (bw > 8 && !ff_width_ok(bits_a, bits_b, bw)) bw--;
    *bw_out = bw;
    const int lo = g_max_lg - 4 > 10 ? g_max_lg - 4 : 10;
    if (force_lg > 0) *lgp_out = force_lg < lo ? lo : (force_lg > g_max_lg ? g_max_lg : force_lg);
    else *lgp_out = split_lg(bits_a, bits_b, bw, sq);
}

void ff_mul_split(const MulItem &it, int force_lg) {
    const bool sq = it.b == nullptr;
    const uint64_t bits_b = sq ? it.bits_a : it.bits_b;
    unsigned bw;
    int lgp;
    ff_split_plan(it.bits_a, bits_b, sq, &bw, &lgp, force_lg);
    const uint64_t Lp = 1ull << lgp, H = Lp / 2;
    const uint64_t ca = (it.bits_a + bw - 1) / bw, cb = (bits_b + bw - 1) / bw;
    const int ka = (int)((ca + H - 1) / H), kb = sq ? ka : (int)((cb + H - 1) / H);
    const uint64_t conv = ca + cb - 1, Lbig = (uint64_t)(ka + kb) * H;
    const int nm = ka + kb - 1;
    SplitPrimes PR;
    ff_prime_consts(PR.p);

    FFPlan pl;
    if (ff_plan_build(&pl, (uint32_t)Lp, (int)bw, lgp, 0, 2 * H - 1, g_inner_log, g_tile_words) != 0) {
        fprintf(stderr, "fastfib_amd: no split schedule for 2^%d\n", lgp);
        exit(1);
    }
    const size_t piece = 3 * Lp;  // words of one piece, all primes
    std::vector<void *> mem;
    auto alloc = [&](size_t bytes) {
        void *p = ff_dalloc(bytes);
        mem.push_back(p);
        return p;
    };
    uint32_t *A = (uint32_t *)alloc((size_t)ka * piece * 4);
    uint32_t *B = sq ? A : (uint32_t *)alloc((size_t)kb * piece * 4);
    uint32_t *C = (uint32_t *)alloc((size_t)nm * piece * 4);
    uint32_t *Y = (uint32_t *)alloc(3 * Lbig * 4);

    // forward: every piece of a (and b), coefficients [i H, (i + 1) H) read
    // in place from the limbs (split_src, mul_split.c:130-143)
    struct Piece {
        const uint32_t *src;
        uint32_t nw;
        uint32_t *data;
    };
    std::vector<Piece> pieces;
    auto add_pieces = [&](const uint64_t *d, uint64_t n, uint64_t nc, int k, uint32_t *dst) {
        for (int i = 0; i < k; i++) {
            const uint64_t limb0 = (uint64_t)i * H * bw / 64;  // H is a multiple of 64
            const uint64_t left = nc - (uint64_t)i * H, cnt = left < H ? left : H;
            uint64_t words = (cnt * bw + 31) / 32;
            const uint64_t avail = 2 * (n - limb0);
            if (words > avail) words = avail;
            pieces.push_back(Piece{(const uint32_t *)(d + limb0), (uint32_t)words, dst + (size_t)i * piece});
        }
    };
    add_pieces(it.a, it.na, ca, ka, A);
    if (!sq) add_pieces(it.b, it.nb, cb, kb, B);

    std::vector<void *> lists;
    auto outer = [&](int pass, bool inv, const std::vector<uint32_t *> &arrays, const std::vector<Piece> *src) {
        const FFPass &ps = pl.pass[pass];
        std::vector<FFOuterItem> its;
        std::vector<uint32_t> start;
        uint32_t blocks = 0;
        for (size_t k = 0; k < arrays.size(); k++) {
            FFOuterItem oi;
            oi.ps = ps;
            oi.x = FFXform{nullptr, 0, arrays[k]};
            if (src) oi.x.src = (*src)[k].src, oi.x.nw = (*src)[k].nw;
            its.push_back(oi);
            start.push_back(blocks);
            blocks += 3 * ps.nblocks * ps.tpb;
        }
        const FFOuterItem *di;
        const uint32_t *ds;
        lists.push_back(upload_list(its, start, &di, &ds));
        FF_CHECK(ff_launch_outer_list(inv, pass == 0, false, ff_outer_shape(&ps, pass == 0), di, ds, (int)its.size(), blocks,
                                      ff_outer_lds_bytes(&ps), 3.0 * 8.0 * Lp * its.size(), g_st));
    };
    auto inner = [&](bool inv, const std::vector<uint32_t *> &arrays) {
        std::vector<FFInnerItem> its;
        std::vector<uint32_t> start;
        uint32_t blocks = 0;
        for (uint32_t *a : arrays) {
            FFInnerItem ii;
            ii.in = pl.inner;
            ii.c = FFConv{nullptr, nullptr, 0, 0, a, nullptr};
            its.push_back(ii);
            start.push_back(blocks);
            blocks += 3 * (uint32_t)(Lp / pl.inner.S);
        }
        const FFInnerItem *di;
        const uint32_t *ds;
        lists.push_back(upload_list(its, start, &di, &ds));
        FF_CHECK(ff_launch_inner_part_list(inv, di, ds, (int)its.size(), blocks, ff_inner_lds_bytes(&pl.inner),
                                           3.0 * 8.0 * Lp * its.size(), g_st));
    };

    std::vector<uint32_t *> fwd;
    for (const Piece &p : pieces) fwd.push_back(p.data);
    if (pl.npass == 0) {
        std::vector<SplitSrc> ss;
        for (const Piece &p : pieces) ss.push_back(SplitSrc{p.src, p.nw, p.data});
        const SplitSrc *dss;
        const uint32_t *unused;
        std::vector<uint32_t> none(1, 0);
        lists.push_back(upload_list(ss, none, &dss, &unused));
        const uint64_t total = (uint64_t)ss.size() * piece;
        hipLaunchKernelGGL(k_split_unpack, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, g_st, dss, Lp,
                           (int)bw, PR, total);
        FF_CHECK(hipGetLastError());
    }
    for (int pass = 0; pass < pl.npass; pass++) outer(pass, false, fwd, pass == 0 ? &pieces : nullptr);
    inner(false, fwd);

    // pointwise sums of every output piece
    {
        const uint64_t total = (uint64_t)nm * piece;
        ff_prof_begin(FK_INNER, g_st);
        hipLaunchKernelGGL(k_split_acc, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, g_st, A, B, C, ka, kb,
                           Lp, PR);
        ff_prof_end(FK_INNER, 4.0 * (double)total * 3, g_st);
        FF_CHECK(hipGetLastError());
    }

    // inverse of every output piece (the top pass applies R^2 / Lp)
    std::vector<uint32_t *> outs;
    for (int m = 0; m < nm; m++) outs.push_back(C + (size_t)m * piece);
    inner(true, outs);
    for (int pass = pl.npass - 1; pass >= 0; pass--) outer(pass, true, outs, nullptr);

    // overlap-add into the product's residues
    {
        const uint64_t total = 3 * Lbig;
        const bool scale = pl.npass == 0;  // no inverse top pass applied R^2 / Lp
        hipLaunchKernelGGL(k_split_add, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, g_st, C, nm, Lp, Lbig,
                           Y, PR, scale ? pl.inner.sc[0] : 0u, scale ? pl.inner.sc[1] : 0u,
                           scale ? pl.inner.sc[2] : 0u);
        FF_CHECK(hipGetLastError());
    }

    // the CRT over the whole product
    const uint64_t nout = it.na + (sq ? it.na : it.nb) + 1;
    FFCrtSet cs{};
    ff_lb_begin();
    cs.j[0] = FFCrtJob{Y, it.out, ff_lb_take(ff_crt_lb_words(nout)), it.lenkey, nout, conv, (uint32_t)Lbig, (int)bw};
    FF_CHECK(ff_launch_crt(&cs, 1, nout, (double)conv, (double)nout, g_st));
    for (void *p : lists) ff_dfree(p);
    for (void *p : mem) ff_dfree(p);
}

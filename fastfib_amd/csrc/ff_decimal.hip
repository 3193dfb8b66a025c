// Binary to decimal on the GPU (the reference's decimal.c, decimal.c:1-281).
//
// Same decomposition as the reference: P_i = 10^(19 * 2^i) by squaring,
// a number below P_i^2 split as q * P_i + r by Barrett division with a
// Newton reciprocal of P_i, q and r each becoming 19 * 2^i digits, and
// pieces below P_5 converted by repeated division by 10^19. The GPU form:
//
//  * the powers and reciprocals depend only on the tree height T and are
//    kept in HBM across calls (the reciprocals made exact, floor(2^2s / P),
//    so Barrett's quotient is at most 2 low);
//  * every level of the tree is one batch: its 2^(T-lvl) nodes sit in
//    fixed-size slots and each step (shift, product, subtraction, the two
//    corrections) is one launch over all of them; the levels below the top
//    and the leaves are captured once per T in a HIP graph;
//  * the two corrections are taken at once: d1 = r - P and d2 = r - 2P are
//    formed side by side and the carry-outs pick r, d1 or d2;
//  * a leaf is one thread doing 2-by-1 divisions by 10^19 with a
//    precomputed reciprocal (Moller and Granlund, "Improved division by
//    invariant integers", IEEE TC 2011, algorithm 4); a leaf that does not
//    reach zero flags an error, which certifies every level above it: each
//    node holds q * P + r exactly, so the digits are right iff every leaf
//    is below 10^608.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <tuple>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/fastfib_amd.h"
#include "ff_bigops.h"
#include "ff_engine.h"
#include "ff_internal.h"
#include "ff_limbs.h"
#include "ff_runtime.h"

#define DEC_LEAF_LG 4  // leaves: below P_5 = 10^608, at most 32 limbs
#define DEC_P0 10000000000000000000ull
#define LEAF_MAXN 33
#define LEAF_T 64

// ------------------------------------------------------------------ leaves

// (u1 : u0) / 10^19 for u1 < 10^19; 10^19 > 2^63 needs no normalization.
__device__ __forceinline__ uint64_t div_p0(uint64_t u1, uint64_t u0, uint64_t *r) {
    const uint64_t d = DEC_P0, v = 0xd83c94fb6d2ac34aull;  // floor((2^128 - 1) / d) - 2^64
    uint64_t q0 = v * u1, q1 = __umul64hi(v, u1);
    const uint64_t s0 = q0 + u0;
    q1 += u1 + (s0 < q0 ? 1 : 0);
    q0 = s0;
    q1 += 1;
    uint64_t rr = u0 - q1 * d;
    if (rr > q0) {
        q1 -= 1;
        rr += d;
    }
    if (rr >= d) {
        q1 += 1;
        rr -= d;
    }
    *r = rr;
    return q1;
}

// Leaf j: x_j (nx limbs at x + j * sx) as chunks * 19 digits at
// out + j * chunks * 19, most significant first (dec_leaf, decimal.c:139-150).
// Books the index of the first nonzero digit (atomicMin) and flags a value
// that does not fit.
__global__ __launch_bounds__(LEAF_T) void k_dec_leaf(const uint64_t *x, uint64_t sx, uint32_t nx, uint32_t N,
                                                     uint32_t chunks, char *out, unsigned long long *first_nz,
                                                     uint32_t *err) {
    __shared__ uint64_t t[LEAF_MAXN][LEAF_T];
    const uint32_t j = blockIdx.x * LEAF_T + threadIdx.x;
    if (j >= N) return;
    const int tid = threadIdx.x;
    int n = 0;
    for (uint32_t i = 0; i < nx; i++) {
        const uint64_t v = x[(uint64_t)j * sx + i];
        t[i][tid] = v;
        if (v) n = (int)i + 1;
    }
    char *o = out + (uint64_t)j * chunks * 19;
    long long first = -1;
    for (int c = (int)chunks - 1; c >= 0; c--) {
        uint64_t rem = 0;
        for (int i = n - 1; i >= 0; i--) t[i][tid] = div_p0(rem, t[i][tid], &rem);
        while (n > 0 && t[n - 1][tid] == 0) n--;
        char dig[19];
        uint64_t v = rem;
        for (int k = 18; k >= 0; k--) {
            const uint64_t q = __umul64hi(v, 0xCCCCCCCCCCCCCCCDull) >> 3;  // v / 10
            dig[k] = (char)('0' + (v - 10 * q));
            v = q;
        }
        for (int k = 0; k < 19; k++) o[19 * c + k] = dig[k];
        if (rem) {
            int lead = 0;
            while (dig[lead] == '0') lead++;
            first = ((long long)j * chunks + c) * 19 + lead;
        }
    }
    if (first >= 0) atomicMin(first_nz, (unsigned long long)first);
    if (n != 0) atomicOr(err, 1u);
}

// ------------------------------------------------------------------ context

namespace {

struct DecCtx {
    std::vector<FFNum> P;   // P_0 .. P_T
    std::vector<FFNum> m;   // m_i = floor(2^(2 s_i) / P_i), i in (DEC_LEAF_LG, T]; empty below
    std::vector<uint64_t> s;
    std::vector<uint64_t *> Y;  // the guarded reciprocals of the chain below (recip_chain)
};

// Levels below the top for one T: fixed slots, batches, one graph.
struct DecLevel {
    int lvl;
    uint32_t N;
    uint64_t s, X, Q1, T1, Qn, QP, R, p;
    const uint64_t *in;
    uint64_t *out;
    ff_batch *mul1 = nullptr, *mul2 = nullptr;
};

struct DecPlan {
    int T = 0;
    uint64_t *top_out = nullptr;  // the top split's q and r (2 slots of R_T)
    uint64_t R_top = 0;
    std::vector<DecLevel> lv;
    // temporaries shared by the levels (stream order keeps them apart)
    uint64_t *q1 = nullptr, *t = nullptr, *q = nullptr, *qp = nullptr, *r = nullptr, *d1 = nullptr, *d2 = nullptr;
    uint8_t *c = nullptr;  // c1 | c2, 2 * max nodes
    void *scr = nullptr;
    std::vector<void *> bufs;
    char *digits = nullptr;
    uint64_t ndigits = 0;
    unsigned long long *first_nz = nullptr;
    uint32_t *err = nullptr;
    const uint64_t *leaf_in = nullptr;
    uint32_t leaf_N = 0;
    uint64_t leaf_slot = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

DecCtx g_dec;
// plans per tree height and schoolbook bound (the bound decides, at plan
// time, which level products skip the transforms)
std::map<std::tuple<int, int, int>, std::unique_ptr<DecPlan>> g_plans;  // by height, schoolbook bound, queue form

uint64_t limbs_of(uint64_t bits) { return bits ? (bits + 63) / 64 : 1; }

uint64_t host_word(const uint64_t *d) {
    uint64_t *h = ff_pinned_results(1);
    FF_CHECK(hipMemcpyAsync(h, d, 8, hipMemcpyDeviceToHost, g_st));
    FF_CHECK(hipStreamSynchronize(g_st));
    return h[0];
}

// big_recip (decimal.c:39-71): floor(2^(2s) / P) minus at most 4, for P of
// exactly s bits; the same steps as the reference, on device numbers.
FFNum recip(const FFNum &P, uint64_t s) {
    if (s <= 62) {
        const unsigned __int128 num = (unsigned __int128)1 << (2 * s);
        return ff_num_u64((uint64_t)(num / host_word(P.d)));
    }
    const uint64_t h = (s + 1) / 2 + 16;
    FFNum Ph = ff_num_shr(P, s - h);
    FFNum mh = recip(Ph, h);
    ff_num_free(Ph);
    FFNum y = ff_num_shl(mh, s - h);
    ff_num_free(mh);
    FFNum Z = ff_num_pow2(2 * s);
    FFNum d = ff_num_mul(P, &y);
    const bool below = ff_num_cmp(d, Z) <= 0;
    FFNum e = below ? ff_num_sub(Z, d) : ff_num_sub(d, Z);
    ff_num_free(d);
    FFNum ye = ff_num_mul(y, &e);
    ff_num_free(e);
    FFNum corr = ff_num_shr(ye, 2 * s);
    ff_num_free(ye);
    FFNum y2 = below ? ff_num_add(y, corr) : ff_num_sub(y, corr);
    ff_num_free(y), ff_num_free(corr), ff_num_free(Z);
    FFNum y3 = ff_num_add_small(y2, -2);
    ff_num_free(y2);
    return y3;
}

// ------------------------------------------------------------------ reciprocal chain
// The decimal levels need the exact m_i = floor(2^(2 s_i) / P_i). They are
// formed bottom-up, every length known in advance, so the whole chain is
// queued without waiting for the device (big_recip's recursion, kept above
// for ff_big_recip, reads lengths and signs back at every step). With
// G guard bits, Y_i approximates T_i* = 2^(2 s_i + G) / P_i from below:
//
//   Y_0 = floor(2^(2 s_0 + G) / P_0) exactly (P_0 = 10^19 < 2^64: a bitwise
//   long division on the host);
//   y0 = Y_{i-1}^2 >> k, k = 4 s' + G - 2 s (P_i = P_{i-1}^2, s' = s_{i-1},
//   s = s_i): with Y_{i-1} = T*' - d, d < 2, y0 = T* - E where
//   0 <= E <= d 2^(2s - 3s' + 2) + 1 <= 2^(s' + 3) + 1 (2s <= 4s');
//   one Newton step y1 = y0 + floor(y0 e / 2^(2s + G)), e = 2^(2s + G) - P y0
//   = P E >= 0: T* - y1 = E^2 / T* + (floor) < 2^(9 - G) + 1, with
//   T* >= 2^(s + G), so Y_i = y1 is T_i or T_i - 1 (T_i = floor(T*));
//   m_c = Y_i >> G is m_i or m_i - 1, and r = 2^(2s) - P m_c in [0, 2P)
//   decides: m_i = m_c + [r >= P].
// e < 2^(s + s' + 4) and r < 2^(s + 1) are formed as negations of P y0 and
// P m_c modulo 2^(64 n) for n limbs below 2s + G and 2s bits respectively
// (so the power of two vanishes). Per level: one square, three products,
// two shifts, two negations, an addition, a comparison and an increment.

constexpr uint64_t RC_G = 16;  // guard bits

// out = a * b (b null: a^2) with operand bit bounds; lengths need not be
// tight (leading zero limbs only make the transform what the bound says).
void chain_mul(const uint64_t *a, uint64_t na, uint64_t bits_a, const uint64_t *b, uint64_t nb, uint64_t bits_b,
               uint64_t *out, unsigned long long *key) {
    const uint64_t bb = b ? bits_b : bits_a;
    const ProdShape sh = ff_product_shape(bits_a, bb);
    MulItem it{a, na, bits_a, b, b ? nb : na, bb, out, sh.conv, key};
    if (sh.split) ff_mul_split(it);
    else ff_mul_group(&it, 1, sh, FFSched{0, 0, 0});
}

// out (n limbs) = -b mod 2^(64 n), or a - b when a is given (na limbs);
// cout, if given, gets the carry-out (1: no borrow).
void chain_sub(const uint64_t *a, uint64_t na, const uint64_t *b, uint64_t nb, uint64_t *out, uint64_t n,
               uint8_t *cout, void *scr) {
    FFAddNodes J{};
    J.a = a, J.na = a ? (na < n ? na : n) : 0;
    J.b = b, J.nb = nb < n ? nb : n;
    J.b_mode = FF_SUB;
    J.out = out, J.n = n, J.N = 1;
    J.cout = cout;
    FF_CHECK(ff_add_nodes(J, scr, g_st));
}

void recip_chain(int T) {
    DecCtx &D = g_dec;
    if ((int)D.m.size() <= T) D.m.resize(T + 1, FFNum{nullptr, 0, 0, 0});
    if (D.Y.empty()) {
        // Y_0 = floor(2^(2 s_0 + G) / P_0) by restoring division
        const uint64_t P0 = DEC_P0;
        const int nbits = (int)(2 * D.s[0] + RC_G);
        unsigned __int128 q = 0, r = 1;  // the numerator's leading 1
        for (int b = 0; b < nbits; b++) {
            r <<= 1;
            q <<= 1;
            if (r >= P0) r -= P0, q |= 1;
        }
        const uint64_t h[2] = {(uint64_t)q, (uint64_t)(q >> 64)};
        uint64_t *y = (uint64_t *)ff_dalloc(16);
        ff_copy(y, h, 16, hipMemcpyHostToDevice);
        D.Y.push_back(y);
    }
    if ((int)D.Y.size() > T) return;
    unsigned long long *key = (unsigned long long *)ff_dalloc(8);
    uint8_t *c = (uint8_t *)ff_dalloc(16);
    FF_CHECK(ff_launch_fill((uint64_t *)c, 2, 0, g_st));  // c[8..15] stay zero: the increment's second carry
    for (int i = (int)D.Y.size(); i <= T; i++) {
        const uint64_t s1 = D.s[i - 1], s = D.s[i];
        const FFNum &P = D.P[i];
        const uint64_t bY = s1 + RC_G + 1, nYp = limbs_of(bY);
        const uint64_t by = s + RC_G + 1, ny = limbs_of(by);
        const uint64_t be = s + s1 + 7, ne = limbs_of(be);
        const uint64_t nm = limbs_of(s + 1), nr = limbs_of(s + 2);
        const uint64_t k = 4 * s1 + RC_G - 2 * s;
        const uint64_t nsq = 2 * nYp + 1, nd = P.n + ny + 1, nt = ny + ne + 1, npm = P.n + nm + 1;
        uint64_t *sq = (uint64_t *)ff_dalloc(nsq * 8), *y0 = (uint64_t *)ff_dalloc(ny * 8);
        uint64_t *d = (uint64_t *)ff_dalloc(nd * 8), *e = (uint64_t *)ff_dalloc(ne * 8);
        uint64_t *t = (uint64_t *)ff_dalloc(nt * 8), *corr = (uint64_t *)ff_dalloc(ny * 8);
        uint64_t *y1 = (uint64_t *)ff_dalloc(ny * 8), *mc = (uint64_t *)ff_dalloc(nm * 8);
        uint64_t *pm = (uint64_t *)ff_dalloc(npm * 8), *r = (uint64_t *)ff_dalloc(nr * 8);
        uint64_t *rp = (uint64_t *)ff_dalloc(nr * 8);
        void *scr = ff_dalloc(ff_add_nodes_scratch(1, std::max(ny, std::max(ne, nr))) + 256);
        chain_mul(D.Y[i - 1], nYp, bY, nullptr, 0, 0, sq, key);
        FF_CHECK(ff_shr_nodes(sq, 0, nsq, k, 1, y0, 0, ny, g_st));
        chain_mul(P.d, P.n, s, y0, ny, by, d, key);
        chain_sub(nullptr, 0, d, nd, e, ne, nullptr, scr);
        chain_mul(y0, ny, by, e, ne, be, t, key);
        FF_CHECK(ff_shr_nodes(t, 0, nt, 2 * s + RC_G, 1, corr, 0, ny, g_st));
        {
            FFAddNodes J{};
            J.a = y0, J.na = ny, J.b = corr, J.nb = ny, J.b_mode = FF_ADD;
            J.out = y1, J.n = ny, J.N = 1;
            FF_CHECK(ff_add_nodes(J, scr, g_st));
        }
        FF_CHECK(ff_shr_nodes(y1, 0, ny, RC_G, 1, mc, 0, nm, g_st));
        chain_mul(P.d, P.n, s, mc, nm, s + 1, pm, key);
        chain_sub(nullptr, 0, pm, npm, r, nr, nullptr, scr);
        chain_sub(r, nr, P.d, P.n, rp, nr, c, scr);  // c[0] = [r >= P]
        FFNum m = ff_num_alloc(nm);
        FF_CHECK(ff_copy_add_small(mc, 0, nm, c, c + 8, 1, m.d, 0, nm, g_st));
        m.n = nm, m.bits = s + 1;  // 2^s < 2^(2s) / P < 2^(s+1): exactly s + 1 bits
        if (i > DEC_LEAF_LG) D.m[i] = m;
        else ff_num_free(m);
        D.Y.push_back(y1);
        for (void *q : {(void *)sq, (void *)y0, (void *)d, (void *)e, (void *)t, (void *)corr, (void *)mc,
                        (void *)pm, (void *)r, (void *)rp, scr})
            ff_dfree(q);
    }
    ff_dfree(key);
    ff_dfree(c);
}

// Powers up to P_T (by squaring, decimal.c:236-244).
void ensure_pow(int T) {
    if (g_dec.P.empty()) {
        g_dec.P.push_back(ff_num_u64(DEC_P0));
        g_dec.s.push_back(g_dec.P[0].bits);
    }
    while ((int)g_dec.P.size() <= T) {
        FFNum sq = ff_num_mul(g_dec.P.back(), nullptr);
        g_dec.s.push_back(sq.bits);
        g_dec.P.push_back(sq);
    }
}

// Exact reciprocals of P_5 .. P_T.
void ensure_recip(int T) { recip_chain(T); }

// The smallest T with 2 (bits(P_T) - 1) >= bits(a) (decimal.c:236-244),
// with the powers and reciprocals it needs.
int tree_height(uint64_t bits) {
    int T = 0;
    ensure_pow(0);
    while (2 * (g_dec.s[T] - 1) < bits) {
        T++;
        ensure_pow(T);
    }
    ensure_recip(T);
    return T;
}

void *palloc(DecPlan &pl, size_t bytes) {
    void *p = ff_dalloc(bytes ? bytes : 8);
    pl.bufs.push_back(p);
    return p;
}

// One Barrett level over N nodes in slots (the levels below the top).
void run_level(DecPlan &pl, const DecLevel &L) {
    const FFNum &P = g_dec.P[L.lvl];
    // q1 = x >> (s - 1)
    FF_CHECK(ff_shr_nodes(L.in, L.X, L.X, L.s - 1, L.N, pl.q1, L.Q1, L.Q1, g_st));
    FF_CHECK(ff_batch_run(L.mul1) == FF_OK ? hipSuccess : hipErrorUnknown);
    // q = t >> (s + 1)
    FF_CHECK(ff_shr_nodes(pl.t, L.T1, L.T1, L.s + 1, L.N, pl.q, L.Qn, L.Qn, g_st));
    FF_CHECK(ff_batch_run(L.mul2) == FF_OK ? hipSuccess : hipErrorUnknown);
    // r = x - q P (mod 2^(64 R)), then r - P and r - 2P side by side
    FFAddNodes A{};
    A.a = L.in, A.sa = L.X, A.na = L.X < L.R ? L.X : L.R;
    A.b = pl.qp, A.sb = L.QP, A.nb = L.QP < L.R ? L.QP : L.R;
    A.b_mode = FF_SUB;
    A.out = pl.r, A.so = L.R, A.n = L.R, A.N = L.N;
    FF_CHECK(ff_add_nodes(A, pl.scr, g_st));
    FFAddNodes B1{};
    B1.a = pl.r, B1.sa = L.R, B1.na = L.R;
    B1.b = P.d, B1.sb = 0, B1.nb = P.n;
    B1.b_mode = FF_SUB;
    B1.out = pl.d1, B1.so = L.R, B1.n = L.R, B1.N = L.N;
    B1.cout = pl.c;
    FF_CHECK(ff_add_nodes(B1, pl.scr, g_st));
    FFAddNodes B2 = B1;
    B2.b_mode = FF_SUB2X;
    B2.out = pl.d2;
    B2.cout = pl.c + L.N;
    FF_CHECK(ff_add_nodes(B2, pl.scr, g_st));
    // children: q + c1 + c2 at 2j, the corrected r at 2j + 1
    FF_CHECK(ff_copy_add_small(pl.q, L.Qn, L.Qn, pl.c, pl.c + L.N, L.N, L.out, 2 * L.R, L.R, g_st));
    FF_CHECK(ff_select3(pl.r, pl.d1, pl.d2, pl.c, pl.c + L.N, L.R, L.R, L.N, L.out + L.R, 2 * L.R, g_st));
}

void enqueue_tail(DecPlan &pl) {
    for (const DecLevel &L : pl.lv) run_level(pl, L);
    FF_CHECK(ff_launch_fill((uint64_t *)pl.first_nz, 1, ~0ull, g_st));
    FF_CHECK(ff_launch_fill((uint64_t *)pl.err, 1, 0, g_st));
    ff_prof_begin(FK_DEC_LEAF, g_st);
    hipLaunchKernelGGL(k_dec_leaf, dim3((pl.leaf_N + LEAF_T - 1) / LEAF_T), dim3(LEAF_T), 0, g_st, pl.leaf_in,
                       pl.leaf_slot, (uint32_t)pl.leaf_slot, pl.leaf_N, 2u << DEC_LEAF_LG, pl.digits, pl.first_nz,
                       pl.err);
    ff_prof_end(FK_DEC_LEAF, 8.0 * pl.leaf_N * pl.leaf_slot + (double)pl.ndigits, g_st);
    FF_CHECK(hipGetLastError());
}

ff_batch *level_batch(const uint64_t *a, uint64_t sa, uint64_t na, uint64_t bits_a, const FFNum &b, uint64_t *out,
                      uint64_t so, uint32_t N) {
    std::vector<ff_prod> it(N);
    for (uint32_t j = 0; j < N; j++)
        it[j] = ff_prod{a + (uint64_t)j * sa, na, b.d, b.n, out + (uint64_t)j * so, 0, bits_a, b.bits};
    ff_batch *B = nullptr;
    std::vector<FFSched> sc(N, FFSched{0, 0, 0});
    if (ff_batch_create_sched(&B, it.data(), N, g_dec_queue ? FF_BATCH_QUEUE : FF_BATCH_GROUPED, sc.data()) != FF_OK) {
        fprintf(stderr, "fastfib_amd: decimal level batch: %s\n", ff_last_error());
        exit(1);
    }
    return B;
}

// The levels and leaves of a plan captured in a HIP graph, for the plan's
// second and later conversions (the first queues them directly: capture,
// instantiation and the graph's first upload cost more than one direct run).
void capture_plan(DecPlan &pl) {
    FF_CHECK(hipStreamSynchronize(g_st));
    ff_prof_pause(1);
    FF_CHECK(hipStreamBeginCapture(g_st, hipStreamCaptureModeRelaxed));
    enqueue_tail(pl);
    FF_CHECK(hipStreamEndCapture(g_st, &pl.graph));
    ff_prof_pause(0);
    FF_CHECK(hipGraphInstantiate(&pl.exec, pl.graph, nullptr, nullptr, 0));
}

// The plan below the top for height T > DEC_LEAF_LG.
DecPlan *get_plan(int T) {
    auto it = g_plans.find(std::make_tuple(T, g_school_limbs, g_dec_queue));
    if (it != g_plans.end()) return it->second.get();
    auto up = std::make_unique<DecPlan>();
    DecPlan &pl = *up;
    pl.T = T;
    pl.R_top = g_dec.P[T].n + 1;
    pl.top_out = (uint64_t *)palloc(pl, 2 * pl.R_top * 8);
    const uint64_t *in = pl.top_out;
    uint64_t X = pl.R_top;
    size_t mx_q1 = 0, mx_t = 0, mx_q = 0, mx_qp = 0, mx_r = 0, mx_scr = 0;
    uint32_t mx_N = 1;
    for (int lvl = T - 1; lvl > DEC_LEAF_LG; lvl--) {
        DecLevel L;
        L.lvl = lvl;
        L.N = 2u << (T - 1 - lvl);
        L.s = g_dec.s[lvl];
        L.p = g_dec.P[lvl].n;
        L.X = X;
        L.in = in;
        L.Q1 = limbs_of(L.s + 1);
        L.T1 = L.Q1 + g_dec.m[lvl].n + 1;
        L.Qn = limbs_of(L.s + 1);
        L.QP = L.Qn + L.p + 1;
        L.R = L.p + 1;
        L.out = (uint64_t *)palloc(pl, (size_t)2 * L.N * L.R * 8);
        mx_q1 = std::max(mx_q1, (size_t)L.N * L.Q1);
        mx_t = std::max(mx_t, (size_t)L.N * L.T1);
        mx_q = std::max(mx_q, (size_t)L.N * L.Qn);
        mx_qp = std::max(mx_qp, (size_t)L.N * L.QP);
        mx_r = std::max(mx_r, (size_t)L.N * L.R);
        mx_scr = std::max(mx_scr, ff_add_nodes_scratch(L.N, L.R));
        mx_N = std::max(mx_N, L.N);
        pl.lv.push_back(L);
        in = L.out;
        X = L.R;
    }
    pl.q1 = (uint64_t *)palloc(pl, mx_q1 * 8);
    pl.t = (uint64_t *)palloc(pl, mx_t * 8);
    pl.q = (uint64_t *)palloc(pl, mx_q * 8);
    pl.qp = (uint64_t *)palloc(pl, mx_qp * 8);
    pl.r = (uint64_t *)palloc(pl, mx_r * 8);
    pl.d1 = (uint64_t *)palloc(pl, mx_r * 8);
    pl.d2 = (uint64_t *)palloc(pl, mx_r * 8);
    pl.c = (uint8_t *)palloc(pl, 2 * (size_t)mx_N + 256);
    pl.scr = palloc(pl, std::max(mx_scr, ff_add_nodes_scratch(1, 2 * pl.R_top)));
    for (DecLevel &L : pl.lv) {
        L.mul1 = level_batch(pl.q1, L.Q1, L.Q1, L.s + 1, g_dec.m[L.lvl], pl.t, L.T1, L.N);
        L.mul2 = level_batch(pl.q, L.Qn, L.Qn, L.s + 1, g_dec.P[L.lvl], pl.qp, L.QP, L.N);
    }
    pl.leaf_in = in;
    pl.leaf_N = 1u << (T - DEC_LEAF_LG);
    pl.leaf_slot = X;
    if (X > LEAF_MAXN) {
        fprintf(stderr, "fastfib_amd: decimal leaves of %llu limbs\n", (unsigned long long)X);
        exit(1);
    }
    pl.ndigits = (uint64_t)pl.leaf_N * (2u << DEC_LEAF_LG) * 19;
    pl.digits = (char *)palloc(pl, pl.ndigits + 64);
    pl.first_nz = (unsigned long long *)palloc(pl, 8);
    pl.err = (uint32_t *)palloc(pl, 8);
    DecPlan *raw = up.get();
    g_plans[std::make_tuple(T, g_school_limbs, g_dec_queue)] = std::move(up);
    return raw;
}

// The top split A = q P_T + r into the plan's first two slots, sized from
// A's bit length (no waiting: every length is a bound known in advance).
void top_split(DecPlan &pl, const uint64_t *a, uint64_t na, uint64_t bits_a) {
    const int T = pl.T;
    const FFNum &P = g_dec.P[T], &m = g_dec.m[T];
    const uint64_t s = g_dec.s[T], R = pl.R_top;
    uint64_t *qs = pl.top_out, *rs = pl.top_out + R;
    if (bits_a < s) {
        FF_CHECK(ff_launch_fill(qs, R, 0, g_st));
        FF_CHECK(ff_launch_fill((uint64_t *)pl.c, 1, 0, g_st));
        FF_CHECK(ff_copy_add_small(a, 0, na < R ? na : R, pl.c, pl.c, 1, rs, 0, R, g_st));
        return;
    }
    std::vector<void *> tmp;
    auto alloc = [&](uint64_t limbs) {
        void *p = ff_dalloc(limbs * 8);
        tmp.push_back(p);
        return (uint64_t *)p;
    };
    unsigned long long *keys = (unsigned long long *)alloc(2);
    const uint64_t bq1 = bits_a - (s - 1), nq1 = limbs_of(bq1);
    uint64_t *q1 = alloc(nq1);
    FF_CHECK(ff_shr_nodes(a, 0, na, s - 1, 1, q1, 0, nq1, g_st));
    const uint64_t nt = nq1 + m.n + 1;
    uint64_t *t = alloc(nt);
    {
        MulItem it{q1, nq1, bq1, m.d, m.n, m.bits, t, 0, keys};
        ProdShape sh = ff_product_shape(bq1, m.bits);
        if (sh.split) {
            fprintf(stderr, "fastfib_amd: decimal top split needs the split multiply\n");
            exit(1);
        }
        it.conv = sh.conv;
        ff_mul_group(&it, 1, sh, FFSched{0, 0, 0});
    }
    // q < 2^(bq1 + bits(m) - s - 1 + 1)
    const uint64_t bq = bq1 + m.bits > s + 1 ? bq1 + m.bits - s : 1, nq = limbs_of(bq);
    uint64_t *q = alloc(nq);
    FF_CHECK(ff_shr_nodes(t, 0, nt, s + 1, 1, q, 0, nq, g_st));
    const uint64_t nqp = nq + P.n + 1;
    uint64_t *qp = alloc(nqp);
    {
        MulItem it{q, nq, bq, P.d, P.n, P.bits, qp, 0, keys + 1};
        ProdShape sh = ff_product_shape(bq, P.bits);
        if (sh.split) {
            fprintf(stderr, "fastfib_amd: decimal top split needs the split multiply\n");
            exit(1);
        }
        it.conv = sh.conv;
        ff_mul_group(&it, 1, sh, FFSched{0, 0, 0});
    }
    uint64_t *r = alloc(R), *d1 = alloc(R), *d2 = alloc(R);
    FFAddNodes A{};
    A.a = a, A.na = na < R ? na : R;
    A.b = qp, A.nb = nqp < R ? nqp : R;
    A.b_mode = FF_SUB;
    A.out = r, A.n = R, A.N = 1;
    FF_CHECK(ff_add_nodes(A, pl.scr, g_st));
    FFAddNodes B1{};
    B1.a = r, B1.na = R;
    B1.b = P.d, B1.nb = P.n;
    B1.b_mode = FF_SUB;
    B1.out = d1, B1.n = R, B1.N = 1;
    B1.cout = pl.c;
    FF_CHECK(ff_add_nodes(B1, pl.scr, g_st));
    FFAddNodes B2 = B1;
    B2.b_mode = FF_SUB2X;
    B2.out = d2;
    B2.cout = pl.c + 1;
    FF_CHECK(ff_add_nodes(B2, pl.scr, g_st));
    FF_CHECK(ff_copy_add_small(q, 0, nq < R ? nq : R, pl.c, pl.c + 1, 1, qs, 0, R, g_st));
    FF_CHECK(ff_select3(r, d1, d2, pl.c, pl.c + 1, 0, R, 1, rs, 0, g_st));
    for (void *p : tmp) ff_dfree(p);
}

// Small conversions (one leaf) keep their digits here.
char *g_small_dig = nullptr;
unsigned long long *g_small_first = nullptr;
uint32_t *g_small_err = nullptr;

// Runs the conversion; the digits stay on the device: *dev points at the
// first significant one, *len digits (decimal.c:229-281).
void decimal_dev(const uint64_t *a, uint64_t na, uint64_t bits, const char **dev, size_t *len) {
    const int T = tree_height(bits);
    const char *dig;
    uint64_t nd;
    unsigned long long *first_nz;
    uint32_t *err;
    if (T <= DEC_LEAF_LG) {
        if (!g_small_dig) {
            g_small_dig = (char *)ff_dalloc((2u << DEC_LEAF_LG) * 19 + 64);
            g_small_first = (unsigned long long *)ff_dalloc(8);
            g_small_err = (uint32_t *)ff_dalloc(8);
        }
        nd = (uint64_t)(2u << T) * 19;
        dig = g_small_dig, first_nz = g_small_first, err = g_small_err;
        FF_CHECK(ff_launch_fill((uint64_t *)first_nz, 1, ~0ull, g_st));
        FF_CHECK(ff_launch_fill((uint64_t *)err, 1, 0, g_st));
        hipLaunchKernelGGL(k_dec_leaf, dim3(1), dim3(LEAF_T), 0, g_st, a, na, (uint32_t)na, 1u, 2u << T, g_small_dig,
                           first_nz, err);
        FF_CHECK(hipGetLastError());
    } else {
        const bool fresh = g_plans.find(std::make_tuple(T, g_school_limbs, g_dec_queue)) == g_plans.end();
        DecPlan *pl = get_plan(T);
        if (!fresh && !pl->exec) capture_plan(*pl);
        top_split(*pl, a, na, bits);
        if (pl->exec) FF_CHECK(hipGraphLaunch(pl->exec, g_st));
        else enqueue_tail(*pl);
        dig = pl->digits;
        nd = pl->ndigits;
        first_nz = pl->first_nz;
        err = pl->err;
    }
    uint64_t *h = ff_pinned_results(3);
    h[2] = 0;
    FF_CHECK(hipMemcpyAsync(h, first_nz, 8, hipMemcpyDeviceToHost, g_st));
    FF_CHECK(hipMemcpyAsync(h + 1, err, 4, hipMemcpyDeviceToHost, g_st));
    // the levels' CRTs resolve carries by look-back: a wait that timed out
    // leaves wrong limbs behind, so the conversion must not report success
    FF_CHECK(hipMemcpyAsync(h + 2, ff_lb_fault(), 4, hipMemcpyDeviceToHost, g_st));
    FF_CHECK(hipStreamSynchronize(g_st));
    if ((uint32_t)h[2]) ff_lb_timeout_fail();
    const uint64_t first = h[0];
    if ((uint32_t)h[1] != 0 || first >= nd) {
        fprintf(stderr, "fastfib_amd: decimal conversion failed its leaf check\n");
        exit(1);
    }
    *dev = dig + first;
    *len = nd - first;
}

// The digits of a device number: a malloc'ed NUL-terminated string and its
// length.
char *to_string_dev(const uint64_t *a, uint64_t na, uint64_t bits, size_t *len) {
    if (bits == 0) {
        char *z = (char *)malloc(2);
        z[0] = '0', z[1] = 0;
        *len = 1;
        return z;
    }
    const char *dev;
    size_t n;
    decimal_dev(a, na, bits, &dev, &n);
    char *s = (char *)malloc(n + 1);
    if (!s) {
        fprintf(stderr, "out of memory (%zu bytes)\n", n + 1);
        exit(1);
    }
    ff_copy(s, dev, n, hipMemcpyDeviceToHost);
    s[n] = 0;
    *len = n;
    return s;
}

}  // namespace

// ------------------------------------------------------------------ C-ABI

extern "C" char *ff_big_to_string(const Big *a) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ff_ensure_init();
    if (!a || !a->d || a->n == 0) {
        char *z = (char *)malloc(2);
        z[0] = '0', z[1] = 0;
        return z;
    }
    size_t n = a->n;
    while (n > 1 && a->d[n - 1] == 0) n--;
    const uint64_t top = a->d[n - 1];
    const uint64_t bits = top ? 64 * (n - 1) + (64 - __builtin_clzll(top)) : 0;
    uint64_t *d = (uint64_t *)ff_dalloc(n * 8);
    FF_CHECK(hipStreamSynchronize(g_st));
    ff_copy(d, a->d, n * 8, hipMemcpyHostToDevice);
    size_t len;
    char *s = to_string_dev(d, n, bits, &len);
    ff_dfree(d);
    return s;
}

extern "C" int ff_dbig_to_string(const ff_dbig *a, char **out, size_t *len) {
    if (!a || !a->d || !out) return FF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ff_ensure_init();
    size_t n;
    *out = to_string_dev(a->d, a->n, a->bits, &n);
    if (len) *len = n;
    return FF_OK;
}

extern "C" int ff_dbig_to_decimal(const ff_dbig *a, const char **dev_digits, size_t *len) {
    if (!a || !a->d || !dev_digits || !len) return FF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ff_ensure_init();
    if (a->bits == 0) {
        static const char zero = '0';
        if (!g_small_dig) {
            g_small_dig = (char *)ff_dalloc((2u << DEC_LEAF_LG) * 19 + 64);
            g_small_first = (unsigned long long *)ff_dalloc(8);
            g_small_err = (uint32_t *)ff_dalloc(8);
        }
        ff_copy(g_small_dig, &zero, 1, hipMemcpyHostToDevice);
        *dev_digits = g_small_dig;
        *len = 1;
        return FF_OK;
    }
    decimal_dev(a->d, a->n, a->bits, dev_digits, len);
    return FF_OK;
}

extern "C" int ff_dec_reciprocal(int i, uint64_t *limbs, size_t cap, size_t *n) {
    if (i <= DEC_LEAF_LG || i > 40 || !n) return FF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ff_ensure_init();
    ensure_pow(i);
    recip_chain(i);
    const FFNum &m = g_dec.m[i];
    uint64_t nn = m.n;
    std::vector<uint64_t> h(nn);
    ff_copy(h.data(), m.d, nn * 8, hipMemcpyDeviceToHost);
    while (nn > 1 && h[nn - 1] == 0) nn--;
    *n = nn;
    if (!limbs) return FF_OK;
    if (cap < nn) return FF_ERR_ARG;
    memcpy(limbs, h.data(), nn * 8);
    return FF_OK;
}

extern "C" Big ff_big_recip(const Big *P, size_t s) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ff_ensure_init();
    size_t n = P->n;
    while (n > 1 && P->d[n - 1] == 0) n--;
    FFNum p = ff_num_alloc(n);
    FF_CHECK(hipStreamSynchronize(g_st));
    ff_copy(p.d, P->d, n * 8, hipMemcpyHostToDevice);
    p.n = n;
    const uint64_t top = P->d[n - 1];
    p.bits = top ? 64 * (n - 1) + (64 - __builtin_clzll(top)) : 0;
    FFNum r = recip(p, s);
    Big out;
    out.n = r.n;
    out.d = (uint64_t *)malloc(r.n * 8);
    FF_CHECK(hipStreamSynchronize(g_st));
    ff_copy(out.d, r.d, r.n * 8, hipMemcpyDeviceToHost);
    ff_num_free(r);
    ff_num_free(p);
    return out;
}

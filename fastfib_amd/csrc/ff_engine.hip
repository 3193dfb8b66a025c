// Host side of fastfib_amd: product schedules, the product pipeline
// (forward passes -> inner convolution -> inverse passes -> CRT + carries),
// device-resident fast doubling, and the C-ABI of include/fastfib_amd.h.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <tuple>
#include <mutex>
#include <utility>
#include <vector>

#include "../../include/fastfib_amd.h"
#include "ff_bigops.h"
#include "ff_engine.h"
#include "ff_internal.h"
#include "ff_limbs.h"
#include "ff_runtime.h"

typedef unsigned __int128 u128;

// ------------------------------------------------------------------ schedule
// The reference's choice of coefficient width and transform length
// (mul_ntt.c:55-118): widths 32..35 while min(ca, cb) * 2^(2w) < p1 p2 p3,
// the shortest L in {2^k, 3 * 2^k} >= ca + cb - 1 (at least 16), the
// narrowest width among equal lengths.

static int width_ok(uint64_t bits_a, uint64_t bits_b, unsigned w) {
    const u128 M = (u128)FF_P1 * FF_P2 * FF_P3;
    const uint64_t ca = (bits_a + w - 1) / w, cb = (bits_b + w - 1) / w;
    return ((u128)(ca < cb ? ca : cb) << (2 * w)) < M;
}

int ff_width_ok(uint64_t bits_a, uint64_t bits_b, unsigned w) { return width_ok(bits_a, bits_b, w); }

static uint64_t ntt_length(uint64_t conv, int *lg_out, int *r3_out) {
    uint64_t L = 16;
    int lg = 4, r3 = 0;
    while (L < conv && lg < 62) L <<= 1, lg++;
    if (L / 4 * 3 >= conv && L / 4 >= 16) L = L / 4 * 3, lg -= 2, r3 = 1;
    *lg_out = lg;
    *r3_out = r3;
    return L;
}

// The next supported length above L (2^k -> 3 * 2^(k-1), 3 * 2^k -> 2^(k+2)).
static uint64_t next_length(uint64_t L, int *lg, int *r3) {
    if (*r3) {
        *r3 = 0;
        *lg += 2;
        return L / 3 * 4;
    }
    *r3 = 1;
    *lg -= 1;
    return L / 2 * 3;
}

ProdShape ff_shape_with(const ProdShape &s0, const FFSched &sc) {
    ProdShape s = s0;
    if (!s.split && sc.len_up) {
        int lg = s.lg, r3 = s.r3;
        const uint64_t L = next_length(s.L, &lg, &r3);
        if (L <= (1u << g_max_lg)) s.L = L, s.lg = lg, s.r3 = r3;
    }
    return s;
}

static ProdShape product_shape(uint64_t bits_a, uint64_t bits_b);

ProdShape ff_product_shape(uint64_t bits_a, uint64_t bits_b) { return product_shape(bits_a, bits_b); }

static ProdShape product_shape(uint64_t bits_a, uint64_t bits_b) {
    ProdShape s{};
    for (unsigned w = 32; w <= 35; w++) {
        if (!width_ok(bits_a, bits_b, w)) break;
        const uint64_t ca = (bits_a + w - 1) / w, cb = (bits_b + w - 1) / w;
        const uint64_t conv = ca + cb > 1 ? ca + cb - 1 : 1;  // a zero operand: one coefficient
        int lg, r3;
        const uint64_t L = ntt_length(conv, &lg, &r3);
        if (s.L == 0 || L < s.L) s.L = L, s.lg = lg, s.r3 = r3, s.bw = w, s.conv = conv;
    }
    if (s.L == 0 || s.L > (1u << g_max_lg)) s.split = 1;
    if (!s.split && g_force_len) {
        // a forced length must be supported (2^k or 3*2^k <= 2^25) and long enough
        const uint64_t F = g_force_len;
        const int r3 = (F % 3 == 0);
        const uint64_t m = r3 ? F / 3 : F;
        if ((m & (m - 1)) == 0 && F >= s.conv && F >= 16 && F <= (1u << 25)) {
            s.L = F;
            s.r3 = r3;
            s.lg = 63 - __builtin_clzll(m);
        }
    }
    return s;
}

// Limbs the CRT writes for a product: na + nb + 1, as ntt_result
// (mul_ntt.c:61-73) allocates (conv * bw bits can overshoot by bw zeros).
static uint64_t out_room(const ProdShape &s, uint64_t na, uint64_t nb) {
    (void)s;
    return na + nb + 1;
}

#define FF_CRT_LIMBS 256  // limbs per CRT block (ff_crt.hip)

// ------------------------------------------------------------------ length slots
// Device words where kernels book a result's normalized length and top-limb
// bit count as ((n << 8) | bits of the top limb), by atomicMax.

static unsigned long long *g_slots = nullptr;
static int g_slot_cap = 0, g_slot_used = 0;

static void slots_reset(int need) {
    if (need > g_slot_cap) {
        if (g_slots) ff_dfree(g_slots);
        g_slot_cap = need < 1024 ? 1024 : need;
        g_slots = (unsigned long long *)ff_dalloc(g_slot_cap * sizeof(unsigned long long));
    }
    g_slot_used = 0;
}

static int slot_new(void) {
    if (g_slot_used >= g_slot_cap) {
        fprintf(stderr, "fastfib_amd: length slots exhausted\n");
        exit(1);
    }
    return g_slot_used++;
}

// The length keys s[0, n) and the look-back time-out word to the pinned
// host buffer h[0, n] in one launch (two copies are two blit launches).
__global__ void k_slots_out(uint64_t *h, const unsigned long long *s, int n, const uint32_t *fault) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) h[i] = s[i];
    if (threadIdx.x == 0) h[n] = *fault;
    __threadfence_system();
}

// Reads n length keys back (waiting for the stream), and with them the
// look-back time-out word, which must still be clear.
static void slots_read(int s0, int n, uint64_t *out) {
    uint64_t *h = ff_pinned_results((size_t)n + 1);
    hipLaunchKernelGGL(k_slots_out, dim3(1), dim3(256), 0, g_st, ff_pinned_results_dev(), g_slots + s0, n,
                       ff_lb_fault());
    FF_CHECK(hipGetLastError());
    FF_CHECK(hipStreamSynchronize(g_st));
    if ((uint32_t)h[n]) ff_lb_timeout_fail();
    memcpy(out, h, (size_t)n * 8);
}

static uint64_t key_of_bits(uint64_t bits) {
    if (bits == 0) return 0;
    const uint64_t n = (bits + 63) / 64;
    return (n << 8) | (bits - 64 * (n - 1));
}

static void key_decode(uint64_t key, uint64_t *n, uint64_t *bits) {
    if (key == 0) {
        *n = 1, *bits = 0;
        return;
    }
    *n = key >> 8;
    *bits = 64 * (*n - 1) + (key & 0xff);
}

// ------------------------------------------------------------------ numbers

struct DNum {
    uint64_t *d;
    uint64_t n, bits, cap;  // n limbs in use (top may be zero only if bits says so)
    int slot;               // where a kernel booked its length (-1: none)
};

static DNum dnum_alloc(uint64_t cap) {
    DNum x{(uint64_t *)ff_dalloc(cap * 8), 1, 0, cap, -1};
    return x;
}

static DNum dnum_u64(uint64_t v) {
    DNum x = dnum_alloc(1);
    FF_CHECK(ff_launch_fill(x.d, 1, v, g_st));
    x.bits = v ? 64 - __builtin_clzll(v) : 0;
    return x;
}

static void dnum_free(DNum &x) {
    ff_dfree(x.d);
    x.d = nullptr;
}

// ------------------------------------------------------------------ products

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

bool ff_school_ok(uint64_t na, uint64_t nb) {
    return na < (uint64_t)g_school_limbs && nb < (uint64_t)g_school_limbs;
}

int ff_sched_inner_log(const FFSched &sc) { return sc.inner_log ? sc.inner_log : g_inner_log; }
int ff_sched_tile_words(const FFSched &sc) { return sc.tile_words ? sc.tile_words : g_tile_words; }

// Runs up to FF_MAXB products that share one schedule, one launch per pass.
// With hand, the CRT is not launched: hand[i] receives product i's residues
// ([3][L] words, CRT-ready) and the buffers holding them are appended to
// keep, for the caller to free after its own use.
void ff_mul_group(const MulItem *it, int cnt, const ProdShape &s, const FFSched &sc, FFCrtJob *hand,
                  std::vector<void *> *keep) {
    FFPlan pl;
    if (ff_plan_build(&pl, (uint32_t)s.L, (int)s.bw, s.lg, s.r3, s.conv, ff_sched_inner_log(sc),
                      ff_sched_tile_words(sc)) != 0) {
        fprintf(stderr, "fastfib_amd: no schedule for L=%llu\n", (unsigned long long)s.L);
        exit(1);
    }
    const size_t words = 3 * (size_t)s.L;
    FFXformSet fwd[2] = {}, inv = {};
    FFConvSet cv = {};
    FFCrtSet crt = {};
    int nfwd = 0;
    std::vector<void *> tmp;
    uint64_t max_nout = 0;
    double opbytes = 0, conv_total = 0, out_limbs = 0;
    int ngeneral = 0;
    if (!hand) ff_lb_begin();
    for (int i = 0; i < cnt; i++) {
        const uint64_t nout = it[i].na + (it[i].b ? it[i].nb : it[i].na) + 1;
        const size_t o_fb = align256(words * 4);
        char *base = (char *)ff_dalloc(o_fb + (it[i].b ? align256(words * 4) : 0) + 256);
        tmp.push_back(base);
        uint32_t *fa = (uint32_t *)base;
        uint32_t *fb = it[i].b ? (uint32_t *)(base + o_fb) : nullptr;
        const uint32_t *asrc = (const uint32_t *)it[i].a, *bsrc = (const uint32_t *)it[i].b;
        fwd[nfwd / FF_MAXB].x[nfwd % FF_MAXB] = FFXform{asrc, (uint32_t)(2 * it[i].na), fa};
        nfwd++;
        if (fb) {
            fwd[nfwd / FF_MAXB].x[nfwd % FF_MAXB] = FFXform{bsrc, (uint32_t)(2 * it[i].nb), fb};
            nfwd++;
        }
        cv.c[i] = FFConv{asrc, bsrc, (uint32_t)(2 * it[i].na), (uint32_t)(2 * it[i].nb), fa, fb};
        inv.x[i] = FFXform{nullptr, 0, fa};
        max_nout = nout > max_nout ? nout : max_nout;
        crt.j[i] = FFCrtJob{fa,       it[i].out, hand ? nullptr : ff_lb_take(ff_crt_lb_words(nout)),
                            it[i].lenkey, nout,   it[i].conv, (uint32_t)s.L, (int)s.bw};
        opbytes += 8.0 * (it[i].na + (it[i].b ? it[i].nb : 0));
        ngeneral += it[i].b ? 1 : 0;
        conv_total += (double)it[i].conv;
        out_limbs += (double)nout;
    }
    for (int f = 0; f * FF_MAXB < nfwd; f++) {
        const int nx = nfwd - f * FF_MAXB < FF_MAXB ? nfwd - f * FF_MAXB : FF_MAXB;
        FF_CHECK(ff_launch_forward(&pl, &fwd[f], nx, opbytes * nx / nfwd, g_st));
    }
    FF_CHECK(ff_launch_conv(&pl, &cv, cnt, ngeneral, opbytes, g_st));
    FF_CHECK(ff_launch_inverse(&pl, &inv, cnt, g_st));
    if (hand) {
        for (int i = 0; i < cnt; i++) hand[i] = crt.j[i];
        keep->insert(keep->end(), tmp.begin(), tmp.end());
        return;
    }
    FF_CHECK(ff_launch_crt(&crt, cnt, max_nout, conv_total, out_limbs, g_st));
    for (void *p : tmp) ff_dfree(p);
}

// out[i] = A[i] * B[i] (B[i] == nullptr or == A[i]: square). Results get a
// length slot each; with pred (bit lengths known in advance) nothing waits
// for the device, otherwise the lengths are read back.
static void mul_many(const DNum *const *A, const DNum *const *B, DNum *out, int cnt, const uint64_t *pred,
                     const FFSched *given = nullptr) {
    std::vector<MulItem> items(cnt);
    std::vector<ProdShape> shapes(cnt);
    std::vector<FFSched> sched(cnt, FFSched{0, 0, 0});
    std::vector<int> zero(cnt, 0);
    for (int i = 0; i < cnt; i++) {
        const DNum *a = A[i], *b = B[i];
        const bool sq = (b == nullptr || b == a);
        const uint64_t bb = sq ? a->bits : b->bits;
        out[i].slot = slot_new();
        if (a->bits == 0 || bb == 0) {
            zero[i] = 1;
            DNum z = dnum_u64(0);
            z.slot = out[i].slot;
            out[i] = z;
            // the slot must read "zero": book nothing, but clear it
            FF_CHECK(ff_launch_fill((uint64_t *)(g_slots + z.slot), 1, 0, g_st));
            continue;
        }
        shapes[i] = product_shape(a->bits, bb);
        if (shapes[i].split) {
            if (given) {
                sched[i] = given[i];
            } else if (g_dispatch) {
                const uint64_t ba = a->bits;
                const uint8_t s1 = sq;
                ff_pick_scheds(&shapes[i], &ba, &bb, &s1, 1, (size_t)cnt, g_dispatch, &sched[i]);
            }
            const uint64_t nb = sq ? a->n : b->n;
            const int slot = out[i].slot;
            out[i] = dnum_alloc(out_room(shapes[i], a->n, nb));
            out[i].slot = slot;
            items[i] = MulItem{a->d, a->n, a->bits, sq ? nullptr : b->d, nb, bb, out[i].d, 0, g_slots + slot};
            continue;
        }
        if (given) {
            sched[i] = given[i];
        } else if (g_dispatch && !g_force_len) {
            const uint64_t ba = a->bits;
            const uint8_t s1 = sq;
            ff_pick_scheds(&shapes[i], &ba, &bb, &s1, 1, (size_t)cnt, g_dispatch, &sched[i]);
        }
        shapes[i] = ff_shape_with(shapes[i], sched[i]);
        const uint64_t nb = sq ? a->n : b->n;
        const int slot = out[i].slot;
        out[i] = dnum_alloc(out_room(shapes[i], a->n, nb));
        out[i].slot = slot;
        items[i] = MulItem{a->d, a->n, a->bits, sq ? nullptr : b->d, nb, bb, out[i].d, shapes[i].conv, g_slots + slot};
    }
    // one group per (length, width), up to FF_MAXB products per launch;
    // products whose operands are both below the NTT threshold go to the
    // schoolbook kernel instead, up to FF_MAXB per launch
    std::vector<int> done(cnt, 0);
    {
        FFSchoolSet ss = {};
        int ns = 0;
        double sbytes = 0;
        for (int i = 0; i < cnt; i++) {
            if (zero[i] || shapes[i].split || !ff_school_ok(items[i].na, items[i].b ? items[i].nb : items[i].na))
                continue;
            const MulItem &m = items[i];
            ss.j[ns++] = FFSchoolItem{m.a, m.b ? m.b : m.a, (uint32_t)m.na, (uint32_t)(m.b ? m.nb : m.na), m.out,
                                      m.lenkey};
            sbytes += 8.0 * (2 * m.na + 2 * (m.b ? m.nb : m.na) + 1);
            done[i] = 1;
            if (ns == FF_MAXB) {
                FF_CHECK(ff_launch_school_set(&ss, ns, sbytes, g_st));
                ns = 0, sbytes = 0;
            }
        }
        if (ns) FF_CHECK(ff_launch_school_set(&ss, ns, sbytes, g_st));
    }
    for (int i = 0; i < cnt; i++) {
        if (zero[i] || done[i]) continue;
        if (shapes[i].split) {
            ff_mul_split(items[i], sched[i].split_lg);
            done[i] = 1;
            continue;
        }
        std::vector<MulItem> grp;
        for (int j = i; j < cnt; j++) {
            if (zero[j] || done[j] || shapes[j].split) continue;
            if (shapes[j].L == shapes[i].L && shapes[j].bw == shapes[i].bw && (int)grp.size() < FF_MAXB &&
                !memcmp(&sched[j], &sched[i], sizeof(FFSched))) {
                grp.push_back(items[j]);
                done[j] = 1;
            }
        }
        ff_mul_group(grp.data(), (int)grp.size(), shapes[i], sched[i]);
    }
    if (pred) {
        for (int i = 0; i < cnt; i++) {
            if (zero[i]) continue;
            out[i].bits = pred[i];
            out[i].n = pred[i] ? (pred[i] + 63) / 64 : 1;
        }
        return;
    }
    std::vector<uint64_t> keys(cnt);
    for (int i = 0; i < cnt; i++) {
        uint64_t k;
        slots_read(out[i].slot, 1, &k);
        key_decode(k, &out[i].n, &out[i].bits);
    }
}

// ------------------------------------------------------------------ fast doubling

// Outer tile of F(n)'s own transforms of L words (0: the library default):
// fib_tile_words up to 2^20 words, where a step's passes are a few hundred
// workgroups and latency-bound, so smaller tiles (more workgroups) win; the
// longer transforms' passes keep the default (narrower rows of their
// multi-pass plans coalesce worse: F(10^9) 5.15 -> 6.04 ms with 2048 there).
static int fib_tile_words(uint64_t L) { return L <= (1u << 20) ? g_fib_tile_words : 0; }


// One linear step (fib.c:78-109) in one launch from s1 = F(k)^2 and
// s2 = F(k-1)^2, given in J as residues or limbs: (lo, hi) = bit ?
// (F(2k), F(2k+1)) : (F(2k-1), F(2k)) on len limbs.
static void step_launch(FFStepJob &J, uint64_t len, int k_odd, int bit, DNum *lo, DNum *hi, const uint64_t *pred,
                        double conv_total) {
    DNum K = dnum_alloc(len), C = dnum_alloc(len);
    K.slot = slot_new();
    C.slot = slot_new();
    J.len = len;
    J.cb0 = k_odd ? -2 : 2;
    J.bit = bit;
    J.K = K.d, J.C = C.d;
    J.keyK = g_slots + K.slot, J.keyC = g_slots + C.slot;
    ff_lb_begin();
    J.lb = ff_lb_take(ff_step_lb_words(len));
    FF_CHECK(ff_launch_step_fused(&J, conv_total, g_st));
    if (pred) {
        // pred[0], pred[1]: bit lengths of (lo, hi)
        const uint64_t bk = bit ? pred[1] : pred[0], bc = bit ? pred[0] : pred[1];
        K.bits = bk, K.n = (bk + 63) / 64;
        C.bits = bc, C.n = (bc + 63) / 64;
    } else {
        uint64_t k[2];
        slots_read(K.slot, 2, k);
        key_decode(k[0], &K.n, &K.bits);
        key_decode(k[1], &C.n, &C.bits);
    }
    if (bit) *lo = C, *hi = K;
    else *lo = K, *hi = C;
}

// A doubling step from (F(k-1), F(k)): the two squarings' transforms (one
// launch per pass when their shapes and schedules agree), then the CRT of
// both squares fused with the linear step (k_step_fused). Squares past the
// longest transform (split products) or of zero go through mul_many and the
// step reads their limbs. ps: predicted bits of the squares; pl: of the new
// pair; given: the squares' schedules.
static void square_step(const DNum &fk, const DNum &fkm1, int k_odd, int bit, DNum *lo, DNum *hi,
                        const uint64_t *ps, const uint64_t *pl, const FFSched *given) {
    ProdShape sh[2];
    sh[0] = product_shape(fk.bits, fk.bits);
    sh[1] = fkm1.bits ? product_shape(fkm1.bits, fkm1.bits) : sh[0];
    if (sh[0].split || sh[1].split || fk.bits == 0 || fkm1.bits == 0) {
        DNum sq[2];
        const DNum *A[2] = {&fk, &fkm1}, *B[2] = {nullptr, nullptr};
        mul_many(A, B, sq, 2, ps, given);
        FFStepJob J{};
        J.s1l = sq[0].d, J.s2l = sq[1].d;
        J.n1 = sq[0].n, J.n2 = sq[1].n;
        step_launch(J, sq[0].n + 1, k_odd, bit, lo, hi, pl, 0.0);
        dnum_free(sq[0]);
        dnum_free(sq[1]);
        return;
    }
    FFSched sc[2] = {FFSched{0, 0, 0}, FFSched{0, 0, 0}};
    const DNum *x[2] = {&fk, &fkm1};
    for (int j = 0; j < 2; j++) {
        if (given) {
            sc[j] = given[j];
        } else if (g_dispatch && !g_force_len) {
            const uint64_t bb = x[j]->bits;
            const uint8_t sq1 = 1;
            ff_pick_scheds(&sh[j], &bb, &bb, &sq1, 1, 2, g_dispatch, &sc[j]);
        } else {
            sc[j].tile_words = fib_tile_words(sh[j].L);
        }
        sh[j] = ff_shape_with(sh[j], sc[j]);
    }
    MulItem it[2];
    for (int j = 0; j < 2; j++)
        it[j] = MulItem{x[j]->d, x[j]->n, x[j]->bits, nullptr, x[j]->n, x[j]->bits, nullptr, sh[j].conv, nullptr};
    FFCrtJob hand[2];
    std::vector<void *> keep;
    if (sh[1].L == sh[0].L && sh[1].bw == sh[0].bw && !memcmp(&sc[0], &sc[1], sizeof(FFSched))) {
        ff_mul_group(it, 2, sh[0], sc[0], hand, &keep);
    } else {
        ff_mul_group(&it[0], 1, sh[0], sc[0], &hand[0], &keep);
        ff_mul_group(&it[1], 1, sh[1], sc[1], &hand[1], &keep);
    }
    FFStepJob J{};
    J.y1 = hand[0].y, J.L1 = hand[0].L, J.bw1 = hand[0].bw, J.conv1 = hand[0].conv;
    J.y2 = hand[1].y, J.L2 = hand[1].L, J.bw2 = hand[1].bw, J.conv2 = hand[1].conv;
    // limbs of s1 plus one (with predicted lengths); else a bound that
    // keeps the results' top limbs among the last three
    const uint64_t len = ps ? (ps[0] + 63) / 64 + 1 : 2 * fk.n + 1;
    step_launch(J, len, k_odd, bit, lo, hi, pl, (double)(sh[0].conv + sh[1].conv));
    for (void *p : keep) ff_dfree(p);
}

// ------------------------------------------------------------------ pair products
// u = cu0 a^2 + cu1 ab + cu2 b^2 and v = cv0 a^2 + cv1 ab + cv2 b^2 (b <= a)
// from one forward transform of each of a and b: the pointwise products are
// weighted and summed before the inverse transforms (k_inner_pair for two
// outputs or degree 4, the general product's middle round for one), and the
// plain CRT writes u and v. A doubling step
// is such a pair, with no linear pass and no constant after its CRT:
//   F(2k-1) = a^2 + b^2, F(2k) = a^2 + 2ab, F(2k+1) = 2a^2 + 2ab + b^2
// for a = F(k), b = F(k-1) (the identities behind fib.c:36-62's step). The
// CRT bound carries the largest weight sum s: s ca 2^(2w) < p1 p2 p3, with
// the reference rule's widths otherwise (32..35, shortest length first).
static ProdShape pair_shape(uint64_t bits_a, unsigned wsum) {
    ProdShape s{};
    const u128 M = (u128)FF_P1 * FF_P2 * FF_P3;
    for (unsigned w = 32; w <= 35; w++) {
        const uint64_t ca = (bits_a + w - 1) / w;
        if ((((u128)ca * wsum) << (2 * w)) >= M) break;
        int lg, r3;
        const uint64_t L = ntt_length(2 * ca - 1, &lg, &r3);
        if (s.L == 0 || L < s.L) s.L = L, s.lg = lg, s.r3 = r3, s.bw = w, s.conv = 2 * ca - 1;
    }
    if (s.L == 0 || s.L > (1u << g_max_lg)) s.split = 1;
    return s;
}

// g_pair_steps bit 0: doubling steps as pairs; bit 1: the last product as a
// pair. Quadrupling (always a pair) follows g_quad_max_lg alone.
static bool pair_free(void) { return !g_dispatch && !g_force_len; }
static bool pair_enabled(int bit) { return pair_free() && (g_pair_steps >> bit & 1); }

// Runs a pair of the given shape when its transforms take outer passes
// (longer than one block, not split); false, with nothing launched,
// otherwise. deg4: the weights apply to x^2, xy, y^2 for x = a^2 + 2ab,
// y = a^2 + b^2. u (and v when cv) get nout limbs and a length slot each;
// pred: their bit lengths. The CRT books a length key only when the top
// nonzero limb is among the last three (lb_keys), so nout must be tight:
// the callers derive it from an upper bound within a few bits of the
// outputs' true lengths.
// Largest quadrupling transform that takes the quad_inner_log blocks: up to
// 2^17 words a quadrupling has fewer than 128 pair workgroups of 2^12 words
// (3 L / 4096), each a serial chain of both operands' forward and both
// outputs' inverse levels; from 2^18 on there are enough of them.
#define FF_QUAD_SMALL_L (1u << 17)
static bool pair_run(const DNum &a, const DNum &b, const ProdShape &s, int deg4, const uint8_t *cu,
                     const uint8_t *cv, uint64_t nout, DNum *u, DNum *v, const uint64_t *pred) {
    if (b.bits == 0 || b.bits > a.bits || s.split || s.L <= FF_TW) return false;
    FFPlan pl;
    // quadruplings of few inner blocks: smaller blocks, more workgroups
    const int il = deg4 && g_quad_inner_log && s.L <= FF_QUAD_SMALL_L ? g_quad_inner_log : g_inner_log;
    const int tw = fib_tile_words(s.L);
    if (ff_plan_build(&pl, (uint32_t)s.L, (int)s.bw, s.lg, s.r3, s.conv, il, tw ? tw : g_tile_words) != 0 ||
        pl.npass == 0)
        return false;
    const int two = cv != nullptr;
    const size_t o_fb = align256(3 * (size_t)s.L * 4);
    char *base = (char *)ff_dalloc(2 * o_fb);
    uint32_t *fa = (uint32_t *)base, *fb = (uint32_t *)(base + o_fb);
    const double opbytes = 8.0 * (a.n + b.n);
    FFXformSet fwd = {};
    fwd.x[0] = FFXform{(const uint32_t *)a.d, (uint32_t)(2 * a.n), fa};
    fwd.x[1] = FFXform{(const uint32_t *)b.d, (uint32_t)(2 * b.n), fb};
    FF_CHECK(ff_launch_forward(&pl, &fwd, 2, opbytes, g_st));
    FFConvSet cs = {};
    cs.c[0] = FFConv{(const uint32_t *)a.d, (const uint32_t *)b.d, (uint32_t)(2 * a.n), (uint32_t)(2 * b.n), fa, fb};
    cs.c[0].pair = (two ? 2 : 1) | (deg4 ? 4 : 0);
    for (int j = 0; j < 3; j++) cs.c[0].cu[j] = cu[j], cs.c[0].cv[j] = two ? cv[j] : 0;
    FF_CHECK(ff_launch_conv(&pl, &cs, 1, 1 + two, opbytes, g_st));
    FFXformSet inv = {};
    inv.x[0] = FFXform{nullptr, 0, fa};
    inv.x[1] = FFXform{nullptr, 0, fb};
    FF_CHECK(ff_launch_inverse(&pl, &inv, 1 + two, g_st));
    DNum *o[2] = {u, v};
    FFCrtSet crt = {};
    ff_lb_begin();
    for (int j = 0; j <= two; j++) {
        *o[j] = dnum_alloc(nout);
        o[j]->slot = slot_new();
        crt.j[j] = FFCrtJob{j ? fb : fa, o[j]->d, ff_lb_take(ff_crt_lb_words(nout)), g_slots + o[j]->slot,
                            nout,        s.conv,  (uint32_t)s.L,                      (int)s.bw};
    }
    FF_CHECK(ff_launch_crt(&crt, 1 + two, nout, (double)s.conv * (1 + two), (double)nout * (1 + two), g_st));
    ff_dfree(base);
    if (pred) {
        for (int j = 0; j <= two; j++) o[j]->bits = pred[j], o[j]->n = pred[j] ? (pred[j] + 63) / 64 : 1;
    } else {
        uint64_t k[2];
        slots_read(u->slot, 1 + two, k);
        for (int j = 0; j <= two; j++) key_decode(k[j], &o[j]->n, &o[j]->bits);
    }
    return true;
}

static unsigned wsum3(const uint8_t *w) { return (unsigned)w[0] + w[1] + w[2]; }

// Output limbs of a doubling pair: with a = F(k) of bits_a bits, F(2k-1),
// F(2k) and F(2k+1) lie in [a^2, 5 a^2], so between 2 bits_a - 1 and
// 2 bits_a + 3 bits; one limb of room above that bound.
static uint64_t pair_limbs(uint64_t bits_a) { return (2 * bits_a + 3 + 63) / 64 + 1; }
// Of a quadrupling: F(4k-1) .. F(4k+3) lie between 2^(4 bits_a - 2) and
// 2^(4 bits_a + 6).
static uint64_t quad_limbs(uint64_t bits_a) { return (4 * bits_a + 6 + 63) / 64 + 1; }

// A doubling step as one pair: (lo, hi) = bit ? (F(2k), F(2k+1)) :
// (F(2k-1), F(2k)); pl: their bit lengths.
static bool pair_step(const DNum &fk, const DNum &fkm1, int bit, DNum *lo, DNum *hi, const uint64_t *pl) {
    static const uint8_t W_LO[2][3] = {{1, 0, 1}, {1, 2, 0}}, W_HI[2][3] = {{1, 2, 0}, {2, 2, 1}};
    const unsigned ws = std::max(wsum3(W_LO[bit]), wsum3(W_HI[bit]));
    return pair_run(fk, fkm1, pair_shape(fk.bits, ws), 0, W_LO[bit], W_HI[bit], pair_limbs(fk.bits), lo, hi, pl);
}

// ------------------------------------------------------------------ quadrupling
// Two bits of n in one transform round trip: with a = F(k), b = F(k-1),
// x = F(2k) = a^2 + 2ab and y = F(2k-1) = a^2 + b^2 are formed pointwise,
// and for m = 4k + d the new pair (F(m-1), F(m)) is a combination of x^2,
// xy, y^2 (F(4k) = x^2 + 2xy, F(4k-1) = x^2 + y^2, then F(j+1) = F(j) +
// F(j-1)):
//   d = 0: (x^2 + y^2,         x^2 + 2xy)
//   d = 1: (x^2 + 2xy,         2x^2 + 2xy + y^2)
//   d = 2: (2x^2 + 2xy + y^2,  3x^2 + 4xy + y^2)
//   d = 3: (3x^2 + 4xy + y^2,  5x^2 + 6xy + 2y^2)
// The outputs are degree 4 in the coefficients of a and b: x has at most
// 2 ca - 1 coefficients below 3 ca B^2 (B = 2^w), y below 2 ca B^2, so a
// coefficient of w0 x^2 + w1 xy + w2 y^2 stays below
// (18 w0 + 12 w1 + 8 w2) ca^3 B^4, which must stay below p1 p2 p3: widths
// near 10 bits. Where a step costs its launches' latency rather than its
// arithmetic, one such step replaces two.
static const uint8_t Q_W[4][2][3] = {{{1, 0, 1}, {1, 2, 0}},
                                     {{1, 2, 0}, {2, 2, 1}},
                                     {{2, 2, 1}, {3, 4, 1}},
                                     {{3, 4, 1}, {5, 6, 2}}};

static unsigned quad_bound(const uint8_t *w) { return 18u * w[0] + 12u * w[1] + 8u * w[2]; }

// The shortest transform over widths 8..35 that keep K ca^3 2^(4w) below
// p1 p2 p3, the widest width among equal lengths (fewest coefficients).
static ProdShape quad_shape(uint64_t bits_a, unsigned K) {
    ProdShape s{};
    const u128 M = (u128)FF_P1 * FF_P2 * FF_P3;
    for (unsigned w = 8; w <= 35 && 4 * w < 93; w++) {
        const uint64_t ca = (bits_a + w - 1) / w;
        if (ca >= (1u << 25)) continue;
        if ((u128)ca * ca * ca * K >= (M >> (4 * w))) break;
        int lg, r3;
        const uint64_t L = ntt_length(4 * ca - 3, &lg, &r3);
        if (s.L == 0 || L <= s.L) s.L = L, s.lg = lg, s.r3 = r3, s.bw = w, s.conv = 4 * ca - 3;
    }
    if (s.L == 0 || s.L > (1u << g_max_lg)) s.split = 1;
    return s;
}

static bool quad_step(const DNum &fk, const DNum &fkm1, int d, DNum *lo, DNum *hi, const uint64_t *pl) {
    if (!g_quad_max_lg || fkm1.bits == 0) return false;
    const unsigned K = std::max(quad_bound(Q_W[d][0]), quad_bound(Q_W[d][1]));
    const ProdShape s = quad_shape(fk.bits, K);
    if (s.split || s.L > (1ull << g_quad_max_lg)) return false;
    return pair_run(fk, fkm1, s, 1, Q_W[d][0], Q_W[d][1], quad_limbs(fk.bits), lo, hi, pl);
}

// A step whose squarings are one-block transforms (L <= 4096), with
// predicted lengths: the two transforms (one launch when their shapes
// agree), then the CRT of both squares and the linear step fused in one
// workgroup (ff_launch_small_step) instead of six launches.
static bool small_fusable(const DNum &fk) {
    if (g_force_len || g_dispatch || fk.bits == 0) return false;
    const ProdShape s = product_shape(fk.bits, fk.bits);
    return !s.split && s.L <= FF_TW && s.bw >= 32;
}

static void small_step(const DNum &fk, const DNum &fkm1, int k_odd, int bit, DNum *lo, DNum *hi,
                       uint64_t bits_sq1, const uint64_t pred[2]) {
    const ProdShape s1 = product_shape(fk.bits, fk.bits);
    const bool z2 = fkm1.bits == 0;
    const ProdShape s2 = z2 ? s1 : product_shape(fkm1.bits, fkm1.bits);
    const uint64_t len = (bits_sq1 + 63) / 64 + 1;  // s2 <= s1, results < 4 s1 + 2
    uint32_t *fa1 = (uint32_t *)ff_dalloc(3 * (size_t)s1.L * 4);
    uint32_t *fa2 = z2 ? nullptr : (uint32_t *)ff_dalloc(3 * (size_t)s2.L * 4);
    FFPlan p1;
    if (ff_plan_build(&p1, (uint32_t)s1.L, (int)s1.bw, s1.lg, s1.r3, s1.conv, g_inner_log, g_tile_words) != 0) {
        fprintf(stderr, "fastfib_amd: no schedule for L=%llu\n", (unsigned long long)s1.L);
        exit(1);
    }
    FFConvSet cv = {};
    cv.c[0] = FFConv{(const uint32_t *)fk.d, nullptr, (uint32_t)(2 * fk.n), 0, fa1, nullptr};
    int nc = 1;
    const bool same = !z2 && s2.L == s1.L && s2.bw == s1.bw;
    if (same) {
        cv.c[1] = FFConv{(const uint32_t *)fkm1.d, nullptr, (uint32_t)(2 * fkm1.n), 0, fa2, nullptr};
        nc = 2;
    }
    FF_CHECK(ff_launch_conv(&p1, &cv, nc, 0, 8.0 * (fk.n + (same ? fkm1.n : 0)), g_st));
    if (!z2 && !same) {
        FFPlan p2;
        if (ff_plan_build(&p2, (uint32_t)s2.L, (int)s2.bw, s2.lg, s2.r3, s2.conv, g_inner_log, g_tile_words) != 0) {
            fprintf(stderr, "fastfib_amd: no schedule for L=%llu\n", (unsigned long long)s2.L);
            exit(1);
        }
        FFConvSet c2 = {};
        c2.c[0] = FFConv{(const uint32_t *)fkm1.d, nullptr, (uint32_t)(2 * fkm1.n), 0, fa2, nullptr};
        FF_CHECK(ff_launch_conv(&p2, &c2, 1, 0, 8.0 * fkm1.n, g_st));
    }
    DNum K = dnum_alloc(len), C = dnum_alloc(len);
    K.slot = slot_new();
    C.slot = slot_new();
    FFSmallStep a{fa1,
                  z2 ? fa1 : fa2,
                  (uint32_t)s1.L,
                  (uint32_t)s2.L,
                  (int)s1.bw,
                  (int)s2.bw,
                  (uint32_t)s1.conv,
                  z2 ? 0u : (uint32_t)s2.conv,
                  (uint32_t)len,
                  k_odd ? -2 : 2,
                  bit,
                  K.d,
                  C.d,
                  g_slots + K.slot,
                  g_slots + C.slot};
    FF_CHECK(ff_launch_small_step(&a, g_st));
    ff_dfree(fa1);
    if (fa2) ff_dfree(fa2);
    // pred[0], pred[1]: bit lengths of (lo, hi)
    const uint64_t bk = bit ? pred[1] : pred[0], bc = bit ? pred[0] : pred[1];
    K.bits = bk, K.n = bk ? (bk + 63) / 64 : 1;
    C.bits = bc, C.n = bc ? (bc + 63) / 64 : 1;
    if (bit) *lo = C, *hi = K;
    else *lo = K, *hi = C;
}

// X = 2u + v and, if Y, Y = 2u - v (fib.c:145-161), in one launch.
static void last_factors(const DNum &u, const DNum &v, DNum *X, DNum *Y, const uint64_t *pred) {
    const uint64_t len = (u.n > v.n ? u.n : v.n) + 1;
    *X = dnum_alloc(len);
    X->slot = slot_new();
    if (Y) {
        *Y = dnum_alloc(len);
        Y->slot = slot_new();
    }
    ff_lb_begin();
    FFLastArgs a{u.d,        v.d, u.n, v.n, len, X->d, Y ? Y->d : nullptr, g_slots + X->slot,
                 Y ? g_slots + Y->slot : nullptr, ff_lb_take(ff_last_lb_words(len))};
    FF_CHECK(ff_launch_last(&a, g_st));
    if (pred) {
        X->bits = pred[0], X->n = (pred[0] + 63) / 64;
        if (Y) Y->bits = pred[1], Y->n = (pred[1] + 63) / 64;
    } else {
        uint64_t k[2];
        slots_read(X->slot, Y ? 2 : 1, k);
        key_decode(k[0], &X->n, &X->bits);
        if (Y) key_decode(k[1], &Y->n, &Y->bits);
    }
}

// Bit lengths of Fibonacci and Lucas numbers, known in advance so that the
// whole doubling chain can be queued without waiting for the device:
//   F(m) = round(phi^m / sqrt 5), L(m) = phi^m + psi^m.
// Exact integers up to m = 93; above, log2 values in long double (about
// 1e-9 absolute error at m = 1e10), refused (-1) within 1e-7 of an
// integer, where the caller falls back to reading lengths back.
enum { FB_F = 0, FB_F2 = 1, FB_L = 2 };

static int fib_bits_pred(uint64_t m, int kind, uint64_t *bits) {
    if (m <= 93) {
        uint64_t a = 0, b = 1;  // F(0), F(1)
        for (uint64_t i = 0; i < m; i++) {
            const uint64_t t = a + b;
            a = b, b = t;
        }
        // a = F(m), b = F(m+1)
        u128 v;
        if (kind == FB_F) v = a;
        else if (kind == FB_F2) v = (u128)a * a;
        else v = (m == 0) ? 2 : (u128)a + 2 * (u128)(b - a);  // L(m) = F(m) + 2F(m-1)
        uint64_t r = 0;
        while (v) v >>= 1, r++;
        *bits = r;
        return 0;
    }
    const long double LOG2PHI = 0.69424191363061730173879026689859522L;
    const long double LOG2SQRT5 = 1.16096404744368117393515971474469509L;
    long double x = (long double)m * LOG2PHI;
    if (kind == FB_F) x -= LOG2SQRT5;
    else if (kind == FB_F2) x = 2 * (x - LOG2SQRT5);
    const long double f = floorl(x);
    const long double frac = x - f;
    if (frac < 1e-7L || frac > 1 - 1e-7L) return -1;
    *bits = (uint64_t)f + 1;
    return 0;
}

struct FibPlanVal {
    int slot;
    uint64_t bits;
};

// fibonacci (fib.c:211-261) with every product and pass on the device: the
// pair (F(k-1), F(k)) stays in HBM; each bit costs two squarings run as one
// launch per pass (same schedule) and one linear step. With predict, all
// lengths come from fib_bits_pred and nothing waits for the device until
// the end, where the lengths the kernels booked are checked against the
// predictions (a mismatch returns false; the caller reruns reading lengths).
// Transform of F(k)^2 in the chain of first steps (ff_launch_fib_early): a
// power-of-2 length, the widest coefficients that fit (as product_shape).
static bool early_shape(uint64_t bits, int *lg, int *bw, uint64_t *conv) {
    for (int w = 35; w >= 32; w--) {
        if (!width_ok(bits, bits, (unsigned)w)) continue;
        const uint64_t cv = 2 * ((bits + w - 1) / w) - 1;
        int g = 4;
        while ((1ull << g) < cv) g++;
        *lg = g, *bw = w, *conv = cv;
        return g <= FF_EARLY_LG;
    }
    return false;
}

// F(k - 1), F(k) exactly on the host for k < 1024 (at most 12 limbs), by the
// doubling of the device steps: s1 = F(k)^2, s2 = F(k-1)^2,
// F(2k-1) = s1 + s2, F(2k+1) = 4 s1 - s2 + 2 (-1)^k, F(2k) = F(2k+1) - F(2k-1).
typedef std::vector<uint64_t> HNum;

static void hn_trim(HNum &a) {
    while (a.size() > 1 && a.back() == 0) a.pop_back();
}

// a^2: the products a_i a_j (i < j) once, doubled, plus the squares a_i^2.
static HNum hn_sqr(const HNum &a) {
    const size_t n = a.size();
    HNum r(2 * n, 0);
    for (size_t i = 0; i + 1 < n; i++) {
        u128 c = 0;
        for (size_t j = i + 1; j < n; j++) {
            const u128 t = (u128)a[i] * a[j] + r[i + j] + c;
            r[i + j] = (uint64_t)t;
            c = t >> 64;
        }
        r[i + n] = (uint64_t)c;
    }
    uint64_t top = 0;
    for (size_t q = 0; q < 2 * n; q++) {  // r <<= 1
        const uint64_t v = r[q];
        r[q] = (v << 1) | top;
        top = v >> 63;
    }
    u128 c = 0;
    for (size_t i = 0; i < n; i++) {
        const u128 sq = (u128)a[i] * a[i];
        u128 t = (u128)r[2 * i] + (uint64_t)sq + c;
        r[2 * i] = (uint64_t)t;
        t = (u128)r[2 * i + 1] + (uint64_t)(sq >> 64) + (t >> 64);
        r[2 * i + 1] = (uint64_t)t;
        c = t >> 64;
    }
    hn_trim(r);
    return r;
}

// a + m b + c (m = 1, or -1 with a >= b; c small, the result nonnegative)
static HNum hn_axpy(const HNum &a, const HNum &b, int m, int64_t c) {
    const size_t n = (a.size() > b.size() ? a.size() : b.size()) + 1;
    HNum r(n, 0);
    __int128 carry = c;
    for (size_t i = 0; i < n; i++) {
        __int128 t = carry + (i < a.size() ? (__int128)a[i] : 0);
        if (i < b.size()) t += m > 0 ? (__int128)b[i] : -(__int128)b[i];
        r[i] = (uint64_t)t;
        carry = t >> 64;  // arithmetic: -1 on a borrow
    }
    hn_trim(r);
    return r;
}

static void hn_fib_pair(uint64_t k, HNum *fkm1, HNum *fk) {
    HNum lo{0}, hi{1};  // F(0), F(1)
    const int top = 63 - __builtin_clzll(k);
    uint64_t kk = 1;
    for (int i = top - 1; i >= 0; i--) {
        const int bit = (int)((k >> i) & 1);
        const HNum s1 = hn_sqr(hi), s2 = hn_sqr(lo);
        const HNum A = hn_axpy(s1, s2, 1, 0);
        const HNum s1x4 = hn_axpy(hn_axpy(s1, s1, 1, 0), hn_axpy(s1, s1, 1, 0), 1, 0);
        const HNum B = hn_axpy(s1x4, s2, -1, (kk & 1) ? -2 : 2);
        const HNum C = hn_axpy(B, A, -1, 0);
        if (bit) lo = C, hi = B;
        else lo = A, hi = C;
        kk = 2 * kk + bit;
    }
    *fkm1 = lo, *fk = hi;
}

static bool dfib(uint64_t n, bool predict, DNum *res) {
    slots_reset(8 * 64 + 16);
    if (n == 0) {
        *res = dnum_u64(0);
        return true;
    }
    std::vector<FibPlanVal> checks;
    DNum fkm1{}, fk{};
    uint64_t k = 1;
    int k_odd = 1;
    const int top = 63 - __builtin_clzll(n);
    auto P = [&](uint64_t m, int kind) {
        uint64_t b = 0;
        if (fib_bits_pred(m, kind, &b) != 0) predict = false;
        return b;
    };
    if (predict) {
        // check every prediction up front; any refusal switches to reading lengths
        uint64_t kk = 1, tmp;
        for (int i = top - 1; i >= 1 && predict; i--) {
            const int bit = (int)((n >> i) & 1);
            if (fib_bits_pred(kk, FB_F2, &tmp) || fib_bits_pred(kk - 1, FB_F2, &tmp) ||
                fib_bits_pred(2 * kk - 1 + bit, FB_F, &tmp) || fib_bits_pred(2 * kk + bit, FB_F, &tmp))
                predict = false;
            kk = 2 * kk + bit;
        }
        if (predict && top > 0) {
            if (fib_bits_pred(kk + 2, FB_F, &tmp) || fib_bits_pred(kk - 1, FB_L, &tmp) ||
                fib_bits_pred(kk, FB_L, &tmp) || fib_bits_pred(n, FB_F, &tmp))
                predict = false;
        }
    }
    // With a tuned or learned dispatcher and predicted lengths, every
    // product's schedule is picked up front in one call (one inference
    // launch): two squarings per step, then the final product.
    // the picks depend only on n and the source: kept for repeated F(n)
    static std::map<std::tuple<uint64_t, int, int>, std::vector<FFSched>> picked;
    std::vector<FFSched> ds;
    auto memo = picked.find(std::make_tuple(n, g_dispatch, g_dispatch_gen));
    if (predict && g_dispatch && !g_force_len && top > 0 && memo != picked.end()) ds = memo->second;
    if (predict && g_dispatch && !g_force_len && top > 0 && ds.empty()) {
        std::vector<ProdShape> s0;
        std::vector<uint64_t> ba, bb;
        std::vector<uint8_t> sqf;
        uint64_t kk = 1, b;
        for (int i = top - 1; i >= 1; i--) {
            for (int j = 0; j < 2; j++) {
                fib_bits_pred(kk - j, FB_F, &b);
                b = b ? b : 1;
                s0.push_back(product_shape(b, b)), ba.push_back(b), bb.push_back(b), sqf.push_back(1);
            }
            kk = 2 * kk + ((n >> i) & 1);
        }
        ds.resize(s0.size() + 1);
        ff_pick_scheds(s0.data(), ba.data(), bb.data(), sqf.data(), s0.size(), 2, g_dispatch, ds.data());
        // the two squarings of a step run as one launch group: the smaller
        // takes the larger one's schedule where their shapes agree
        for (size_t j = 0; j + 1 < s0.size(); j += 2) {
            if (s0[j + 1].L == s0[j].L && s0[j + 1].bw == s0[j].bw) ds[j + 1] = ds[j];
        }
        uint64_t bx, by;
        if (n & 1) fib_bits_pred(kk + 2, FB_F, &bx), fib_bits_pred(kk - 1, FB_L, &by);
        else fib_bits_pred(kk, FB_F, &bx), fib_bits_pred(kk, FB_L, &by);
        bx = bx ? bx : 1, by = by ? by : 1;
        const ProdShape sf = product_shape(bx, by);
        const uint8_t gen = 0;
        if (!sf.split) ff_pick_scheds(&sf, &bx, &by, &gen, 1, 1, g_dispatch, &ds[s0.size()]);
        picked[std::make_tuple(n, g_dispatch, g_dispatch_gen)] = ds;
    }
    int i = top - 1, step = 0;
    if (predict && !g_force_len && top >= g_host_bits) {
        // The top host_bits bits on the host (F(k0) for k0 < 2^host_bits:
        // at most 89 limbs), then the steps while F(k)^2 is a transform of
        // at most 2^10 words in one workgroup, the pair never leaving LDS.
        const int i0 = top - (g_host_bits - 1);
        const uint64_t k0 = n >> i0;
        HNum fa, fb;
        hn_fib_pair(k0, &fa, &fb);
        FFEarlyArgs ea;
        memset(&ea, 0, sizeof ea);
        for (size_t q = 0; q < fa.size() && q < FF_EARLY_F0; q++) ea.f[0][q] = fa[q];
        for (size_t q = 0; q < fb.size() && q < FF_EARLY_F0; q++) ea.f[1][q] = fb[q];
        uint64_t kk = k0, blo = 0, bhi = 0;
        int ii = i0 - 1, kodd = (int)(k0 & 1);
        std::vector<uint64_t> pred;
        while (ii >= 1 && ea.nsteps < FF_EARLY_MAXS) {
            const uint64_t b1 = P(kk, FB_F), b2 = P(kk - 1, FB_F), bsq = P(kk, FB_F2);
            int lg, bw;
            uint64_t conv;
            const uint64_t len = (bsq + 63) / 64 + 1;
            if (!early_shape(b1, &lg, &bw, &conv) || len > FF_EARLY_LIMBS) break;
            const int bit = (int)((n >> ii) & 1);
            FFEarlyStep &e = ea.st[ea.nsteps++];
            e.lg = lg, e.bw = bw;
            e.conv1 = (uint32_t)conv;
            e.nw1 = (uint32_t)((b1 + 31) / 32), e.nw2 = (uint32_t)((b2 + 31) / 32);
            e.len = (uint32_t)len;
            e.bit = bit;
            blo = P(2 * kk - 1 + bit, FB_F), bhi = P(2 * kk + bit, FB_F);
            pred.push_back(blo), pred.push_back(bhi);
            kk = 2 * kk + bit, kodd = bit, ii--;
        }
        if (predict && ea.nsteps > 0) {
            for (int lg = 4; lg <= FF_EARLY_LG; lg++)
                for (int q = 0; q < 3; q++) ea.sc[lg][q] = ff_inv_scale(1u << lg, q);
            const int slot0 = slot_new();
            for (size_t q = 1; q < pred.size(); q++) slot_new();
            for (size_t q = 0; q < pred.size(); q++) checks.push_back({slot0 + (int)q, pred[q]});
            fkm1 = dnum_alloc(blo ? (blo + 63) / 64 : 1);
            fk = dnum_alloc((bhi + 63) / 64);
            fkm1.bits = blo, fkm1.n = blo ? (blo + 63) / 64 : 1;
            fk.bits = bhi, fk.n = (bhi + 63) / 64;
            ea.nlo = (uint32_t)fkm1.n, ea.nhi = (uint32_t)fk.n;
            ea.lo = fkm1.d, ea.hi = fk.d;
            ea.keys = g_slots + slot0;
            FF_CHECK(ff_launch_fib_early(&ea, g_st));
            k = kk, k_odd = kodd;
            step = (top - 1) - ii;
            i = ii;
        }
    }
    if (!fk.d) fkm1 = dnum_u64(0), fk = dnum_u64(1);
    for (; i >= 1; i--, step++) {
        if (!g_force_len && i >= 2) {
            // bits i and i - 1 at once (own shapes and the default schedule,
            // whatever the schedule source: the up-front picks of the two
            // doubling steps it replaces go unused)
            const int d = (int)((n >> (i - 1)) & 3);
            uint64_t pq[2] = {0, 0};
            if (predict) pq[0] = P(4 * k + d - 1, FB_F), pq[1] = P(4 * k + d, FB_F);
            DNum nlo, nhi;
            if (quad_step(fk, fkm1, d, &nlo, &nhi, predict ? pq : nullptr)) {
                if (predict) {
                    checks.push_back({nlo.slot, pq[0]});
                    checks.push_back({nhi.slot, pq[1]});
                }
                dnum_free(fk);
                dnum_free(fkm1);
                fkm1 = nlo, fk = nhi;
                k = 4 * k + d;
                k_odd = d & 1;
                i--, step++;
                continue;
            }
        }
        if (predict && small_fusable(fk)) {
            const int bit = (int)((n >> i) & 1);
            const uint64_t bsq = P(k, FB_F2);
            uint64_t pl[2] = {P(2 * k - 1 + bit, FB_F), P(2 * k + bit, FB_F)};
            DNum nlo, nhi;
            small_step(fk, fkm1, k_odd, bit, &nlo, &nhi, bsq, pl);
            checks.push_back({nlo.slot, pl[0]});
            checks.push_back({nhi.slot, pl[1]});
            dnum_free(fk);
            dnum_free(fkm1);
            fkm1 = nlo, fk = nhi;
            k = 2 * k + bit;
            k_odd = bit;
            continue;
        }
        const int bit = (int)((n >> i) & 1);
        uint64_t ps[2] = {0, 0}, pl[2] = {0, 0};
        if (predict) {
            ps[0] = P(k, FB_F2), ps[1] = P(k - 1, FB_F2);
            pl[0] = P(2 * k - 1 + bit, FB_F), pl[1] = P(2 * k + bit, FB_F);
        }
        DNum nlo, nhi;
        if (!(pair_enabled(0) && pair_step(fk, fkm1, bit, &nlo, &nhi, predict ? pl : nullptr)))
            square_step(fk, fkm1, k_odd, bit, &nlo, &nhi, predict ? ps : nullptr, predict ? pl : nullptr,
                        ds.empty() ? nullptr : &ds[2 * step]);
        if (predict) {
            checks.push_back({nlo.slot, pl[0]});
            checks.push_back({nhi.slot, pl[1]});
        }
        dnum_free(fk);
        dnum_free(fkm1);
        fkm1 = nlo, fk = nhi;
        k = 2 * k + bit;
        k_odd = bit;
    }
    if (top == 0) {
        dnum_free(fkm1);
        *res = fk;
        return true;
    }
    DNum result;
    // the last step as a pair with one output: F(2k) = a^2 + 2ab,
    // F(2k+1) = 2a^2 + 2ab + b^2
    static const uint8_t W_LAST[2][3] = {{1, 2, 0}, {2, 2, 1}};
    uint64_t pr_last = predict ? P(n, FB_F) : 0;
    if (pair_enabled(1) && pair_run(fk, fkm1, pair_shape(fk.bits, wsum3(W_LAST[n & 1])), 0, W_LAST[n & 1], nullptr,
                                   pair_limbs(fk.bits), &result, nullptr, predict ? &pr_last : nullptr)) {
        if (predict) checks.push_back({result.slot, pr_last});
        dnum_free(fk);
        dnum_free(fkm1);
    } else if (n & 1) {
        // F(2k+1) = (2F(k) + F(k-1)) (2F(k) - F(k-1)) + 2(-1)^k, the
        // factors being F(k+2) and L(k-1)
        DNum x, y;
        uint64_t pf[2] = {0, 0};
        if (predict) pf[0] = P(k + 2, FB_F), pf[1] = P(k - 1, FB_L);
        last_factors(fk, fkm1, &x, &y, predict ? pf : nullptr);
        if (predict) checks.push_back({x.slot, pf[0]}), checks.push_back({y.slot, pf[1]});
        dnum_free(fk);
        dnum_free(fkm1);
        const DNum *A[1] = {&x}, *B[1] = {&y};
        uint64_t pr = predict ? P(n, FB_F) : 0;
        mul_many(A, B, &result, 1, predict ? &pr : nullptr, ds.empty() ? nullptr : &ds.back());
        dnum_free(x);
        dnum_free(y);
        const uint64_t len = result.n + 1 <= result.cap ? result.n + 1 : result.cap;
        FF_CHECK(ff_launch_add_small(result.d, len, k_odd ? -2 : 2, g_slots + result.slot, g_st));
        if (predict) {
            checks.push_back({result.slot, pr});
        } else {
            uint64_t kk;
            slots_read(result.slot, 1, &kk);
            key_decode(kk, &result.n, &result.bits);
        }
    } else {
        // F(2k) = F(k) (F(k) + 2F(k-1)), the factor being L(k)
        DNum x;
        uint64_t pf[1] = {0};
        if (predict) pf[0] = P(k, FB_L);
        last_factors(fkm1, fk, &x, nullptr, predict ? pf : nullptr);
        if (predict) checks.push_back({x.slot, pf[0]});
        dnum_free(fkm1);
        const DNum *A[1] = {&fk}, *B[1] = {&x};
        uint64_t pr = predict ? P(n, FB_F) : 0;
        mul_many(A, B, &result, 1, predict ? &pr : nullptr, ds.empty() ? nullptr : &ds.back());
        if (predict) checks.push_back({result.slot, pr});
        dnum_free(x);
        dnum_free(fk);
    }
    *res = result;
    if (!predict || checks.empty()) return true;
    std::vector<uint64_t> keys(g_slot_used);
    slots_read(0, g_slot_used, keys.data());
    for (const auto &c : checks) {
        if (keys[c.slot] != key_of_bits(c.bits)) return false;
    }
    return true;
}

static DNum dfib_checked(uint64_t n) {
    DNum r;
    if (n >= 200 && dfib(n, true, &r)) return r;
    if (n >= 200) {
        dnum_free(r);
        fprintf(stderr, "fastfib_amd: predicted lengths disagreed for F(%llu); recomputing with read-back\n",
                (unsigned long long)n);
    }
    dfib(n, false, &r);
    return r;
}

// a * b for the limb helpers (ff_bigops.h): lengths read back.
FFNum ff_num_mul(const FFNum &a, const FFNum *b) {
    slots_reset(16);
    DNum A{a.d, a.n, a.bits, a.cap, -1}, B{};
    const bool sq = !b || b == &a || b->d == a.d;
    if (!sq) B = DNum{b->d, b->n, b->bits, b->cap, -1};
    const DNum *pa[1] = {&A}, *pb[1] = {sq ? nullptr : &B};
    DNum R;
    mul_many(pa, pb, &R, 1, nullptr);
    return FFNum{R.d, R.n, R.bits, R.cap};
}

// ------------------------------------------------------------------ C-ABI

static size_t norm_len(const uint64_t *d, size_t n) {
    while (n > 1 && d[n - 1] == 0) n--;
    return n ? n : 1;
}

static DNum upload(const uint64_t *h, size_t n) {
    n = h ? norm_len(h, n) : 1;
    DNum x = dnum_alloc(n);
    const uint64_t zero = 0;
    FF_CHECK(hipStreamSynchronize(g_st));
    ff_copy(x.d, h ? h : &zero, n * 8, hipMemcpyHostToDevice);
    x.n = n;
    const uint64_t top = h ? h[n - 1] : 0;
    x.bits = top ? 64 * (n - 1) + (64 - __builtin_clzll(top)) : 0;
    return x;
}

static Big download(DNum &x) {
    Big r;
    r.n = x.n;
    r.d = (uint64_t *)malloc(r.n * 8);
    if (!r.d) {
        fprintf(stderr, "out of memory (%zu bytes)\n", (size_t)(r.n * 8));
        exit(1);
    }
    FF_CHECK(hipStreamSynchronize(g_st));
    ff_copy(r.d, x.d, r.n * 8, hipMemcpyDeviceToHost);
    return r;
}

extern "C" void ff_ntt_init(size_t max_limbs) {
    (void)max_limbs;  // the device tables cover every length up to 2^25 at once
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ff_ensure_init();
}

static Big host_mul(const Big *a, const Big *b, bool square) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ff_ensure_init();
    slots_reset(16);
    DNum A = upload(a->d, a->n), B{};
    if (!square) B = upload(b->d, b->n);
    const DNum *pa[1] = {&A}, *pb[1] = {square ? nullptr : &B};
    DNum R;
    mul_many(pa, pb, &R, 1, nullptr);
    Big r = download(R);
    dnum_free(R);
    dnum_free(A);
    if (!square) dnum_free(B);
    return r;
}

extern "C" int ff_product_plan(uint64_t bits_a, uint64_t bits_b, int square, ff_plan_info *out) {
    if (!out) return FF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    memset(out, 0, sizeof *out);
    if (square) bits_b = bits_a;
    if (bits_a == 0 || bits_b == 0) {
        out->zero = 1;  // written as one zero limb, no transform
        return FF_OK;
    }
    const ProdShape s = product_shape(bits_a, bits_b);
    out->width = s.bw;
    out->length = s.L;
    out->conv = s.conv;
    out->radix3 = s.r3;
    out->split = s.split;
    if (s.split) {
        unsigned bw;
        int lgp;
        ff_split_plan(bits_a, bits_b, square != 0, &bw, &lgp);
        const uint64_t H = 1ull << (lgp - 1);
        const uint64_t ca = (bits_a + bw - 1) / bw, cb = (bits_b + bw - 1) / bw;
        out->width = bw;
        out->length = 1ull << lgp;
        out->conv = ca + cb - 1;
        out->radix3 = 0;
        out->pieces_a = (ca + H - 1) / H;
        out->pieces_b = square ? out->pieces_a : (cb + H - 1) / H;
    }
    return FF_OK;
}

extern "C" Big ff_big_mul(const Big *a, const Big *b) { return host_mul(a, b, a == b); }
extern "C" Big ff_big_sqr(const Big *a) { return host_mul(a, a, true); }
extern "C" Big ff_big_mul_ntt(const Big *a, const Big *b) { return host_mul(a, b, a == b); }

extern "C" Big ff_fibonacci(uint64_t n) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ff_ensure_init();
    DNum R = dfib_checked(n);
    Big r = download(R);
    dnum_free(R);
    return r;
}

extern "C" void ff_big_free(Big *x) {
    free(x->d);
    x->d = nullptr;
    x->n = 0;
}

extern "C" void ff_fib_step_linear(const Big *s1, const Big *s2, int k_odd, int bit, Big *lo, Big *hi) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ff_ensure_init();
    slots_reset(16);
    DNum S1 = upload(s1->d, s1->n), S2 = upload(s2->d, s2->n), L, H;
    FFStepJob J{};
    J.s1l = S1.d, J.s2l = S2.d;
    J.n1 = S1.n, J.n2 = S2.n;
    step_launch(J, S1.n + 1, k_odd, bit, &L, &H, nullptr, 0.0);
    *lo = download(L);
    *hi = download(H);
    dnum_free(S1), dnum_free(S2), dnum_free(L), dnum_free(H);
}

extern "C" void ff_last_factors(const Big *u, const Big *v, Big *x, Big *y) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ff_ensure_init();
    slots_reset(16);
    DNum U = upload(u->d, u->n), V = upload(v->d, v->n), X, Y;
    last_factors(U, V, &X, y ? &Y : nullptr, nullptr);
    *x = download(X);
    dnum_free(X);
    if (y) {
        *y = download(Y);
        dnum_free(Y);
    }
    dnum_free(U), dnum_free(V);
}

static DNum from_dbig(const ff_dbig *x) { return DNum{x->d, x->n, x->bits, x->cap, -1}; }

static void to_dbig(const DNum &x, ff_dbig *o) {
    o->d = x.d;
    o->n = x.n;
    o->bits = x.bits;
    o->cap = x.cap;
}

extern "C" int ff_dbig_alloc(ff_dbig *x, size_t cap) {
    if (!x || cap == 0) return FF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ff_ensure_init();
    DNum d = dnum_alloc(cap);
    FF_CHECK(ff_launch_fill(d.d, 1, 0, g_st));
    to_dbig(d, x);
    return FF_OK;
}

extern "C" void ff_dbig_release(ff_dbig *x) {
    if (!x || !x->d) return;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ff_dfree(x->d);
    x->d = nullptr;
    x->n = x->cap = 0;
    x->bits = 0;
}

extern "C" int ff_dbig_upload(ff_dbig *x, const uint64_t *limbs, size_t n) {
    if (!x || (!limbs && n)) return FF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ff_ensure_init();
    const size_t m = limbs ? norm_len(limbs, n) : 1;
    if (x->d && x->cap >= m) {
        // in place: device pointers held elsewhere (a batch) stay valid
        const uint64_t zero = 0;
        FF_CHECK(hipStreamSynchronize(g_st));
        ff_copy(x->d, limbs ? limbs : &zero, m * 8, hipMemcpyHostToDevice);
        const uint64_t top = limbs ? limbs[m - 1] : 0;
        x->n = m;
        x->bits = top ? 64 * (m - 1) + (64 - __builtin_clzll(top)) : 0;
        return FF_OK;
    }
    ff_dbig_release(x);
    DNum d = upload(limbs, n);
    to_dbig(d, x);
    return FF_OK;
}

extern "C" int ff_dbig_download(const ff_dbig *x, uint64_t *limbs, size_t cap) {
    if (!x || !x->d || !limbs || cap < x->n) return FF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (hipStreamSynchronize(g_st) != hipSuccess) return FF_ERR_DEVICE;
    if (hipMemcpy(limbs, x->d, x->n * 8, hipMemcpyDeviceToHost) != hipSuccess) return FF_ERR_DEVICE;
    return FF_OK;
}

extern "C" int ff_dmul(ff_dbig *out, const ff_dbig *a, const ff_dbig *b) {
    if (!out || !a || !a->d) return FF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ff_ensure_init();
    slots_reset(16);
    DNum A = from_dbig(a), B = b ? from_dbig(b) : DNum{};
    const bool sq = (b == nullptr || b == a || b->d == a->d);
    const DNum *pa[1] = {&A}, *pb[1] = {sq ? nullptr : &B};
    // out's old block goes back to the pool after the product is queued (the
    // pool reuses blocks in stream order, so the product still reads it
    // first) - also when out is the operand itself (out == a squares in
    // place), but not when another handle (a or b) still names that block
    uint64_t *old = out->d;
    const bool keep_old = (out != a && old == a->d) || (b && out != b && old == b->d);
    DNum R;
    mul_many(pa, pb, &R, 1, nullptr);
    to_dbig(R, out);
    if (old && old != R.d && !keep_old) ff_dfree(old);
    return FF_OK;
}

extern "C" int ff_dfibonacci(uint64_t n, ff_dbig *out) {
    if (!out) return FF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ff_ensure_init();
    DNum R = dfib_checked(n);
    to_dbig(R, out);
    return FF_OK;
}

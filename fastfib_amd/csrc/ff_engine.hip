// Host side of fastfib_amd: device memory, product schedules, the product
// pipeline (forward passes -> inner convolution -> inverse passes -> CRT ->
// carries), device-resident fast doubling, and the C-ABI of
// include/fastfib_amd.h.
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/fastfib_amd.h"
#include "ff_internal.h"
#include "ff_limbs.h"

// ------------------------------------------------------------------ state

static int g_inited = 0;
static int g_device = 0;
static hipStream_t g_st = nullptr;
static std::recursive_mutex g_mu;  // the drop-in calls may come from several threads
static char g_err[512];

static int g_inner_log = 12;
static int g_tile_words = 16384;
static uint64_t g_force_len = 0;

static void set_err(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

extern "C" const char *ff_last_error(void) { return g_err; }

// Drop-in failure behaviour: a message and exit(1) (alloc.c:142-146).
extern "C" void ff_fail(const char *file, int line, const char *what, hipError_t e) {
    fprintf(stderr, "fastfib_amd: %s failed at %s:%d: %s\n", what, file, line, hipGetErrorString(e));
    exit(1);
}

// ------------------------------------------------------------------ device memory
// Freed blocks are kept by size and reused (stream-ordered: every user of a
// block runs on g_st, so a block freed after enqueueing its last use can be
// handed to the next caller at once). hipMalloc only on a miss.

static std::multimap<size_t, void *> g_free;
static std::unordered_map<void *, size_t> g_sizes;
static size_t g_pool_bytes = 0;

static size_t size_class(size_t b) {
    if (b < 4096) return 4096;
    // round up to 1/8 of the next power of two
    size_t p = 1;
    while (p < b) p <<= 1;
    size_t step = p / 8;
    return (b + step - 1) / step * step;
}

static void *dalloc(size_t bytes) {
    const size_t sz = size_class(bytes);
    auto it = g_free.lower_bound(sz);
    if (it != g_free.end() && it->first <= sz + sz / 2) {
        void *p = it->second;
        g_free.erase(it);
        return p;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, sz);
    if (e != hipSuccess) {
        // drop the cache and retry once
        (void)hipStreamSynchronize(g_st);
        for (auto &kv : g_free) {
            (void)hipFree(kv.second);
            g_pool_bytes -= kv.first;
            g_sizes.erase(kv.second);
        }
        g_free.clear();
        (void)hipGetLastError();
        e = hipMalloc(&p, sz);
        if (e != hipSuccess) {
            fprintf(stderr, "out of memory (%zu bytes)\n", bytes);
            exit(1);
        }
    }
    g_sizes[p] = sz;
    g_pool_bytes += sz;
    return p;
}

static void dfree(void *p) {
    if (!p) return;
    auto it = g_sizes.find(p);
    if (it == g_sizes.end()) return;
    g_free.emplace(it->second, p);
}

// Pinned staging ring for the small descriptor arrays the kernels read.
static char *g_ring = nullptr;
static const size_t kRingBytes = 8u << 20;
static size_t g_ring_pos = 0;
static uint64_t *g_res_h = nullptr;  // pinned: normalized lengths read back
static const int kResMax = 4096;

template <class T>
static T *stage(const T *src, size_t n) {
    const size_t bytes = (n * sizeof(T) + 255) / 256 * 256;
    if (bytes > kRingBytes) {
        set_err("descriptor array too large");
        return nullptr;
    }
    if (g_ring_pos + bytes > kRingBytes) {
        FF_CHECK(hipStreamSynchronize(g_st));
        g_ring_pos = 0;
    }
    memcpy(g_ring + g_ring_pos, src, n * sizeof(T));
    T *d = (T *)dalloc(bytes);
    FF_CHECK(hipMemcpyAsync(d, g_ring + g_ring_pos, bytes, hipMemcpyHostToDevice, g_st));
    g_ring_pos += bytes;
    return d;
}

extern "C" int ff_init(int device) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (g_inited) return FF_OK;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        set_err("no HIP device");
        return FF_ERR_DEVICE;
    }
    if (device >= 0) {
        if (hipSetDevice(device) != hipSuccess) {
            set_err("hipSetDevice(%d) failed", device);
            return FF_ERR_DEVICE;
        }
    }
    (void)hipGetDevice(&g_device);
    if (hipStreamCreateWithFlags(&g_st, hipStreamNonBlocking) != hipSuccess) {
        set_err("stream creation failed");
        return FF_ERR_DEVICE;
    }
    if (hipHostMalloc((void **)&g_ring, kRingBytes, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&g_res_h, kResMax * 2 * sizeof(uint64_t), hipHostMallocDefault) != hipSuccess) {
        set_err("pinned staging allocation failed");
        return FF_ERR_NOMEM;
    }
    if (ff_tables_init() != 0) {
        set_err("table upload failed");
        return FF_ERR_DEVICE;
    }
    g_inited = 1;
    return FF_OK;
}

static void ensure_init(void) {
    if (!g_inited && ff_init(-1) != FF_OK) {
        fprintf(stderr, "fastfib_amd: %s\n", g_err);
        exit(1);
    }
}

extern "C" void *ff_stream(void) {
    ensure_init();
    return (void *)g_st;
}

extern "C" int ff_sync(void) {
    ensure_init();
    return hipStreamSynchronize(g_st) == hipSuccess ? FF_OK : FF_ERR_DEVICE;
}

extern "C" int ff_set_option(const char *name, long long v) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (!strcmp(name, "inner_log")) {
        if (v < 4 || v > 12) return FF_ERR_ARG;
        g_inner_log = (int)v;
    } else if (!strcmp(name, "tile_words")) {
        if (v < 1024 || v > 16384) return FF_ERR_ARG;
        g_tile_words = (int)v;
    } else if (!strcmp(name, "force_len")) {
        if (v < 0) return FF_ERR_ARG;
        g_force_len = (uint64_t)v;
    } else {
        return FF_ERR_ARG;
    }
    return FF_OK;
}

// ------------------------------------------------------------------ schedule
// The reference's choice of coefficient width and transform length
// (mul_ntt.c:55-118): widths 32..35 while min(ca, cb) * 2^(2w) < p1 p2 p3,
// the shortest L in {2^k, 3 * 2^k} >= ca + cb - 1 (at least 16), the
// narrowest width among equal lengths.

typedef unsigned __int128 u128;

static int width_ok(uint64_t bits_a, uint64_t bits_b, unsigned w) {
    const u128 M = (u128)FF_P1 * FF_P2 * FF_P3;
    const uint64_t ca = (bits_a + w - 1) / w, cb = (bits_b + w - 1) / w;
    return ((u128)(ca < cb ? ca : cb) << (2 * w)) < M;
}

static uint64_t ntt_length(uint64_t conv, int *lg_out, int *r3_out) {
    uint64_t L = 16;
    int lg = 4, r3 = 0;
    while (L < conv) L <<= 1, lg++;
    if (L / 4 * 3 >= conv && L / 4 >= 16) L = L / 4 * 3, lg -= 2, r3 = 1;
    *lg_out = lg;
    *r3_out = r3;
    return L;
}

struct ProdShape {
    unsigned bw;
    uint64_t L;
    int lg, r3;
    uint64_t conv;
    int split;  // the product needs pieces (L > 2^25)
};

static ProdShape product_shape(uint64_t bits_a, uint64_t bits_b) {
    ProdShape s{};
    for (unsigned w = 32; w <= 35; w++) {
        if (!width_ok(bits_a, bits_b, w)) break;
        const uint64_t ca = (bits_a + w - 1) / w, cb = (bits_b + w - 1) / w;
        int lg, r3;
        const uint64_t L = ntt_length(ca + cb - 1, &lg, &r3);
        if (s.L == 0 || L < s.L) s.L = L, s.lg = lg, s.r3 = r3, s.bw = w, s.conv = ca + cb - 1;
    }
    if (s.L == 0 || s.L > (1u << 25)) s.split = 1;
    if (!s.split && g_force_len) {
        // a forced length must be supported (2^k or 3*2^k <= 2^25) and long enough
        uint64_t F = g_force_len;
        int r3 = (F % 3 == 0);
        uint64_t m = r3 ? F / 3 : F;
        if ((m & (m - 1)) == 0 && F >= s.conv && F >= 16 && F <= (1u << 25)) {
            s.L = F;
            s.r3 = r3;
            s.lg = 63 - __builtin_clzll(m);
        }
    }
    return s;
}

// Limbs the CRT writes for a product: whole chunks of 64 coefficients
// (bw limbs each) plus two limbs of final carry, at least na + nb + 1.
static uint64_t out_room(const ProdShape &s, uint64_t na, uint64_t nb) {
    const uint64_t nch = (s.conv + 63) / 64;
    const uint64_t a = nch * s.bw + 2, b = na + nb + 1;
    return a > b ? a : b;
}

// ------------------------------------------------------------------ products

struct MulItem {
    const uint64_t *a;
    uint64_t na, bits_a;
    const uint64_t *b;  // nullptr: square
    uint64_t nb, bits_b;
    uint64_t *out;      // room for out_room() limbs
    uint64_t conv;      // product coefficients at the group's width
};

// Runs products that share one schedule, as one launch per pass.
static void mul_group(MulItem *it, int cnt, const ProdShape &s) {
    FFPlan pl;
    if (ff_plan_build(&pl, (uint32_t)s.L, (int)s.bw, s.lg, s.r3, s.conv, g_inner_log, g_tile_words) != 0) {
        fprintf(stderr, "fastfib_amd: no schedule for L=%llu\n", (unsigned long long)s.L);
        exit(1);
    }
    const size_t words = 3 * (size_t)s.L;
    std::vector<FFXform> fwd, xa(cnt), xb(cnt), inv(cnt);
    std::vector<FFConv> cv(cnt);
    std::vector<FFCrtJob> crt(cnt);
    std::vector<void *> tmp;
    uint64_t max_chunks = 0;
    for (int i = 0; i < cnt; i++) {
        uint32_t *fa = (uint32_t *)dalloc(words * 4);
        uint32_t *fb = it[i].b ? (uint32_t *)dalloc(words * 4) : nullptr;
        tmp.push_back(fa);
        if (fb) tmp.push_back(fb);
        xa[i] = FFXform{(const uint32_t *)it[i].a, (uint32_t)(2 * it[i].na), fa};
        xb[i] = FFXform{it[i].b ? (const uint32_t *)it[i].b : nullptr, (uint32_t)(2 * it[i].nb), fb};
        fwd.push_back(xa[i]);
        if (fb) fwd.push_back(xb[i]);
        cv[i] = FFConv{fa, fb};
        inv[i] = FFXform{nullptr, 0, fa};
        const uint64_t nch = (it[i].conv + 63) / 64;
        max_chunks = nch > max_chunks ? nch : max_chunks;
        FFChunk *ch = (FFChunk *)dalloc(nch * sizeof(FFChunk));
        u128 *cin = (u128 *)dalloc((nch + 1) * sizeof(u128));
        tmp.push_back(ch);
        tmp.push_back(cin);
        const uint64_t nout = it[i].na + (it[i].b ? it[i].nb : it[i].na) + 1;
        crt[i] = FFCrtJob{fa, it[i].out, ch, cin, nout, it[i].conv, (uint32_t)s.L, (int)s.bw};
    }
    FFXform *d_fwd = stage(fwd.data(), fwd.size());
    FFXform *d_xa = stage(xa.data(), cnt);
    FFXform *d_xb = stage(xb.data(), cnt);
    FFXform *d_inv = stage(inv.data(), cnt);
    FFConv *d_cv = stage(cv.data(), cnt);
    FFCrtJob *d_crt = stage(crt.data(), cnt);
    FF_CHECK(ff_launch_forward(&pl, d_fwd, (int)fwd.size(), g_st));
    FF_CHECK(ff_launch_conv(&pl, d_xa, d_xb, d_cv, cnt, g_st));
    FF_CHECK(ff_launch_inverse(&pl, d_inv, cnt, g_st));
    FF_CHECK(ff_launch_crt(d_crt, cnt, max_chunks, g_st));
    for (void *p : tmp) dfree(p);
    dfree(d_fwd), dfree(d_xa), dfree(d_xb), dfree(d_inv), dfree(d_cv), dfree(d_crt);
}

// Reads back the normalized lengths and bit counts of n arrays.
static void read_lengths(const FFNormJob *jobs, int n, uint64_t *res) {
    FFNormJob *dj = stage(jobs, n);
    uint64_t *dres = (uint64_t *)dalloc(2 * n * sizeof(uint64_t));
    FF_CHECK(ff_launch_normlen_multi(dj, n, dres, g_st));
    FF_CHECK(hipMemcpyAsync(g_res_h, dres, 2 * n * sizeof(uint64_t), hipMemcpyDeviceToHost, g_st));
    FF_CHECK(hipStreamSynchronize(g_st));
    memcpy(res, g_res_h, 2 * n * sizeof(uint64_t));
    dfree(dj);
    dfree(dres);
}

struct DNum {
    uint64_t *d;
    uint64_t n, bits, cap;
};

static DNum dnum_alloc(uint64_t cap) {
    DNum x{(uint64_t *)dalloc(cap * 8), 1, 0, cap};
    return x;
}

static DNum dnum_u64(uint64_t v) {
    DNum x = dnum_alloc(1);
    FF_CHECK(hipMemcpyAsync(x.d, &v, 8, hipMemcpyHostToDevice, g_st));
    FF_CHECK(hipStreamSynchronize(g_st));  // v lives on this stack frame
    x.n = 1;
    x.bits = v ? 64 - __builtin_clzll(v) : 0;
    return x;
}

static void dnum_free(DNum &x) {
    dfree(x.d);
    x.d = nullptr;
}

// Products of device numbers: groups the items by schedule, runs each group,
// and reads every result's length. Zero operands give zero.
static void mul_many(const DNum *const *A, const DNum *const *B, DNum *out, int cnt) {
    std::vector<MulItem> items(cnt);
    std::vector<ProdShape> shapes(cnt);
    std::vector<int> zero(cnt, 0);
    for (int i = 0; i < cnt; i++) {
        const DNum *a = A[i], *b = B[i];
        const bool sq = (b == nullptr || b == a);
        const uint64_t bb = sq ? a->bits : b->bits;
        if (a->bits == 0 || bb == 0) {
            zero[i] = 1;
            out[i] = dnum_alloc(1);
            FF_CHECK(hipMemsetAsync(out[i].d, 0, 8, g_st));
            out[i].n = 1, out[i].bits = 0;
            continue;
        }
        shapes[i] = product_shape(a->bits, bb);
        if (shapes[i].split) {
            fprintf(stderr, "fastfib_amd: products past 2^25-word transforms are not supported yet\n");
            exit(1);
        }
        const uint64_t nb = sq ? a->n : b->n;
        const uint64_t room = out_room(shapes[i], a->n, nb);
        out[i] = dnum_alloc(room);
        items[i] = MulItem{a->d, a->n, a->bits, sq ? nullptr : b->d, nb, bb, out[i].d, shapes[i].conv};
    }
    // one group per (length, width): the passes only depend on those, the
    // CRT reads each product's own conv
    std::vector<int> done(cnt, 0);
    for (int i = 0; i < cnt; i++) {
        if (zero[i] || done[i]) continue;
        std::vector<MulItem> grp;
        for (int j = i; j < cnt; j++) {
            if (zero[j] || done[j]) continue;
            if (shapes[j].L == shapes[i].L && shapes[j].bw == shapes[i].bw) {
                grp.push_back(items[j]);
                done[j] = 1;
            }
        }
        mul_group(grp.data(), (int)grp.size(), shapes[i]);
    }
    std::vector<FFNormJob> nj;
    std::vector<int> which;
    for (int i = 0; i < cnt; i++) {
        if (zero[i]) continue;
        const uint64_t nb = items[i].b ? items[i].nb : items[i].na;
        nj.push_back(FFNormJob{out[i].d, items[i].na + nb + 1});
        which.push_back(i);
    }
    if (!nj.empty()) {
        std::vector<uint64_t> res(2 * nj.size());
        read_lengths(nj.data(), (int)nj.size(), res.data());
        for (size_t k = 0; k < which.size(); k++) out[which[k]].n = res[2 * k], out[which[k]].bits = res[2 * k + 1];
    }
}

static void mul2(const DNum *a, const DNum *b, DNum *out) {
    const DNum *A[1] = {a}, *B[1] = {b};
    mul_many(A, B, out, 1);
}

// ------------------------------------------------------------------ fast doubling

// One linear step (fib.c:78-109): from s1 = F(k)^2, s2 = F(k-1)^2 makes
// (lo, hi) = bit ? (F(2k), F(2k+1)) : (F(2k-1), F(2k)).
static void fib_step(const DNum &s1, const DNum &s2, int k_odd, int bit, DNum *lo, DNum *hi) {
    const uint64_t len = s1.n + 1;  // s2 <= s1, and each result is < 4 s1 + 2
    DNum A{}, B{}, C = dnum_alloc(len);
    if (bit) B = dnum_alloc(len);
    else A = dnum_alloc(len);
    const uint64_t nch = (len + FF_CH - 1) / FF_CH;
    FFCarrySum *sums = (FFCarrySum *)dalloc(2 * nch * sizeof(FFCarrySum));
    int8_t *cin = (int8_t *)dalloc(2 * nch);
    FFStepArgs a{s1.d, s2.d, s1.n, s2.n, len, bit ? nullptr : A.d, bit ? B.d : nullptr, C.d, k_odd ? -2 : 2, sums};
    FF_CHECK(ff_launch_step(&a, g_st));
    FFCarryStreams cs{{bit ? B.d : A.d, C.d}, len, nch, 2, sums, cin};
    FF_CHECK(ff_launch_carry(&cs, g_st));
    dfree(sums);
    dfree(cin);
    FFNormJob nj[2] = {{bit ? B.d : A.d, len}, {C.d, len}};
    uint64_t res[4];
    read_lengths(nj, 2, res);
    DNum &K = bit ? B : A;
    K.n = res[0], K.bits = res[1];
    C.n = res[2], C.bits = res[3];
    if (bit) *lo = C, *hi = B;
    else *lo = A, *hi = C;
}

// X = 2u + v and, if Y, Y = 2u - v (fib.c:145-161).
static void last_factors(const DNum &u, const DNum &v, DNum *X, DNum *Y) {
    const uint64_t len = (u.n > v.n ? u.n : v.n) + 1;
    *X = dnum_alloc(len);
    if (Y) *Y = dnum_alloc(len);
    const uint64_t nch = (len + FF_CH - 1) / FF_CH;
    FFCarrySum *sums = (FFCarrySum *)dalloc(2 * nch * sizeof(FFCarrySum));
    int8_t *cin = (int8_t *)dalloc(2 * nch);
    FFLastArgs a{u.d, v.d, u.n, v.n, len, X->d, Y ? Y->d : nullptr, sums};
    FF_CHECK(ff_launch_last(&a, g_st));
    FFCarryStreams cs{{X->d, Y ? Y->d : nullptr}, len, nch, Y ? 2 : 1, sums, cin};
    FF_CHECK(ff_launch_carry(&cs, g_st));
    dfree(sums);
    dfree(cin);
    FFNormJob nj[2] = {{X->d, len}, {Y ? Y->d : X->d, len}};
    uint64_t res[4];
    read_lengths(nj, Y ? 2 : 1, res);
    X->n = res[0], X->bits = res[1];
    if (Y) Y->n = res[2], Y->bits = res[3];
}

// fibonacci (fib.c:211-261) with every product and pass on the device: the
// pair (F(k-1), F(k)) stays in HBM; each bit costs two squarings run as one
// launch per pass (same schedule) and one linear step.
static DNum dfib(uint64_t n) {
    if (n == 0) return dnum_u64(0);
    DNum fkm1 = dnum_u64(0), fk = dnum_u64(1);
    int k_odd = 1;
    const int top = 63 - __builtin_clzll(n);
    for (int i = top - 1; i >= 1; i--) {
        DNum sq[2];
        const DNum *A[2] = {&fk, &fkm1}, *B[2] = {nullptr, nullptr};
        mul_many(A, B, sq, 2);
        dnum_free(fk);
        dnum_free(fkm1);
        const int bit = (int)((n >> i) & 1);
        fib_step(sq[0], sq[1], k_odd, bit, &fkm1, &fk);
        dnum_free(sq[0]);
        dnum_free(sq[1]);
        k_odd = bit;
    }
    if (top == 0) {
        dnum_free(fkm1);
        return fk;
    }
    DNum result;
    if (n & 1) {
        DNum x, y;
        last_factors(fk, fkm1, &x, &y);
        dnum_free(fk);
        dnum_free(fkm1);
        mul2(&x, &y, &result);
        dnum_free(x);
        dnum_free(y);
        FF_CHECK(ff_launch_add_small(result.d, result.n + 1, k_odd ? -2 : 2, g_st));
        FFNormJob nj = {result.d, result.n + 1};
        uint64_t res[2];
        read_lengths(&nj, 1, res);
        result.n = res[0], result.bits = res[1];
    } else {
        DNum x;
        last_factors(fkm1, fk, &x, nullptr);
        dnum_free(fkm1);
        mul2(&fk, &x, &result);
        dnum_free(x);
        dnum_free(fk);
    }
    return result;
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
    FF_CHECK(hipMemcpy(x.d, h ? h : &zero, n * 8, hipMemcpyHostToDevice));
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
    FF_CHECK(hipMemcpy(r.d, x.d, r.n * 8, hipMemcpyDeviceToHost));
    return r;
}

extern "C" void ff_ntt_init(size_t max_limbs) {
    (void)max_limbs;  // the device tables cover every length up to 2^25 at once
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ensure_init();
}

static Big host_mul(const Big *a, const Big *b, bool square) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ensure_init();
    DNum A = upload(a->d, a->n), B{};
    if (!square) B = upload(b->d, b->n);
    DNum R;
    mul2(&A, square ? nullptr : &B, &R);
    Big r = download(R);
    dnum_free(R);
    dnum_free(A);
    if (!square) dnum_free(B);
    return r;
}

extern "C" Big ff_big_mul(const Big *a, const Big *b) { return host_mul(a, b, a == b); }
extern "C" Big ff_big_sqr(const Big *a) { return host_mul(a, a, true); }
extern "C" Big ff_big_mul_ntt(const Big *a, const Big *b) { return host_mul(a, b, a == b); }

extern "C" Big ff_fibonacci(uint64_t n) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ensure_init();
    DNum R = dfib(n);
    Big r = download(R);
    dnum_free(R);
    return r;
}

extern "C" void ff_big_free(Big *x) {
    free(x->d);
    x->d = nullptr;
    x->n = 0;
}

static DNum from_dbig(const ff_dbig *x) { return DNum{x->d, x->n, x->bits, x->cap}; }

static void to_dbig(const DNum &x, ff_dbig *o) {
    o->d = x.d;
    o->n = x.n;
    o->bits = x.bits;
    o->cap = x.cap;
}

extern "C" int ff_dbig_alloc(ff_dbig *x, size_t cap) {
    if (!x || cap == 0) return FF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ensure_init();
    DNum d = dnum_alloc(cap);
    FF_CHECK(hipMemsetAsync(d.d, 0, 8, g_st));
    to_dbig(d, x);
    return FF_OK;
}

extern "C" void ff_dbig_release(ff_dbig *x) {
    if (!x || !x->d) return;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    dfree(x->d);
    x->d = nullptr;
    x->n = x->cap = 0;
    x->bits = 0;
}

extern "C" int ff_dbig_upload(ff_dbig *x, const uint64_t *limbs, size_t n) {
    if (!x || (!limbs && n)) return FF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ensure_init();
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
    ensure_init();
    DNum A = from_dbig(a), B = b ? from_dbig(b) : DNum{};
    const bool sq = (b == nullptr || b == a || b->d == a->d);
    DNum R;
    mul2(&A, sq ? nullptr : &B, &R);
    if (out->d && out->d != a->d && (!b || out->d != b->d)) ff_dbig_release(out);
    to_dbig(R, out);
    return FF_OK;
}

extern "C" int ff_dfibonacci(uint64_t n, ff_dbig *out) {
    if (!out) return FF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ensure_init();
    DNum R = dfib(n);
    to_dbig(R, out);
    return FF_OK;
}

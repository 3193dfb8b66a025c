// Schedule choice for products: the reference rule, the autotuner's table
// or the learned dispatcher.
//
// A schedule candidate is (transform length: the shortest or the next
// supported one up, which swaps the radix-3 layer in or out; inner block
// 2^10 or 2^12 words; outer tile 2048, 4096, 8192 or 16384 LDS words). The
// autotuner (ff_tune_measure) times every candidate for a product shape as
// the per-product device time of a grouped batch of copies of it, which is
// the product's marginal cost inside a batch of that size. Its table maps a
// size class (shortest length, radix, square, batch size) to the fastest
// candidate; the learned dispatcher predicts every candidate's relative
// log-time from the product's features and takes the argmin.
#include <dlfcn.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <chrono>
#include <mutex>
#include <vector>

#include "../../include/fastfib_amd.h"
#include "ff_engine.h"
#include "ff_internal.h"
#include "ff_limbs.h"
#include "ff_mlp.h"
#include "ff_runtime.h"

namespace {

// Candidates 0..15 are single-transform schedules; 16..20 serve products
// past the longest transform (the split multiply of mul_split.c): piece
// transforms of 2^(max_lg - 4) .. 2^max_lg words, the range the
// reference's cost model chooses from (mul_split.c:160-170).
constexpr int K_SINGLE = 16, K_SPLIT = 5, K_CAND = K_SINGLE + K_SPLIT;
constexpr int kInner[2] = {10, 12};
constexpr int kTile[4] = {2048, 4096, 8192, 16384};
constexpr int K_DEFAULT = 5;  // the library default: shortest length (the reference rule), inner 2^12, tile 4096

FFSched cand_sched(int k) {
    if (k >= K_SINGLE) return FFSched{0, 0, 0, g_max_lg - 4 + (k - K_SINGLE)};
    return FFSched{k / 8, kInner[(k / 4) % 2], kTile[k % 4], 0};
}

// size classes of the table: lg of the shortest length (4..28; above 25
// the product splits), radix-3, square, floor(log2(batch size)) clamped to
// 0..12
constexpr int T_LG = 25, T_CNT = 13, T_SIZE = T_LG * 2 * 2 * T_CNT;
std::vector<int32_t> g_table;  // empty: not loaded
bool g_auto_tried = false;

int lg_count(size_t count) {
    int c = 0;
    while (c < T_CNT - 1 && (2ull << c) <= count) c++;
    return c;
}

int table_class(const ProdShape &s, bool sq, size_t count) {
    const int lg = s.lg < 4 ? 4 : (s.lg > 28 ? 28 : s.lg);
    return (((lg - 4) * 2 + s.r3) * 2 + (sq ? 1 : 0)) * T_CNT + lg_count(count);
}

// HBM available to the library as a fraction of the whole: free memory
// plus the blocks its pool holds for reuse. The driver query is made at
// most once a second (it costs more than an inference).
double free_hbm_fraction() {
    static size_t fr = 0, tot = 0;
    static std::chrono::steady_clock::time_point when;
    const auto now = std::chrono::steady_clock::now();
    if (tot == 0 || now - when >= std::chrono::seconds(1)) {
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = 0, tot = 0;
        when = now;
    }
    if (tot == 0) return 1.0;
    const double f = (double)(fr + ff_pool_cached()) / (double)tot;
    return f < 1.0 ? f : 1.0;
}

void features(uint64_t bits_a, uint64_t bits_b, bool sq, size_t count, const ProdShape &s, double hbm, float *x) {
    x[0] = (float)log2((double)bits_a);
    x[1] = (float)log2((double)(sq ? bits_a : bits_b));
    x[2] = sq ? 1.0f : 0.0f;
    x[3] = (float)log2((double)(count ? count : 1));
    x[4] = (float)hbm;
    x[5] = (float)log2((double)s.L);
    x[6] = (float)s.r3;
    x[7] = (float)((int)s.bw - 32);
}

// Candidates that apply to a shape: a longer length must stay supported;
// split products take the piece-length candidates (pieces of at least 2^10
// words), the others the single-transform ones.
uint32_t cand_mask(const ProdShape &s) {
    uint32_t m = 0;
    if (s.split) {
        for (int k = K_SINGLE; k < K_CAND; k++)
            if (cand_sched(k).split_lg >= 10) m |= 1u << k;
        return m;
    }
    for (int k = 0; k < K_SINGLE; k++) {
        const FFSched sc = cand_sched(k);
        if (sc.len_up && ff_shape_with(s, sc).L == s.L) continue;
        m |= 1u << k;
    }
    return m;
}

// The reference rule's candidate for a shape: K_DEFAULT, or for a split
// product the piece length of the reference's cost model.
int rule_candidate(uint64_t bits_a, uint64_t bits_b, bool sq, const ProdShape &s) {
    if (!s.split) return K_DEFAULT;
    unsigned bw;
    int lgp;
    ff_split_plan(bits_a, bits_b, sq, &bw, &lgp);
    const int k = K_SINGLE + lgp - (g_max_lg - 4);
    return k >= K_SINGLE && k < K_CAND ? k : K_CAND - 1;
}

// The dispatcher file: "FFDISP1\0", int32 K, int32 table entries, the
// table, int32 parameter count, the parameters (float32).
int load_file(const char *path) {
    FILE *f = fopen(path, "rb");
    if (!f) {
        ff_set_err("cannot open %s", path);
        return FF_ERR_ARG;
    }
    char magic[8];
    int32_t K = 0, nt = 0, np = 0;
    int rc = FF_ERR_ARG;
    std::vector<int32_t> table;
    std::vector<float> params;
    if (fread(magic, 1, 8, f) == 8 && memcmp(magic, "FFDISP1", 8) == 0 && fread(&K, 4, 1, f) == 1 &&
        fread(&nt, 4, 1, f) == 1 && nt == T_SIZE) {
        table.resize(nt);
        if (fread(table.data(), 4, nt, f) == (size_t)nt && fread(&np, 4, 1, f) == 1 && np >= 0) {
            params.resize(np);
            if (fread(params.data(), 4, np, f) == (size_t)np) rc = FF_OK;
        }
    }
    fclose(f);
    if (rc != FF_OK || K != K_CAND) {
        ff_set_err("%s is not a dispatcher file for %d candidates", path, K_CAND);
        return FF_ERR_ARG;
    }
    g_table = table;
    g_dispatch_gen++;
    if (np && ff_mlp_set(K, params.data(), params.size()) != FF_OK) return FF_ERR_ARG;
    return FF_OK;
}

// dispatch.bin next to the library, loaded on first use of a tuned or
// learned schedule when nothing was set explicitly.
void auto_load() {
    if (g_auto_tried) return;
    g_auto_tried = true;
    if (!g_table.empty() && ff_mlp_ready()) return;
    Dl_info info;
    if (!dladdr((void *)&auto_load, &info) || !info.dli_fname) return;
    std::string p(info.dli_fname);
    const size_t slash = p.rfind('/');
    p = (slash == std::string::npos ? std::string(".") : p.substr(0, slash)) + "/dispatch.bin";
    FILE *f = fopen(p.c_str(), "rb");
    if (!f) return;
    fclose(f);
    (void)load_file(p.c_str());
}

}  // namespace

void ff_pick_scheds(const ProdShape *s0, const uint64_t *bits_a, const uint64_t *bits_b, const uint8_t *square,
                    size_t n, size_t count, int flags, FFSched *out) {
    for (size_t i = 0; i < n; i++) out[i] = FFSched{0, 0, 0};
    if (!(flags & (FF_BATCH_TUNED | FF_BATCH_LEARNED)) || n == 0) return;
    auto_load();
    if ((flags & FF_BATCH_LEARNED) && ff_mlp_ready() == K_CAND) {
        const double hbm = free_hbm_fraction();
        std::vector<float> x(n * FF_MLP_IN);
        std::vector<uint32_t> mask(n);
        for (size_t i = 0; i < n; i++) {
            features(bits_a[i], bits_b[i], square[i], count, s0[i], hbm, &x[i * FF_MLP_IN]);
            mask[i] = cand_mask(s0[i]);
        }
        // features, masks and choices staged through pinned memory
        const size_t words = n * (FF_MLP_IN + 2);
        uint64_t *h = ff_pinned_results((words + 1) / 2);
        float *hx = (float *)h;
        uint32_t *hm = (uint32_t *)(hx + n * FF_MLP_IN);
        memcpy(hx, x.data(), x.size() * 4);
        memcpy(hm, mask.data(), n * 4);
        char *buf = (char *)ff_dalloc(words * 4 + 512);
        float *dx = (float *)buf;
        uint32_t *dm = (uint32_t *)(dx + n * FF_MLP_IN);
        int *dc = (int *)(dm + n);
        FF_CHECK(hipMemcpyAsync(dx, hx, (n * FF_MLP_IN + n) * 4, hipMemcpyHostToDevice, g_st));
        std::vector<int> choice(n, -1);
        if (ff_mlp_infer_dev(dx, n, dm, nullptr, dc, FF_MLP_MARGIN, K_DEFAULT) == FF_OK) {
            int *hc = (int *)(hm + n);
            FF_CHECK(hipMemcpyAsync(hc, dc, n * 4, hipMemcpyDeviceToHost, g_st));
            FF_CHECK(hipStreamSynchronize(g_st));
            memcpy(choice.data(), hc, n * 4);
        }
        ff_dfree(buf);
        // split products keep the reference's cost model: on the GPU it was
        // within 0.05% of the exhaustive optimum on held-out samples, while
        // the network's features do not resolve the piece-count rounding it
        // follows (its picks cost 1.9% there)
        for (size_t i = 0; i < n; i++) {
            const int k = !s0[i].split && choice[i] >= 0 && ((mask[i] >> choice[i]) & 1)
                              ? choice[i]
                              : rule_candidate(bits_a[i], bits_b[i], square[i], s0[i]);
            out[i] = cand_sched(k);
        }
        return;
    }
    if (g_table.size() == (size_t)T_SIZE) {
        for (size_t i = 0; i < n; i++) {
            int k = g_table[table_class(s0[i], square[i], count)];
            if (k < 0 || k >= K_CAND || !((cand_mask(s0[i]) >> k) & 1))
                k = rule_candidate(bits_a[i], bits_b[i], square[i], s0[i]);
            out[i] = cand_sched(k);
        }
    }
}

// ------------------------------------------------------------------ C-ABI

extern "C" int ff_tune_candidates(void) { return K_CAND; }

extern "C" int ff_tune_candidate(int k, int *len_up, int *inner_log, int *tile_words) {
    if (k < 0 || k >= K_CAND) return FF_ERR_ARG;
    const FFSched s = cand_sched(k);
    if (len_up) *len_up = s.len_up;
    if (inner_log) *inner_log = s.split_lg ? -s.split_lg : s.inner_log;
    if (tile_words) *tile_words = s.tile_words;
    return FF_OK;
}

extern "C" int ff_tune_default_candidate(void) { return K_DEFAULT; }

extern "C" int ff_tune_features(uint64_t bits_a, uint64_t bits_b, int square, size_t count, float *x,
                                uint32_t *mask) {
    if (!x || bits_a == 0 || (!square && bits_b == 0)) return FF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ff_ensure_init();
    const ProdShape s = ff_product_shape(bits_a, square ? bits_a : bits_b);
    features(bits_a, bits_b, square != 0, count, s, free_hbm_fraction(), x);
    if (mask) *mask = cand_mask(s);
    return FF_OK;
}

extern "C" int ff_tune_class(uint64_t bits_a, uint64_t bits_b, int square, size_t count) {
    if (bits_a == 0 || (!square && bits_b == 0)) return FF_ERR_ARG;
    const ProdShape s = ff_product_shape(bits_a, square ? bits_a : bits_b);
    return table_class(s, square != 0, count);
}

extern "C" int ff_tune_rule_candidate(uint64_t bits_a, uint64_t bits_b, int square) {
    if (bits_a == 0 || (!square && bits_b == 0)) return FF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    const uint64_t bb = square ? bits_a : bits_b;
    return rule_candidate(bits_a, bb, square != 0, ff_product_shape(bits_a, bb));
}

extern "C" int ff_tune_table_size(void) { return T_SIZE; }

extern "C" int ff_tune_set_table(const int32_t *table, size_t n) {
    if (n != (size_t)T_SIZE || !table) return FF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    g_table.assign(table, table + n);
    g_dispatch_gen++;
    return FF_OK;
}

extern "C" int ff_dispatch_load(const char *path) {
    if (!path) return FF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    g_auto_tried = true;
    return load_file(path);
}

// Uniform limbs of exactly `bits` bits (top bit set) from a counter hash.
__global__ void k_fill_operand(uint64_t *d, uint64_t n, uint64_t bits, uint64_t seed) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;  // splitmix64 (Steele, Lea, Flood 2014)
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    if (i == n - 1) {
        const unsigned top = (unsigned)(bits - 64 * (n - 1));
        if (top < 64) z &= (1ull << top) - 1;
        z |= 1ull << (top - 1);
    }
    d[i] = z;
}

extern "C" int ff_tune_measure(uint64_t bits_a, uint64_t bits_b, int square, size_t copies, int reps,
                               uint64_t seed, float *us) {
    if (!us || bits_a == 0 || (!square && bits_b == 0) || copies == 0 || copies > 65536 || reps < 1)
        return FF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ff_ensure_init();
    const uint64_t bb = square ? bits_a : bits_b;
    const ProdShape s0 = ff_product_shape(bits_a, bb);
    if (s0.split) copies = 1;  // one product past the longest transform fills the GPU
    const uint64_t na = (bits_a + 63) / 64, nb = (bb + 63) / 64;
    std::vector<void *> mem;
    std::vector<ff_prod> items(copies);
    for (size_t c = 0; c < copies; c++) {
        uint64_t *a = (uint64_t *)ff_dalloc(na * 8);
        hipLaunchKernelGGL(k_fill_operand, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, g_st, a, na, bits_a,
                           seed + 2 * c);
        uint64_t *b = nullptr;
        if (!square) {
            b = (uint64_t *)ff_dalloc(nb * 8);
            hipLaunchKernelGGL(k_fill_operand, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, g_st, b, nb, bb,
                               seed + 2 * c + 1);
        }
        uint64_t *o = (uint64_t *)ff_dalloc((na + nb + 1) * 8);
        mem.push_back(a), mem.push_back(o);
        if (b) mem.push_back(b);
        items[c] = ff_prod{a, na, b, square ? 0 : nb, o, 0, bits_a, square ? 0 : bb};
    }
    FF_CHECK(hipGetLastError());
    hipEvent_t e0, e1;
    FF_CHECK(hipEventCreate(&e0));
    FF_CHECK(hipEventCreate(&e1));
    const uint32_t mask = cand_mask(s0);
    std::vector<FFSched> sc(copies);
    int rc = FF_OK;
    for (int k = 0; k < K_CAND && rc == FF_OK; k++) {
        if (!((mask >> k) & 1)) {
            us[k] = NAN;
            continue;
        }
        if (s0.split) {
            // split products run alone (batches hold single transforms only):
            // one warm run, then reps timed runs of the forced piece length
            unsigned long long *key = (unsigned long long *)ff_dalloc(8);
            MulItem it{items[0].a, na, bits_a, square ? nullptr : items[0].b, nb, bb, items[0].out, s0.conv, key};
            const int lg = cand_sched(k).split_lg;
            ff_mul_split(it, lg);
            FF_CHECK(hipEventRecord(e0, g_st));
            for (int r = 0; r < reps; r++) ff_mul_split(it, lg);
            FF_CHECK(hipEventRecord(e1, g_st));
            FF_CHECK(hipEventSynchronize(e1));
            float ms = 0;
            FF_CHECK(hipEventElapsedTime(&ms, e0, e1));
            us[k] = ms * 1000.0f / (float)reps;
            ff_dfree(key);
            continue;
        }
        for (auto &x : sc) x = cand_sched(k);
        ff_batch *B = nullptr;
        rc = ff_batch_create_sched(&B, items.data(), copies, FF_BATCH_GROUPED, sc.data());
        if (rc != FF_OK) break;
        // warm run (code objects, caches), timed to size the repetitions:
        // at least reps, more for short runs, up to about a millisecond
        FF_CHECK(hipEventRecord(e0, g_st));
        ff_batch_run(B);
        FF_CHECK(hipEventRecord(e1, g_st));
        FF_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        FF_CHECK(hipEventElapsedTime(&ms, e0, e1));
        int nrep = reps;
        if (ms > 0 && ms < 1.0f) {
            const int want = (int)(1.0f / ms);
            nrep = want > nrep ? (want < 32 ? want : 32) : nrep;
        }
        FF_CHECK(hipEventRecord(e0, g_st));
        for (int r = 0; r < nrep; r++) ff_batch_run(B);
        FF_CHECK(hipEventRecord(e1, g_st));
        FF_CHECK(hipEventSynchronize(e1));
        FF_CHECK(hipEventElapsedTime(&ms, e0, e1));
        us[k] = ms * 1000.0f / (float)nrep / (float)copies;
        ff_batch_destroy(B);
    }
    FF_CHECK(hipEventDestroy(e0));
    FF_CHECK(hipEventDestroy(e1));
    for (void *p : mem) ff_dfree(p);
    return rc;
}

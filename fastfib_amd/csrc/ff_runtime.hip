// Runtime of fastfib_amd: device selection and stream, the device memory
// pool, the live kernel timer, options and error reporting.
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
#include "ff_runtime.h"

static int g_inited = 0;
static int g_device = 0;
hipStream_t g_st = nullptr;
std::recursive_mutex g_mu;  // the drop-in calls may come from several threads
static char g_err[512];

int g_inner_log = 12;
int g_tile_words = 4096;
int g_fib_tile_words = 2048;
int g_outer_rl = 0;
int g_pair_rl = 0;
long long g_crt_wave_limbs = 1280ll * 256;
int g_outer_rl3 = (1 << 5) | (1 << 6) | (1 << 9);
uint64_t g_force_len = 0;
int g_dispatch = 0;
int g_max_lg = 25;
int g_pair_steps = 2;
int g_quad_max_lg = 18;
int g_quad_inner_log = 10;
int g_host_bits = 13;
int g_queue_waves = 4;
int g_dec_queue = 0;
int g_queue_parts = 8;
int g_queue_chunk = 8;
int g_school_limbs = 32;  // measured: the one-block NTT is as fast or faster from 32 limbs up
int g_dispatch_gen = 0;

void ff_set_err(const char *fmt, ...) {
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

// Every device or pinned-host allocation of the library goes through
// ff_hip_malloc / ff_host_malloc, which count them (a caller checks that
// repeated work allocates nothing new) and can be told to fail (the
// "test_fail_alloc" option: after that many more allocations the next ones
// report out-of-memory, so the failure path runs without exhausting HBM).
static uint64_t g_n_dev_mallocs = 0, g_n_host_mallocs = 0;
static long long g_fail_after = -1;

static bool alloc_injected_failure(void) {
    if (g_fail_after < 0) return false;
    if (g_fail_after == 0) return true;
    g_fail_after--;
    return false;
}

hipError_t ff_hip_malloc(void **p, size_t bytes) {
    if (alloc_injected_failure()) return hipErrorOutOfMemory;
    const hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess) g_n_dev_mallocs++;
    return e;
}

hipError_t ff_host_malloc(void **p, size_t bytes) {
    if (alloc_injected_failure()) return hipErrorOutOfMemory;
    const hipError_t e = hipHostMalloc(p, bytes, hipHostMallocDefault);
    if (e == hipSuccess) g_n_host_mallocs++;
    return e;
}

static std::multimap<size_t, void *> g_free;
static std::unordered_map<void *, size_t> g_sizes;
static size_t g_cached = 0;  // bytes held in g_free

size_t ff_pool_cached(void) { return g_cached; }

static size_t size_class(size_t b) {
    if (b < 4096) return 4096;
    size_t p = 1;
    while (p < b) p <<= 1;
    const size_t step = p / 8;  // within 1/8 of the next power of two
    return (b + step - 1) / step * step;
}

void *ff_dalloc(size_t bytes) {
    const size_t sz = size_class(bytes);
    auto it = g_free.lower_bound(sz);
    if (it != g_free.end() && it->first <= sz + sz / 2) {
        void *p = it->second;
        g_cached -= it->first;
        g_free.erase(it);
        return p;
    }
    void *p = nullptr;
    hipError_t e = ff_hip_malloc(&p, sz);
    if (e != hipSuccess) {
        // give the cached blocks back and retry once
        (void)hipStreamSynchronize(g_st);
        for (auto &kv : g_free) {
            (void)hipFree(kv.second);
            g_sizes.erase(kv.second);
        }
        g_free.clear();
        g_cached = 0;
        (void)hipGetLastError();
        e = ff_hip_malloc(&p, sz);
        if (e != hipSuccess) {
            fprintf(stderr, "out of memory (%zu bytes)\n", bytes);
            exit(1);
        }
    }
    g_sizes[p] = sz;
    return p;
}

void ff_dfree(void *p) {
    if (!p) return;
    auto it = g_sizes.find(p);
    if (it == g_sizes.end()) return;
    g_free.emplace(it->second, p);
    g_cached += it->second;
}

void ff_copy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
    if (!bytes) return;
    FF_CHECK(hipMemcpyAsync(dst, src, bytes, kind, g_st));
    FF_CHECK(hipStreamSynchronize(g_st));
}

// ------------------------------------------------------------------ look-back workspace
// Zeroed device words for the single-pass carry kernels (ff_lookback.h):
// every job's region is left zeroed by its last workgroup, so regions are
// handed out from a bump pointer that restarts with each launch set. A
// larger need allocates a new zeroed buffer; the old ones stay allocated
// (captured graphs may still point into them). One separate word collects
// look-back time-outs.

static uint32_t *g_lb = nullptr, *g_lb_fault = nullptr;
static size_t g_lb_cap = 0, g_lb_off = 0;
static std::vector<uint32_t *> g_lb_old;

void ff_lb_begin(void) { g_lb_off = 0; }

uint32_t *ff_lb_take(size_t words) {
    words = (words + 15) / 16 * 16;
    if (g_lb_off + words > g_lb_cap) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        FF_CHECK(hipStreamIsCapturing(g_st, &cs));
        if (cs != hipStreamCaptureStatusNone) {
            fprintf(stderr, "fastfib_amd: look-back workspace grown during graph capture\n");
            exit(1);
        }
        size_t cap = g_lb_cap ? 2 * g_lb_cap : (1u << 20);
        while (cap < words) cap *= 2;
        uint32_t *p = nullptr;
        FF_CHECK(ff_hip_malloc((void **)&p, cap * 4));
        FF_CHECK(hipMemsetAsync(p, 0, cap * 4, g_st));
        if (g_lb) g_lb_old.push_back(g_lb);
        g_lb = p, g_lb_cap = cap, g_lb_off = 0;
    }
    uint32_t *r = g_lb + g_lb_off;
    g_lb_off += words;
    return r;
}

// Workgroups of a kernel that the GPU holds at once (occupancy per CU times
// the CUs), memoized per kernel.
uint64_t ff_resident_blocks(const void *kernel, int threads) {
    static std::map<const void *, uint64_t> memo;
    auto it = memo.find(kernel);
    if (it != memo.end()) return it->second;
    int per_cu = 0, cus = 0;
    FF_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0));
    FF_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g_device));
    const uint64_t r = (uint64_t)per_cu * (uint64_t)cus;
    memo[kernel] = r;
    return r;
}

uint32_t *ff_lb_fault(void) {
    if (!g_lb_fault) {
        FF_CHECK(ff_hip_malloc((void **)&g_lb_fault, 256));
        FF_CHECK(hipMemsetAsync(g_lb_fault, 0, 256, g_st));
    }
    return g_lb_fault;
}

void ff_lb_timeout_fail(void) {
    fprintf(stderr, "fastfib_amd: a carry look-back wait timed out\n");
    exit(1);
}

static uint64_t *g_res_h = nullptr, *g_res_d = nullptr;
static size_t g_res_words = 0;

uint64_t *ff_pinned_results(size_t words) {
    if (words > g_res_words) {
        if (g_res_h) (void)hipHostFree(g_res_h);
        g_res_words = words < 4096 ? 4096 : words;
        FF_CHECK(ff_host_malloc((void **)&g_res_h, g_res_words * sizeof(uint64_t)));
        FF_CHECK(hipHostGetDevicePointer((void **)&g_res_d, g_res_h, 0));
    }
    return g_res_h;
}

uint64_t *ff_pinned_results_dev(void) { return g_res_d; }

extern "C" int ff_mem_stats(uint64_t *device_allocs, uint64_t *host_allocs, size_t *cached_bytes) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (device_allocs) *device_allocs = g_n_dev_mallocs;
    if (host_allocs) *host_allocs = g_n_host_mallocs;
    if (cached_bytes) *cached_bytes = g_cached;
    return FF_OK;
}

extern "C" int ff_test_set_lb_fault(void) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ff_ensure_init();
    const uint32_t one = 1;
    if (hipMemcpyAsync(ff_lb_fault(), &one, 4, hipMemcpyHostToDevice, g_st) != hipSuccess ||
        hipStreamSynchronize(g_st) != hipSuccess)
        return FF_ERR_DEVICE;
    return FF_OK;
}

// ------------------------------------------------------------------ live kernel timer
// When enabled, every launch is bracketed by two HIP events on its stream;
// ff_prof_read syncs and books the elapsed time and the launch's
// algorithmic bytes per kernel kind.

struct ProfRec {
    int kind;
    double bytes, mults;
    hipEvent_t a, b;
};
static int g_prof_on = 0;
static std::vector<ProfRec> g_prof_pending;
static std::vector<hipEvent_t> g_prof_spare;
static double g_prof_ms[FK_COUNT], g_prof_bytes[FK_COUNT], g_prof_mults[FK_COUNT];
static double g_prof_next_mults = 0;  // the open launch's algorithmic 32-bit multiplies
static long g_prof_n[FK_COUNT];
static hipEvent_t g_prof_open = nullptr;
static int g_prof_paused = 0;

void ff_prof_pause(int pause) { g_prof_paused = pause; }

static hipEvent_t prof_event(void) {
    if (!g_prof_spare.empty()) {
        hipEvent_t e = g_prof_spare.back();
        g_prof_spare.pop_back();
        return e;
    }
    hipEvent_t e;
    FF_CHECK(hipEventCreate(&e));
    return e;
}

extern "C" void ff_prof_begin(int kind, hipStream_t st) {
    (void)kind;
    if (!g_prof_on || g_prof_paused) return;
    g_prof_open = prof_event();
    FF_CHECK(hipEventRecord(g_prof_open, st));
}

extern "C" void ff_prof_mults(double mults) { g_prof_next_mults = mults; }

extern "C" void ff_prof_end(int kind, double bytes, hipStream_t st) {
    const double mults = g_prof_next_mults;
    g_prof_next_mults = 0;
    if (!g_prof_on || g_prof_paused || !g_prof_open) return;
    hipEvent_t e = prof_event();
    FF_CHECK(hipEventRecord(e, st));
    g_prof_pending.push_back(ProfRec{kind, bytes, mults, g_prof_open, e});
    g_prof_open = nullptr;
}

static void prof_collect(void) {
    if (g_prof_pending.empty()) return;
    FF_CHECK(hipStreamSynchronize(g_st));
    for (auto &r : g_prof_pending) {
        float ms = 0;
        FF_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
        g_prof_ms[r.kind] += ms;
        g_prof_bytes[r.kind] += r.bytes;
        g_prof_mults[r.kind] += r.mults;
        g_prof_n[r.kind]++;
        g_prof_spare.push_back(r.a);
        g_prof_spare.push_back(r.b);
    }
    g_prof_pending.clear();
}

static const char *kKindNames[FK_COUNT] = {"ntt_fwd_top", "ntt_fwd", "ntt_inner", "ntt_small", "ntt_inv",
                                           "ntt_inv_top", "crt_pack", "crt_scan", "crt_apply", "fib_step",
                                           "carry", "normlen", "bigops", "dec_leaf", "fib_early", "step_crt",
                                           "school", "batch_queue"};

extern "C" int ff_prof_enable(int on) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ff_ensure_init();
    prof_collect();
    g_prof_on = on;
    return FF_OK;
}

extern "C" void ff_prof_reset(void) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ff_ensure_init();
    prof_collect();
    memset(g_prof_ms, 0, sizeof g_prof_ms);
    memset(g_prof_bytes, 0, sizeof g_prof_bytes);
    memset(g_prof_mults, 0, sizeof g_prof_mults);
    memset(g_prof_n, 0, sizeof g_prof_n);
}

extern "C" int ff_prof_read_mults(int k, double *mults) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (k < 0 || k >= FK_COUNT) return FF_ERR_ARG;
    ff_ensure_init();
    prof_collect();
    *mults = g_prof_mults[k];
    return FF_OK;
}

extern "C" int ff_prof_read(int k, const char **name, long *launches, double *ms, double *bytes) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (k < 0 || k >= FK_COUNT) return FF_ERR_ARG;
    ff_ensure_init();
    prof_collect();
    *name = kKindNames[k];
    *launches = g_prof_n[k];
    *ms = g_prof_ms[k];
    *bytes = g_prof_bytes[k];
    return FF_OK;
}

// ------------------------------------------------------------------ init

extern "C" int ff_init(int device) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (g_inited) return FF_OK;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        ff_set_err("no HIP device");
        return FF_ERR_DEVICE;
    }
    if (device >= 0 && hipSetDevice(device) != hipSuccess) {
        ff_set_err("hipSetDevice(%d) failed", device);
        return FF_ERR_DEVICE;
    }
    (void)hipGetDevice(&g_device);
    if (hipStreamCreateWithFlags(&g_st, hipStreamNonBlocking) != hipSuccess) {
        ff_set_err("stream creation failed");
        return FF_ERR_DEVICE;
    }
    if (ff_tables_init() != 0) {
        ff_set_err("table upload failed");
        return FF_ERR_DEVICE;
    }
    // the carry look-back workspace (1M words) and its time-out word
    (void)ff_lb_take(1u << 20);
    ff_lb_begin();
    (void)ff_lb_fault();
    (void)ff_pinned_results(4096);
    if (ff_noalloc_setup(g_device) != 0) {
        ff_set_err("scratch array set-up failed");
        return FF_ERR_DEVICE;
    }
    g_inited = 1;
    return FF_OK;
}

int ff_is_inited(void) { return g_inited; }
int ff_device(void) { return g_device; }

void ff_ensure_init(void) {
    if (!g_inited && ff_init(-1) != FF_OK) {
        fprintf(stderr, "fastfib_amd: %s\n", g_err);
        exit(1);
    }
}

extern "C" void *ff_stream(void) {
    ff_ensure_init();
    return (void *)g_st;
}

extern "C" int ff_sync(void) {
    ff_ensure_init();
    return hipStreamSynchronize(g_st) == hipSuccess ? FF_OK : FF_ERR_DEVICE;
}

extern "C" int ff_get_option(const char *name, long long *v) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (!name || !v) return FF_ERR_ARG;
    const struct {
        const char *name;
        long long value;
    } opts[] = {{"inner_log", g_inner_log},       {"tile_words", g_tile_words},     {"force_len", (long long)g_force_len},
                {"max_lg", g_max_lg},             {"pair_steps", g_pair_steps},     {"quad_max_lg", g_quad_max_lg},
                {"quad_inner_log", g_quad_inner_log}, {"host_bits", g_host_bits}, {"school_limbs", g_school_limbs},
                {"queue_waves", g_queue_waves},   {"dec_queue", g_dec_queue},
                {"queue_parts", g_queue_parts},   {"queue_chunk", g_queue_chunk},       {"dispatch", g_dispatch},
                {"fib_tile_words", g_fib_tile_words}, {"outer_rl", g_outer_rl}, {"outer_rl3", g_outer_rl3}, {"pair_rl", g_pair_rl}, {"crt_wave_limbs", g_crt_wave_limbs}};
    for (const auto &o : opts) {
        if (!strcmp(name, o.name)) {
            *v = o.value;
            return FF_OK;
        }
    }
    return FF_ERR_ARG;
}

extern "C" int ff_set_option(const char *name, long long v) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (!name) return FF_ERR_ARG;
    if (!strcmp(name, "inner_log")) {
        if (v < 4 || v > 12) return FF_ERR_ARG;
        g_inner_log = (int)v;
    } else if (!strcmp(name, "tile_words")) {
        if (v < 1024 || v > 16384) return FF_ERR_ARG;
        g_tile_words = (int)v;
    } else if (!strcmp(name, "fib_tile_words")) {
        if (v != 0 && (v < 1024 || v > 16384)) return FF_ERR_ARG;
        g_fib_tile_words = (int)v;
    } else if (!strcmp(name, "outer_rl")) {
        if (v != 0 && (v < 2 || v > 4)) return FF_ERR_ARG;
        g_outer_rl = (int)v;
    } else if (!strcmp(name, "crt_wave_limbs")) {
        if (v < 0 || v > (1ll << 40)) return FF_ERR_ARG;
        g_crt_wave_limbs = v;
    } else if (!strcmp(name, "pair_rl")) {
        if (v != 0 && (v < 2 || v > 4)) return FF_ERR_ARG;
        g_pair_rl = (int)v;
    } else if (!strcmp(name, "outer_rl3")) {
        if (v < 0 || v >= (1 << 26)) return FF_ERR_ARG;
        g_outer_rl3 = (int)v;
    } else if (!strcmp(name, "force_len")) {
        if (v < 0) return FF_ERR_ARG;
        g_force_len = (uint64_t)v;
    } else if (!strcmp(name, "max_lg")) {
        if (v < 12 || v > 25) return FF_ERR_ARG;
        g_max_lg = (int)v;
    } else if (!strcmp(name, "pair_steps")) {
        if (v < 0 || v > 3) return FF_ERR_ARG;
        g_pair_steps = (int)v;
    } else if (!strcmp(name, "quad_max_lg")) {
        if (v != 0 && (v < 13 || v > 25)) return FF_ERR_ARG;
        g_quad_max_lg = (int)v;
    } else if (!strcmp(name, "quad_inner_log")) {
        if (v != 0 && (v < 8 || v > 12)) return FF_ERR_ARG;
        g_quad_inner_log = (int)v;
    } else if (!strcmp(name, "host_bits")) {
        if (v < 10 || v > 13) return FF_ERR_ARG;
        g_host_bits = (int)v;
    } else if (!strcmp(name, "school_limbs")) {
        if (v < 0 || v > FF_SCHOOL_LIMBS) return FF_ERR_ARG;
        g_school_limbs = (int)v;
    } else if (!strcmp(name, "queue_waves")) {
        if (v != 1 && v != 3 && v != 4) return FF_ERR_ARG;
        g_queue_waves = (int)v;
    } else if (!strcmp(name, "queue_parts")) {
        if (v < 1 || v > FQ_MAXQ) return FF_ERR_ARG;
        g_queue_parts = (int)v;
    } else if (!strcmp(name, "queue_chunk")) {
        if (v < 1 || v > 64) return FF_ERR_ARG;
        g_queue_chunk = (int)v;
    } else if (!strcmp(name, "dec_queue")) {
        if (v != 0 && v != 1) return FF_ERR_ARG;
        g_dec_queue = (int)v;
    } else if (!strcmp(name, "test_fail_alloc")) {
        if (v < -1) return FF_ERR_ARG;
        g_fail_after = v;
    } else if (!strcmp(name, "dispatch")) {
        if (v != 0 && v != FF_BATCH_TUNED && v != FF_BATCH_LEARNED) return FF_ERR_ARG;
        g_dispatch = (int)v;
    } else {
        return FF_ERR_ARG;
    }
    return FF_OK;
}

// A product that allocates nothing (ff_dmul_noalloc): the same pipeline as
// ff_dmul (forward passes -> inner convolution -> inverse passes -> CRT with
// look-back carries), with every scratch buffer - the two transform arrays,
// the carry look-back words and the length word - inside one fixed device
// array, `scratch`, that each GPU holds from library start until the process
// ends. For callers that may not call hipMalloc: code under HIP graph
// capture, tight serving loops, harnesses that forbid allocation.
//
// Layout of the array (bytes):
//   [0, 256)                  the product's length key (CRT-booked)
//   [256, 256 + 4 kLbWords)   look-back words, zero between calls (the CRT's
//                             last workgroup re-zeroes what it used)
//   then two arrays of 3 * 2^20 words: a's and b's transforms ([3][L]),
//   and 256 bytes of slack after them, as the pooled path allocates
// 24 MiB + 65 KiB per GPU. Products whose transform is longer than 2^20
// words are refused (FF_ERR_UNSUPPORTED) before anything is launched.
#include <stdio.h>
#include <string.h>

#include "../../include/fastfib_amd.h"
#include "ff_engine.h"
#include "ff_internal.h"
#include "ff_limbs.h"
#include "ff_runtime.h"

namespace {

constexpr int kMaxLg = 20;
constexpr size_t kArrWords = 3ull << kMaxLg;  // one transform array, all primes
constexpr size_t kLbWords = 16384;
constexpr size_t kKeyBytes = 256;
constexpr size_t kLbOff = kKeyBytes;
constexpr size_t kArrOff = kLbOff + 4 * kLbWords;
constexpr size_t kScratchBytes = kArrOff + 2 * 4 * kArrWords + 256;
static_assert(kArrOff % 256 == 0, "transform arrays stay 256-byte aligned");
static_assert(kScratchBytes == FF_NOALLOC_SCRATCH_BYTES, "header and layout agree");

constexpr int kMaxDevices = 64;
// Per device, the array's device address (hipGetSymbolAddress on that device).
unsigned char *g_scratch_dev[kMaxDevices];

}  // namespace

// The array itself. In host code its name is a host-side stand-in, never a
// device address: only g_scratch_dev[device] may be passed to kernels.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Called by ff_init once per device, with that device current: looks up the
// array's address on it and zeroes the key and look-back words.
int ff_noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return -1;
    if (g_scratch_dev[device]) return 0;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -1;
    if (ff_launch_fill((uint64_t *)p, kArrOff / 8, 0, g_st) != hipSuccess) return -1;
    if (hipStreamSynchronize(g_st) != hipSuccess) return -1;
    g_scratch_dev[device] = (unsigned char *)p;
    return 0;
}

extern "C" int ff_dmul_noalloc(ff_dbig *out, const ff_dbig *a, const ff_dbig *b) {
    if (!out || !out->d || !a || !a->d) return FF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (!ff_is_inited()) {
        ff_set_err("ff_dmul_noalloc: the library is not initialized (call ff_init first)");
        return FF_ERR_ARG;
    }
    unsigned char *base = g_scratch_dev[ff_device()];
    if (!base) {
        ff_set_err("ff_dmul_noalloc: no scratch array on this device");
        return FF_ERR_DEVICE;
    }
    const bool sq = !b || b == a || b->d == a->d;
    const uint64_t nb = sq ? a->n : b->n, bits_b = sq ? a->bits : b->bits;
    const uint64_t nout = a->n + nb + 1;
    if (out->cap < nout) {
        ff_set_err("ff_dmul_noalloc: out has room for %zu limbs, the product needs %llu", out->cap,
                   (unsigned long long)nout);
        return FF_ERR_ARG;
    }
    unsigned long long *key = (unsigned long long *)base;
    if (a->bits == 0 || bits_b == 0) {
        FF_CHECK(ff_launch_fill(out->d, 1, 0, g_st));
        FF_CHECK(hipStreamSynchronize(g_st));
        out->n = 1, out->bits = 0;
        return FF_OK;
    }
    const ProdShape s = ff_product_shape(a->bits, bits_b);
    if (s.split || s.L > (1ull << kMaxLg) || ff_crt_lb_words(nout) > kLbWords) {
        ff_set_err("ff_dmul_noalloc: product of %llu and %llu bits needs a transform longer than 2^%d words",
                   (unsigned long long)a->bits, (unsigned long long)bits_b, kMaxLg);
        return FF_ERR_UNSUPPORTED;
    }
    FFPlan pl;
    if (ff_plan_build(&pl, (uint32_t)s.L, (int)s.bw, s.lg, s.r3, s.conv, g_inner_log, g_tile_words) != 0) {
        ff_set_err("ff_dmul_noalloc: no schedule for L=%llu", (unsigned long long)s.L);
        return FF_ERR_UNSUPPORTED;
    }
    uint32_t *fa = (uint32_t *)(base + kArrOff);
    uint32_t *fb = sq ? nullptr : fa + kArrWords;
    uint32_t *lb = (uint32_t *)(base + kLbOff);
    const uint32_t *asrc = (const uint32_t *)a->d, *bsrc = sq ? nullptr : (const uint32_t *)b->d;
    const double opbytes = 8.0 * (a->n + (sq ? 0 : nb));

    FF_CHECK(ff_launch_fill((uint64_t *)key, 1, 0, g_st));  // the CRT books the key by atomicMax
    FFXformSet fwd = {};
    fwd.x[0] = FFXform{asrc, (uint32_t)(2 * a->n), fa};
    if (!sq) fwd.x[1] = FFXform{bsrc, (uint32_t)(2 * nb), fb};
    FF_CHECK(ff_launch_forward(&pl, &fwd, sq ? 1 : 2, opbytes, g_st));
    FFConvSet cv = {};
    cv.c[0] = FFConv{asrc, bsrc, (uint32_t)(2 * a->n), (uint32_t)(2 * nb), fa, fb};
    FF_CHECK(ff_launch_conv(&pl, &cv, 1, sq ? 0 : 1, opbytes, g_st));
    FFXformSet inv = {};
    inv.x[0] = FFXform{nullptr, 0, fa};
    FF_CHECK(ff_launch_inverse(&pl, &inv, 1, g_st));
    FFCrtSet crt = {};
    crt.j[0] = FFCrtJob{fa, out->d, lb, key, nout, s.conv, (uint32_t)s.L, (int)s.bw};
    FF_CHECK(ff_launch_crt(&crt, 1, nout, (double)s.conv, (double)nout, g_st));

    // the length key and the look-back time-out word, through the pinned
    // buffer the library holds from its start
    uint64_t *h = ff_pinned_results(2);
    h[1] = 0;
    FF_CHECK(hipMemcpyAsync(h, key, 8, hipMemcpyDeviceToHost, g_st));
    FF_CHECK(hipMemcpyAsync(h + 1, ff_lb_fault(), 4, hipMemcpyDeviceToHost, g_st));
    FF_CHECK(hipStreamSynchronize(g_st));
    if ((uint32_t)h[1]) ff_lb_timeout_fail();
    const uint64_t k = h[0];
    if (k == 0) {
        out->n = 1, out->bits = 0;
    } else {
        out->n = k >> 8;
        out->bits = 64 * (out->n - 1) + (k & 0xff);
    }
    return FF_OK;
}

// Modular arithmetic for the three NTT primes on gfx950 (32-bit lanes).
//
// The primes and generators are those of ntt_arith.c:44-52 (reference):
// p = c * 2^k + 1 < 2^31, with 3 * 2^25 | p - 1, so every transform length
// 2^k and 3 * 2^k up to 2^25 has its roots of unity. Values are canonical in
// [0, p). Multiplication is Montgomery with R = 2^32 computed from the high
// halves of 32x32 products (v_mul_hi_u32): with m = lo(a*b) * p^-1 mod 2^32,
// a*b - m*p is divisible by 2^32 and the low halves cancel exactly, so
//   mont(a, b) = hi(a*b) - hi(m*p)   in (-p, p)
// followed by one conditional add. It accepts any 32-bit a (b < p), which
// lets callers pass x - y + p without reducing it first.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#define FF_P1 2113929217u  // 63 * 2^25 + 1, generator 5
#define FF_P2 2013265921u  // 15 * 2^27 + 1, generator 31
#define FF_P3 1811939329u  // 27 * 2^26 + 1, generator 13

// Device pointers known to address global memory. Accesses through them
// compile to global_* instructions (an SGPR base plus a 32-bit lane offset
// where the base is uniform) instead of flat ones, which need a 64-bit
// address per lane and also count against the LDS wait counter, so that an
// LDS wait would wait for them too. Pointers loaded from a work item in
// memory are otherwise flat: the compiler cannot tell their address space.
template <class T>
using ff_gptr = __attribute__((address_space(1))) T *;
template <class T>
__device__ __forceinline__ ff_gptr<T> ff_g(T *p) {
    return (ff_gptr<T>)p;
}

// Loads and stores of data that other workgroups of the same launch write
// or read (the batch work queue, k_queue): with COH, agent-scope relaxed
// accesses, which the L2 of one XCD does not hold stale or dirty (the XCDs'
// L2s are not coherent with each other), so only the order of those
// accesses needs care, not a write-back or invalidation of whole caches.
template <bool COH, class P>
__device__ __forceinline__ auto ff_ld(P p) {
    typedef typename std::remove_cv<typename std::remove_reference<decltype(*p)>::type>::type T;
    if (COH) return __hip_atomic_load((const T *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return (T)*p;
}
template <bool COH, class P, class T>
__device__ __forceinline__ void ff_st(P p, T v) {
    typedef typename std::remove_reference<decltype(*p)>::type U;
    if (COH) __hip_atomic_store((U *)p, (U)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *p = (U)v;
}

// Every supported length divides FF_NBASE = 3 * 2^25, so w_L = w_NBASE^(NBASE/L).
#define FF_NBASE (3u << 25)
// Two-level table of w_NBASE^x: x = hi * FF_TL_SIZE + lo.
#define FF_TL_BITS 14
#define FF_TL_SIZE (1u << FF_TL_BITS)
#define FF_TH_SIZE (FF_NBASE >> FF_TL_BITS)
// Longest LDS-resident sub-transform (radix-2 levels with table twiddles).
#define FF_TW 4096

struct FFPrime {
    uint32_t p;     // the prime
    uint32_t pinv;  // p^-1 mod 2^32
    uint32_t r2;    // R^2 mod p (mont(h, r2) = h * 2^32 mod p)
    uint32_t w3;    // primitive cube root of unity (Montgomery form)
    uint32_t w3p;   // w3 * pinv
    uint32_t iw3;   // w3^-1 = w3^2 (Montgomery form)
    uint32_t iw3p;
    uint32_t one;   // R mod p: 1 in Montgomery form
};

// Device tables, filled once by ff_tables_init (all in Montgomery form):
//   rt[pi][h + j]  = w_2h^j      for h = 1, 2, .., FF_TW/2, j < h
//   irt[pi][h + j] = w_2h^-j
//   rtp / irtp     = the same times pinv (mod 2^32), for mont_w
//   tl[pi][x]      = w_NBASE^x        x < FF_TL_SIZE
//   th[pi][x]      = w_NBASE^(x * FF_TL_SIZE)
struct FFTables {
    uint32_t rt[3][FF_TW], rtp[3][FF_TW], irt[3][FF_TW], irtp[3][FF_TW];
    uint32_t tl[3][FF_TL_SIZE];
    uint32_t th[3][FF_TH_SIZE];
};

// hi - mh mod p for hi, mh < p: r = hi - mh wraps to above 2^32 - p when
// hi < mh, where r + p wraps back below p; otherwise r < p <= r + p < 2^32
// (p < 2^31). So the result is min(r, r + p): three instructions, no
// compare-and-select.
__device__ __forceinline__ uint32_t ff_mont_fix(uint32_t hi, uint32_t mh, uint32_t p) {
    const uint32_t r = hi - mh;
    const uint32_t rp = r + p;
    return rp < r ? rp : r;
}

// a * b * 2^-32 mod p, any 32-bit a, b < p.
__device__ __forceinline__ uint32_t ff_mont(uint32_t a, uint32_t b, uint32_t p, uint32_t pinv) {
    uint32_t lo = a * b;
    uint32_t hi = __umulhi(a, b);
    uint32_t m = lo * pinv;
    return ff_mont_fix(hi, __umulhi(m, p), p);
}

// The same with wp = w * pinv precomputed: m = a * wp needs no lo(a*w).
__device__ __forceinline__ uint32_t ff_mont_w(uint32_t a, uint32_t w, uint32_t wp, uint32_t p) {
    uint32_t hi = __umulhi(a, w);
    uint32_t m = a * wp;
    return ff_mont_fix(hi, __umulhi(m, p), p);
}

// (x - y) * w * 2^-32 mod p for x, y < p, with wp = w * pinv: the
// difference d = x - y in (-p, p) enters as a signed 32-bit value (p < 2^31),
// so it needs no "+ p". With m = lo(d * wp) taken signed, d*w - m*p is
// divisible by 2^32 (its low halves agree) and the signed high halves give
// r = (d*w - m*p) / 2^32 exactly, |r| < (p^2 + 2^31 p) / 2^32 < p; the
// correction of ff_mont_fix then lands in [0, p). One instruction fewer than
// ff_mont_w(x - y + p, w, wp, p), with the same (canonical) result.
__device__ __forceinline__ uint32_t ff_mont_w_diff(uint32_t x, uint32_t y, uint32_t w, uint32_t wp, uint32_t p) {
    const int32_t d = (int32_t)(x - y);
    const int32_t hi = __mulhi(d, (int32_t)w);
    const int32_t m = (int32_t)((uint32_t)d * wp);
    return ff_mont_fix((uint32_t)hi, (uint32_t)__mulhi(m, (int32_t)p), p);
}

__device__ __forceinline__ uint32_t ff_add(uint32_t x, uint32_t y, uint32_t p) {
    uint32_t s = x + y;  // < 2p < 2^32
    uint32_t t = s - p;
    return t < s ? t : s;  // min(s, s - p) as unsigned
}

__device__ __forceinline__ uint32_t ff_sub(uint32_t x, uint32_t y, uint32_t p) {
    uint32_t d = x - y;
    uint32_t e = d + p;
    return e < d ? e : d;  // min(d, d + p)
}

// Reduces any 32-bit value to [0, p) (p > 2^30, so at most 3 steps).
__device__ __forceinline__ uint32_t ff_red32(uint32_t x, uint32_t p) {
    uint32_t t = x - p;
    x = t < x ? t : x;
    t = x - p;
    x = t < x ? t : x;
    t = x - p;
    return t < x ? t : x;
}

// w_NBASE^x (Montgomery form) for x < FF_NBASE.
__device__ __forceinline__ uint32_t ff_root(const FFTables *T, int pi, uint32_t x, uint32_t p, uint32_t pinv) {
    return ff_mont(T->th[pi][x >> FF_TL_BITS], T->tl[pi][x & (FF_TL_SIZE - 1)], p, pinv);
}

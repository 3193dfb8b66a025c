/*
 * src/mul_gpu.c - the binding a maintainer adds to the reference
 * (shwwwa/fastfib) to route its multiply, F(n) and decimal output through
 * libfastfib_amd.so. Not part of the library: this file is compiled inside
 * the reference's unity build (fastfib.c), which includes it after
 * src/alloc.c (it needs Big and xmalloc) and before src/mul_basic.c:
 *
 *     #include "src/alloc.c"
 *     #include "src/mul_gpu.c"     (the added line)
 *     #include "src/pool.c"
 *
 * Build flags: -I<repo>/include -L<repo>/fastfib_amd -lfastfib_amd.
 * The reference's dispatch then calls the _gpu functions, e.g. as the first
 * statement of big_mul (src/mul_basic.c:64):
 *
 *     #ifdef FASTFIB_GPU
 *         return big_mul_gpu(a, b);
 *     #endif
 *
 * and likewise big_sqr (src/mul_basic.c:109), fibonacci (src/fib.c:211) and
 * big_to_string (src/decimal.c:229); ntt_init (src/ntt_tables.c:72) becomes
 * ff_ntt_init. big_mul_ntt_split needs no binding: ff_big_mul splits
 * products past the 2^25-word transform itself.
 */
#pragma once
#define FF_CALLER_BIG /* the reference's Big (src/common.h:29-33) is in scope */
#include "fastfib_amd.h"

_Static_assert(sizeof(Big) == sizeof(uint64_t *) + sizeof(size_t), "Big must be {uint64_t *d; size_t n}");

/* The library returns limbs from malloc (freed by ff_big_free); the reference
   frees with its arena's xfree, so results move into an arena block and every
   later big_free stays valid. */
static Big gpu_adopt(Big r) {
    Big out = {xmalloc(r.n * sizeof(uint64_t)), r.n};
    memcpy(out.d, r.d, r.n * sizeof(uint64_t));
    ff_big_free(&r);
    return out;
}

/* a == b (pointer identity) squares, as in the reference (mul_ntt.c:36,125). */
static Big big_mul_gpu(const Big *a, const Big *b) { return gpu_adopt(ff_big_mul(a, b)); }
static Big big_sqr_gpu(const Big *a) { return gpu_adopt(ff_big_sqr(a)); }
static Big fibonacci_gpu(uint64_t n) { return gpu_adopt(ff_fibonacci(n)); }
/* malloc'ed like the reference's (decimal.c:219-225): release with free(). */
static char *big_to_string_gpu(const Big *a) { return ff_big_to_string(a); }

/*
 * C-ABI harness over the reference's own sources (shwwwa/fastfib), built into
 * oracle/_ref/libffref.so by oracle/Makefile from the files where they lie in
 * the reference checkout. Test infrastructure only: the tests use it as a
 * second, independent checker of the oracle and of the GPU library, and the
 * benchmark times it as the CPU baseline. The GPU library never links it.
 *
 * The reference is a unity build whose functions are all static
 * (fastfib.c:16-21, common.h:8-11); including fastfib.c here with
 * UNIT_TEST_NO_MAIN is how its own test harnesses reach them
 * (tests/nttfuzz.c:7-8, tests/hexfib.c:1-3). The functions below only move
 * limbs between caller buffers and the reference's Big.
 */
#define UNIT_TEST_NO_MAIN
#include "fastfib.c"

#define EXPORT __attribute__((visibility("default")))

/* Builds a Big view of caller limbs (normalized by length only). */
static Big view_of(const uint64_t *d, size_t n) {
    Big v = {(uint64_t *)d, n ? normalize_len(d, n) : 1};
    if (!n) v.d = &g_zero_limb;
    return v;
}

/* Copies r into a malloc'ed buffer for the caller; frees r. */
static uint64_t *export_big(Big *r, size_t *n_out) {
    uint64_t *o = (uint64_t *)malloc(r->n * sizeof(uint64_t));
    memcpy(o, r->d, r->n * sizeof(uint64_t));
    *n_out = r->n;
    big_free(r);
    return o;
}

/** ntt_init (ntt_tables.c:72): grow-only tables for max_limbs-limb products. */
EXPORT void ref_ntt_init(size_t max_limbs) { ntt_init(max_limbs); }

/** Largest transform the tables allow right now (g_ntt_max_L, mul_basic.c:26). */
EXPORT size_t ref_ntt_max_len(void) { return g_ntt_max_L; }

/** big_mul / big_sqr (mul_basic.c:64,109): the full dispatch. square = 1
    passes the same pointer twice, the reference's squaring signal. */
EXPORT uint64_t *ref_mul(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, int square, size_t *n_out) {
    Big A = view_of(a, na), B = view_of(b, nb);
    Big r = square ? big_sqr(&A) : big_mul(&A, &B);
    return export_big(&r, n_out);
}

/** big_mul_ntt (mul_ntt.c:97) called directly, as tests/nttfuzz.c:50 does;
    the caller must have run ref_ntt_init large enough. */
EXPORT uint64_t *ref_mul_ntt(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, int square, size_t *n_out) {
    Big A = view_of(a, na), B = view_of(b, nb);
    Big r = big_mul_ntt(&A, square ? &A : &B);
    return export_big(&r, n_out);
}

/** big_mul with the NTT hidden (g_ntt_max_L = 0): schoolbook, Karatsuba and
    Toom-3 only, the independent product tests/nttfuzz.c:32-39 checks with. */
EXPORT uint64_t *ref_mul_nontt(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, int square,
                               size_t *n_out) {
    size_t saved = g_ntt_max_L;
    g_ntt_max_L = 0;
    Big A = view_of(a, na), B = view_of(b, nb);
    Big r = square ? big_sqr(&A) : big_mul(&A, &B);
    g_ntt_max_L = saved;
    return export_big(&r, n_out);
}

/** fibonacci (fib.c:211). */
EXPORT uint64_t *ref_fibonacci(uint64_t n, size_t *n_out) {
    Big r = fibonacci(n);
    return export_big(&r, n_out);
}

/** big_to_string (decimal.c:229); free the result with ref_free. */
EXPORT char *ref_to_string(const uint64_t *a, size_t na) {
    Big A = view_of(a, na);
    return big_to_string(&A);
}

/** Seconds for one fibonacci(n) call (the F(n)-only time main.c:88-91
    prints), result discarded. */
EXPORT double ref_time_fibonacci(uint64_t n) {
    double t0 = plat_seconds();
    Big r = fibonacci(n);
    double t = plat_seconds() - t0;
    big_free(&r);
    return t;
}

/** Seconds for one big_mul_ntt of the given operands (tables must be built). */
EXPORT double ref_time_mul_ntt(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, int square) {
    Big A = view_of(a, na), B = view_of(b, nb);
    double t0 = plat_seconds();
    Big r = big_mul_ntt(&A, square ? &A : &B);
    double t = plat_seconds() - t0;
    big_free(&r);
    return t;
}

/** Seconds for big_mul/big_sqr over a list of products (the reference has no
    batch API, so this is its loop over the list). Operand k is a[off_a[k] ..
    off_a[k] + na[k]) and likewise for b; sq[k] = 1 squares a. */
EXPORT double ref_time_mul_list(const uint64_t *a, const size_t *off_a, const size_t *na, const uint64_t *b,
                                const size_t *off_b, const size_t *nb, const int *sq, size_t count) {
    double t0 = plat_seconds();
    for (size_t k = 0; k < count; k++) {
        Big A = view_of(a + off_a[k], na[k]), B = view_of(b + off_b[k], nb[k]);
        Big r = sq[k] ? big_sqr(&A) : big_mul(&A, &B);
        big_free(&r);
    }
    return plat_seconds() - t0;
}

/** big_mul/big_sqr over the same kind of list, keeping every product:
    product k is written to out + off_out[k] (room for na[k] + nb[k] + 1
    limbs, what ntt_result allocates, mul_ntt.c:61-73) and its normalized
    length to n_out[k]. */
EXPORT void ref_mul_list(const uint64_t *a, const size_t *off_a, const size_t *na, const uint64_t *b,
                         const size_t *off_b, const size_t *nb, const int *sq, size_t count, uint64_t *out,
                         const size_t *off_out, size_t *n_out) {
    for (size_t k = 0; k < count; k++) {
        Big A = view_of(a + off_a[k], na[k]), B = view_of(b + off_b[k], nb[k]);
        Big r = sq[k] ? big_sqr(&A) : big_mul(&A, &B);
        memcpy(out + off_out[k], r.d, r.n * sizeof(uint64_t));
        n_out[k] = r.n;
        big_free(&r);
    }
}

/** Seconds for one big_to_string (decimal.c:229), string discarded. */
EXPORT double ref_time_to_string(const uint64_t *a, size_t na) {
    Big A = view_of(a, na);
    double t0 = plat_seconds();
    char *s = big_to_string(&A);
    double t = plat_seconds() - t0;
    free(s);
    return t;
}

/** Threads the reference pool uses (pool_cpus, pool.c:125). */
EXPORT int ref_pool_cpus(void) { return pool_cpus(); }

EXPORT void ref_free(void *p) { free(p); }

/*
 * The reference's unity build with the GPU binding
 * (fastfib_amd/integration/mul_gpu.c) included where a maintainer adds it,
 * linked to libfastfib_amd.so: the drop-in exercised from the reference's own
 * code. Test infrastructure, built by oracle/Makefile into oracle/_ref.
 *
 *     ref_gpu_binding n...
 *
 * For each n: F(n) by the reference's fibonacci (fib.c:211) and by
 * fibonacci_gpu; F(n)^2 by big_sqr and big_sqr_gpu and by big_mul(&f, &f)
 * and big_mul_gpu(&f, &f) (pointer identity squares); F(n) F(n+1) by
 * big_mul and big_mul_gpu; the decimal digits by big_to_string and
 * big_to_string_gpu. Prints "n <n> ok" or "n <n> differs: <what>"; exit
 * status 0 when every n agrees.
 */
#define UNIT_TEST_NO_MAIN
#include "src/common.h"
#include "src/platform.h"
#include "src/alloc.c"
#include "mul_gpu.c"
#include "fastfib.c"

static int same(const Big *a, const Big *b) {
    return a->n == b->n && !memcmp(a->d, b->d, a->n * sizeof(uint64_t));
}

int main(int argc, char **argv) {
    int bad = 0;
    for (int i = 1; i < argc; i++) {
        const uint64_t n = strtoull(argv[i], NULL, 10);
        ntt_init(2 * (size_t)(n * 0.6942419136306174 / 64 + 8));
        Big f = fibonacci(n), g = fibonacci_gpu(n);
        Big f1 = fibonacci(n + 1);
        const char *what = NULL;
        if (!same(&f, &g)) what = "fibonacci";
        Big s_ref = big_sqr(&f), s_gpu = big_sqr_gpu(&f);
        Big m_ref = big_mul(&f, &f), m_gpu = big_mul_gpu(&f, &f);
        Big p_ref = big_mul(&f, &f1), p_gpu = big_mul_gpu(&f, &f1);
        if (!what && !same(&s_ref, &s_gpu)) what = "big_sqr";
        if (!what && !same(&m_ref, &m_gpu)) what = "big_mul square";
        if (!what && !same(&p_ref, &p_gpu)) what = "big_mul";
        char *d_ref = big_to_string(&f), *d_gpu = big_to_string_gpu(&f);
        if (!what && strcmp(d_ref, d_gpu)) what = "big_to_string";
        if (what) {
            printf("n %llu differs: %s\n", (unsigned long long)n, what);
            bad = 1;
        } else {
            printf("n %llu ok\n", (unsigned long long)n);
        }
        free(d_ref), free(d_gpu);
        big_free(&f), big_free(&g), big_free(&f1);
        big_free(&s_ref), big_free(&s_gpu), big_free(&m_ref), big_free(&m_gpu), big_free(&p_ref), big_free(&p_gpu);
    }
    return bad;
}

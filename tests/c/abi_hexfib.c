/*
 * A C caller of libfastfib_amd.so through include/fastfib_amd.h only, the
 * way the reference's own harness tests/hexfib.c:5-15 calls fibonacci and
 * prints F(n) in hex (top limb "%llx", then every lower limb "%016llx").
 *
 *     abi_hexfib n     prints, one per line:
 *       hex of F(n)                       (ff_fibonacci)
 *       "square ok" / "square differs"    ff_big_mul(&f, &f) (pointer identity
 *                                          squares) against ff_big_sqr(&f) and
 *                                          against ff_big_mul(&f, &g) for a
 *                                          copy g of f (a general product)
 *       decimal digits of F(n)            (ff_big_to_string)
 *
 * Built and run by tests/test_cli.py.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fastfib_amd.h"

static void print_hex(const Big *x) {
    printf("%llx", (unsigned long long)x->d[x->n - 1]);
    for (size_t i = x->n - 1; i-- > 0;) printf("%016llx", (unsigned long long)x->d[i]);
    printf("\n");
}

static int same(const Big *a, const Big *b) {
    return a->n == b->n && !memcmp(a->d, b->d, a->n * sizeof(uint64_t));
}

int main(int argc, char **argv) {
    if (argc != 2) return 2;
    const uint64_t n = strtoull(argv[1], NULL, 10);
    ff_ntt_init(0);
    Big f = ff_fibonacci(n);
    print_hex(&f);

    Big copy = {malloc(f.n * sizeof(uint64_t)), f.n};
    memcpy(copy.d, f.d, f.n * sizeof(uint64_t));
    Big s1 = ff_big_mul(&f, &f), s2 = ff_big_sqr(&f), s3 = ff_big_mul(&f, &copy), s4 = ff_big_mul_ntt(&f, &f);
    printf(same(&s1, &s2) && same(&s1, &s3) && same(&s1, &s4) ? "square ok\n" : "square differs\n");

    char *dec = ff_big_to_string(&f);
    printf("%s\n", dec);
    free(dec);
    ff_big_free(&s1), ff_big_free(&s2), ff_big_free(&s3), ff_big_free(&s4);
    free(copy.d);
    ff_big_free(&f);
    return 0;
}

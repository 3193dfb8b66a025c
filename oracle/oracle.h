/*
 * ORACLE — test infrastructure, not product code.
 *
 * A plain-C, scalar restatement of the reference's exact-multiply path
 * (shwwwa/fastfib: three-prime NTT multiply, CRT + carry, split multiply,
 * fast-doubling F(n), Barrett decimal conversion). Only tests/, the smoke
 * check and bench.py's cpu_baseline leg load it, and only to check or time;
 * the GPU library (fastfib_amd) never links or calls it.
 *
 * Numbers are little-endian arrays of 64-bit limbs. Returned arrays are
 * malloc'ed; free them with ffo_free. Lengths returned are normalized
 * (at least 1 limb; zero is one 0 limb).
 *
 * Parity pins (tests/test_oracle.py): every function here is checked against
 * the reference build (oracle/_ref), Python integers, closed forms
 * (all-ones products, tests/bound.c:15-26), textbook O(L^2) DFTs, and
 * F(n) fingerprints by fast doubling modulo fixed primes.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The three NTT primes and generators (ntt_arith.c:44-52). */
extern const uint32_t ffo_prime[3];
extern const uint32_t ffo_gen[3];

/* Transform schedule of one product under the reference rule
   (mul_ntt.c:103-118): width bw, length L = 2^lg or 3*2^lg (r3), conv
   coefficients; split = 1 when L exceeds the longest transform (then the
   split fields follow mul_split.c:151-177). */
typedef struct {
    unsigned bw;
    size_t L, conv;
    int lg, r3, split;
    /* split only */
    int lgp, ka, kb;
} ffo_plan;

/* Longest transform is 2^max_lg (default 25, common.h:46-48); a smaller
   value forces the split multiply, like the reference's -DNTT_MAX_LG. */
void ffo_set_max_lg(int max_lg);
int ffo_get_max_lg(void);

int ffo_width_ok(size_t bits_a, size_t bits_b, unsigned w);
size_t ffo_ntt_length(size_t conv, int *lg, int *r3);
void ffo_plan_product(size_t bits_a, size_t bits_b, int square, ffo_plan *pl);

/* Cyclic convolution of a and b (L coefficients each, < p) modulo prime
   index pi, natural order, normalized (divided by L): out[i] =
   sum_{j+k = i mod L} a[j] b[k] mod p. L = 2^k or 3*2^k. */
void ffo_cyclic_conv(int pi, const uint32_t *a, const uint32_t *b, size_t L, uint32_t *out);

/* Forward transform (natural order in, natural order out) X[k] =
   sum_n x[n] w^(nk) with w = g^((p-1)/L); for the textbook DFT check. */
void ffo_ntt_forward(int pi, uint32_t *x, size_t L);

/* Garner CRT of three residue arrays and packing at bit i*bw with carry
   (crt.c:33-194): out gets nout limbs. */
void ffo_crt_pack(const uint32_t *const y[3], size_t conv, unsigned bw, uint64_t *out, size_t nout);

/* Coefficient k of a number read as bw-bit fields (ntt_kernels.c:40-74). */
uint64_t ffo_field(const uint64_t *d, size_t n, size_t k, unsigned bw);

/* a*b by the NTT path (mul_ntt.c:97, mul_split.c:149). square: b ignored. */
uint64_t *ffo_mul_ntt(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, int square, size_t *nout);
/* a*b with the reference's threshold dispatch (mul_basic.c:64,109):
   schoolbook below NTT_THRESHOLD_LIMBS = 200 limbs, NTT otherwise. */
uint64_t *ffo_mul(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, int square, size_t *nout);
/* Schoolbook product (bigint.c:257), for tiny cross-checks. */
uint64_t *ffo_mul_school(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, size_t *nout);

/* F(n) by top-down fast doubling (fib.c:211-261). */
uint64_t *ffo_fibonacci(uint64_t n, size_t *nout);
/* One linear step (fib.c:36-62): from s1 = F(k)^2, s2 = F(k-1)^2 builds
   (lo, hi) = bit ? (F(2k), F(2k+1)) : (F(2k-1), F(2k)). */
void ffo_fib_step(const uint64_t *s1, size_t n1, const uint64_t *s2, size_t n2, int k_odd, int bit,
                  uint64_t **lo, size_t *nlo, uint64_t **hi, size_t *nhi);

/* Decimal digits, no leading zeros (decimal.c:229-281). */
char *ffo_to_string(const uint64_t *a, size_t na);
/* floor(2^(2s)/P) - e with 0 <= e <= 4 (decimal.c:39-71). */
uint64_t *ffo_recip(const uint64_t *P, size_t np, size_t s, size_t *nout);

/* Fingerprint of a number: bit length, low 64 bits and the value modulo
   the four primes of tests/fingerprint.py:22 (out[0..3]). */
void ffo_fingerprint(const uint64_t *d, size_t n, uint64_t *bits, uint64_t *low64, uint64_t out[4]);

void ffo_free(void *p);

#ifdef __cplusplus
}
#endif

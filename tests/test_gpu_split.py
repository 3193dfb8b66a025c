"""Products past the longest transform (the split multiply, mul_split.c).

With the transform bound lowered ("max_lg", the reference's NTT_MAX_LG)
products split at small sizes, where Python ints check them exactly, over
piece lengths with and without outer passes, squares and general products,
balanced and unbalanced piece counts. At the real bound (2^25) the split
products behind F(n) for n past 1.644 * 10^9 are checked against the
reference's own F(n) bit for bit, and single split products through
residues and the all-ones closed form.
"""
import random

import numpy as np
import pytest

from oracle.oracle import FINGERPRINT_PRIMES, fib_fingerprint, from_limbs, to_limbs

pytestmark = pytest.mark.gpu


def _rand(rng, bits):
    return rng.getrandbits(bits) | (1 << (bits - 1))


@pytest.mark.parametrize("max_lg", [12, 13, 14, 16])
def test_split_small_bound(ff, max_lg):
    rng = random.Random(max_lg)
    ff.set_option("max_lg", max_lg)
    try:
        for bits_a, bits_b in [(1 << max_lg) * 20, (1 << max_lg) * 20], [(1 << max_lg) * 50, 3000], \
                [(1 << max_lg) * 33 + 7, (1 << max_lg) * 41 - 5], [(1 << max_lg) * 90, (1 << max_lg) * 17]:
            a, b = _rand(rng, bits_a), _rand(rng, bits_b)
            assert from_limbs(ff.big_mul_ntt(a, b)) == a * b, (max_lg, bits_a, bits_b)
            assert from_limbs(ff.big_mul_ntt(a)) == a * a, (max_lg, bits_a)
        # all ones, the carry worst case
        x = (1 << ((1 << max_lg) * 40)) - 1
        assert from_limbs(ff.big_mul_ntt(x)) == x * x
    finally:
        ff.set_option("max_lg", 25)


def test_fibonacci_with_small_bound(ff):
    ff.set_option("max_lg", 14)
    try:
        for n in (10 ** 6, 3 * 10 ** 6 + 7):
            x = from_limbs(ff.fibonacci(n))
            fp = fib_fingerprint(n)
            assert x.bit_length() == fp["bits"]
            assert x % (1 << 64) == fp["low64"]
            assert all(x % q == r for q, r in zip(FINGERPRINT_PRIMES, fp["mods"]))
    finally:
        ff.set_option("max_lg", 25)


PR = (2305843009213693951, 1152921504606846883)


def test_split_at_full_bound(ff):
    """Operands of 600M and 900M bits (the 2^25 transform holds about
    1.07G-bit products): residues of the product against the operands'."""
    rng = np.random.default_rng(11)
    A = rng.integers(0, 1 << 64, size=600_000_000 // 64, dtype=np.uint64)
    B = rng.integers(0, 1 << 64, size=900_000_000 // 64, dtype=np.uint64)
    A[-1] |= np.uint64(1 << 63)
    B[-1] |= np.uint64(1 << 63)
    a, b = from_limbs(A), from_limbs(B)
    prod = from_limbs(ff.big_mul_ntt(A, B))
    assert prod.bit_length() in (a.bit_length() + b.bit_length(), a.bit_length() + b.bit_length() - 1)
    for q in PR:
        assert prod % q == (a % q) * (b % q) % q
    sq = from_limbs(ff.big_mul_ntt(B))
    for q in PR:
        assert sq % q == (b % q) ** 2 % q


def test_split_all_ones_full_bound(ff):
    """(2^x - 1)^2 = 2^(2x) - 2^(x+1) + 1 with x = 600012345 bits (35M
    coefficients of 34 bits: past the 2^25 transform)."""
    x = 600_012_345
    n = (x + 63) // 64
    A = np.full(n, np.uint64(0xFFFFFFFFFFFFFFFF))
    A[-1] = np.uint64((1 << (x - 64 * (n - 1))) - 1)
    got = ff.big_mul_ntt(A)
    # expected limbs: 1 at bit 0, zeros up to bit x+1, ones from bit x+1 to 2x
    want = np.zeros(2 * n, dtype=np.uint64)
    want[0] = 1
    ones_lo, ones_hi = x + 1, 2 * x  # bits [x+1, 2x) set
    lo_limb, hi_limb = ones_lo // 64, ones_hi // 64
    want[lo_limb + 1:hi_limb] = np.uint64(0xFFFFFFFFFFFFFFFF)
    want[lo_limb] |= np.uint64((0xFFFFFFFFFFFFFFFF << (ones_lo % 64)) & 0xFFFFFFFFFFFFFFFF)
    if ones_hi % 64:
        want[hi_limb] |= np.uint64((1 << (ones_hi % 64)) - 1)
    want = want[:max(1, int(np.max(np.nonzero(want)[0])) + 1)]
    assert got.size == want.size
    assert np.array_equal(got, want)


def test_fibonacci_past_bound_vs_reference(ff, ref):
    """F(1.7 * 10^9): its last products split (the first split is near
    n = 1.644 * 10^9); bit for bit against the reference."""
    n = 1_700_000_000
    got = ff.dfibonacci(n).limbs()
    want = ref.fibonacci(n)
    assert got.size == want.size
    assert np.array_equal(got, want)

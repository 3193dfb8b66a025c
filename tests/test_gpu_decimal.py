"""Decimal output on the GPU (ff_big_to_string, ff_big_recip), exact.

Reciprocals are compared bit for bit with the oracle's restatement of
big_recip (decimal.c:39-71). Strings are compared with Python's own int to
str at sizes it finishes quickly, and with the reference's big_to_string
(oracle/_ref) beyond. Cases cover every tree height from a single leaf up,
the boundaries of the tree (10^(19 * 2^i) and its neighbours, 10^k - 1 with
all nines, powers of two) and F(n).
"""
import hashlib
import json
import os

import numpy as np
import pytest

from fastfib_amd.workloads import operand_limbs
from oracle.oracle import FINGERPRINT_PRIMES, decimal_mod, fib_digits, fib_fingerprint, fib_int, from_limbs, to_limbs

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.json")))


def test_recip_matches_oracle(ff, oracle):
    rng = np.random.default_rng(7)
    sizes = [2, 30, 62, 63, 64, 65, 127, 128, 129, 1000, 2020, 4039, 20000, 100000, 400000]
    for s in sizes:
        for _ in range(3):
            P = from_limbs(operand_limbs(rng, s))
            got = from_limbs(ff.big_recip(P, s))
            want = from_limbs(oracle.recip(to_limbs(P), s))
            assert got == want, s
            exact = (1 << (2 * s)) // P
            assert exact - 4 <= got <= exact, s


def test_recip_of_decimal_powers(ff, oracle):
    for i in range(0, 14):
        P = 10 ** (19 * 2 ** i)
        s = P.bit_length()
        got = from_limbs(ff.big_recip(P, s))
        assert got == from_limbs(oracle.recip(to_limbs(P), s)), i


EDGE = [0, 1, 9, 10, 99, 10 ** 18, 10 ** 19 - 1, 10 ** 19, 10 ** 19 + 1, (1 << 64) - 1, 1 << 64, (1 << 128) + 1]


def test_small_values(ff):
    for x in EDGE:
        assert ff.to_string(x) == str(x), x


def test_tree_boundaries(ff):
    """10^(19 * 2^i) and 10^(38 * 2^i) with their neighbours, all-nines and
    powers of two, for every tree height up to 11."""
    for i in range(0, 12):
        for k in (19 * 2 ** i, 38 * 2 ** i):
            P = 10 ** k
            for x in (P - 1, P, P + 1, 2 * P - 1, P * P - 1 if k < 20000 else P - 2):
                assert ff.to_string(x) == str(x), (k, x.bit_length())
        b = 64 * 2 ** i
        for x in ((1 << b) - 1, 1 << b, (1 << (2 * b - 2)) - 1):
            assert ff.to_string(x) == str(x), b


@pytest.mark.parametrize("bits", [100, 1000, 2019, 2021, 4040, 9000, 30000, 70000, 150000, 330000])
def test_random_vs_python(ff, bits):
    rng = np.random.default_rng(bits)
    for _ in range(3):
        x = from_limbs(operand_limbs(rng, bits))
        assert ff.to_string(x) == str(x), bits


@pytest.mark.parametrize("bits", [1_000_000, 4_000_000, 25_000_000])
def test_large_vs_reference(ff, ref, bits):
    """Against the reference build, and independently of it: the string's
    value modulo the four fingerprint primes equals the number's."""
    rng = np.random.default_rng(bits)
    A = operand_limbs(rng, bits)
    s = ff.to_string(A)
    assert s == ref.to_string(A)
    x = from_limbs(A)
    assert [decimal_mod(s, p) for p in FINGERPRINT_PRIMES] == [x % p for p in FINGERPRINT_PRIMES]


def test_fib_1e8_digits(ff):
    """The 20,898,764 digits of F(10^8) (tree height 20), converted on the
    device from the device-resident F(10^8): the reference's big_to_string
    digest (tests/golden), the digit count from Binet's formula and the
    string's value modulo four primes against F(n) mod p by fast doubling."""
    n = 10 ** 8
    g = GOLDEN["decimal"]["fib_100000000"]
    s = ff.dfibonacci(n).to_string()
    assert len(s) == g["digits"] == fib_digits(n) == 20898764
    assert s[:40] == g["head"] and s[-40:] == g["tail"]
    assert [decimal_mod(s, p) for p in FINGERPRINT_PRIMES] == fib_fingerprint(n)["mods"]
    assert hashlib.sha256(s.encode()).hexdigest() == g["sha256"]


def test_fibonacci_digits(ff, ref):
    """F(10^6) against Python, F(10^7) against the reference, both through
    the device-resident entry (the number never leaves the GPU)."""
    d = ff.dfibonacci(10 ** 6)
    assert d.to_string() == str(fib_int(10 ** 6))
    d = ff.dfibonacci(10 ** 7)
    assert d.to_string() == ref.to_string(d.limbs())


def test_repeat_and_mixed_heights(ff):
    """Cached powers, reciprocals and level graphs: conversions of different
    heights interleaved, each repeated, stay exact."""
    rng = np.random.default_rng(3)
    xs = [from_limbs(operand_limbs(rng, b)) for b in (5000, 200000, 3000, 200000, 60000)]
    want = [str(x) for x in xs]
    for _ in range(2):
        assert [ff.to_string(x) for x in xs] == want


def _dec_recip(ff, i):
    import ctypes
    L = ff.lib()
    L.ff_dec_reciprocal.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    n = ctypes.c_size_t()
    assert L.ff_dec_reciprocal(i, None, 0, ctypes.byref(n)) == 0
    out = np.zeros(n.value, dtype=np.uint64)
    assert L.ff_dec_reciprocal(i, out.ctypes.data, out.size, ctypes.byref(n)) == 0
    return from_limbs(out)


def test_decimal_reciprocal_chain(ff):
    """The exact reciprocals m_i = floor(2^(2s) / 10^(19 * 2^i)) the Barrett
    levels use (formed bottom-up from squares and one Newton step each, no
    length read back): equal to Python's floor division where Python is
    quick (i <= 12), and beyond, up to F(10^8)'s height 20, the defining
    inequalities P m <= 2^(2s) < P (m + 1) with P and the products formed
    exactly by the library's multiply."""
    for i in range(5, 13):
        P = 10 ** (19 * 2 ** i)
        s = P.bit_length()
        assert _dec_recip(ff, i) == (1 << (2 * s)) // P, i
    P = to_limbs(10 ** (19 * 2 ** 12))
    for i in range(13, 21):
        P = ff.big_sqr(P)  # P_i = P_(i-1)^2, exact
        Pv = from_limbs(P)
        s = Pv.bit_length()
        m = _dec_recip(ff, i)
        assert m.bit_length() == s + 1, i
        lo = from_limbs(ff.big_mul(P, to_limbs(m)))
        hi = from_limbs(ff.big_mul(P, to_limbs(m + 1)))
        assert lo <= (1 << (2 * s)) < hi, i


def test_levels_as_work_queues(ff, ref):
    """The level batches run as persistent work queues (option dec_queue):
    values of every tree height against Python and the reference build, and
    F(10^8)'s 20,898,764 digits by their digest, each conversion repeated
    (the second runs the captured levels)."""
    g = GOLDEN["decimal"]["fib_100000000"]
    rng = np.random.default_rng(19)
    ff.set_option("dec_queue", 1)
    try:
        for bits in (100, 2021, 70000, 330000):
            x = from_limbs(operand_limbs(rng, bits))
            assert ff.to_string(x) == ff.to_string(x) == str(x), bits
        A = operand_limbs(rng, 4_000_000)
        want = ref.to_string(A)
        assert ff.to_string(A) == want and ff.to_string(A) == want
        d = ff.dfibonacci(10 ** 8)
        for _ in range(2):
            s = d.to_string()
            assert len(s) == g["digits"]
            assert hashlib.sha256(s.encode()).hexdigest() == g["sha256"]
    finally:
        ff.set_option("dec_queue", 0)


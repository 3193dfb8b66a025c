"""CPU tests of the oracle (oracle/oracle.c), pinned from several sides.

* Textbook mathematics: the NTT against an O(L^2) DFT mod p, the cyclic
  convolution against brute force, 2^k and 3*2^k lengths, roots of the
  right order for each prime.
* Python integers: products (random, all-ones, sparse, squares) through the
  plain path and the split path (forced longest transform), F(n), decimal.
* Closed forms: (2^x - 1)(2^y - 1) (tests/bound.c:15-26).
* The reference's own C sources (oracle/_ref): products, F(n), decimal.
* The committed golden fixtures (tests/golden/golden.json) and fingerprints
  by fast doubling on residues + Binet's formula.
A planted mistake in any of the restated steps (a dropped Garner term, a
wrong root, an off-by-one width or length) fails one of these.
"""
import hashlib
import json
import os
import random

import numpy as np
import pytest

from oracle.oracle import FINGERPRINT_PRIMES, fib_fingerprint, fib_int, fib_mod, from_limbs, to_limbs

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))
PRIMES = [2113929217, 2013265921, 1811939329]
GENS = [5, 31, 13]


def test_primes_and_generators():
    # p = c * 2^k + 1 with 3 * 2^25 | p - 1 (ntt_arith.c:44-52); g has full order
    for p, g, fac in zip(PRIMES, GENS, [(2, 3, 7), (2, 3, 5), (2, 3)]):
        assert (p - 1) % (3 << 25) == 0 and p < 2 ** 31
        assert all(pow(g, (p - 1) // q, p) != 1 for q in fac)
    assert PRIMES[0] * PRIMES[1] * PRIMES[2] > 2 ** 92


@pytest.mark.parametrize("L", [16, 32, 48, 64, 96, 192, 256, 384])
def test_ntt_forward_is_dft(oracle, L):
    rng = random.Random(L)
    for pi, p in enumerate(PRIMES):
        x = [rng.randrange(p) for _ in range(L)]
        w = pow(GENS[pi], (p - 1) // L, p)
        want = [sum(x[n] * pow(w, n * k, p) for n in range(L)) % p for k in range(L)]
        assert list(oracle.ntt_forward(pi, x)) == want


@pytest.mark.parametrize("L", [16, 48, 64, 96, 128, 3 * 64])
def test_cyclic_conv_brute_force(oracle, L):
    rng = random.Random(L + 1)
    for pi, p in enumerate(PRIMES):
        a = [rng.randrange(p) for _ in range(L)]
        b = [rng.randrange(p) for _ in range(L)]
        want = [0] * L
        for i in range(L):
            for j in range(L):
                want[(i + j) % L] = (want[(i + j) % L] + a[i] * b[j]) % p
        assert list(oracle.cyclic_conv(pi, a, b)) == want


def test_width_and_length_rule(oracle):
    """mul_ntt.c:55-118 restated: widths 32..35 under the CRT bound, shortest
    length 2^k / 3*2^k, narrowest width on ties."""
    assert oracle.ntt_length(1) == (16, 4, 0)
    assert oracle.ntt_length(48) == (48, 4, 1)
    assert oracle.ntt_length(49) == (64, 6, 0)
    assert oracle.ntt_length(3 << 20) == (3 << 20, 20, 1)
    assert oracle.ntt_length((3 << 20) + 1) == (1 << 22, 22, 0)
    # schedules quoted from the reference rule: the 2^26-bit product runs at
    # L = 2^22, b = 32; F(10^8)'s final product at 2^21, b = 34; the square
    # of a 3*2^25-bit operand at 3*2^21, b = 32
    pl = oracle.plan(1 << 26, 1 << 26)
    assert (pl["L"], pl["bw"], pl["r3"], pl["split"]) == (1 << 22, 32, 0, 0)
    fb = fib_fingerprint(10 ** 8)["bits"]
    pl = oracle.plan(fb // 2 + 1, fb // 2 + 2)
    assert (pl["L"], pl["bw"]) == (1 << 21, 34)
    pl = oracle.plan(3 << 25, 3 << 25, True)
    assert (pl["L"], pl["bw"], pl["r3"]) == (3 << 21, 32, 1)
    # the width bound at its edge: the widest width whose bound holds for a
    # product of two c-coefficient operands is the one the reference picks
    # among equal lengths (narrowest), so a width refused here must leave
    # coefficients that can exceed p1 p2 p3: c (2^w - 1)^2 >= p1 p2 p3 at the
    # first refused width of each size (checked on the all-ones worst case)
    M = PRIMES[0] * PRIMES[1] * PRIMES[2]
    for bits in (1000, 10 ** 6, 2 ** 26, 10 ** 9, 3 * 10 ** 9):
        ok = [w for w in range(32, 36) if oracle.width_ok(bits, bits, w)]
        assert ok == list(range(32, 32 + len(ok))), bits
        if len(ok) < 4:
            w = 32 + len(ok)
            c = (bits + w - 1) // w
            assert c * ((1 << w) - 1) ** 2 >= M // 4, bits


def _rand(rng, n, kind):
    if kind == 0:
        x = rng.getrandbits(64 * n)
    elif kind == 1:
        x = (1 << (64 * n)) - 1
    else:
        x = sum(rng.getrandbits(64) << (64 * i) for i in range(n) if rng.randrange(7) == 0)
    return x | (1 << (64 * n - 1))


def test_products_vs_python(oracle):
    rng = random.Random(2)
    for t in range(150):
        na = rng.randint(1, 3000)
        nb = rng.randint(1, na)
        a, b = _rand(rng, na, t % 3), _rand(rng, nb, (t + 1) % 3)
        sq = t % 3 == 0
        got = from_limbs(oracle.mul_ntt(to_limbs(a), None if sq else to_limbs(b), square=sq))
        assert got == (a * a if sq else a * b), t


@pytest.mark.parametrize("max_lg", [10, 12])
def test_split_products_vs_python(oracle, max_lg):
    """mul_split.c restated, forced like the reference's -DNTT_MAX_LG=10|12
    variants (tests/test.ps1:29,38-39)."""
    oracle.set_max_lg(max_lg)
    try:
        rng = random.Random(max_lg)
        for t in range(12):
            lo = (1 << max_lg) // 3 + 50  # conv above 2^max_lg: the split path
            na = rng.randint(lo, 4000)
            nb = rng.randint(lo, na)
            a, b = _rand(rng, na, t % 3), _rand(rng, nb, 0)
            sq = t % 2 == 0
            pl = oracle.plan(a.bit_length(), (a if sq else b).bit_length(), sq)
            assert pl["split"] == 1
            got = from_limbs(oracle.mul_ntt(to_limbs(a), None if sq else to_limbs(b), square=sq))
            assert got == (a * a if sq else a * b), t
    finally:
        oracle.set_max_lg(25)


@pytest.mark.parametrize("x,y", [(64 * 1000, 64 * 1000), (64 * 20000 + 17, 64 * 20000 + 17), (64 * 3000, 64 * 50)])
def test_all_ones_closed_form(oracle, x, y):
    a, b = (1 << x) - 1, (1 << y) - 1
    got = from_limbs(oracle.mul_ntt(to_limbs(a), None if x == y else to_limbs(b), square=x == y))
    assert got == (1 << (x + y)) - (1 << x) - (1 << y) + 1


def test_fib_vs_python(oracle):
    for n in list(range(120)) + [1000, 12345, 99999, 300001]:
        assert from_limbs(oracle.fibonacci(n)) == fib_int(n), n


def test_fib_step(oracle):
    """The oracle's linear step (fib.c:36-62, 81-109) from s1 = F(k)^2 and
    s2 = F(k-1)^2 must give (F(2k-1), F(2k)) for bit 0 and (F(2k), F(2k+1))
    for bit 1, checked against F by Python-int fast doubling; k covers both
    signs of 2(-1)^k, sizes where s1 is one limb longer than s2, and a k
    whose F(k)^2 ends in a long run of one bits (a borrow chain)."""
    ks = [2, 3, 5, 64, 93, 94, 1000, 1001, 4096, 12345]
    for k in ks:
        f = {i: fib_int(i) for i in (k - 1, k, 2 * k - 1, 2 * k, 2 * k + 1)}
        s1, s2 = to_limbs(f[k] ** 2), to_limbs(f[k - 1] ** 2)
        for bit in (0, 1):
            lo, hi = oracle.fib_step(s1, s2, k & 1, bit)
            want = (f[2 * k], f[2 * k + 1]) if bit else (f[2 * k - 1], f[2 * k])
            assert (from_limbs(lo), from_limbs(hi)) == want, (k, bit)
    # the step is linear in (s1, s2): all-ones limbs force every carry and
    # borrow through the whole length (A = s1 + s2, B = 4 s1 - s2 + c, C = B - A)
    s1, s2 = (1 << 6400) - 1, (1 << 6336) - 1
    for k_odd in (0, 1):
        c = -2 if k_odd else 2
        for bit in (0, 1):
            lo, hi = oracle.fib_step(to_limbs(s1), to_limbs(s2), k_odd, bit)
            A, B = s1 + s2, 4 * s1 - s2 + c
            want = (B - A, B) if bit else (A, B - A)
            assert (from_limbs(lo), from_limbs(hi)) == want, (k_odd, bit)


def test_fib_golden(oracle):
    for n, g in GOLDEN["fib"].items():
        n = int(n)
        if n > 10 ** 7:
            continue  # F(10^8) is checked on the GPU side and by the reference test
        x = from_limbs(oracle.fibonacci(n))
        assert hashlib.sha256(format(x, "x").encode()).hexdigest() == g["hex_sha256"]
        assert oracle.fingerprint(to_limbs(x)) == {"bits": g["bits"], "low64": g["low64"], "mods": g["mods"]}


def test_fingerprint_method():
    """fib_mod / Binet bits against exact values on small n."""
    for n in (1, 2, 10, 93, 94, 1000, 4321):
        x = fib_int(n)
        assert fib_mod(n, 2 ** 64) == x % 2 ** 64
        assert all(fib_mod(n, p) == x % p for p in FINGERPRINT_PRIMES)
        assert fib_fingerprint(n)["bits"] == x.bit_length()


def test_decimal_vs_python(oracle):
    rng = random.Random(3)
    nums = [0, 1, 9, 10, 2 ** 64 - 1, 2 ** 64, 10 ** 19 - 1, 10 ** 19, 10 ** 19 + 1]
    for i in range(0, 8):
        for k in (19 << i, (19 << i) + 1, (19 << i) - 1, 38 << i, (38 << i) - 1):
            for d in (-1, 0, 1):
                nums.append(10 ** k + d)
    for _ in range(40):
        bits = int(2 ** rng.uniform(1, 17))
        nums.append(rng.getrandbits(bits))
    for x in nums:
        assert oracle.to_string(to_limbs(x)) == str(x), x


def test_decimal_golden(oracle):
    g = GOLDEN["decimal"]
    s = oracle.to_string(oracle.fibonacci(10 ** 5))
    assert hashlib.sha256(s.encode()).hexdigest() == g["fib_100000"]["sha256"]
    for h, d in g["edges_hex_to_decimal"].items():
        assert oracle.to_string(to_limbs(int(h, 16))) == d


def test_recip_window(oracle):
    """decimal.c:35-38: floor(2^(2s)/P) - 4 <= recip <= floor(2^(2s)/P)."""
    rng = random.Random(4)
    for s in (50, 63, 64, 130, 1000, 20000):
        P = rng.getrandbits(s) | (1 << (s - 1))
        y = from_limbs(oracle.recip(to_limbs(P), s))
        q = (1 << (2 * s)) // P
        assert q - 4 <= y <= q, s


def test_product_golden(oracle):
    """The 2^26-bit workload's digests, from the oracle (plain path)."""
    g = GOLDEN["product_2p26"]
    s = 88172645463325252
    limbs = []
    for _ in range(2):
        out = np.empty(1 << 20, dtype=np.uint64)
        for i in range(1 << 20):
            s ^= (s << 13) & 0xFFFFFFFFFFFFFFFF
            s ^= s >> 7
            s ^= (s << 17) & 0xFFFFFFFFFFFFFFFF
            out[i] = s
        out[-1] |= np.uint64(1 << 63)
        limbs.append(out)
    a, b = limbs
    assert hashlib.sha256(a.tobytes()).hexdigest() == g["a_sha256"]
    assert hashlib.sha256(oracle.mul_ntt(a, square=True).tobytes()).hexdigest() == g["square_sha256"]


# -------------------------------------------------- against the reference build


def test_oracle_vs_reference_products(oracle, ref):
    rng = random.Random(5)
    for t in range(60):
        na = rng.randint(1, 20000)
        nb = rng.randint(1, na)
        a = to_limbs(_rand(rng, na, t % 3))
        b = to_limbs(_rand(rng, nb, (t + 2) % 3))
        sq = t % 3 == 0
        assert (oracle.mul_ntt(a, None if sq else b, square=sq) == ref.mul_ntt(a, None if sq else b, square=sq)).all()


def test_oracle_vs_reference_fib_and_decimal(oracle, ref):
    for n in (0, 1, 2, 93, 94, 200, 4097, 65537, 10 ** 6):
        f = oracle.fibonacci(n)
        assert (f == ref.fibonacci(n)).all(), n
        if n <= 65537:
            assert oracle.to_string(f) == ref.to_string(f)

"""GPU products through the C-ABI, bit-exact.

Small and medium sizes are checked against exact Python-int products, large
ones against the reference's own big_mul_ntt (oracle/_ref) and closed forms.
Sizes sweep every schedule kind: one-block transforms (L <= 4096, 2^k and
3*2^k), one and two outer passes, the radix-3 top pass, widths 32..35, and
forced lengths. Operands follow the reference's fuzzer (tests/nttfuzz.c:
random, all-ones and sparse limbs with the top bit set, a third squares).
"""
import random

import numpy as np
import pytest

from oracle.oracle import from_limbs, to_limbs

pytestmark = pytest.mark.gpu


class XorShift:
    """tests/nttfuzz.c:10-18's generator and seed."""

    def __init__(self, seed=88172645463325252):
        self.s = seed

    def __call__(self):
        s = self.s
        s ^= (s << 13) & 0xFFFFFFFFFFFFFFFF
        s ^= s >> 7
        s ^= (s << 17) & 0xFFFFFFFFFFFFFFFF
        self.s = s
        return s


def _operand(rng, n, kind):
    """n limbs, top bit set; kind 0 random, 1 all ones, 2 sparse (nttfuzz.c:22-30)."""
    x = 0
    for i in range(n):
        if kind == 0:
            v = rng()
        elif kind == 1:
            v = (1 << 64) - 1
        else:
            v = rng() if rng() % 7 == 0 else 0
        x |= v << (64 * i)
    return x | (1 << (64 * n - 1))


def test_nttfuzz_products(ff):
    """The reference fuzzer's 500 small trials (tests/nttfuzz.c:44-61), sizes
    up to 2000 limbs, a third squares, against Python ints."""
    rng = XorShift()
    for t in range(500):
        na = 1 + rng() % 2000
        nb = 1 + rng() % na if rng() % 3 == 0 else na
        kind = rng() % 3
        a = _operand(rng, na, kind)
        b = _operand(rng, nb, rng() % 3)
        square = rng() % 3 == 0
        want = a * a if square else a * b
        got = from_limbs(ff.big_mul_ntt(a) if square else ff.big_mul_ntt(a, b))
        assert got == want, f"trial {t}: na={na} nb={nb} kind={kind} square={square}"


@pytest.mark.parametrize("na,nb", [(1, 1), (3, 2), (17, 17), (100, 64), (200, 200), (513, 300),
                                   (1000, 1000), (1025, 7), (2047, 2047), (3000, 2500)])
def test_small_products(ff, na, nb):
    rng = XorShift(na * 7919 + nb)
    for kind in range(3):
        a, b = _operand(rng, na, kind), _operand(rng, nb, (kind + 1) % 3)
        assert from_limbs(ff.big_mul(a, b)) == a * b
        assert from_limbs(ff.big_sqr(a)) == a * a


@pytest.mark.parametrize("na", [4000, 6000, 12000, 24000, 50000, 100000, 300000])
def test_medium_products(ff, na):
    rng = random.Random(na)
    a = rng.getrandbits(64 * na) | (1 << (64 * na - 1))
    b = rng.getrandbits(64 * (na - 3)) | (1 << (64 * (na - 3) - 1))
    assert from_limbs(ff.big_mul(a, b)) == a * b
    assert from_limbs(ff.big_sqr(a)) == a * a


@pytest.mark.parametrize("na", [600000, 1 << 20, 3 << 19, 1500000, 3000000])
def test_large_products_vs_reference(ff, ref, na):
    """2^k and 3*2^k transforms up to 2^23, one and two outer passes, against
    the reference's big_mul_ntt on the same operands."""
    rng = random.Random(na + 1)
    a = to_limbs(rng.getrandbits(64 * na) | (1 << (64 * na - 1)))
    b = to_limbs(rng.getrandbits(64 * (na - 1)) | 1)
    assert (ff.big_mul_ntt(a) == ref.mul_ntt(a, square=True)).all()
    assert (ff.big_mul_ntt(a, b) == ref.mul_ntt(a, b)).all()


@pytest.mark.parametrize("nb", [1, 2, 65, 4097, 65537, (1 << 18) - 1, (1 << 18) + 1, 400000])
def test_uneven_products_vs_reference(ff, ref, nb):
    """A 2^19-limb operand times operands of many sizes, against the
    reference's big_mul_ntt: the smaller operand's forward top-pass tile is
    zero from some row on, and those rows are stored without loads in batches
    of rows; the sizes put that row at many places against the batches, at
    widths from 32 to 35 bits (the narrower the smaller operand, the wider)."""
    rng = random.Random(nb + 7)
    na = 1 << 19
    a = to_limbs(rng.getrandbits(64 * na) | (1 << (64 * na - 1)))
    b = to_limbs(rng.getrandbits(64 * nb) | (1 << (64 * nb - 1)))
    assert (ff.big_mul_ntt(a, b) == ref.mul_ntt(a, b)).all()
    assert (ff.big_mul_ntt(b, a) == ref.mul_ntt(b, a)).all()


def xorshift_limbs(rng, n):
    """n random limbs from the generator, top bit set (nttfuzz.c:22-30)."""
    out = np.fromiter((rng() for _ in range(n)), dtype=np.uint64, count=n)
    out[-1] |= np.uint64(1 << 63)
    return out


def test_config2_product(ff, ref):
    """Two seeded 2^26-bit operands (2^20 limbs, top bit set), square and
    general: L = 2^22 at b = 32 under the reference rule, and forced 3*2^21;
    against the reference build and the committed digests."""
    import hashlib
    import json
    import os
    golden = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))["product_2p26"]
    rng = XorShift()
    n = 1 << 20
    a = xorshift_limbs(rng, n)
    b = xorshift_limbs(rng, n)
    assert hashlib.sha256(a.tobytes()).hexdigest() == golden["a_sha256"]
    assert hashlib.sha256(b.tobytes()).hexdigest() == golden["b_sha256"]
    want_sq, want_ab = ref.mul_ntt(a, square=True), ref.mul_ntt(a, b)
    assert hashlib.sha256(want_sq.tobytes()).hexdigest() == golden["square_sha256"]
    assert hashlib.sha256(want_ab.tobytes()).hexdigest() == golden["general_sha256"]
    assert (ff.big_mul_ntt(a) == want_sq).all()
    assert (ff.big_mul_ntt(a, b) == want_ab).all()
    ff.set_option("force_len", 3 << 21)
    try:
        assert (ff.big_mul_ntt(a) == want_sq).all()
        assert (ff.big_mul_ntt(a, b) == want_ab).all()
    finally:
        ff.set_option("force_len", 0)


@pytest.mark.parametrize("force", [64, 96, 3 << 10, 1 << 13, 3 << 12, 1 << 15, 3 << 16])
def test_forced_lengths(ff, force):
    """Any supported length at least conv gives the same product."""
    rng = random.Random(force)
    na = max(1, force // 8)  # conv about force / 2: the forced length is longer than the natural one
    a = rng.getrandbits(64 * na) | (1 << (64 * na - 1))
    b = rng.getrandbits(64 * na) | 1
    ff.set_option("force_len", force)
    try:
        assert from_limbs(ff.big_mul_ntt(a, b)) == a * b
        assert from_limbs(ff.big_mul_ntt(a)) == a * a
    finally:
        ff.set_option("force_len", 0)


@pytest.mark.parametrize("inner_log,tile", [(4, 1024), (7, 4096), (10, 16384), (12, 2048)])
def test_schedule_knobs(ff, inner_log, tile):
    """Inner block and tile sizes change the passes, never the product."""
    rng = random.Random(inner_log * 100 + tile)
    for na in (300, 5000, 70000):
        a = rng.getrandbits(64 * na) | (1 << (64 * na - 1))
        b = rng.getrandbits(64 * na - 17) | 1
        ff.set_option("inner_log", inner_log)
        ff.set_option("tile_words", tile)
        try:
            assert from_limbs(ff.big_mul_ntt(a, b)) == a * b
            assert from_limbs(ff.big_mul_ntt(a)) == a * a
        finally:
            ff.set_option("inner_log", 12)
            ff.set_option("tile_words", 4096)


@pytest.mark.parametrize("x,y", [(64 * 1000, 64 * 1000), (64 * 100000, 64 * 100000), (64 * 542377, 64 * 542377),
                                 (64 * 1100000, 64 * 1100000), (64 * 2000000 + 17, 64 * 2000000 + 17),
                                 (64 * 1100000, 64 * 1000000 + 3), (64 * 3000000, 64 * 50000),
                                 (64 * 2500000 + 9, 64 * 2400000)])
def test_all_ones_closed_form(ff, x, y):
    """tests/bound.c:15-58: (2^x - 1)(2^y - 1) = 2^(x+y) - 2^x - 2^y + 1, the
    CRT bound's worst case (every coefficient at its maximum) and long
    all-ones carry runs."""
    a, b = (1 << x) - 1, (1 << y) - 1
    got = from_limbs(ff.big_mul_ntt(a)) if x == y else from_limbs(ff.big_mul_ntt(a, b))
    assert got == (1 << (x + y)) - (1 << x) - (1 << y) + 1


def test_zero_and_one(ff):
    assert from_limbs(ff.big_mul(0, 12345)) == 0
    assert from_limbs(ff.big_mul((1 << 200) - 1, 0)) == 0
    assert from_limbs(ff.big_mul(1, (1 << 640) + 7)) == (1 << 640) + 7
    assert from_limbs(ff.big_sqr(1)) == 1


def test_device_resident(ff):
    """ff_dmul on numbers kept in device memory."""
    rng = random.Random(5)
    a = rng.getrandbits(64 * 40000) | 1
    b = rng.getrandbits(64 * 30000) | 1
    A, B = ff.Device(a), ff.Device(b)
    C = ff.dmul(A, B)
    assert C.value() == a * b
    assert C.bits == (a * b).bit_length()
    D = ff.dmul(C)
    assert D.value() == (a * b) ** 2

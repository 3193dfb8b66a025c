"""GPU products through the C-ABI against exact Python-int products.

Sizes sweep every schedule kind: one-block transforms (L <= 4096, 2^k and
3*2^k), one and two outer passes, the radix-3 top pass, widths 32..35, and
forced lengths. Operands follow the reference's fuzzer (tests/nttfuzz.c:
random, all-ones and sparse limbs with the top bit set, a third squares).
"""
import random

import pytest

from oracle.oracle import from_limbs, to_limbs

pytestmark = pytest.mark.gpu


def _operand(rng, n, kind):
    if kind == 0:
        x = rng.getrandbits(64 * n)
    elif kind == 1:
        x = (1 << (64 * n)) - 1
    else:
        x = 0
        for i in range(n):
            if rng.randrange(7) == 0:
                x |= rng.getrandbits(64) << (64 * i)
    return x | (1 << (64 * n - 1))


@pytest.mark.parametrize("na,nb", [(1, 1), (3, 2), (17, 17), (100, 64), (200, 200), (513, 300),
                                   (1000, 1000), (1025, 7), (2047, 2047), (3000, 2500)])
def test_small_products(ff, na, nb):
    rng = random.Random(na * 7919 + nb)
    for kind in range(3):
        a, b = _operand(rng, na, kind), _operand(rng, nb, (kind + 1) % 3)
        print(f"small na={na} nb={nb} kind={kind}", flush=True)
        assert from_limbs(ff.big_mul(a, b)) == a * b
        assert from_limbs(ff.big_sqr(a)) == a * a


@pytest.mark.parametrize("na", [4000, 6000, 12000, 24000, 50000, 100000, 300000])
def test_large_products(ff, na):
    rng = random.Random(na)
    a, b = _operand(rng, na, 0), _operand(rng, na - 3, 0)
    print(f"large na={na}", flush=True)
    assert from_limbs(ff.big_mul(a, b)) == a * b
    assert from_limbs(ff.big_sqr(a)) == a * a


def test_all_ones_closed_form(ff):
    # (2^x - 1)(2^y - 1) = 2^(x+y) - 2^x - 2^y + 1 (tests/bound.c:15-26)
    for x, y in [(64 * 1000, 64 * 1000), (64 * 100000, 64 * 100000), (64 * 2000 + 17, 64 * 1500 - 5)]:
        a, b = (1 << x) - 1, (1 << y) - 1
        print(f"ones x={x} y={y}", flush=True)
        got = from_limbs(ff.big_mul(a, b)) if x != y else from_limbs(ff.big_sqr(a))
        assert got == (1 << (x + y)) - (1 << x) - (1 << y) + 1


def test_fibonacci_small(ff):
    from oracle.oracle import fib_int
    for n in list(range(0, 120)) + [1000, 12345, 100000, 1000000]:
        got = from_limbs(ff.fibonacci(n))
        assert got == fib_int(n), n

"""Products below the NTT threshold by the schoolbook kernel (both operands
under 200 limbs: the reference's non-NTT range, mul_basic.c:64-106), against
Python integers: random, all-ones (every column sum and carry at its
maximum), sparse and zero-run operands, squares, every length pair at the
edges, through the drop-in entry, the device entry and every batch form,
and the same products with the schoolbook path switched off (the one-block
NTT) for identical limbs."""
import random

import numpy as np
import pytest

from oracle.oracle import from_limbs, to_limbs

pytestmark = pytest.mark.gpu


def _rand(rng, n, kind):
    if kind == 0:
        x = rng.getrandbits(64 * n)
    elif kind == 1:
        x = (1 << (64 * n)) - 1
    elif kind == 2:
        x = sum(rng.getrandbits(64) << (64 * i) for i in range(n) if rng.randrange(5) == 0)
    else:
        x = ((1 << (64 * n)) - 1) ^ (((1 << (64 * (n // 2))) - 1) << 32)  # a long zero run inside ones
    return x | (1 << (64 * n - 1))


@pytest.fixture(autouse=True)
def _school_bound(ff):
    """Every product under 200 limbs through the schoolbook kernels (the
    library's default bound is lower, where they measured faster)."""
    ff.set_option("school_limbs", 200)
    yield
    ff.set_option("school_limbs", 32)


def test_school_products_vs_python(ff):
    rng = random.Random(11)
    sizes = [1, 2, 3, 63, 64, 65, 127, 128, 150, 198, 199]
    for na in sizes:
        for nb in sizes:
            for kind in range(4):
                a, b = _rand(rng, na, kind), _rand(rng, nb, (kind + 1) % 4)
                assert from_limbs(ff.big_mul(a, b)) == a * b, (na, nb, kind)
        for kind in range(4):
            a = _rand(rng, na, kind)
            assert from_limbs(ff.big_sqr(a)) == a * a, (na, kind)


def test_school_matches_ntt(ff):
    """The same products with the schoolbook bound at 0 (every size through
    the one-block NTT) give identical limbs."""
    rng = random.Random(12)
    pairs = [(_rand(rng, rng.randint(1, 199), rng.randrange(4)), _rand(rng, rng.randint(1, 199), rng.randrange(4)))
             for _ in range(60)]
    school = [ff.big_mul(a, b) for a, b in pairs]
    ff.set_option("school_limbs", 0)
    try:
        ntt = [ff.big_mul(a, b) for a, b in pairs]
    finally:
        ff.set_option("school_limbs", 200)
    assert all(np.array_equal(x, y) for x, y in zip(school, ntt))


def test_school_in_batches(ff):
    rng = random.Random(13)
    pairs, want = [], []
    for t in range(300):
        na = rng.randint(1, 199)
        a = _rand(rng, na, t % 4)
        da = ff.Device(a)
        if t % 3 == 0:
            pairs.append((da, None))
            want.append(a * a)
        else:
            # small partners and, every fifth, a large one (the NTT path) in the same batch
            nb = rng.randint(1, 199) if t % 5 else rng.randint(300, 3000)
            b = _rand(rng, nb, (t + 1) % 4)
            pairs.append((da, ff.Device(b)))
            want.append(a * b)
    for flags in (ff.BATCH_PLAIN, ff.BATCH_GROUPED, ff.BATCH_GRAPH, ff.BATCH_QUEUE):
        bt = ff.Batch(pairs, flags)
        bt.run()
        bt.finish()
        assert bt.results() == want, flags
        assert [o.n for o in bt.outs] == [max(1, (w.bit_length() + 63) // 64) for w in want], flags
        bt.close()

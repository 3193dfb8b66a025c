"""Batched products (ff_batch_*) on the GPU, bit-exact.

Every run form (one product at a time, grouped work lists, the captured
graph, the persistent work queue alone and captured) is checked against exact Python-int products on mixed batches that
cover every schedule kind at once: one-block transforms, one and two outer
passes, radix-3 lengths, widths 32..35, squares by pointer identity, zero
operands. The full 4096-product uneven batch is checked per product through
residues modulo 61-bit primes (the product's residue is the product of the
operands' residues).
"""
import numpy as np
import pytest

from fastfib_amd.workloads import operand_limbs, uneven_mix
from oracle.oracle import from_limbs

pytestmark = pytest.mark.gpu

PRIMES = (2305843009213693951, 2305843009213693921, 1152921504606846883)


def _mixed_pairs(ff, seed, count, lo, hi):
    rng = np.random.default_rng(seed)
    mix = uneven_mix(seed, count, lo, hi)
    pairs, want = [], []
    for ba, bb in mix:
        A = operand_limbs(rng, ba)
        da = ff.Device(A)
        if bb is None:
            pairs.append((da, None))
            want.append(from_limbs(A) ** 2)
        else:
            B = operand_limbs(rng, bb)
            pairs.append((da, ff.Device(B)))
            want.append(from_limbs(A) * from_limbs(B))
    return pairs, want


@pytest.mark.parametrize("flags", [0, 1, 2, 16, 18])
def test_mixed_batch_exact(ff, flags):
    """80 products, operands of 2^6 .. 2^21 bits, against Python ints."""
    pairs, want = _mixed_pairs(ff, 11 + flags, 80, 6.0, 21.0)
    bt = ff.Batch(pairs, flags)
    bt.run()
    bt.finish()
    got = bt.results()
    bad = [i for i in range(len(want)) if got[i] != want[i]]
    assert not bad, f"products {bad[:5]} differ (flags {flags})"
    for i, o in enumerate(bt.outs):
        assert o.n == max(1, (want[i].bit_length() + 63) // 64)
    bt.close()


@pytest.mark.parametrize("waves", [1, 3, 4])
def test_queue_register_budgets(ff, waves):
    """The work-queue kernel under each register budget (unbounded, 3 and 4
    waves per SIMD) on a mixed batch, twice in a row (its counters are
    cleared by the launch itself)."""
    pairs, want = _mixed_pairs(ff, 41, 60, 6.0, 21.0)
    ff.set_option("queue_waves", waves)
    try:
        bt = ff.Batch(pairs, ff.BATCH_QUEUE)
        for rep in range(2):
            bt.run()
            bt.finish()
            assert bt.results() == want, (waves, rep)
        bt.close()
    finally:
        ff.set_option("queue_waves", 4)


def test_batch_rerun_new_values(ff):
    """A built batch re-reads its operands on each run: new values written in
    place (same limb counts, bit lengths not above the declared ones) give
    new products; graph replays match the grouped launch."""
    rng = np.random.default_rng(5)
    sizes = [(3000, 2000), (70, None), (40000, 39000), (1 << 17, None), (9000, 50)]
    pairs = []
    for ba, bb in sizes:
        da = ff.Device(operand_limbs(rng, ba))
        db = None if bb is None else ff.Device(operand_limbs(rng, bb))
        pairs.append((da, db))
    ptrs = [(da.h.d, None if db is None else db.h.d) for da, db in pairs]
    g = ff.Batch(pairs, 2)
    p = ff.Batch(pairs, 1)
    q = ff.Batch(pairs, 18)
    for rep in range(3):
        vals = []
        for (ba, bb), (da, db) in zip(sizes, pairs):
            A = operand_limbs(rng, ba)
            B = None if bb is None else operand_limbs(rng, bb)
            da.assign(A)
            if db is not None:
                db.assign(B)
            vals.append(from_limbs(A) ** 2 if B is None else from_limbs(A) * from_limbs(B))
        assert [(da.h.d, None if db is None else db.h.d) for da, db in pairs] == ptrs
        for _ in range(2):
            g.run()
            g.finish()
            assert g.results() == vals, f"graph, values {rep}"
        p.run()
        p.finish()
        assert p.results() == vals, f"grouped, values {rep}"
        for _ in range(2):
            q.run()
            q.finish()
            assert q.results() == vals, f"queue graph, values {rep}"
    g.close()
    p.close()
    q.close()


@pytest.mark.parametrize("flags", [2, 18])
def test_graph_replays_without_rebuild(ff, flags):
    """The captured graph (of the grouped launches, or of the one queue
    launch) replayed many times over the same operands, with other library
    work in between, gives the same products each time."""
    pairs, want = _mixed_pairs(ff, 23, 40, 8.0, 20.0)
    bt = ff.Batch(pairs, flags)
    x = ff.Device(12345678901234567890123 ** 50)
    for rep in range(4):
        bt.run()
        bt.finish()
        assert bt.results() == want, f"replay {rep}"
        y = ff.dmul(x, x)
        assert y.value() == (12345678901234567890123 ** 50) ** 2
    bt.close()


def test_zero_and_tiny_items(ff):
    zero = ff.Device(0)
    one = ff.Device(1)
    big = ff.Device((1 << 4000) - 1)
    pairs = [(zero, big), (big, zero), (zero, None), (one, one), (one, big), (big, None)]
    want = [0, 0, 0, 1, (1 << 4000) - 1, ((1 << 4000) - 1) ** 2]
    for flags in (0, 1, 2, 16):
        bt = ff.Batch(pairs, flags)
        bt.run()
        bt.finish()
        assert bt.results() == want, flags
        assert [o.n for o in bt.outs] == [1, 1, 1, 1, 63, 125]
        bt.close()


def test_uneven_4096_residues(ff):
    """The full uneven batch (4096 products, operands 2^12 .. 2^24 bits, 30%
    squares, seed 0) checked per product modulo three 61-bit primes, grouped,
    graph and work-queue forms."""
    rng = np.random.default_rng(0)
    mix = uneven_mix(0, 4096)
    pairs, res = [], []
    for ba, bb in mix:
        A = operand_limbs(rng, ba)
        a = from_limbs(A)
        ra = [a % p for p in PRIMES]
        da = ff.Device(A)
        if bb is None:
            pairs.append((da, None))
            res.append([x * x % p for x, p in zip(ra, PRIMES)])
        else:
            B = operand_limbs(rng, bb)
            b = from_limbs(B)
            pairs.append((da, ff.Device(B)))
            res.append([x * (b % p) % p for x, p in zip(ra, PRIMES)])
    for flags in (1, 2, 16):
        bt = ff.Batch(pairs, flags)
        bt.run()
        bt.finish()
        for i, o in enumerate(bt.outs):
            v = o.value()
            assert [v % p for p in PRIMES] == res[i], f"product {i} {mix[i]} flags {flags}"
        bt.close()


def test_uneven_4096_vs_reference(ff, ref):
    """The full uneven batch (4096 products, operands 2^12 .. 2^24 bits
    log-uniform, 30% squares, seed 0) limb for limb against the reference
    build's big_mul / big_sqr over the same operands, in every form: one
    product at a time, grouped work lists, the captured graph and the
    captured work queue."""
    rng = np.random.default_rng(0)
    mix = uneven_mix(0, 4096)
    items, pairs = [], []
    for ba, bb in mix:
        A = operand_limbs(rng, ba)
        B = None if bb is None else operand_limbs(rng, bb)
        items.append((A, B, B is None))
        pairs.append((ff.Device(A), None if B is None else ff.Device(B)))
    want = ref.mul_list(items)
    for flags in (0, 1, 2, 18):
        bt = ff.Batch(pairs, flags)
        bt.run()
        bt.finish()
        for i, o in enumerate(bt.outs):
            assert np.array_equal(o.limbs(), want[i]), f"product {i} {mix[i]} flags {flags}"
        bt.close()


def test_shared_operand_one_block(ff):
    """Products short enough for one-block transforms (L <= 4096, radix 2
    and 3) sharing their b: one spectrum of b per group, computed ahead of
    the products and read by each; groups of one, other operands and
    squares mixed in; every form, then new values in place."""
    rng = np.random.default_rng(77)
    Bs = [operand_limbs(rng, 9000), operand_limbs(rng, 3000)]
    dbs = [ff.Device(x) for x in Bs]
    sizes = [int(s) for s in np.exp(rng.uniform(np.log(1000), np.log(60000), 24))]
    As = [operand_limbs(rng, s) for s in sizes]
    das = [ff.Device(a) for a in As]
    pairs = [(das[k], dbs[k % 2]) for k in range(len(As))] + [(das[0], None), (dbs[1], None)]
    for flags in (0, 1, 2, 16):
        want = [from_limbs(As[k]) * from_limbs(Bs[k % 2]) for k in range(len(As))]
        want += [from_limbs(As[0]) ** 2, from_limbs(Bs[1]) ** 2]
        bt = ff.Batch(pairs, flags)
        bt.run()
        bt.finish()
        assert bt.results() == want, flags
        # new values of the same lengths in place: the spectra are formed again
        for k in range(2):
            Bs[k] = operand_limbs(rng, 9000 if k == 0 else 3000)
            dbs[k].assign(Bs[k])
        want = [from_limbs(As[k]) * from_limbs(Bs[k % 2]) for k in range(len(As))]
        want += [from_limbs(As[0]) ** 2, from_limbs(Bs[1]) ** 2]
        bt.run()
        bt.finish()
        assert bt.results() == want, ("rerun", flags)
        bt.close()


def test_shared_operand(ff):
    """Many products with one common b (transformed once for all of them),
    mixed with squares and other products, in every form."""
    rng = np.random.default_rng(31)
    B = operand_limbs(rng, 300000)
    db = ff.Device(B)
    b = from_limbs(B)
    pairs, want = [], []
    for bits in (1000, 5000, 90000, 300000, 700000, 300000, 64):
        A = operand_limbs(rng, bits)
        pairs.append((ff.Device(A), db))
        want.append(from_limbs(A) * b)
    pairs.append((db, None))
    want.append(b * b)
    for flags in (0, 1, 2, 16):
        bt = ff.Batch(pairs, flags)
        bt.run()
        bt.finish()
        assert bt.results() == want, flags
        bt.close()

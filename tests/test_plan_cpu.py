"""The library's product schedule on the host (ff_product_plan, no device
needed) against the oracle's restatement of big_mul_ntt's rule
(mul_ntt.c:52-118) and of the split plan (mul_split.c:146-177), including
the cases a planner can trip on: zero operands, one-bit and one-limb
operands, sizes at the width and length boundaries, split sizes and forced
small transform bounds (the reference's -DNTT_MAX_LG)."""
import ctypes
import random

import pytest

import fastfib_amd


class PlanInfo(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint), ("length", ctypes.c_uint64), ("conv", ctypes.c_uint64),
                ("radix3", ctypes.c_int), ("split", ctypes.c_int), ("pieces_a", ctypes.c_uint64),
                ("pieces_b", ctypes.c_uint64), ("zero", ctypes.c_int)]


@pytest.fixture(scope="module")
def plan():
    L = fastfib_amd.lib()
    L.ff_product_plan.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(PlanInfo)]

    def f(a, b, sq=False):
        p = PlanInfo()
        assert L.ff_product_plan(a, b, int(sq), ctypes.byref(p)) == 0
        return p
    return f


def _same(p, q):
    assert (p.width, p.conv, p.radix3, p.split) == (q["bw"], q["conv"], q["r3"], q["split"])
    if q["split"]:
        # the piece transform length; the oracle's L is the residue array's
        assert (p.length, p.pieces_a, p.pieces_b) == (1 << q["lgp"], q["ka"], q["kb"])
        assert q["L"] == (p.pieces_a + p.pieces_b) * p.length // 2
    else:
        assert p.length == q["L"]


def test_zero_and_tiny_operands(plan, oracle):
    for a, b in ((0, 0), (0, 5000), (7000, 0)):
        assert plan(a, b).zero == 1
    assert plan(0, 123, True).zero == 1
    for a in (1, 2, 63, 64, 65, 127, 128, 129):
        for b in (1, 64, 65, 3000):
            p = plan(a, b)
            assert p.zero == 0 and p.length >= 16 and p.length >= p.conv
            _same(p, oracle.plan(a, b))
    p = plan(1, 1)
    assert (p.width, p.length, p.conv) == (32, 16, 1)


def test_random_sizes(plan, oracle):
    rng = random.Random(5)
    for _ in range(3000):
        a = int(2 ** rng.uniform(0, 30))
        b = int(2 ** rng.uniform(0, 30))
        sq = rng.random() < 0.3
        _same(plan(a, b if not sq else a, sq), oracle.plan(a, b if not sq else a, sq))


def test_length_and_width_edges(plan, oracle):
    for L in (16, 48, 64, 3 << 9, 1 << 12, 3 << 12, 1 << 20, 3 << 21, 1 << 22, 3 << 23, 1 << 25):
        for w in (32, 33, 34, 35):
            for conv in (L - 1, L, L + 1):
                ca = (conv + 1) // 2
                bits = ca * w
                for d in (-1, 0, 1):
                    _same(plan(bits + d, bits + d, True), oracle.plan(bits + d, bits + d, True))


@pytest.mark.parametrize("max_lg", [12, 16, 25])
def test_split_plans(plan, oracle, max_lg):
    fastfib_amd.set_option("max_lg", max_lg)
    oracle.set_max_lg(max_lg)
    try:
        rng = random.Random(max_lg)
        top = 33 + max_lg
        for _ in range(400):
            a = int(2 ** rng.uniform(max_lg + 3, top))
            b = int(2 ** rng.uniform(max_lg + 3, top))
            sq = rng.random() < 0.3
            p = plan(a, a if sq else b, sq)
            q = oracle.plan(a, a if sq else b, sq)
            _same(p, q)
    finally:
        fastfib_amd.set_option("max_lg", 25)
        oracle.set_max_lg(25)

"""The allocation-free product (ff_dmul_noalloc) against ff_dmul.

Both entry points run the same product pipeline; the allocation-free one
takes its scratch from the fixed per-GPU array instead of the pool. Two
calls in a row, the second smaller (the array keeps the first call's
contents), general products and squares, one-block and multi-pass
transforms, radix-3 lengths, the largest accepted length, a zero operand,
and the refusals (too long, too small an output) - each result limb for
limb equal to ff_dmul's, with no allocation counted across the calls.
"""
import ctypes

import numpy as np
import pytest

from fastfib_amd.workloads import operand_limbs

pytestmark = pytest.mark.gpu

FF_ERR_ARG, FF_ERR_UNSUPPORTED = -3, -4


def _noalloc(ff):
    L = ff.lib()
    L.ff_dmul_noalloc.argtypes = [ctypes.POINTER(ff.DBig), ctypes.POINTER(ff.DBig), ctypes.POINTER(ff.DBig)]
    L.ff_dmul_noalloc.restype = ctypes.c_int
    return L.ff_dmul_noalloc


def _out(ff, cap):
    o = ff.Device()
    assert ff.lib().ff_dbig_alloc(ctypes.byref(o.h), cap) == 0
    return o


# (bits of a, bits of b or None for a square): large first, then smaller
CASES = [
    [(9_000_000, 7_000_000), (300_000, 200_000)],
    [(16_000_000, None), (5_000, None)],
    [(1_000_000, 900_000), (40_000, 3_000)],
    [(60_000, 60_000), (100, 64)],
    [(16_700_000, 16_700_000), (3_000_000, None)],
]


@pytest.mark.parametrize("case", CASES, ids=[str(c) for c in CASES])
def test_noalloc_matches_dmul(ff, case):
    fn = _noalloc(ff)
    rng = np.random.default_rng(sum(a for a, _ in case))
    ops = []
    for ba, bb in case:
        A = ff.Device(operand_limbs(rng, ba))
        B = None if bb is None else ff.Device(operand_limbs(rng, bb))
        want = ff.dmul(A, B).limbs()  # the existing entry point (allocates)
        out = _out(ff, A.n + (A.n if B is None else B.n) + 1)
        ops.append((A, B, out, want))
    before = ff.mem_stats()["device_allocs"]
    for A, B, out, want in ops:
        rc = fn(ctypes.byref(out.h), ctypes.byref(A.h), None if B is None else ctypes.byref(B.h))
        assert rc == 0, ff.lib().ff_last_error()
        assert out.n == want.size and out.bits == int(want[-1]).bit_length() + 64 * (want.size - 1)
        assert np.array_equal(out.limbs(), want)
    assert ff.mem_stats()["device_allocs"] == before


def test_noalloc_square_in_place_and_zero(ff):
    fn = _noalloc(ff)
    rng = np.random.default_rng(3)
    v = operand_limbs(rng, 2_000_000)
    X = ff.Device(v)
    want = ff.dmul(X).limbs()
    # x = x * x with out == a: the forward pass reads x before the CRT writes
    Y = _out(ff, 2 * X.n + 1)
    Y.assign(v)
    assert fn(ctypes.byref(Y.h), ctypes.byref(Y.h), None) == 0
    assert np.array_equal(Y.limbs(), want)
    Z = ff.Device(0)
    out = _out(ff, X.n + 2)
    assert fn(ctypes.byref(out.h), ctypes.byref(X.h), ctypes.byref(Z.h)) == 0
    assert out.n == 1 and out.bits == 0 and int(out.limbs()[0]) == 0


def test_noalloc_refusals(ff):
    fn = _noalloc(ff)
    rng = np.random.default_rng(4)
    big = ff.Device(operand_limbs(rng, 40_000_000))  # needs a 2^22-word transform
    out = _out(ff, 2 * big.n + 1)
    before = ff.mem_stats()["device_allocs"]
    assert fn(ctypes.byref(out.h), ctypes.byref(big.h), None) == FF_ERR_UNSUPPORTED
    small = ff.Device(operand_limbs(rng, 10_000))
    tight = _out(ff, 2 * small.n)  # one limb short of na + nb + 1
    after_alloc = ff.mem_stats()["device_allocs"]
    assert fn(ctypes.byref(tight.h), ctypes.byref(small.h), None) == FF_ERR_ARG
    assert ff.mem_stats()["device_allocs"] == after_alloc
    assert after_alloc >= before

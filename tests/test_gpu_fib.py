"""F(n) on the GPU through the C-ABI, bit-exact.

The n lists follow the reference's own checks: every n < 200, 40 seeded
random n up to 3e6 and the values around 5e6 and 2^22 (tests/check.py:17-18),
the large list of tests/bigcheck.ps1:8 up to 1e8 (against the reference
build), and the fingerprint of tests/fingerprint.py / memrun.c (bit length,
low 64 bits, F(n) mod four primes) at full size. The linear step and the
last-step factors are also checked on their own with carry-adversarial
operands (all-ones and all-zero limb runs), which force the exact carry scan.
"""
import hashlib
import json
import os
import random

import numpy as np
import pytest

from oracle.oracle import FINGERPRINT_PRIMES, fib_fingerprint, fib_int, from_limbs, to_limbs

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.json")))


def _check_fingerprint(x: int, n: int) -> None:
    fp = fib_fingerprint(n)
    assert x.bit_length() == fp["bits"], n
    assert x % (1 << 64) == fp["low64"], n
    assert [x % p for p in FINGERPRINT_PRIMES] == fp["mods"], n


def test_fib_small_n(ff):
    for n in range(200):
        assert from_limbs(ff.fibonacci(n)) == fib_int(n), n


def test_fib_check_py_list(ff):
    rng = random.Random(1)
    ns = [rng.randint(200, 3_000_000) for _ in range(40)]
    ns += [4_999_999, 5_000_000, 2 ** 22 - 1, 2 ** 22, 2 ** 22 + 1]
    for n in ns:
        assert from_limbs(ff.fibonacci(n)) == fib_int(n), n


@pytest.mark.parametrize("n", [20_000_000, 33_554_431, 50_000_001, 77_777_777, 100_000_000])
def test_fib_large_vs_reference(ff, ref, n):
    """tests/bigcheck.ps1:8's n (up to 1e8) against the reference build, and
    independently of it by the fingerprint (bit length from Binet's formula,
    low 64 bits and F(n) mod four primes by fast doubling on residues)."""
    got = ff.fibonacci(n)
    assert np.array_equal(got, ref.fibonacci(n)), n
    _check_fingerprint(from_limbs(got), n)


@pytest.mark.parametrize("mode", [0, 4, 8], ids=["reference_rule", "tuned", "learned"])
def test_fib_1e9_schedule_sources(ff, mode):
    """F(10^9), the learned-dispatcher workload, with every product's
    schedule from the reference rule, the autotuner table or the MLP (the
    committed dispatch.bin): the same number whatever the source, against
    the reference build's hex digest and the fingerprint."""
    from fastfib_amd import dispatch as dp

    dp.load(os.path.join(os.path.dirname(dp.__file__), "dispatch.bin"))
    ff.set_option("dispatch", mode)
    try:
        d = ff.dfibonacci(10 ** 9)
        x = d.value()
    finally:
        ff.set_option("dispatch", 0)
    g = GOLDEN["fib_large"]["1000000000"]
    assert d.bits == g["bits"] == 694241913
    _check_fingerprint(x, 10 ** 9)
    assert hashlib.sha256(format(x, "x").encode()).hexdigest() == g["hex_sha256"]


def test_fib_1e8_fingerprint(ff):
    fp = fib_fingerprint(10 ** 8)
    d = ff.dfibonacci(10 ** 8)
    x = d.value()
    assert d.bits == fp["bits"] == x.bit_length() == 69424191
    assert x % (1 << 64) == fp["low64"]
    from oracle.oracle import FINGERPRINT_PRIMES
    assert [x % p for p in FINGERPRINT_PRIMES] == fp["mods"]


@pytest.mark.parametrize("n", [264_889, 563_764, 2_387_614])
def test_fib_step_lengths_predicted(ff, n, capfd):
    """Each doubling or quadrupling step books its results' lengths from
    their top limbs, which must fall among the last three limbs the step
    writes. For these n a quadrupling step's result is up to 4 limbs shorter
    than four times its input, so the step's limb count has to be tight:
    the values are exact and the predicted lengths are confirmed without the
    read-back rerun."""
    assert from_limbs(ff.fibonacci(n)) == fib_int(n), n
    assert "predicted lengths disagreed" not in capfd.readouterr().err


@pytest.mark.parametrize("mode", [0, 1, 3])
def test_fib_pair_steps(ff, mode):
    """Doubling steps and the last product as weighted pointwise pairs
    (option pair_steps), against exact values and the F(10^8) fingerprint,
    with quadrupling off and on."""
    from oracle.oracle import FINGERPRINT_PRIMES
    ns = [200_003, 563_764, 2_387_614, 5_000_000]
    try:
        ff.set_option("pair_steps", mode)
        for q in (0, 18):
            ff.set_option("quad_max_lg", q)
            for n in ns:
                assert from_limbs(ff.fibonacci(n)) == fib_int(n), (mode, q, n)
        fp = fib_fingerprint(10 ** 8)
        x = ff.dfibonacci(10 ** 8).value()
        assert x.bit_length() == fp["bits"] and x % (1 << 64) == fp["low64"]
        assert [x % p for p in FINGERPRINT_PRIMES] == fp["mods"]
    finally:
        ff.set_option("pair_steps", 2)
        ff.set_option("quad_max_lg", 18)


def test_fib_host_bits(ff):
    """The top bits of n on the host (10..13) before the device chain, at
    sizes where the chain starts right after them and further on."""
    try:
        for hb in (10, 11, 12, 13):
            ff.set_option("host_bits", hb)
            for n in [8_191, 8_192, 16_385, 563_764, 5_000_000]:
                assert from_limbs(ff.fibonacci(n)) == fib_int(n), (hb, n)
    finally:
        ff.set_option("host_bits", 13)


def test_fib_quad_off(ff):
    """Doubling only (quad_max_lg 0), the path of the reference's fib.c."""
    try:
        ff.set_option("quad_max_lg", 0)
        for n in [200_003, 563_764, 5_000_000]:
            assert from_limbs(ff.fibonacci(n)) == fib_int(n), n
    finally:
        ff.set_option("quad_max_lg", 18)


def test_fib_fingerprint_sweep(ff, capfd):
    """250 seeded n, log-uniform from 2^13 to 2^27, odd and even: every mix of
    host bits, one-workgroup steps, quadruplings, doubling steps and the
    last product's two forms. Each result's bit length, low 64 bits and
    residues mod four primes against the fingerprint (fast doubling on
    residues), and every predicted length confirmed (no read-back rerun)."""
    import math

    from oracle.oracle import FINGERPRINT_PRIMES
    rng = random.Random(20261203)
    ns = sorted({int(math.exp(rng.uniform(math.log(2 ** 13), math.log(2 ** 27)))) for _ in range(250)})
    for n in ns:
        x = ff.dfibonacci(n).value()
        fp = fib_fingerprint(n)
        assert x.bit_length() == fp["bits"], n
        assert x % (1 << 64) == fp["low64"], n
        assert [x % p for p in FINGERPRINT_PRIMES] == fp["mods"], n
    assert "predicted lengths disagreed" not in capfd.readouterr().err


def test_fib_repeatable(ff):
    """The sync-free path (lengths predicted on the host) and repeated runs
    agree with the reference values."""
    want = fib_int(1_234_567)
    for _ in range(3):
        assert from_limbs(ff.fibonacci(1_234_567)) == want


def _step_ref(s1, s2, k_odd, bit):
    A = s1 + s2
    B = 4 * s1 - s2 + (-2 if k_odd else 2)
    C = B - A
    return (C, B) if bit else (A, C)


@pytest.mark.parametrize("nl", [1, 2, 15, 16, 17, 33, 100, 5000, 70000])
def test_fib_step_linear_adversarial(ff, nl):
    """fib.c:36-62 on operands whose chunks are all ones or all zero, so that
    carries and borrows run across chunk boundaries (the exact carry path)."""
    rng = random.Random(nl)
    ones = (1 << (64 * nl)) - 1
    cases = [
        (ones, 1),
        (ones, ones),
        (1 << (64 * nl), (1 << (64 * nl)) - 1),
        (1 << (64 * nl - 1), 3),
        (rng.getrandbits(64 * nl) | (1 << (64 * nl - 1)), rng.getrandbits(64 * nl - 5)),
        ((ones << 64) | 2, 3),
    ]
    for s1, s2 in cases:
        assert s2 <= s1
        for k_odd in (0, 1):
            for bit in (0, 1):
                lo, hi = ff.fib_step_linear(s1, s2, k_odd, bit)
                want = _step_ref(s1, s2, k_odd, bit)
                assert (from_limbs(lo), from_limbs(hi)) == want, (nl, k_odd, bit)


@pytest.mark.parametrize("nl", [1, 16, 17, 40, 3000])
def test_last_factors_adversarial(ff, nl):
    ones = (1 << (64 * nl)) - 1
    for u, v in [(ones, ones), (ones, 1), (1 << (64 * nl), 1), (12345 << (64 * nl - 20), 777)]:
        x, y = ff.last_factors(u, v)
        assert from_limbs(x) == 2 * u + v
        assert from_limbs(y) == 2 * u - v
        x, none = ff.last_factors(v, u, with_y=False)
        assert from_limbs(x) == 2 * v + u and none is None


def test_lookback_timeout_fails_loudly():
    """A carry look-back wait that ran out (forced from the host: the testing
    hook sets the time-out word, no kernel waits) makes the next length
    read-back print a message and exit(1), in a process of its own."""
    import subprocess
    import sys

    code = ("import fastfib_amd as ff; ff.init(); "
            "assert ff.lib().ff_test_set_lb_fault() == 0; "
            "ff.dfibonacci(10 ** 6); print('returned')")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300,
                       cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    assert r.returncode == 1, (r.returncode, r.stdout[-400:], r.stderr[-400:])
    assert "look-back wait timed out" in r.stderr
    assert "returned" not in r.stdout

"""The device memory pool and the failure paths of the library, on the GPU.

The reference's arena keeps warm pages for reuse and, when memory runs out,
prints "out of memory (N bytes)" and exits with status 1
(alloc.c:19-23,74-146). Here the device pool must reach a steady state in
which repeating the same work (F(n), a batch replay, a decimal conversion,
in-place products) allocates nothing new, and an allocation failure -
injected from the host, so no memory is actually exhausted - must end the
process the same way. A carry look-back wait that timed out (its word set
from the host, no kernel waits) must fail loudly from every entry point that
reads results back, each case in a process of its own.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from fastfib_amd.workloads import operand_limbs, uneven_mix
from oracle.oracle import from_limbs

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _allocs(ff):
    s = ff.mem_stats()
    return s["device_allocs"], s["host_allocs"]


def test_fibonacci_steady_state(ff):
    ff.dfibonacci(10 ** 8).release()
    ff.dfibonacci(10 ** 8).release()
    before = _allocs(ff)
    for _ in range(3):
        d = ff.dfibonacci(10 ** 8)
        assert d.bits == 69424191
        d.release()
    assert _allocs(ff) == before


def test_batch_replay_steady_state(ff):
    rng = np.random.default_rng(11)
    pairs, want = [], []
    for ba, bb in uneven_mix(11, 256, 12.0, 20.0):
        A = operand_limbs(rng, ba)
        B = None if bb is None else operand_limbs(rng, bb)
        pairs.append((ff.Device(A), None if B is None else ff.Device(B)))
        want.append(from_limbs(A) * (from_limbs(A) if B is None else from_limbs(B)))
    for flags in (ff.BATCH_GROUPED, ff.BATCH_GRAPH, ff.BATCH_QUEUE | ff.BATCH_GRAPH):
        bt = ff.Batch(pairs, flags)
        bt.run()
        bt.finish()
        before = _allocs(ff)
        for _ in range(5):
            bt.run()
        bt.finish()
        assert _allocs(ff) == before, flags
        assert bt.results() == want, flags
        bt.close()


def test_decimal_steady_state(ff):
    rng = np.random.default_rng(12)
    x = ff.Device(operand_limbs(rng, 3_000_000))
    s0 = x.to_string()
    before = _allocs(ff)
    for _ in range(3):
        assert x.to_string() == s0
    assert _allocs(ff) == before
    assert s0 == str(x.value())


def test_in_place_products_return_their_blocks(ff):
    """ff_dmul(&x, &x, NULL) squares x in place; the old block goes back to
    the pool, so running the same sequence again allocates nothing."""
    rng = np.random.default_rng(13)
    v = from_limbs(operand_limbs(rng, 40000))
    x = ff.Device(v)
    for _ in range(5):
        ff.dmul_into(x, x)
    assert x.value() == v ** 32
    before = _allocs(ff)
    for _ in range(2):
        x.assign(v)
        for _ in range(5):
            ff.dmul_into(x, x)
    assert x.value() == v ** 32
    assert _allocs(ff) == before


def _run(code: str):
    return subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300, cwd=ROOT)


def test_out_of_memory_exits():
    """An injected allocation failure: the pool gives back its cached blocks,
    retries once, then prints the reference's message and exits 1."""
    code = ("import fastfib_amd as ff; ff.init(); ff.set_option('test_fail_alloc', 0); "
            "ff.big_mul((1 << 300000) - 1, (1 << 200000) + 1); print('returned')")
    r = _run(code)
    assert r.returncode == 1, (r.returncode, r.stdout[-400:], r.stderr[-400:])
    assert "out of memory (" in r.stderr and "bytes)" in r.stderr
    assert "returned" not in r.stdout


@pytest.mark.parametrize("entry", ["batch", "queue", "decimal"])
def test_lookback_timeout_fails_loudly_everywhere(entry):
    """The time-out word set from the host: a batch's finish (grouped, or a
    work queue launched after it is set) and a decimal conversion (both carry their products by look-back) end the process
    with the message and status 1 instead of returning wrong limbs."""
    setup = ("import fastfib_amd as ff; ff.init(); "
             "a = ff.Device((1 << 200000) - 3); b = ff.Device((1 << 150000) + 7); ")
    if entry == "batch":
        work = ("bt = ff.Batch([(a, b), (a, None)], ff.BATCH_GROUPED); bt.run(); "
                "assert ff.lib().ff_test_set_lb_fault() == 0; bt.finish(); print('returned')")
    elif entry == "queue":
        # set before the launch: the work queue's dependency waits give up
        # at once (every task still runs and the kernel ends)
        work = ("bt = ff.Batch([(a, b), (a, None)], ff.BATCH_QUEUE); "
                "assert ff.lib().ff_test_set_lb_fault() == 0; bt.run(); bt.finish(); print('returned')")
    else:
        work = "assert ff.lib().ff_test_set_lb_fault() == 0; a.to_string(); print('returned')"
    r = _run(setup + work)
    assert r.returncode == 1, (r.returncode, r.stdout[-400:], r.stderr[-400:])
    assert "look-back wait timed out" in r.stderr
    assert "returned" not in r.stdout

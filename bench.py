#!/usr/bin/env python3
"""Benchmark: F(10^8) computed bit-exact on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--cpu-seconds S]

A step is one complete F(n) (n = 10^8 by default: 25 doubling steps with
two squarings each, plus the final product) computed by fast doubling with
every product, CRT and carry pass on the GPU, the result left in HBM. W
untimed steps, then K timed steps between device synchronizations; the
result of the last step is checked bit-exact against an independent
fingerprint (bit length by Binet's formula, the low 64 bits and F(n) modulo
four fixed primes by fast doubling on residues).

Besides the timing the JSON line carries:
  roofline      the dominant kernel's algorithmic bytes per launch over its
                average launch time, both measured live with HIP events
                around each launch (one extra profiled step), against the
                8 TB/s HBM peak;
  cpu_baseline  the reference's own F(n) (its C sources built by
                oracle/Makefile), timed on this host's cores in the same run.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
if hasattr(sys, "set_int_max_str_digits"):
    sys.set_int_max_str_digits(0)

METRIC = "F(1e8) ms bit-exact; batched uneven-mul Gbit/s vs plain GPU; HBM % of peak"
HBM_PEAK_GBS = 8000.0
FP_PRIMES = (18446744073709551557, 18446744073709551533, 9223372036854775783, 4611686018427387847)


def fib_mod(n: int, m: int) -> int:
    a, b = 0, 1 % m
    for bit in bin(n)[2:]:
        c = a * ((2 * b - a) % m) % m
        d = (a * a + b * b) % m
        a, b = (d, (c + d) % m) if bit == "1" else (c, d)
    return a


def fib_bits(n: int) -> int:
    if n == 0:
        return 0
    from mpmath import floor, log, mp, sqrt
    mp.dps = 60
    phi = (1 + sqrt(5)) / 2
    return int(floor(n * log(phi, 2) - log(sqrt(5), 2))) + 1


def fingerprint_ok(n: int, limbs) -> bool:
    x = int.from_bytes(limbs.tobytes(), "little")
    if x.bit_length() != fib_bits(n):
        return False
    if x % (1 << 64) != fib_mod(n, 1 << 64):
        return False
    return all(x % p == fib_mod(n, p) for p in FP_PRIMES)


def _profiled(file: str, tag: str, kind: str):
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", file)
    try:
        with open(path) as f:
            for line in f:
                w = line.split()
                if len(w) == 3 and w[0] == tag and w[1] == kind:
                    return int(w[2])
    except OSError:
        pass
    return None


def profiled_traffic(kind: str):
    """HBM bytes per launch of a kernel kind from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (calibrated against a plain copy), or
    None when the summary does not cover it."""
    return _profiled("fib_1e8_hbm_traffic.md", "traffic_bytes_per_launch", kind)


# One wave64 VALU instruction per 4 cycles per SIMD: 256 CUs x 4 SIMDs x 2.4 GHz / 4.
VALU_PEAK_G = 256 * 4 * 2.4 / 4
# 32-bit integer multiplies at the VALU issue rate, one per lane per cycle
# (256 CUs x 64 lanes x 2.4 GHz): the spec figure beside the measured probe.
INT32_MUL_SPEC_T = 256 * 64 * 2.4e9 / 1e12


def int32_probe(ff) -> dict:
    """The measured INT32-multiply peak: the library's probe kernel
    (v_mul_lo_u32 / v_mul_hi_u32 chains on every SIMD), best of three runs."""
    L = ff.lib()
    L.ff_probe_int32_mul.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    rate, secs = ctypes.c_double(), ctypes.c_double()
    ff._check(L.ff_probe_int32_mul(ctypes.byref(rate), ctypes.byref(secs)))
    return {"measured_T_per_s": round(rate.value / 1e12, 2), "launch_ms": round(secs.value * 1e3, 3),
            "spec_T_per_s": round(INT32_MUL_SPEC_T, 1)}


def copy_probe(ff, mib: int = 256) -> dict:
    """The measured HBM bandwidth: the library's plain device copy of mib
    MiB (k_copy_probe, 4 launches, the best of the last three), bytes read
    plus written per second. Its dispatches are also the calibration of the
    profiler's FETCH_SIZE / WRITE_SIZE counters."""
    L = ff.lib()
    L.ff_probe_copy.argtypes = [ctypes.c_size_t, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    rate, secs = ctypes.c_double(), ctypes.c_double()
    ff._check(L.ff_probe_copy(mib << 20, ctypes.byref(rate), ctypes.byref(secs)))
    return {"measured_GB_per_s": round(rate.value / 1e9, 1), "copy_MiB": mib, "launch_ms": round(secs.value * 1e3, 4)}


def profiled_valu(kind: str):
    """VALU wave-instructions per launch of a kernel kind from the committed
    rocprofv3 SQ_INSTS_VALU summary, or None."""
    return _profiled("fib_1e8_valu.md", "valu_insts_per_launch", kind)


def kernel_stats(ff) -> dict:
    L = ff.lib()
    L.ff_prof_read.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_long),
                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    out = {}
    k = 0
    while True:
        name, n, ms, b = ctypes.c_char_p(), ctypes.c_long(), ctypes.c_double(), ctypes.c_double()
        if L.ff_prof_read(k, ctypes.byref(name), ctypes.byref(n), ctypes.byref(ms), ctypes.byref(b)) != 0:
            break
        if n.value:
            m = ctypes.c_double()
            L.ff_prof_read_mults(k, ctypes.byref(m))
            out[name.value.decode()] = {"launches": n.value, "ms": ms.value, "bytes": b.value, "mults": m.value}
        k += 1
    return out


def host_cpu() -> dict:
    """The host's CPU as the baseline runs on it: model, logical CPUs of the
    machine, CPUs this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count(), "affinity": affinity}


REF_BUILDS = (("x86-64-v3", "libffref.so"), ("native", "libffref_native.so"))
REF_CHOSEN = None  # the build the F(n) baseline found fastest; the other legs time it too


def reference():
    from oracle.oracle import Reference
    return Reference(REF_CHOSEN) if REF_CHOSEN else Reference()


def _ref_build_runs(path: str) -> bool:
    """Whether a build of the reference runs on this host (a build for
    another CPU's instruction set would fault): F(10^5) in a child process,
    checked by its fingerprint."""
    import subprocess
    code = ("import sys; sys.path.insert(0, %r); from oracle.oracle import Reference, fib_fingerprint, from_limbs; "
            "x = from_limbs(Reference(%r).fibonacci(100000)); fp = fib_fingerprint(100000); "
            "assert x.bit_length() == fp['bits'] and x %% (1 << 64) == fp['low64']; print('ok')") % (ROOT, path)
    try:
        r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    except subprocess.TimeoutExpired:
        return False
    return r.returncode == 0 and r.stdout.strip().endswith("ok")


def cpu_baseline(n: int, seconds: float) -> dict | None:
    """The reference's fibonacci(n) on this host (oracle/_ref), best of runs
    within a time budget (the README method), for each build of it that runs
    here: the portable AVX2 build (-march=x86-64-v3) and the build for this
    host's own instruction set (gcc -march=native as the GPU host's gcc
    expands it), when present. The pool uses every CPU this process may run
    on, capped at 16 (FF_REF_THREADS): the reference starts one spinning
    worker per CPU, and a GPU host shares its cores."""
    host = host_cpu()
    threads = min(host["affinity"], 16)
    os.environ["FF_REF_THREADS"] = str(threads)
    try:
        from oracle.oracle import REF_SO, Reference
    except Exception as e:  # not built: report nothing rather than a guess
        print(f"cpu baseline unavailable: {e}", file=sys.stderr)
        return None
    builds = {}
    best_build, best, pool = None, None, None
    for name, so in REF_BUILDS:
        path = os.path.join(os.path.dirname(REF_SO), so)
        if not os.path.exists(path) or (name != "x86-64-v3" and not _ref_build_runs(path)):
            continue
        ref = Reference(path)
        times = []
        t_end = time.perf_counter() + seconds / 2
        ref.time_fibonacci(min(n, 10 ** 6))  # start the pool
        while time.perf_counter() < t_end or len(times) < 3:
            times.append(ref.time_fibonacci(n))
            if len(times) >= 100:
                break
        b = min(times)
        builds[name] = {"best_ms": round(b * 1e3, 3), "median_ms": round(sorted(times)[len(times) // 2] * 1e3, 3),
                        "runs": len(times)}
        pool = ref.pool_cpus()
        if best is None or b < best:
            best, best_build = b, name
    if best is None:
        return None
    global REF_CHOSEN
    REF_CHOSEN = os.path.join(os.path.dirname(REF_SO), dict(REF_BUILDS)[best_build])
    return {"value": round(best * 1e3, 3), "unit": "ms", "cores": pool, "kind": "reference",
            "sample": f"fibonacci({n}), best of runs in one process (README method: best time), "
                      f"build {best_build}; pool threads = cores",
            "host": host, "threads": pool, "builds": builds}


BATCH_PRIMES = (2305843009213693951, 1152921504606846883)


def batch_leg(ff, count: int, seed: int, steps: int, cpu_seconds: float, cpu: bool = True, only=None) -> dict:
    """The uneven batch: count products, operand bits log-uniform in
    [2^12, 2^24], 30% squares, uniform limbs (fastfib_amd.workloads), one
    step = the whole batch. Run forms: one product at a time (plain), grouped
    work lists, the grouped launches replayed from a HIP graph, the
    persistent work queue (one launch per run; alone, captured, and under
    the unbounded and 3-waves-per-SIMD register budgets). Every form's
    outputs are checked: the first form's per product modulo two primes, the
    others byte-equal to it."""
    import numpy as np
    from fastfib_amd.workloads import operand_bits, operand_limbs, uneven_mix

    mix = uneven_mix(seed, count)
    gbit = operand_bits(mix) / 1e9
    rng = np.random.default_rng(seed)
    pairs, res, host = [], [], []
    for ba, bb in mix:
        A = operand_limbs(rng, ba)
        B = None if bb is None else operand_limbs(rng, bb)
        a = int.from_bytes(A.tobytes(), "little")
        b = a if B is None else int.from_bytes(B.tobytes(), "little")
        res.append(tuple(a % p * (b % p) % p for p in BATCH_PRIMES))
        pairs.append((ff.Device(A), None if B is None else ff.Device(B)))
        host.append((A, B))
    ff.sync()
    forms = [("plain", ff.BATCH_PLAIN, 1), ("grouped", ff.BATCH_GROUPED, 1), ("graph", ff.BATCH_GRAPH, 1),
             ("graph_tuned", ff.BATCH_GRAPH | ff.BATCH_TUNED, 1),
             ("graph_learned", ff.BATCH_GRAPH | ff.BATCH_LEARNED, 1),
             ("queue", ff.BATCH_QUEUE, 4), ("graph_queue", ff.BATCH_GRAPH | ff.BATCH_QUEUE, 4),
             ("queue_w1", ff.BATCH_QUEUE, 1), ("queue_w3", ff.BATCH_QUEUE, 3)]
    if only:
        forms = [f for f in forms if f[0] in only]
    # the reference's own big_mul / big_sqr over the same products
    want = None
    try:
        if not cpu:
            raise RuntimeError("skipped in a profiling run")
        from oracle.oracle import Reference
        want = Reference().mul_list([(A, B, B is None) for A, B in host])
    except Exception as e:
        print(f"reference products unavailable, residues only: {e}", file=sys.stderr)
    out, ref_limbs, ok = {}, None, True
    for name, flags, waves in forms:
        ff.set_option("queue_waves", waves)
        bt = ff.Batch(pairs, flags)
        bt.run()
        bt.finish()
        n_steps = steps if name != "plain" else max(1, min(steps, 3))
        ff.sync()
        t0 = time.perf_counter()
        for _ in range(n_steps):
            bt.run()
        bt.finish()
        dt = (time.perf_counter() - t0) / n_steps
        limbs = [o.limbs() for o in bt.outs]
        if want is not None:
            good = all(np.array_equal(x, y) for x, y in zip(limbs, want))
        elif ref_limbs is None:
            good = all(tuple(int.from_bytes(x.tobytes(), "little") % p for p in BATCH_PRIMES) == r
                       for x, r in zip(limbs, res))
            ref_limbs = limbs
        else:
            good = all(np.array_equal(x, y) for x, y in zip(limbs, ref_limbs))
        ok &= good
        out[name] = {"ms": round(dt * 1e3, 3), "gbit_products_per_s": round(gbit / dt, 2), "bit_exact": good,
                     "steps": n_steps}
        bt.close()
    ff.set_option("queue_waves", 4)
    leg = {"workload": f"{count} uneven products, operands 2^12..2^24 bits log-uniform, 30% squares, seed {seed}",
           "operand_gbit": round(gbit, 3), "forms": out, "bit_exact": ok,
           "checked_against": "reference big_mul/big_sqr, limb for limb" if want is not None else "residues"}
    if "plain" in out:
        plain = out["plain"]["gbit_products_per_s"]
        leg["speedup_vs_plain"] = {k: round(v["gbit_products_per_s"] / plain, 2) for k, v in out.items()
                                   if k != "plain"}
    if cpu:
        leg["cpu_baseline"] = batch_cpu_baseline(host, gbit, cpu_seconds)
    return leg


def dispatcher_regret(live_mixes: int = 40) -> dict:
    """Regret of the learned dispatcher (inference by the library's HIP
    kernel), the autotuner table and the reference rule against the
    exhaustive best: on the committed held-out samples (timed when the
    dispatcher was trained, on the build stamped in its report), and on a
    subset of those mixes timed again now, on this build and GPU."""
    import numpy as np
    from fastfib_amd import dispatch as dp

    here = os.path.dirname(os.path.abspath(dp.__file__))
    with np.load(os.path.join(here, "dispatch_data", "heldout.npz")) as z:
        ho = {k: z[k] for k in z.files}
    K, table, params = dp.read_bin(os.path.join(here, "dispatch.bin"))
    dp.set_params(K, params)
    y, _ = dp.predict(ho["X"], ho["mask"])
    rule = dp.rule_candidates(ho)
    choice = dp.dispatch_choice(y, ho["mask"], dp.default_candidate(), rule=rule)
    tab = dp.table_choice(table, ho)

    def summary(T, sel):
        r = {"learned": dp.regret(T, choice[sel], ho["mix"][sel]),
             "table": dp.regret(T, tab[sel], ho["mix"][sel]),
             "reference_rule": dp.regret(T, rule[sel], ho["mix"][sel])}
        return {k: {"mean": round(v["mean"], 4), "aggregate": round(v["aggregate"], 4), "p95": round(v["p95"], 4)}
                for k, v in r.items()} | {"mixes": int(np.unique(ho["mix"][sel]).size)}

    allrows = np.ones(ho["X"].shape[0], dtype=bool)
    out = {"committed_timings": summary(ho["T_us"], allrows)}
    mixes = np.unique(ho["mix"])
    pick = set(mixes[:: max(1, len(mixes) // live_mixes)].tolist())
    sel = np.array([m in pick for m in ho["mix"]])
    rows = np.flatnonzero(sel)
    T = np.array([dp.measure(int(ho["bits_a"][i]), int(ho["bits_b"][i]), int(ho["square"][i]),
                             int(ho["count"][i]), seed=7 * int(i) + 1) for i in rows])
    out["retimed_now"] = summary(T, sel)
    try:
        with open(os.path.join(here, "dispatch_data", "report.json")) as f:
            out["trained_on_build"] = json.load(f).get("build")
    except OSError:
        pass
    return out


DEC_PRIMES = (2147483647, 2147483629, 2147483587, 1000000007)


def digits_residue(s: str, p: int) -> int:
    """The value of a decimal string modulo p < 2^31, by pairwise combination
    of 9-digit chunks (numpy, no big integers)."""
    import numpy as np

    d = np.frombuffer(s.encode(), dtype=np.uint8).astype(np.int64) - 48
    pad = (-d.size) % 9
    d = np.concatenate([np.zeros(pad, np.int64), d]).reshape(-1, 9)
    v = (d * (10 ** np.arange(8, -1, -1, dtype=np.int64))).sum(1) % p
    m = 10 ** 9 % p
    while v.size > 1:
        if v.size % 2:
            v = np.concatenate([np.zeros(1, np.int64), v])
        v = (v[0::2] * m + v[1::2]) % p
        m = m * m % p
    return int(v[0])


def fib_digits(n: int) -> int:
    from mpmath import floor, log10, mp, sqrt

    mp.dps = 50
    phi = (1 + sqrt(5)) / 2
    return int(floor(n * log10(phi) - log10(sqrt(5)))) + 1


def decimal_leg(ff, n: int, steps: int, cpu_seconds: float, cpu: bool = True) -> dict:
    """Decimal digits of F(n) (device-resident input): the first conversion
    (powers, reciprocals and level graphs built), then warm conversions with
    the digits left in HBM, and one with the string copied to the host. The
    string is checked by its length (Binet) and its value modulo four primes
    against F(n) modulo them (fast doubling on residues)."""
    L = ff.lib()
    ff._dec_sigs(L)
    L.ff_dbig_to_decimal.argtypes = [ctypes.POINTER(ff.DBig), ctypes.POINTER(ctypes.c_void_p),
                                     ctypes.POINTER(ctypes.c_size_t)]
    x = ff.dfibonacci(n)
    ff.sync()
    p, ln = ctypes.c_void_p(), ctypes.c_size_t()
    t0 = time.perf_counter()
    ff._check(L.ff_dbig_to_decimal(ctypes.byref(x.h), ctypes.byref(p), ctypes.byref(ln)))
    cold = time.perf_counter() - t0
    # the second conversion of a size captures its levels in a graph: untimed
    ff._check(L.ff_dbig_to_decimal(ctypes.byref(x.h), ctypes.byref(p), ctypes.byref(ln)))
    ff.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        ff._check(L.ff_dbig_to_decimal(ctypes.byref(x.h), ctypes.byref(p), ctypes.byref(ln)))
    warm = (time.perf_counter() - t0) / steps
    # the level batches in each form: grouped launches, persistent work
    # queues (one launch per batch)
    default = ff.get_option("dec_queue")
    forms = {}
    for name, q in (("grouped", 0), ("queue", 1)):
        ff.set_option("dec_queue", q)
        for _ in range(2):  # plan built, then captured
            ff._check(L.ff_dbig_to_decimal(ctypes.byref(x.h), ctypes.byref(p), ctypes.byref(ln)))
        ff.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            ff._check(L.ff_dbig_to_decimal(ctypes.byref(x.h), ctypes.byref(p), ctypes.byref(ln)))
        forms[name] = round((time.perf_counter() - t0) / steps * 1e3, 3)
    ff.set_option("dec_queue", default)
    t0 = time.perf_counter()
    s = x.to_string()
    host = time.perf_counter() - t0
    ok = len(s) == fib_digits(n) and s[0] != "0" and all(digits_residue(s, q) == fib_mod(n, q) for q in DEC_PRIMES)
    leg = {"workload": f"decimal digits of F({n})", "digits": len(s), "bit_exact": ok,
           "cold_ms": round(cold * 1e3, 3),
           "cold_note": "first conversion of this size: powers 10^(19*2^i), their reciprocals and the level "
                        "graph built, as the reference's big_to_string does on every call; the figure "
                        "compared with the CPU baseline",
           "warm_device_ms_cached": round(warm * 1e3, 3),
           "warm_device_ms_cached_by_level_form": forms,
           "level_form_default": "queue" if default else "grouped",
           "warm_to_host_string_ms_cached": round(host * 1e3, 3),
           "warm_note": "powers/reciprocals cached from the first conversion; digits left in HBM, or copied "
                        "to a host string"}
    limbs = x.limbs()
    x.release()
    leg["cpu_baseline"] = decimal_cpu_baseline(limbs, n, cpu_seconds) if cpu else None
    if leg["cpu_baseline"]:
        leg["speedup_cold_vs_cpu"] = round(leg["cpu_baseline"]["value"] / leg["cold_ms"], 2)
    return leg


def product_leg(ff, steps: int) -> dict:
    """The single large product: two 2^26-bit operands (2^20 limbs from the
    xorshift64 generator and seed of the reference's fuzzer, top bit set),
    square and general, device-resident; L = 2^22 at b = 32 by the reference
    rule, and the forced 3 * 2^21. Results checked against the committed
    digests of the reference build's products; per-pass GB/s from the live
    event timer."""
    import hashlib

    import numpy as np

    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "golden.json")) as f:
        golden = json.load(f)["product_2p26"]
    state = [88172645463325252]

    def limbs(n):
        out = np.empty(n, dtype=np.uint64)
        s = state[0]
        for i in range(n):
            s ^= (s << 13) & 0xFFFFFFFFFFFFFFFF
            s ^= s >> 7
            s ^= (s << 17) & 0xFFFFFFFFFFFFFFFF
            out[i] = s
        state[0] = s
        out[-1] |= np.uint64(1 << 63)
        return out

    A, B = limbs(1 << 20), limbs(1 << 20)
    da, db = ff.Device(A), ff.Device(B)
    L = ff.lib()
    out = {"workload": "2^26-bit operands (2^20 limbs), square and general product, device-resident"}
    ok = True
    for force in (0, 3 << 21):
        ff.set_option("force_len", force)
        try:
            for name, other, key in (("square", None, "square_sha256"), ("general", db, "general_sha256")):
                r = ff.dmul(da, other)
                ok &= hashlib.sha256(r.limbs().tobytes()).hexdigest() == golden[key]
                r.release()
                ff.sync()
                t0 = time.perf_counter()
                for _ in range(steps):
                    ff.dmul(da, other).release()
                ff.sync()
                ms = (time.perf_counter() - t0) / steps * 1e3
                L.ff_prof_reset()
                L.ff_prof_enable(1)
                ff.dmul(da, other).release()
                ff.sync()
                st = kernel_stats(ff)
                L.ff_prof_enable(0)
                passes = {k: round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) for k, v in st.items()
                          if v["ms"] > 0 and v["bytes"] > 0}
                out[f"{name}_L{force or 1 << 22}"] = {"ms": round(ms, 4), "GBps_per_kernel": passes}
        finally:
            ff.set_option("force_len", 0)
    out["bit_exact"] = ok
    da.release()
    db.release()
    return out


def large_fib_leg(ff, n: int, steps: int, cpu_seconds: float, cpu: bool = True) -> dict:
    """F(n) at a large n (10^9: 29 NTT products up to 3 * 2^23 words) with
    the reference rule, the autotuner table and the learned dispatcher
    choosing every product's schedule, run interleaved (one of each per
    round, best of the rounds); each result checked by its bit length and
    residues (fast doubling on residues)."""
    modes = (("reference_rule", 0), ("tuned", ff.BATCH_TUNED), ("learned", ff.BATCH_LEARNED))
    best = {name: float("inf") for name, _ in modes}
    good = {name: True for name, _ in modes}
    try:
        for r in range(steps + 1):
            for name, mode in modes:
                ff.set_option("dispatch", mode)
                ff.sync()
                t0 = time.perf_counter()
                x = ff.dfibonacci(n)
                ff.sync()
                dt = time.perf_counter() - t0
                if r == 0:  # warm-up round, checked
                    good[name] = fingerprint_ok(n, x.limbs())
                else:
                    best[name] = min(best[name], dt)
                x.release()
    finally:
        ff.set_option("dispatch", 0)
    out = {"workload": f"F({n}) fast doubling, device-resident", "n": n, "rounds": steps}
    for name, _ in modes:
        out[name] = {"ms": round(best[name] * 1e3, 3), "bit_exact": good[name]}
    out["bit_exact"] = all(good.values())
    out["learned_vs_reference_rule"] = round(best["reference_rule"] / best["learned"], 3)
    if cpu:
        out["cpu_baseline"] = large_fib_cpu_baseline(n, cpu_seconds)
    return out


def large_fib_cpu_baseline(n: int, seconds: float) -> dict | None:
    """The reference's fibonacci(n) on this host, best of a few runs."""
    try:
        ref = reference()
    except Exception as e:
        print(f"cpu baseline unavailable: {e}", file=sys.stderr)
        return None
    times = []
    t_end = time.perf_counter() + seconds
    while not times or time.perf_counter() < t_end:
        times.append(ref.time_fibonacci(n))
        if len(times) >= 5:
            break
    return {"value": round(min(times) * 1e3, 2), "unit": "ms", "cores": ref.pool_cpus(), "kind": "reference",
            "sample": f"fibonacci({n}) x {len(times)}, best"}


def decimal_cpu_baseline(limbs, n: int, seconds: float) -> dict | None:
    """The reference's big_to_string of the same number on this host."""
    try:
        ref = reference()
    except Exception as e:
        print(f"cpu baseline unavailable: {e}", file=sys.stderr)
        return None
    times = []
    t_end = time.perf_counter() + seconds
    while not times or time.perf_counter() < t_end:
        times.append(ref.time_to_string(limbs))
        if len(times) >= 5:
            break
    return {"value": round(min(times) * 1e3, 2), "unit": "ms", "cores": ref.pool_cpus(), "kind": "reference",
            "sample": f"big_to_string(F({n})) x {len(times)}, best"}


def batch_cpu_baseline(host, gbit: float, seconds: float) -> dict | None:
    """The reference's big_mul / big_sqr over the same products, in the
    batch's order, as far as the time budget goes (whole passes)."""
    try:
        ref = reference()
    except Exception as e:
        print(f"cpu baseline unavailable: {e}", file=sys.stderr)
        return None
    items = [(A, B, B is None) for A, B in host]
    times = []
    t_end = time.perf_counter() + seconds
    while not times or time.perf_counter() < t_end:
        times.append(ref.time_mul_list(items))
        if len(times) >= 5:
            break
    best = min(times)
    return {"value": round(gbit / best, 3), "unit": "Gbit-products/s", "cores": ref.pool_cpus(),
            "kind": "reference", "sample": f"the whole batch through big_mul/big_sqr x {len(times)}, best "
                                           f"{best * 1e3:.1f} ms"}


def small_leg(ff, seed: int, steps: int) -> dict:
    """Products below the NTT threshold (both operands under 200 limbs) by
    the schoolbook kernel against the one-block NTT they took before
    (school_limbs 0): (a) the uneven batch's products of that kind (the same
    draw as the batch leg, grouped form), (b) one whole warm decimal
    conversion of F(10^7), whose low tree levels multiply such numbers by
    the thousand. Results are compared between the two paths."""
    import numpy as np
    from fastfib_amd.workloads import operand_limbs, uneven_mix

    rng = np.random.default_rng(seed)
    pairs = []
    for ba, bb in uneven_mix(seed, 4096):
        A = operand_limbs(rng, ba)
        B = None if bb is None else operand_limbs(rng, bb)
        if A.size < 200 and (B is None or B.size < 200):
            pairs.append((ff.Device(A), None if B is None else ff.Device(B)))
    out, limbs = {}, {}
    for name, bound in (("schoolbook", 200), ("schoolbook_below_64", 64), ("schoolbook_below_32", 32),
                        ("one_block_ntt", 0)):
        ff.set_option("school_limbs", bound)
        try:
            bt = ff.Batch(pairs, ff.BATCH_GROUPED)
            bt.run()
            bt.finish()
            ff.sync()
            t0 = time.perf_counter()
            for _ in range(steps):
                bt.run()
            bt.finish()
            out[f"batch_{name}_us"] = round((time.perf_counter() - t0) / steps * 1e6, 1)
            limbs[name] = [o.limbs() for o in bt.outs]
            bt.close()
            x = ff.dfibonacci(10 ** 7)
            s0 = x.to_string()
            x.to_string()  # the plan's graph
            ff.sync()
            t0 = time.perf_counter()
            for _ in range(steps):
                s = x.to_string()
            out[f"decimal_F1e7_{name}_ms"] = round((time.perf_counter() - t0) / steps * 1e3, 3)
            limbs[name + "_dec"] = s
            x.release()
        finally:
            ff.set_option("school_limbs", 32)  # the library default
    same = True
    for name in ("schoolbook_below_64", "schoolbook_below_32", "one_block_ntt"):
        same &= all(np.array_equal(a, b) for a, b in zip(limbs["schoolbook"], limbs[name]))
        same &= limbs["schoolbook_dec"] == limbs[name + "_dec"]
    return {"workload": f"{len(pairs)} products of the uneven batch (seed {seed}) with both operands under 200 "
                        "limbs; warm decimal conversion of F(10^7)", "products": len(pairs), **out,
            "identical": bool(same)}


def run_legs_only(ff, args, legs) -> int:
    """Only the named legs, no timed F(n) (profiling runs of one workload):
    one JSON line with each leg's result, no CPU baselines."""
    out, ok = {"legs": sorted(legs)}, True
    if "batch" in legs:
        out["batch"] = batch_leg(ff, args.batch_count, args.batch_seed, args.batch_steps, 0.0, cpu=False,
                                 only=set(args.batch_forms.split(",")) if args.batch_forms else None)
        ok &= out["batch"]["bit_exact"]
    if "regret" in legs:
        out["dispatcher_regret"] = dispatcher_regret()
    if "product" in legs:
        out["product_2p26"] = product_leg(ff, args.product_steps)
        ok &= out["product_2p26"]["bit_exact"]
    if "large" in legs:
        out["large_fib"] = large_fib_leg(ff, args.large_n, args.large_steps, 0.0, cpu=False)
        ok &= out["large_fib"]["bit_exact"]
    if "decimal" in legs:
        out["decimal"] = decimal_leg(ff, args.n, args.decimal_steps, 0.0, cpu=False)
        ok &= out["decimal"]["bit_exact"]
    if "probe" in legs:
        out["int32_probe"] = int32_probe(ff)
    if "small" in legs:
        out["small_products"] = small_leg(ff, args.batch_seed, max(3, args.batch_steps))
        ok &= out["small_products"]["identical"]
    print(json.dumps(out))
    return 0 if ok else 1


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=10 ** 8)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--batch-count", type=int, default=4096, help="uneven batch size (0: skip the batch leg)")
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--batch-seed", type=int, default=0)
    ap.add_argument("--batch-forms", default="", help="comma list of batch forms for a profiling run "
                    "(plain, grouped, graph, graph_tuned, graph_learned, queue, graph_queue, queue_w1, queue_w3; "
                    "default: all)")
    ap.add_argument("--decimal-steps", type=int, default=5, help="warm decimal conversions of F(n) (0: skip)")
    ap.add_argument("--large-n", type=int, default=10 ** 9, help="F(n) leg with dispatcher schedules (0: skip)")
    ap.add_argument("--large-steps", type=int, default=5)
    ap.add_argument("--product-steps", type=int, default=10, help="2^26-bit product leg (0: skip)")
    ap.add_argument("--legs", default="fib,batch,product,large,decimal,small",
                    help="comma list of legs to run (fib: the timed F(n) and its roofline; a subset serves "
                         "profiling runs of one workload)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 256 MiB device copy three times (the reference dispatch for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes), then exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        import fastfib_amd as ff
        ff.init()
        print(json.dumps(copy_probe(ff)))
        return 0
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (products shard by prime or "
              "by item only in design); run with --gpus 1", file=sys.stderr)
        return 2

    import fastfib_amd as ff
    ff.init()
    n = args.n
    legs = set(args.legs.split(","))
    if "fib" not in legs:
        return run_legs_only(ff, args, legs)
    for _ in range(args.warmup):
        ff.dfibonacci(n).release()
    ff.sync()
    t0 = time.perf_counter()
    last = None
    for i in range(args.steps):
        if last is not None:
            last.release()
        last = ff.dfibonacci(n)
    ff.sync()
    elapsed = time.perf_counter() - t0
    ms = elapsed / args.steps * 1e3
    ok = fingerprint_ok(n, last.limbs())
    bits = last.bits
    last.release()

    # one extra step with every launch bracketed by events
    L = ff.lib()
    L.ff_prof_reset()
    L.ff_prof_enable(1)
    ff.dfibonacci(n).release()
    ff.sync()
    stats = kernel_stats(ff)
    L.ff_prof_enable(0)
    ntt = {k: v for k, v in stats.items() if k.startswith("ntt") or k.startswith("crt")}
    dom = max(ntt, key=lambda k: ntt[k]["ms"])
    d = stats[dom]
    per_launch_ms = d["ms"] / d["launches"]
    achieved = d["bytes"] / d["launches"] / (per_launch_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": profiled_traffic(dom),
                "algorithmic_bytes_per_launch": round(d["bytes"] / d["launches"]),
                "avg_launch_us": round(per_launch_ms * 1e3, 2), "launches": d["launches"]}
    if d.get("mults"):
        # algorithmic multiplies booked by the launcher (3 per twiddled
        # butterfly, 4 per pointwise Montgomery product) over the launch time,
        # against the probe's measured multiply rate
        probe = int32_probe(ff)
        peak = probe["measured_T_per_s"]
        mr = d["mults"] / d["launches"] / (per_launch_ms * 1e-3) / 1e12
        roofline["int32_mul"] = {"mults_per_launch": round(d["mults"] / d["launches"]), "achieved": round(mr, 2),
                                 "peak": peak, "unit": "T int32 mul/s", "frac": round(mr / peak, 4),
                                 "peak_source": "measured by ff_probe_int32_mul on this GPU", "probe": probe}
    # the measured HBM bandwidth of a plain copy on this GPU, beside the spec peak
    roofline["hbm_probe"] = copy_probe(ff)
    vi = profiled_valu(dom)
    if vi:
        # the transform kernels' actual bound: VALU issue (profiled instructions
        # per launch over this run's launch time)
        rate = vi / (per_launch_ms * 1e-3) / 1e9
        roofline["valu"] = {"insts_per_launch": vi, "achieved": round(rate, 1), "peak": VALU_PEAK_G,
                            "unit": "G wave64 VALU instr/s", "frac": round(rate / VALU_PEAK_G, 4)}

    cpu = cpu_baseline(n, args.cpu_seconds)
    batch = None
    if args.batch_count > 0 and "batch" in legs:
        batch = batch_leg(ff, args.batch_count, args.batch_seed, args.batch_steps, args.cpu_seconds)
        batch["dispatcher_regret"] = dispatcher_regret()
        ok &= batch["bit_exact"]
    product = None
    if args.product_steps > 0 and "product" in legs:
        product = product_leg(ff, args.product_steps)
        ok &= product["bit_exact"]
    large = None
    if args.large_n > 0 and "large" in legs:
        large = large_fib_leg(ff, args.large_n, args.large_steps, args.cpu_seconds)
        ok &= large["bit_exact"]
    small = None
    if "small" in legs:
        small = small_leg(ff, args.batch_seed, max(3, args.batch_steps))
        ok &= small["identical"]
    decimal = None
    if args.decimal_steps > 0 and "decimal" in legs:
        decimal = decimal_leg(ff, n, args.decimal_steps, args.cpu_seconds)
        ok &= decimal["bit_exact"]
    line = {
        "metric": METRIC, "value": round(ms, 4), "unit": "ms", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic (F(n) has no input)",
        "config": {"workload": f"F({n}) fast doubling, device-resident", "n": n, "result_bits": bits,
                   "bit_exact": ok, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": round(cpu["value"] / ms, 2) if cpu else None,
        "batch": batch,
        "decimal": decimal,
        "large_fib": large,
        "product_2p26": product,
        "small_products": small,
        "kernels": {k: {"launches": v["launches"], "ms": round(v["ms"], 4),
                        "GBps": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 else None}
                    for k, v in stats.items()},
    }
    print(json.dumps(line))
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())

"""Writes tests/golden/golden.json: known answers for the parity tests.

Every value comes from the reference's own C sources built by oracle/Makefile
(oracle/_ref/libffref.so) and, where Python can afford it, is cross-checked
here against Python integers or an independent method before it is written:

* F(n) for the CLI known answers (tests/cli.ps1:15-16) and F(n) for
  n in {10^6, 10^7, 10^8}: SHA-256 of the lowercase hex string
  (tests/hexfib.c:9-10's format), bit length, low 64 bits and F(n) modulo the
  four primes of tests/fingerprint.py:22 (checked against fast doubling on
  residues and Binet's formula);
* the 2^26-bit product workload: SHA-256 of the little-endian limb bytes of
  a*a and a*b for the xorshift64 operands of tests/nttfuzz.c:10-30 (seed
  88172645463325252, 2^20 limbs, top bit forced);
* F(10^9): SHA-256 of the hex string plus the fingerprint (checked as above);
* decimal strings: SHA-256 and length of big_to_string(F(10^5)) and of a few
  edge values of tests/dectest.py:8 (checked against str()), and of
  big_to_string(F(10^8)) (checked by its digit count from Binet's formula and
  its value modulo the fingerprint primes).

Run it from the repo root where the reference checkout exists:
    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import (FINGERPRINT_PRIMES, Reference, decimal_mod, fib_digits, fib_fingerprint,  # noqa: E402
                           fib_int, from_limbs)

import numpy as np  # noqa: E402


def xorshift_limbs(n, state):
    out = np.empty(n, dtype=np.uint64)
    s = state
    for i in range(n):
        s ^= (s << 13) & 0xFFFFFFFFFFFFFFFF
        s ^= s >> 7
        s ^= (s << 17) & 0xFFFFFFFFFFFFFFFF
        out[i] = s
    out[-1] |= np.uint64(1 << 63)
    return out, s


def hexstr(limbs):
    x = from_limbs(limbs)
    return format(x, "x")


def main():
    ref = Reference()
    g = {"source": "shwwwa/fastfib C sources built by oracle/Makefile; see make_golden.py"}

    cli = {}
    for n in (0, 1, 2, 7, 10, 12, 93, 94):
        s = ref.to_string(ref.fibonacci(n))
        assert s == str(fib_int(n))
        cli[str(n)] = s
    assert cli["93"] == "12200160415121876738" and cli["94"] == "19740274219868223167"
    g["fib_decimal_small"] = cli

    fib = {}
    for n in (10 ** 6, 10 ** 7, 10 ** 8):
        limbs = ref.fibonacci(n)
        x = from_limbs(limbs)
        fp = fib_fingerprint(n)
        assert x.bit_length() == fp["bits"]
        assert x % (1 << 64) == fp["low64"]
        assert [x % p for p in FINGERPRINT_PRIMES] == fp["mods"]
        if n <= 10 ** 6:
            assert x == fib_int(n)
        fib[str(n)] = {"hex_sha256": hashlib.sha256(format(x, "x").encode()).hexdigest(),
                       "limbs": int(limbs.size), **fp}
    g["fib"] = fib

    # F(10^9): the workload of the learned-dispatcher runs; the hex digest of
    # the reference's result, cross-checked by the fingerprint
    n = 10 ** 9
    limbs = ref.fibonacci(n)
    x = from_limbs(limbs)
    fp = fib_fingerprint(n)
    assert x.bit_length() == fp["bits"]
    assert x % (1 << 64) == fp["low64"]
    assert [x % p for p in FINGERPRINT_PRIMES] == fp["mods"]
    g["fib_large"] = {str(n): {"hex_sha256": hashlib.sha256(format(x, "x").encode()).hexdigest(),
                               "limbs": int(limbs.size), **fp}}
    del x, limbs

    a, s = xorshift_limbs(1 << 20, 88172645463325252)
    b, s = xorshift_limbs(1 << 20, s)
    sq = ref.mul_ntt(a, square=True)
    ab = ref.mul_ntt(a, b)
    # an independent check of the products: their residues mod a few primes
    A, B = from_limbs(a), from_limbs(b)
    for p in FINGERPRINT_PRIMES:
        assert from_limbs(sq) % p == (A % p) * (A % p) % p
        assert from_limbs(ab) % p == (A % p) * (B % p) % p
    g["product_2p26"] = {
        "generator": "xorshift64 seed 88172645463325252, a then b, 2^20 limbs each, top bit set",
        "a_sha256": hashlib.sha256(a.tobytes()).hexdigest(),
        "b_sha256": hashlib.sha256(b.tobytes()).hexdigest(),
        "square_sha256": hashlib.sha256(sq.tobytes()).hexdigest(),
        "square_limbs": int(sq.size),
        "general_sha256": hashlib.sha256(ab.tobytes()).hexdigest(),
        "general_limbs": int(ab.size),
    }

    dec = {}
    f5 = ref.fibonacci(10 ** 5)
    s5 = ref.to_string(f5)
    assert s5 == str(fib_int(10 ** 5))
    dec["fib_100000"] = {"sha256": hashlib.sha256(s5.encode()).hexdigest(), "digits": len(s5)}
    edges = {}
    for x in (0, 1, 9, 10, 2 ** 64 - 1, 2 ** 64, 10 ** 19 - 1, 10 ** 19, 10 ** 19 + 1, 10 ** 608 - 1, 10 ** 608):
        from oracle.oracle import to_limbs
        s = ref.to_string(to_limbs(x))
        assert s == str(x)
        edges[format(x, "x")] = s
    dec["edges_hex_to_decimal"] = edges

    # the decimal digits of F(10^8) by the reference's big_to_string, checked
    # independently: the digit count from Binet's formula and the string's
    # value modulo the four fingerprint primes against F(n) mod p
    n = 10 ** 8
    s8 = ref.to_string(ref.fibonacci(n))
    assert len(s8) == fib_digits(n)
    assert [decimal_mod(s8, p) for p in FINGERPRINT_PRIMES] == fib_fingerprint(n)["mods"]
    dec["fib_100000000"] = {"sha256": hashlib.sha256(s8.encode()).hexdigest(), "digits": len(s8),
                            "head": s8[:40], "tail": s8[-40:]}
    g["decimal"] = dec

    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json")
    with open(out, "w") as f:
        json.dump(g, f, indent=1, sort_keys=True)
    print("wrote", out)


if __name__ == "__main__":
    main()

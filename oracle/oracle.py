"""ORACLE — test infrastructure, not product code.

ctypes access to the two CPU checkers built by oracle/Makefile:

* ``Oracle`` wraps liboracle.so, this repo's scalar C restatement of the
  reference path (oracle.c).
* ``Reference`` wraps _ref/libffref.so, the reference's own C sources
  (shwwwa/fastfib) compiled with a Linux platform layer.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg use
this module, as the checker or the timed CPU baseline; the GPU library never
imports it. Numbers cross the boundary as Python ints or numpy uint64 limb
arrays (little-endian limbs).
"""
from __future__ import annotations

import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libffref.so")

if hasattr(sys, "set_int_max_str_digits"):
    sys.set_int_max_str_digits(0)

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_szp = ctypes.POINTER(ctypes.c_size_t)


def to_limbs(x: int) -> np.ndarray:
    """Python int (>= 0) -> normalized uint64 limb array (zero is [0])."""
    if x < 0:
        raise ValueError("negative")
    n = max(1, (x.bit_length() + 63) // 64)
    return np.frombuffer(x.to_bytes(8 * n, "little"), dtype=np.uint64).copy()


def from_limbs(d) -> int:
    a = np.ascontiguousarray(d, dtype=np.uint64)
    return int.from_bytes(a.tobytes(), "little")


def _arr(d) -> tuple[np.ndarray, _u64p, int]:
    a = np.ascontiguousarray(d, dtype=np.uint64)
    if a.size == 0:
        a = np.zeros(1, dtype=np.uint64)
    return a, a.ctypes.data_as(_u64p), a.size


class _Lib:
    def __init__(self, path: str):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run `make -C oracle all ref`)")
        self.lib = ctypes.CDLL(path)

    def _take(self, ptr, n: int, free) -> np.ndarray:
        out = np.ctypeslib.as_array(ptr, shape=(n,)).copy()
        free(ptr)
        return out


class Oracle(_Lib):
    """The C restatement (oracle.c)."""

    def __init__(self, path: str = ORACLE_SO):
        super().__init__(path)
        L = self.lib
        for name in ("ffo_mul_ntt", "ffo_mul", ):
            f = getattr(L, name)
            f.restype = _u64p
            f.argtypes = [_u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t, ctypes.c_int, _szp]
        L.ffo_mul_school.restype = _u64p
        L.ffo_mul_school.argtypes = [_u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t, _szp]
        L.ffo_fibonacci.restype = _u64p
        L.ffo_fibonacci.argtypes = [ctypes.c_uint64, _szp]
        L.ffo_to_string.restype = ctypes.c_void_p
        L.ffo_to_string.argtypes = [_u64p, ctypes.c_size_t]
        L.ffo_recip.restype = _u64p
        L.ffo_recip.argtypes = [_u64p, ctypes.c_size_t, ctypes.c_size_t, _szp]
        L.ffo_free.argtypes = [ctypes.c_void_p]
        L.ffo_set_max_lg.argtypes = [ctypes.c_int]
        L.ffo_get_max_lg.restype = ctypes.c_int
        L.ffo_width_ok.argtypes = [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_uint]
        L.ffo_ntt_length.restype = ctypes.c_size_t
        L.ffo_ntt_length.argtypes = [ctypes.c_size_t, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.ffo_cyclic_conv.argtypes = [ctypes.c_int, _u32p, _u32p, ctypes.c_size_t, _u32p]
        L.ffo_ntt_forward.argtypes = [ctypes.c_int, _u32p, ctypes.c_size_t]
        L.ffo_field.restype = ctypes.c_uint64
        L.ffo_field.argtypes = [_u64p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_uint]
        L.ffo_fingerprint.argtypes = [_u64p, ctypes.c_size_t, _u64p, _u64p, _u64p]
        L.ffo_plan_product.argtypes = [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
        L.ffo_fib_step.argtypes = [_u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(_u64p), _szp, ctypes.POINTER(_u64p), _szp]
        self.primes = [2113929217, 2013265921, 1811939329]

    def fib_step(self, s1, s2, k_odd: int, bit: int) -> tuple[np.ndarray, np.ndarray]:
        """fib_step_linear (fib.c:81-109) from s1 = F(k)^2, s2 = F(k-1)^2."""
        A, pa, na = _arr(s1)
        B, pb, nb = _arr(s2)
        lo, hi = _u64p(), _u64p()
        nlo, nhi = ctypes.c_size_t(), ctypes.c_size_t()
        self.lib.ffo_fib_step(pa, na, pb, nb, int(k_odd), int(bit), ctypes.byref(lo), ctypes.byref(nlo),
                              ctypes.byref(hi), ctypes.byref(nhi))
        return self._ret(lo, nlo.value), self._ret(hi, nhi.value)

    def _ret(self, ptr, n) -> np.ndarray:
        return self._take(ptr, n, self.lib.ffo_free)

    def mul_ntt(self, a, b=None, square=False) -> np.ndarray:
        A, pa, na = _arr(a)
        B, pb, nb = _arr(a if b is None else b)
        n = ctypes.c_size_t()
        return self._ret(self.lib.ffo_mul_ntt(pa, na, pb, nb, int(square), ctypes.byref(n)), n.value)

    def mul(self, a, b=None, square=False) -> np.ndarray:
        A, pa, na = _arr(a)
        B, pb, nb = _arr(a if b is None else b)
        n = ctypes.c_size_t()
        return self._ret(self.lib.ffo_mul(pa, na, pb, nb, int(square), ctypes.byref(n)), n.value)

    def mul_school(self, a, b) -> np.ndarray:
        A, pa, na = _arr(a)
        B, pb, nb = _arr(b)
        n = ctypes.c_size_t()
        return self._ret(self.lib.ffo_mul_school(pa, na, pb, nb, ctypes.byref(n)), n.value)

    def fibonacci(self, n: int) -> np.ndarray:
        k = ctypes.c_size_t()
        return self._ret(self.lib.ffo_fibonacci(n, ctypes.byref(k)), k.value)

    def to_string(self, a) -> str:
        A, pa, na = _arr(a)
        p = self.lib.ffo_to_string(pa, na)
        s = ctypes.string_at(p).decode()
        self.lib.ffo_free(p)
        return s

    def recip(self, P, s: int) -> np.ndarray:
        A, pa, na = _arr(P)
        k = ctypes.c_size_t()
        return self._ret(self.lib.ffo_recip(pa, na, s, ctypes.byref(k)), k.value)

    def set_max_lg(self, lg: int) -> None:
        self.lib.ffo_set_max_lg(lg)

    def width_ok(self, bits_a: int, bits_b: int, w: int) -> bool:
        return bool(self.lib.ffo_width_ok(bits_a, bits_b, w))

    def ntt_length(self, conv: int) -> tuple[int, int, int]:
        lg, r3 = ctypes.c_int(), ctypes.c_int()
        L = self.lib.ffo_ntt_length(conv, ctypes.byref(lg), ctypes.byref(r3))
        return L, lg.value, r3.value

    def plan(self, bits_a: int, bits_b: int, square: bool = False) -> dict:
        class P(ctypes.Structure):
            _fields_ = [("bw", ctypes.c_uint), ("L", ctypes.c_size_t), ("conv", ctypes.c_size_t),
                        ("lg", ctypes.c_int), ("r3", ctypes.c_int), ("split", ctypes.c_int),
                        ("lgp", ctypes.c_int), ("ka", ctypes.c_int), ("kb", ctypes.c_int)]
        p = P()
        self.lib.ffo_plan_product(bits_a, bits_b, int(square), ctypes.byref(p))
        return {f: getattr(p, f) for f, _ in P._fields_}

    def cyclic_conv(self, pi: int, a, b) -> np.ndarray:
        A = np.ascontiguousarray(a, dtype=np.uint32)
        B = np.ascontiguousarray(b, dtype=np.uint32)
        out = np.zeros(A.size, dtype=np.uint32)
        self.lib.ffo_cyclic_conv(pi, A.ctypes.data_as(_u32p), B.ctypes.data_as(_u32p), A.size,
                                 out.ctypes.data_as(_u32p))
        return out

    def ntt_forward(self, pi: int, x) -> np.ndarray:
        X = np.ascontiguousarray(x, dtype=np.uint32).copy()
        self.lib.ffo_ntt_forward(pi, X.ctypes.data_as(_u32p), X.size)
        return X

    def field(self, a, k: int, bw: int) -> int:
        A, pa, na = _arr(a)
        return int(self.lib.ffo_field(pa, na, k, bw))

    def fingerprint(self, a) -> dict:
        A, pa, na = _arr(a)
        bits, low = ctypes.c_uint64(), ctypes.c_uint64()
        mods = (ctypes.c_uint64 * 4)()
        self.lib.ffo_fingerprint(pa, na, ctypes.byref(bits), ctypes.byref(low), mods)
        return {"bits": bits.value, "low64": low.value, "mods": [int(m) for m in mods]}


class Reference(_Lib):
    """The reference's own sources (oracle/_ref/libffref.so)."""

    def __init__(self, path: str = REF_SO):
        super().__init__(path)
        L = self.lib
        for name in ("ref_mul", "ref_mul_ntt", "ref_mul_nontt"):
            f = getattr(L, name)
            f.restype = _u64p
            f.argtypes = [_u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t, ctypes.c_int, _szp]
        L.ref_fibonacci.restype = _u64p
        L.ref_fibonacci.argtypes = [ctypes.c_uint64, _szp]
        L.ref_to_string.restype = ctypes.c_void_p
        L.ref_to_string.argtypes = [_u64p, ctypes.c_size_t]
        L.ref_free.argtypes = [ctypes.c_void_p]
        L.ref_ntt_init.argtypes = [ctypes.c_size_t]
        L.ref_ntt_max_len.restype = ctypes.c_size_t
        L.ref_time_fibonacci.restype = ctypes.c_double
        L.ref_time_fibonacci.argtypes = [ctypes.c_uint64]
        L.ref_time_mul_ntt.restype = ctypes.c_double
        L.ref_time_mul_ntt.argtypes = [_u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t, ctypes.c_int]
        L.ref_time_to_string.restype = ctypes.c_double
        L.ref_time_to_string.argtypes = [_u64p, ctypes.c_size_t]
        L.ref_time_mul_list.restype = ctypes.c_double
        L.ref_time_mul_list.argtypes = [_u64p, _szp, _szp, _u64p, _szp, _szp, ctypes.POINTER(ctypes.c_int),
                                        ctypes.c_size_t]
        L.ref_pool_cpus.restype = ctypes.c_int
        L.ref_mul_list.restype = None
        L.ref_mul_list.argtypes = [_u64p, _szp, _szp, _u64p, _szp, _szp, ctypes.POINTER(ctypes.c_int), ctypes.c_size_t,
                                   _u64p, _szp, _szp]

    def _ret(self, ptr, n) -> np.ndarray:
        return self._take(ptr, n, self.lib.ref_free)

    def ntt_init(self, max_limbs: int) -> None:
        self.lib.ref_ntt_init(max_limbs)

    def _mul(self, fn, a, b, square):
        A, pa, na = _arr(a)
        B, pb, nb = _arr(a if b is None else b)
        n = ctypes.c_size_t()
        return self._ret(fn(pa, na, pb, nb, int(square), ctypes.byref(n)), n.value)

    def mul(self, a, b=None, square=False) -> np.ndarray:
        A, _, na = _arr(a)
        self.ntt_init(na + (na if b is None else _arr(b)[2]))
        return self._mul(self.lib.ref_mul, a, b, square)

    def mul_ntt(self, a, b=None, square=False) -> np.ndarray:
        A, _, na = _arr(a)
        self.ntt_init(na + (na if b is None else _arr(b)[2]))
        return self._mul(self.lib.ref_mul_ntt, a, b, square)

    def mul_nontt(self, a, b=None, square=False) -> np.ndarray:
        return self._mul(self.lib.ref_mul_nontt, a, b, square)

    def fibonacci(self, n: int) -> np.ndarray:
        k = ctypes.c_size_t()
        return self._ret(self.lib.ref_fibonacci(n, ctypes.byref(k)), k.value)

    def to_string(self, a) -> str:
        A, pa, na = _arr(a)
        p = self.lib.ref_to_string(pa, na)
        s = ctypes.string_at(p).decode()
        self.lib.ref_free(p)
        return s

    def time_fibonacci(self, n: int) -> float:
        return float(self.lib.ref_time_fibonacci(n))

    def time_mul_ntt(self, a, b=None, square=False) -> float:
        A, pa, na = _arr(a)
        B, pb, nb = _arr(a if b is None else b)
        self.ntt_init(na + nb)
        return float(self.lib.ref_time_mul_ntt(pa, na, pb, nb, int(square)))

    def time_to_string(self, a) -> float:
        A, pa, na = _arr(a)
        return float(self.lib.ref_time_to_string(pa, na))

    def _list_args(self, items):
        a_parts, b_parts, oa, ob, la, lb, sq = [], [], [], [], [], [], []
        pos_a = pos_b = 0
        maxl = 0
        for a, b, s in items:
            A = np.ascontiguousarray(a, dtype=np.uint64)
            B = A if (b is None or s) else np.ascontiguousarray(b, dtype=np.uint64)
            a_parts.append(A), b_parts.append(B)
            oa.append(pos_a), ob.append(pos_b), la.append(A.size), lb.append(B.size), sq.append(int(s))
            pos_a += A.size
            pos_b += B.size
            maxl = max(maxl, A.size + B.size)
        self.ntt_init(maxl)
        return np.concatenate(a_parts), np.concatenate(b_parts), oa, ob, la, lb, sq

    def mul_list(self, items) -> list:
        """big_mul / big_sqr over items [(a_limbs, b_limbs_or_None, square)];
        returns the products' limb arrays."""
        AA, BB, oa, ob, la, lb, sq = self._list_args(items)
        room = [x + y + 1 for x, y in zip(la, lb)]
        oo = np.concatenate([[0], np.cumsum(room)[:-1]]).astype(np.uint64).tolist()
        out = np.zeros(int(sum(room)), dtype=np.uint64)
        nout = (ctypes.c_size_t * len(items))()
        sz = lambda v: (ctypes.c_size_t * len(v))(*[int(x) for x in v])  # noqa: E731
        self.lib.ref_mul_list(AA.ctypes.data_as(_u64p), sz(oa), sz(la), BB.ctypes.data_as(_u64p), sz(ob), sz(lb),
                              (ctypes.c_int * len(sq))(*sq), len(items), out.ctypes.data_as(_u64p), sz(oo), nout)
        return [out[int(o):int(o) + int(n)] for o, n in zip(oo, nout)]

    def time_mul_list(self, items) -> float:
        """items: list of (a_limbs, b_limbs_or_None, square)."""
        a_parts, b_parts, oa, ob, la, lb, sq = [], [], [], [], [], [], []
        pos_a = pos_b = 0
        maxl = 0
        for a, b, s in items:
            A = np.ascontiguousarray(a, dtype=np.uint64)
            B = A if (b is None or s) else np.ascontiguousarray(b, dtype=np.uint64)
            a_parts.append(A), b_parts.append(B)
            oa.append(pos_a), ob.append(pos_b), la.append(A.size), lb.append(B.size), sq.append(int(s))
            pos_a += A.size
            pos_b += B.size
            maxl = max(maxl, A.size + B.size)
        self.ntt_init(maxl)
        AA = np.concatenate(a_parts)
        BB = np.concatenate(b_parts)
        sz = lambda v: (ctypes.c_size_t * len(v))(*v)  # noqa: E731
        return float(self.lib.ref_time_mul_list(
            AA.ctypes.data_as(_u64p), sz(oa), sz(la), BB.ctypes.data_as(_u64p), sz(ob), sz(lb),
            (ctypes.c_int * len(sq))(*sq), len(items)))

    def pool_cpus(self) -> int:
        return int(self.lib.ref_pool_cpus())


def fib_int(n: int) -> int:
    """F(n) by Python-int fast doubling (the method of the reference's
    fibonacci.py:7-18, written independently)."""
    a, b = 0, 1  # F(0), F(1)
    for bit in bin(n)[2:]:
        c = a * (2 * b - a)
        d = a * a + b * b
        a, b = (d, c + d) if bit == "1" else (c, d)
    return a


def fib_mod(n: int, m: int) -> int:
    """F(n) mod m by fast doubling on residues (tests/fingerprint.py:6-13)."""
    a, b = 0, 1 % m
    for bit in bin(n)[2:]:
        c = a * ((2 * b - a) % m) % m
        d = (a * a + b * b) % m
        a, b = (d, (c + d) % m) if bit == "1" else (c, d)
    return a


FINGERPRINT_PRIMES = (18446744073709551557, 18446744073709551533, 9223372036854775783, 4611686018427387847)


def fib_bits(n: int) -> int:
    """Bit length of F(n) from Binet's formula at high precision
    (tests/fingerprint.py:16-19)."""
    if n < 100:  # Binet's rounding term matters only for tiny n
        return fib_int(n).bit_length()
    from mpmath import floor, log, mp, sqrt
    mp.dps = 60
    phi = (1 + sqrt(5)) / 2
    return int(floor(n * log(phi, 2) - log(sqrt(5), 2))) + 1


def fib_digits(n: int) -> int:
    """Decimal digits of F(n) from Binet's formula (exact for n >= 100)."""
    if n < 100:
        return len(str(fib_int(n)))
    from mpmath import floor, log10, mp, sqrt
    mp.dps = 60
    phi = (1 + sqrt(5)) / 2
    return int(floor(n * log10(phi) - log10(sqrt(5)))) + 1


def decimal_mod(s: str, p: int) -> int:
    """The value of the decimal string s modulo p (Horner over 18-digit chunks)."""
    v = 0
    head = len(s) % 18
    if head:
        v = int(s[:head]) % p
    m = 10 ** 18 % p
    for i in range(head, len(s), 18):
        v = (v * m + int(s[i:i + 18])) % p
    return v


def fib_fingerprint(n: int) -> dict:
    """bits, low 64 bits and F(n) mod the four primes, without computing F(n)."""
    return {"bits": fib_bits(n), "low64": fib_mod(n, 2 ** 64), "mods": [fib_mod(n, p) for p in FINGERPRINT_PRIMES]}

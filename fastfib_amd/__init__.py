"""fastfib_amd: exact big-integer products and F(n) on MI355X (gfx950).

Python access to libfastfib_amd.so over its C-ABI (include/fastfib_amd.h).
The library does all arithmetic on the GPU; this module only moves limbs.
It raises at import-time use if the library is not built (there is no CPU
fallback): build it with ``python -m fastfib_amd.build``.

Numbers cross as Python ints or numpy uint64 limb arrays (little-endian).
"""
from __future__ import annotations

import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfastfib_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "fastfib_amd.h")

if hasattr(sys, "set_int_max_str_digits"):
    sys.set_int_max_str_digits(0)

_u64p = ctypes.POINTER(ctypes.c_uint64)


class Big(ctypes.Structure):
    _fields_ = [("d", _u64p), ("n", ctypes.c_size_t)]


class DBig(ctypes.Structure):
    _fields_ = [("d", ctypes.c_void_p), ("n", ctypes.c_size_t), ("cap", ctypes.c_size_t),
                ("bits", ctypes.c_uint64)]


_lib = None


def lib() -> ctypes.CDLL:
    """The loaded library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m fastfib_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    L.ff_init.argtypes = [ctypes.c_int]
    L.ff_init.restype = ctypes.c_int
    L.ff_stream.restype = ctypes.c_void_p
    L.ff_sync.restype = ctypes.c_int
    L.ff_last_error.restype = ctypes.c_char_p
    L.ff_ntt_init.argtypes = [ctypes.c_size_t]
    for name in ("ff_big_mul", "ff_big_mul_ntt"):
        getattr(L, name).argtypes = [ctypes.POINTER(Big), ctypes.POINTER(Big)]
        getattr(L, name).restype = Big
    L.ff_big_sqr.argtypes = [ctypes.POINTER(Big)]
    L.ff_big_sqr.restype = Big
    L.ff_fibonacci.argtypes = [ctypes.c_uint64]
    L.ff_fibonacci.restype = Big
    L.ff_big_free.argtypes = [ctypes.POINTER(Big)]
    L.ff_dbig_alloc.argtypes = [ctypes.POINTER(DBig), ctypes.c_size_t]
    L.ff_dbig_release.argtypes = [ctypes.POINTER(DBig)]
    L.ff_dbig_upload.argtypes = [ctypes.POINTER(DBig), _u64p, ctypes.c_size_t]
    L.ff_dbig_download.argtypes = [ctypes.POINTER(DBig), _u64p, ctypes.c_size_t]
    L.ff_dmul.argtypes = [ctypes.POINTER(DBig), ctypes.POINTER(DBig), ctypes.POINTER(DBig)]
    L.ff_dfibonacci.argtypes = [ctypes.c_uint64, ctypes.POINTER(DBig)]
    L.ff_set_option.argtypes = [ctypes.c_char_p, ctypes.c_longlong]
    L.ff_get_option.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_longlong)]
    L.ff_fib_step_linear.argtypes = [ctypes.POINTER(Big), ctypes.POINTER(Big), ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(Big), ctypes.POINTER(Big)]
    L.ff_last_factors.argtypes = [ctypes.POINTER(Big), ctypes.POINTER(Big), ctypes.POINTER(Big),
                                  ctypes.POINTER(Big)]
    L.ff_prof_enable.argtypes = [ctypes.c_int]
    L.ff_prof_read.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_long),
                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.ff_prof_read_mults.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != 0:
        raise RuntimeError(f"fastfib_amd error {rc}: {lib().ff_last_error().decode()}")


def to_limbs(x: int) -> np.ndarray:
    n = max(1, (x.bit_length() + 63) // 64)
    return np.frombuffer(x.to_bytes(8 * n, "little"), dtype=np.uint64).copy()


def from_limbs(d) -> int:
    return int.from_bytes(np.ascontiguousarray(d, dtype=np.uint64).tobytes(), "little")


def _as_big(a) -> tuple[Big, np.ndarray]:
    arr = to_limbs(a) if isinstance(a, int) else np.ascontiguousarray(a, dtype=np.uint64)
    if arr.size == 0:
        arr = np.zeros(1, dtype=np.uint64)
    return Big(arr.ctypes.data_as(_u64p), arr.size), arr


def _take(b: Big) -> np.ndarray:
    out = np.ctypeslib.as_array(b.d, shape=(b.n,)).copy()
    lib().ff_big_free(ctypes.byref(b))
    return out


def init(device: int = -1) -> None:
    _check(lib().ff_init(device))


def sync() -> None:
    _check(lib().ff_sync())


def set_option(name: str, value: int) -> None:
    _check(lib().ff_set_option(name.encode(), int(value)))


def get_option(name: str) -> int:
    v = ctypes.c_longlong()
    _check(lib().ff_get_option(name.encode(), ctypes.byref(v)))
    return v.value


def big_mul(a, b) -> np.ndarray:
    """ff_big_mul (general product; pass the same object twice to square)."""
    A, ka = _as_big(a)
    if b is a:
        return _take(lib().ff_big_mul(ctypes.byref(A), ctypes.byref(A)))
    B, kb = _as_big(b)
    return _take(lib().ff_big_mul(ctypes.byref(A), ctypes.byref(B)))


def big_sqr(a) -> np.ndarray:
    A, ka = _as_big(a)
    return _take(lib().ff_big_sqr(ctypes.byref(A)))


def big_mul_ntt(a, b=None) -> np.ndarray:
    """ff_big_mul_ntt; b None squares (the reference's a == b)."""
    A, ka = _as_big(a)
    if b is None:
        return _take(lib().ff_big_mul_ntt(ctypes.byref(A), ctypes.byref(A)))
    B, kb = _as_big(b)
    return _take(lib().ff_big_mul_ntt(ctypes.byref(A), ctypes.byref(B)))


_libc = ctypes.CDLL(None)
_libc.free.argtypes = [ctypes.c_void_p]


def _dec_sigs(L) -> None:
    if getattr(L, "_dec_sigs", False):
        return
    L.ff_big_to_string.argtypes = [ctypes.POINTER(Big)]
    L.ff_big_to_string.restype = ctypes.c_void_p
    L.ff_dbig_to_string.argtypes = [ctypes.POINTER(DBig), ctypes.POINTER(ctypes.c_void_p),
                                    ctypes.POINTER(ctypes.c_size_t)]
    L.ff_big_recip.argtypes = [ctypes.POINTER(Big), ctypes.c_size_t]
    L.ff_big_recip.restype = Big
    L._dec_sigs = True


def to_string(a) -> str:
    """ff_big_to_string: decimal digits of a (int or limbs)."""
    L = lib()
    _dec_sigs(L)
    A, ka = _as_big(a)
    p = L.ff_big_to_string(ctypes.byref(A))
    try:
        return ctypes.string_at(p).decode()
    finally:
        _libc.free(p)


def big_recip(P, s: int) -> np.ndarray:
    """ff_big_recip: floor(2^(2s) / P) minus at most 4, P of exactly s bits."""
    L = lib()
    _dec_sigs(L)
    A, ka = _as_big(P)
    return _take(L.ff_big_recip(ctypes.byref(A), s))


def fibonacci(n: int) -> np.ndarray:
    return _take(lib().ff_fibonacci(n))


def fib_step_linear(s1, s2, k_odd: int, bit: int) -> tuple[np.ndarray, np.ndarray]:
    """ff_fib_step_linear: (lo, hi) from s1 = F(k)^2, s2 = F(k-1)^2."""
    A, ka = _as_big(s1)
    B, kb = _as_big(s2)
    lo, hi = Big(), Big()
    lib().ff_fib_step_linear(ctypes.byref(A), ctypes.byref(B), int(k_odd), int(bit), ctypes.byref(lo),
                             ctypes.byref(hi))
    return _take(lo), _take(hi)


def last_factors(u, v, with_y: bool = True):
    """ff_last_factors: x = 2u + v and, if with_y, y = 2u - v."""
    U, ku = _as_big(u)
    V, kv = _as_big(v)
    x, y = Big(), Big()
    lib().ff_last_factors(ctypes.byref(U), ctypes.byref(V), ctypes.byref(x), ctypes.byref(y) if with_y else None)
    return (_take(x), _take(y)) if with_y else (_take(x), None)


def kernel_stats() -> dict:
    """Per-kernel-kind launches, milliseconds and algorithmic bytes booked by
    the live timer (enable with lib().ff_prof_enable(1))."""
    out = {}
    k = 0
    while True:
        name, n = ctypes.c_char_p(), ctypes.c_long()
        ms, b = ctypes.c_double(), ctypes.c_double()
        if lib().ff_prof_read(k, ctypes.byref(name), ctypes.byref(n), ctypes.byref(ms), ctypes.byref(b)) != 0:
            break
        if n.value:
            out[name.value.decode()] = {"launches": n.value, "ms": ms.value, "bytes": b.value}
        k += 1
    return out


class Device:
    """A number held in device memory (ff_dbig)."""

    def __init__(self, value=None):
        self.h = DBig()
        if value is not None:
            arr = to_limbs(value) if isinstance(value, int) else np.ascontiguousarray(value, dtype=np.uint64)
            _check(lib().ff_dbig_upload(ctypes.byref(self.h), arr.ctypes.data_as(_u64p), arr.size))

    def assign(self, value) -> None:
        """New value, written into this number's own memory when it fits."""
        arr = to_limbs(value) if isinstance(value, int) else np.ascontiguousarray(value, dtype=np.uint64)
        _check(lib().ff_dbig_upload(ctypes.byref(self.h), arr.ctypes.data_as(_u64p), arr.size))

    @property
    def n(self) -> int:
        return int(self.h.n)

    @property
    def bits(self) -> int:
        return int(self.h.bits)

    def limbs(self) -> np.ndarray:
        out = np.zeros(max(1, self.h.n), dtype=np.uint64)
        _check(lib().ff_dbig_download(ctypes.byref(self.h), out.ctypes.data_as(_u64p), out.size))
        return out

    def value(self) -> int:
        return from_limbs(self.limbs())

    def to_string(self) -> str:
        """Decimal digits (ff_dbig_to_string)."""
        L = lib()
        _dec_sigs(L)
        p, n = ctypes.c_void_p(), ctypes.c_size_t()
        _check(L.ff_dbig_to_string(ctypes.byref(self.h), ctypes.byref(p), ctypes.byref(n)))
        try:
            return ctypes.string_at(p, n.value).decode()
        finally:
            _libc.free(p)

    def release(self) -> None:
        lib().ff_dbig_release(ctypes.byref(self.h))

    def __del__(self):
        try:
            if self.h.d and _lib is not None:
                self.release()
        except Exception:
            pass


def dmul(a: Device, b: Device | None = None) -> Device:
    out = Device()
    _check(lib().ff_dmul(ctypes.byref(out.h), ctypes.byref(a.h), None if b is None else ctypes.byref(b.h)))
    return out


def dmul_into(out: Device, a: Device, b: Device | None = None) -> None:
    """out = a * b in place (out may be a itself: ff_dmul(&x, &x, NULL))."""
    _check(lib().ff_dmul(ctypes.byref(out.h), ctypes.byref(a.h), None if b is None else ctypes.byref(b.h)))


def mem_stats() -> dict:
    """The library's allocation counters (ff_mem_stats)."""
    L = lib()
    L.ff_mem_stats.argtypes = [_u64p, _u64p, ctypes.POINTER(ctypes.c_size_t)]
    dev, host, cached = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_size_t()
    _check(L.ff_mem_stats(ctypes.byref(dev), ctypes.byref(host), ctypes.byref(cached)))
    return {"device_allocs": dev.value, "host_allocs": host.value, "cached_bytes": cached.value}


def dfibonacci(n: int) -> Device:
    out = Device()
    _check(lib().ff_dfibonacci(n, ctypes.byref(out.h)))
    return out


class Prod(ctypes.Structure):
    """ff_prod: one product of a batch (device pointers)."""
    _fields_ = [("a", ctypes.c_void_p), ("na", ctypes.c_size_t), ("b", ctypes.c_void_p), ("nb", ctypes.c_size_t),
                ("out", ctypes.c_void_p), ("nout", ctypes.c_size_t), ("bits_a", ctypes.c_uint64),
                ("bits_b", ctypes.c_uint64)]


BATCH_PLAIN, BATCH_GROUPED, BATCH_GRAPH, BATCH_TUNED, BATCH_LEARNED, BATCH_QUEUE = 0, 1, 2, 4, 8, 16


def _batch_sigs(L) -> None:
    if getattr(L, "_batch_sigs", False):
        return
    L.ff_batch_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(Prod), ctypes.c_size_t,
                                  ctypes.c_int]
    L.ff_batch_run.argtypes = [ctypes.c_void_p]
    L.ff_batch_finish.argtypes = [ctypes.c_void_p, ctypes.POINTER(Prod)]
    L.ff_batch_destroy.argtypes = [ctypes.c_void_p]
    L._batch_sigs = True


class Batch:
    """A fixed list of device products run together (ff_batch_*).

    pairs: (a, b) of Device numbers; b None or b is a squares. Outputs are
    allocated here with room for na + nb + 1 limbs."""

    def __init__(self, pairs, flags: int = BATCH_GROUPED):
        L = lib()
        _batch_sigs(L)
        self.pairs = list(pairs)
        self.outs = []
        self.items = (Prod * max(1, len(self.pairs)))()
        for i, (a, b) in enumerate(self.pairs):
            sq = b is None or b is a
            o = Device()
            _check(L.ff_dbig_alloc(ctypes.byref(o.h), a.n + (a.n if sq else b.n) + 1))
            self.outs.append(o)
            it = self.items[i]
            it.a, it.na, it.bits_a = a.h.d, a.n, a.bits
            if not sq:
                it.b, it.nb, it.bits_b = b.h.d, b.n, b.bits
            it.out = o.h.d
        self.h = ctypes.c_void_p()
        _check(L.ff_batch_create(ctypes.byref(self.h), self.items, len(self.pairs), flags))

    def run(self) -> None:
        _check(lib().ff_batch_run(self.h))

    def finish(self) -> None:
        _check(lib().ff_batch_finish(self.h, self.items))
        for i, o in enumerate(self.outs):
            o.h.n = self.items[i].nout

    def results(self) -> list:
        return [o.value() for o in self.outs]

    def close(self) -> None:
        if self.h:
            lib().ff_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            if self.h and _lib is not None:
                self.close()
        except Exception:
            pass

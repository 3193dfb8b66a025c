"""Seeded synthetic workloads for batched products.

uneven_mix() draws the uneven batch: count products, each operand's bit
length floor(2^(lo + (hi - lo) u)) with u uniform, a fraction of squares,
the two operands of a general product sized independently. operand_limbs()
fills an operand of a given bit length with uniform 64-bit limbs and its top
bit set, so its bit length is exactly the drawn one.
"""
from __future__ import annotations

import numpy as np


def uneven_mix(seed: int, count: int = 4096, lo: float = 12.0, hi: float = 24.0,
               square_frac: float = 0.3) -> list[tuple[int, int | None]]:
    """[(bits_a, bits_b or None for a square)] * count."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        ba = int(np.floor(2.0 ** (lo + (hi - lo) * rng.random())))
        if rng.random() < square_frac:
            out.append((ba, None))
        else:
            out.append((ba, int(np.floor(2.0 ** (lo + (hi - lo) * rng.random())))))
    return out


def operand_limbs(rng: np.random.Generator, bits: int) -> np.ndarray:
    """ceil(bits / 64) uniform limbs, masked to bits, top bit set."""
    n = (bits + 63) // 64
    d = rng.integers(0, 1 << 64, size=n, dtype=np.uint64, endpoint=False)
    top = bits - 64 * (n - 1)
    if top < 64:
        d[-1] &= np.uint64((1 << top) - 1)
    d[-1] |= np.uint64(1 << (top - 1))
    return d


def operand_bits(mix) -> int:
    """Operand bits of a mix (a square counts its operand twice)."""
    return sum(a + (a if b is None else b) for a, b in mix)

"""Autotuner samples, training and evaluation of the learned dispatcher.

    python -m fastfib_amd.dispatch collect --mixes 1000 --per-mix 20 --seed 1 --out DIR/train.npz
    python -m fastfib_amd.dispatch train --train DIR/train.npz --heldout DIR/heldout.npz \\
                                         --out fastfib_amd/dispatch.bin

A sample is one product drawn from a seeded random size mix (batch size,
operand bit range, square fraction), with the autotuner's measured
per-product time of every schedule candidate (ff_tune_measure: a grouped
batch of copies of that product, copies = min(batch size, 1024, 2^29 bits /
product bits): a product's marginal cost in a batch that keeps the GPU busy). Training fits the library's MLP (ff_mlp_train, on the GPU)
to the candidates' relative log-times; the autotuner table takes, per size
class, the candidate with the least mean log-time. Regret on held-out mixes
compares, per mix, the summed time of the chosen candidates with the summed
time of the exhaustive autotuner's (the measured best per product).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import struct
import sys
import time

import numpy as np

from . import _check, lib

NI = 8
MAX_COPIES = 1024
FILL_BITS = 1 << 29


def _sigs(L) -> None:
    if getattr(L, "_tune_sigs", False):
        return
    L.ff_tune_features.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_size_t,
                                   ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_uint32)]
    L.ff_tune_measure.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_size_t, ctypes.c_int,
                                  ctypes.c_uint64, ctypes.POINTER(ctypes.c_float)]
    L.ff_tune_class.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_size_t]
    L.ff_tune_rule_candidate.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int]
    L.ff_tune_set_table.argtypes = [ctypes.POINTER(ctypes.c_int32), ctypes.c_size_t]
    L.ff_dispatch_load.argtypes = [ctypes.c_char_p]
    L.ff_mlp_param_count.argtypes = [ctypes.c_int]
    L.ff_mlp_param_count.restype = ctypes.c_size_t
    fp = ctypes.POINTER(ctypes.c_float)
    L.ff_mlp_set.argtypes = [ctypes.c_int, fp, ctypes.c_size_t]
    L.ff_mlp_get.argtypes = [ctypes.POINTER(ctypes.c_int), fp, ctypes.c_size_t]
    L.ff_mlp_grad.argtypes = [fp, fp, ctypes.c_size_t, fp, fp]
    L.ff_mlp_train.argtypes = [fp, fp, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                               ctypes.c_uint64, fp]
    L.ff_mlp_predict.argtypes = [fp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint32), fp,
                                 ctypes.POINTER(ctypes.c_int)]
    L._tune_sigs = True


def _L():
    L = lib()
    _sigs(L)
    return L


def _fp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def candidates() -> int:
    return int(_L().ff_tune_candidates())


def default_candidate() -> int:
    return int(_L().ff_tune_default_candidate())


def candidate(k: int) -> dict:
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(_L().ff_tune_candidate(k, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
    if b.value < 0:
        return {"split_piece_lg": -b.value}
    return {"len_up": a.value, "inner_log": b.value, "tile_words": c.value}


def rule_candidates(d: dict) -> np.ndarray:
    """Per sample, the candidate the reference rule amounts to (the default
    schedule, or the split cost model's piece length)."""
    L = _L()
    return np.array([L.ff_tune_rule_candidate(int(a), int(b), int(q))
                     for a, b, q in zip(d["bits_a"], d["bits_b"], d["square"])], dtype=np.int64)


# ------------------------------------------------------------------ sampling

def draw_samples(seed: int, mixes: int, per_mix: int, lo_lg: float = 6.0, hi_lg: float = 29.0,
                 mix0: int = 0) -> list[tuple[int, int, int, int, int]]:
    """[(bits_a, bits_b, square, count, mix)]: per mix a batch size
    floor(2^(12 u)), an operand range 2^lo .. 2^hi inside 2^lo_lg .. 2^hi_lg
    (by default 2^6 .. 2^29; 2^29 .. 2^31.5 reaches the split multiply) and
    a square fraction, then per_mix products drawn from it; mixes are
    numbered from mix0."""
    rng = np.random.default_rng(seed)
    out = []
    for m in range(mix0, mix0 + mixes):
        count = int(np.floor(2.0 ** (12.0 * rng.random())))
        lo, hi = sorted(lo_lg + (hi_lg - lo_lg) * rng.random(2))
        sqf = rng.random()
        for _ in range(per_mix):
            ba = int(np.floor(2.0 ** (lo + (hi - lo) * rng.random())))
            if rng.random() < sqf:
                out.append((ba, ba, 1, count, m))
            else:
                out.append((ba, int(np.floor(2.0 ** (lo + (hi - lo) * rng.random()))), 0, count, m))
    return out


def copies_for(bits_a: int, bits_b: int, count: int) -> int:
    return int(max(1, min(count, MAX_COPIES, FILL_BITS // (bits_a + bits_b))))


def features(bits_a: int, bits_b: int, square: int, count: int) -> tuple[np.ndarray, int]:
    x = np.zeros(NI, dtype=np.float32)
    m = ctypes.c_uint32()
    _check(_L().ff_tune_features(bits_a, bits_b, square, count, _fp(x), ctypes.byref(m)))
    return x, int(m.value)


def measure(bits_a: int, bits_b: int, square: int, count: int, reps: int = 2, seed: int = 1) -> np.ndarray:
    K = candidates()
    us = np.zeros(K, dtype=np.float32)
    _check(_L().ff_tune_measure(bits_a, bits_b, square, copies_for(bits_a, bits_b, count), reps, seed, _fp(us)))
    return us


def collect(samples, reps: int = 2, progress: int = 500) -> dict:
    L = _L()
    K = candidates()
    n = len(samples)
    X = np.zeros((n, NI), dtype=np.float32)
    T = np.zeros((n, K), dtype=np.float32)
    mask = np.zeros(n, dtype=np.uint32)
    cls = np.zeros(n, dtype=np.int32)
    t0 = time.time()
    for i, (ba, bb, sq, cnt, _) in enumerate(samples):
        X[i], mask[i] = features(ba, bb, sq, cnt)
        cls[i] = L.ff_tune_class(ba, bb, sq, cnt)
        T[i] = measure(ba, bb, sq, cnt, reps, seed=1000003 * i + 7)
        if progress and (i + 1) % progress == 0:
            print(f"collect: {i + 1}/{n} samples, {time.time() - t0:.0f} s", flush=True)
    arr = np.array(samples, dtype=np.int64)
    return {"X": X, "T_us": T, "mask": mask, "cls": cls, "bits_a": arr[:, 0], "bits_b": arr[:, 1],
            "square": arr[:, 2], "count": arr[:, 3], "mix": arr[:, 4]}


# ------------------------------------------------------------------ training

def targets(T_us: np.ndarray) -> np.ndarray:
    """Relative log-times: ln t - mean over the row's applicable candidates;
    a candidate that does not apply gets the row's worst + ln 2."""
    lt = np.log(T_us.astype(np.float64))
    ok = np.isfinite(lt)
    mean = np.nanmean(np.where(ok, lt, np.nan), axis=1, keepdims=True)
    worst = np.nanmax(np.where(ok, lt, np.nan), axis=1, keepdims=True)
    lt = np.where(ok, lt, worst + np.log(2.0))
    return (lt - mean).astype(np.float32)


def init_params(K: int, X: np.ndarray, seed: int = 0) -> np.ndarray:
    """Normalization from the data, He-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    mu = X.mean(0)
    # a feature that barely varies in the samples (free HBM) must not blow
    # up when it does vary in use: at least 0.1 (features are logs, flags
    # and fractions)
    sd = np.maximum(X.std(0), 0.1)
    parts = [mu, sd]
    for fan_in, shape in ((NI, (64, NI)), (None, (64,)), (64, (64, 64)), (None, (64,)), (64, (K, 64)),
                          (None, (K,))):
        if fan_in is None:
            parts.append(np.zeros(shape))
        else:
            lim = np.sqrt(6.0 / fan_in)
            parts.append(rng.uniform(-lim, lim, size=shape))
    p = np.concatenate([np.ravel(v) for v in parts]).astype(np.float32)
    assert p.size == _L().ff_mlp_param_count(K)
    return p


def set_params(K: int, p: np.ndarray) -> None:
    p = np.ascontiguousarray(p, dtype=np.float32)
    _check(_L().ff_mlp_set(K, _fp(p), p.size))


def get_params() -> tuple[int, np.ndarray]:
    L = _L()
    K = ctypes.c_int()
    L.ff_mlp_get(ctypes.byref(K), None, 0)
    p = np.zeros(L.ff_mlp_param_count(K.value), dtype=np.float32)
    _check(L.ff_mlp_get(ctypes.byref(K), _fp(p), p.size))
    return K.value, p


def train(X: np.ndarray, Y: np.ndarray, epochs: int = 300, batch: int = 256, lr: float = 2e-3,
          seed: int = 0) -> float:
    """Trains the library's MLP (current parameters) on the GPU; returns the
    last epoch's mean loss."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.ascontiguousarray(Y, dtype=np.float32)
    loss = ctypes.c_float()
    _check(_L().ff_mlp_train(_fp(X), _fp(Y), X.shape[0], epochs, batch, lr, seed, ctypes.byref(loss)))
    return float(loss.value)


def predict(X: np.ndarray, mask: np.ndarray | None = None) -> tuple[np.ndarray, np.ndarray]:
    K = candidates()
    X = np.ascontiguousarray(X, dtype=np.float32)
    n = X.shape[0]
    y = np.zeros((n, K), dtype=np.float32)
    c = np.zeros(n, dtype=np.int32)
    mp = None
    if mask is not None:
        mask = np.ascontiguousarray(mask, dtype=np.uint32)
        mp = mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    _check(_L().ff_mlp_predict(_fp(X), n, mp, _fp(y), c.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
    return y, c


def build_table(d: dict, size: int, K: int, default: int) -> np.ndarray:
    """Per size class, the candidate with the least mean log-time."""
    lt = np.log(np.where(np.isfinite(d["T_us"]), d["T_us"], np.inf).astype(np.float64))
    table = np.full(size, default, dtype=np.int32)
    for c in np.unique(d["cls"]):
        rows = lt[d["cls"] == c]
        m = np.mean(rows, axis=0)
        if np.all(~np.isfinite(m)):
            continue
        table[c] = int(np.argmin(np.where(np.isfinite(m), m, np.inf)))
    return table


MARGIN = 0.02  # the library's FF_MLP_MARGIN


def dispatch_choice(y: np.ndarray, mask: np.ndarray, default: int, margin: float = MARGIN,
                    rule: np.ndarray | None = None) -> np.ndarray:
    """The dispatcher's pick from the network's outputs: the argmin over the
    applicable candidates, unless the default is predicted within margin of
    it (the rule of the library's inference kernel); products past the
    longest transform take the reference rule's candidate (rule), as the
    library does."""
    K = y.shape[1]
    m = np.asarray(mask, dtype=np.int64)
    allowed = ((m[:, None] >> np.arange(K)) & 1) == 1
    yy = np.where(allowed, y, np.inf)
    best = np.argmin(yy, axis=1)
    ok_def = allowed[:, default]
    gain = np.where(ok_def, yy[np.arange(len(yy)), default] - yy[np.arange(len(yy)), best], np.inf)
    pick = np.where(ok_def & (gain <= margin), default, best)
    if rule is not None:
        pick = np.where((m >> 16) != 0, rule, pick)
    return pick


def table_choice(table: np.ndarray, d: dict) -> np.ndarray:
    """The table's candidate per sample, the reference rule where it does not apply."""
    tc = table[d["cls"]]
    ok = ((d["mask"].astype(np.int64) >> tc) & 1) == 1
    return np.where(ok, tc, rule_candidates(d))


def regret(T_us: np.ndarray, choice: np.ndarray, mix: np.ndarray) -> dict:
    """Per-mix regret of a choice against the measured best per product."""
    T = np.where(np.isfinite(T_us), T_us, np.inf).astype(np.float64)
    best = T.min(1)
    got = T[np.arange(T.shape[0]), choice]
    per = []
    for m in np.unique(mix):
        sel = mix == m
        per.append(got[sel].sum() / best[sel].sum() - 1.0)
    per = np.array(per)
    return {"mean": float(per.mean()), "median": float(np.median(per)), "p95": float(np.quantile(per, 0.95)),
            "aggregate": float(got.sum() / best.sum() - 1.0), "mixes": int(per.size)}


def ho_split(d: dict) -> np.ndarray:
    """Samples whose products take the split multiply (piece candidates only)."""
    return (d["mask"].astype(np.int64) >> 16) != 0


def build_stamp() -> dict:
    """The library the samples were measured with: its SHA-256 and the
    sources' newest commit."""
    import hashlib
    import subprocess
    from . import LIB_PATH
    with open(LIB_PATH, "rb") as f:
        h = hashlib.sha256(f.read()).hexdigest()[:16]
    c = os.environ.get("FF_BUILD_COMMIT", "")
    if not c:
        try:
            c = subprocess.run(["git", "log", "-1", "--format=%h", "--", "fastfib_amd/csrc"], capture_output=True,
                               text=True, cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__)))).stdout
            c = c.strip()
        except OSError:
            c = ""
    return {"library_sha256_16": h, "csrc_commit": c or None}


def write_bin(path: str, K: int, table: np.ndarray, params: np.ndarray) -> None:
    with open(path, "wb") as f:
        f.write(b"FFDISP1\0")
        f.write(struct.pack("<ii", K, table.size))
        f.write(np.ascontiguousarray(table, dtype="<i4").tobytes())
        f.write(struct.pack("<i", params.size))
        f.write(np.ascontiguousarray(params, dtype="<f4").tobytes())


def read_bin(path: str) -> tuple[int, np.ndarray, np.ndarray]:
    with open(path, "rb") as f:
        assert f.read(8) == b"FFDISP1\0"
        K, nt = struct.unpack("<ii", f.read(8))
        table = np.frombuffer(f.read(4 * nt), dtype="<i4").copy()
        (np_,) = struct.unpack("<i", f.read(4))
        params = np.frombuffer(f.read(4 * np_), dtype="<f4").copy()
    return K, table, params


def load(path: str) -> None:
    _check(_L().ff_dispatch_load(path.encode()))


def _save(path: str, d: dict) -> None:
    np.savez_compressed(path, **d)


def _load(path: str) -> dict:
    with np.load(path) as z:
        return {k: z[k] for k in z.files}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--mixes", type=int, default=1000)
    c.add_argument("--per-mix", type=int, default=20)
    c.add_argument("--seed", type=int, default=1)
    c.add_argument("--reps", type=int, default=2)
    c.add_argument("--lo-lg", type=float, default=6.0)
    c.add_argument("--hi-lg", type=float, default=29.0)
    c.add_argument("--mix0", type=int, default=0)
    c.add_argument("--append", default=None, help="samples file to extend")
    c.add_argument("--out", required=True)
    t = sub.add_parser("train")
    t.add_argument("--train", required=True)
    t.add_argument("--heldout", required=True)
    t.add_argument("--epochs", type=int, default=300)
    t.add_argument("--batch", type=int, default=256)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--out", required=True)
    t.add_argument("--report", default=None)
    a = ap.parse_args(argv)
    from . import init
    init()
    if a.cmd == "collect":
        s = draw_samples(a.seed, a.mixes, a.per_mix, a.lo_lg, a.hi_lg, a.mix0)
        print(f"collect: {len(s)} samples from {a.mixes} mixes (seed {a.seed})", flush=True)
        d = collect(s, a.reps)
        if a.append:
            old = _load(a.append)
            d = {k: np.concatenate([old[k], d[k]]) for k in d}
        _save(a.out, d)
        print(f"wrote {a.out}", flush=True)
        return 0
    tr, ho = _load(a.train), _load(a.heldout)
    K = candidates()
    Ytr = targets(tr["T_us"])
    # products past the longest transform are few among the samples (they
    # are drawn from a range of their own) but cost the most: repeated so
    # that they make up about a twentieth of the training rows
    sp = ho_split(tr)
    rep = max(1, int(round(0.05 * np.sum(~sp) / max(1, np.sum(sp)))))
    Xt = np.concatenate([tr["X"]] + [tr["X"][sp]] * (rep - 1))
    Yt = np.concatenate([Ytr] + [Ytr[sp]] * (rep - 1))
    t0 = time.time()
    set_params(K, init_params(K, Xt, a.seed))
    loss = train(Xt, Yt, a.epochs, a.batch, a.lr, a.seed)
    dt = time.time() - t0
    K, params = get_params()
    size = int(_L().ff_tune_table_size())
    table = build_table(tr, size, K, default_candidate())
    y, _ = predict(ho["X"], ho["mask"])
    choice = dispatch_choice(y, ho["mask"], default_candidate(), rule=rule_candidates(ho))
    rep = {"train_samples": int(tr["X"].shape[0]), "heldout_samples": int(ho["X"].shape[0]),
           "train_mixes": int(np.unique(tr["mix"]).size), "epochs": a.epochs, "batch": a.batch, "lr": a.lr,
           "train_seconds": round(dt, 2), "final_loss": loss,
           "regret_learned": regret(ho["T_us"], choice, ho["mix"]),
           "regret_table": regret(ho["T_us"], table_choice(table, ho), ho["mix"]),
           "regret_reference_rule": regret(ho["T_us"], rule_candidates(ho), ho["mix"]),
           "candidates": K, "split_samples": {"train": int(np.sum(sp)), "heldout": int(np.sum(ho_split(ho))),
                                              "train_repeats": rep},
           "build": build_stamp()}
    write_bin(a.out, K, table, params)
    print(json.dumps(rep), flush=True)
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""CPU checks of the dispatcher's oracle and of the committed dispatcher.

oracle/mlp.py is pinned by mathematics: its backpropagated gradient against
central finite differences of its own loss, and its forward pass against a
per-sample evaluation written out neuron by neuron.
"""
import numpy as np
import pytest

from oracle import mlp as om

K = 12


def _problem(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 8)) * 3 + 10
    T = rng.normal(size=(n, K))
    p = om.init_params(K, X.mean(0), X.std(0), seed).astype(np.float64)
    L = om.layout(K)
    for name in ("b1", "b2", "b3"):
        o, s = L[name]
        p[o:o + s] = rng.normal(scale=0.1, size=s)
    return X, T, p


def test_gradient_matches_finite_differences():
    X, T, p = _problem(16, 0)
    g = om.grad(p, X, T, K)
    L = om.layout(K)
    rng = np.random.default_rng(1)
    idx = []
    for name in ("W1", "b1", "W2", "b2", "W3", "b3"):
        o, s = L[name]
        idx += list(o + rng.choice(s, size=min(s, 12), replace=False))
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp = om.loss(q, X, T, K)
        q[i] -= 2 * h
        lm = om.loss(q, X, T, K)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])
    # normalization constants are not trained
    for name in ("mu", "sd"):
        o, s = L[name]
        assert np.all(g[o:o + s] == 0)


def test_forward_neuron_by_neuron():
    X, T, p = _problem(5, 2)
    g = om.unpack(p, K)
    y = om.forward(p, X, K)[3]
    for b in range(X.shape[0]):
        xn = [(X[b, i] - g["mu"][i]) / g["sd"][i] for i in range(8)]
        h1 = [max(0.0, g["b1"][o] + sum(xn[i] * g["W1"][o, i] for i in range(8))) for o in range(64)]
        h2 = [max(0.0, g["b2"][o] + sum(h1[i] * g["W2"][o, i] for i in range(64))) for o in range(64)]
        for k in range(K):
            v = g["b3"][k] + sum(h2[i] * g["W3"][k, i] for i in range(64))
            assert abs(v - y[b, k]) < 1e-9 * max(1.0, abs(v))


def test_choose_respects_mask():
    X, T, p = _problem(50, 3)
    mask = np.full(50, 0b101, dtype=np.uint32)
    c = om.choose(p, X, K, mask)
    assert set(np.unique(c)) <= {0, 2}


# ------------------------------------------------------------------ the committed dispatcher

import os

from fastfib_amd import dispatch as dp

DATA = os.path.join(os.path.dirname(dp.__file__), "dispatch_data")
BIN = os.path.join(os.path.dirname(dp.__file__), "dispatch.bin")


def _load(name):
    with np.load(os.path.join(DATA, name)) as z:
        return {k: z[k] for k in z.files}


def test_training_set_size():
    tr = _load("train.npz")
    assert tr["X"].shape[0] >= 20000
    assert np.unique(tr["mix"]).size >= 1000
    assert tr["T_us"].shape == (tr["X"].shape[0], dp.candidates())
    # every applicable candidate was timed
    applicable = ((tr["mask"][:, None].astype(np.int64) >> np.arange(dp.candidates())) & 1) == 1
    assert np.all(np.isfinite(tr["T_us"][applicable]) & (tr["T_us"][applicable] > 0))
    assert np.all(np.isnan(tr["T_us"][~applicable]))


def test_committed_dispatcher_regret():
    """The committed network, evaluated by the oracle's float64 forward pass,
    stays within 5% of the exhaustive autotuner on the 1000 held-out mixes,
    and beats the reference rule; so does the autotuner table."""
    K, table, params = dp.read_bin(BIN)
    assert K == dp.candidates()
    assert params.size == om.layout(K)["total"]
    ho = _load("heldout.npz")
    assert np.unique(ho["mix"]).size >= 1000
    y = om.forward(params, ho["X"], K)[3]
    rule = dp.rule_candidates(ho)
    choice = dp.dispatch_choice(y, ho["mask"], dp.default_candidate(), rule=rule)
    learned = dp.regret(ho["T_us"], choice, ho["mix"])
    raw = dp.regret(ho["T_us"], om.choose(params, ho["X"], K, ho["mask"]), ho["mix"])
    assert raw["mean"] <= 0.05, raw
    ref_rule = dp.regret(ho["T_us"], rule, ho["mix"])
    tab = dp.regret(ho["T_us"], dp.table_choice(table, ho), ho["mix"])
    assert learned["mean"] <= 0.05, learned
    assert tab["mean"] <= 0.05, tab
    assert learned["mean"] < ref_rule["mean"], (learned, ref_rule)
    assert learned["aggregate"] < ref_rule["aggregate"], (learned, ref_rule)

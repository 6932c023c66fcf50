"""Pins the oracle width MLP: finite-difference gradients, the loss formula
and the vectorized forward pass."""
import math

import numpy as np

from oracle import mlp as M


def _data(n=40, seed=0):
    r = np.random.default_rng(seed)
    X = r.standard_normal((n, M.N_IN)) * 3.0 + 1.0
    y = (r.random(n) < 0.5).astype(np.float64)
    norm = np.concatenate([np.full(M.N_IN, 1.0), np.full(M.N_IN, 1.0 / 3.0)])
    return X, y, norm


def test_gradient_matches_central_differences():
    X, y, norm = _data()
    p = M.init_params(3).astype(np.float64)
    p[M.N_HID * M.N_IN:M.N_HID * M.N_IN + M.N_HID] = 0.05  # biases off zero
    g = M.gradient(p, norm, X, y)
    r = np.random.default_rng(1)
    idx = r.choice(M.N_PARAMS, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (M.loss(p + e, norm, X, y) - M.loss(p - e, norm, X, y)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_loss_and_forward_against_per_sample_loop():
    X, y, norm = _data(12, seed=2)
    p = M.init_params(5).astype(np.float64)
    W1, b1, W2, b2, W3, b3 = M.unpack(p)
    z, _ = M.forward(p, norm, X)
    tot = 0.0
    for i in range(X.shape[0]):
        x = (X[i] - norm[:12]) * norm[12:]
        h1 = [max(0.0, sum(W1[j, k] * x[k] for k in range(12)) + b1[j]) for j in range(64)]
        h2 = [max(0.0, sum(W2[j, k] * h1[k] for k in range(64)) + b2[j]) for j in range(64)]
        zi = sum(W3[j] * h2[j] for j in range(64)) + b3
        assert abs(zi - z[i]) <= 1e-12 * max(1.0, abs(zi))
        s = 1.0 / (1.0 + math.exp(-zi))
        tot += -(y[i] * math.log(s) + (1 - y[i]) * math.log(1 - s))
    assert abs(tot / X.shape[0] - M.loss(p, norm, X, y)) <= 1e-12
    assert M.N_PARAMS == 5057 and M.init_params(0).shape == (5057,)


def test_shipped_width_model_reproduces_its_training_accuracy():
    """The shipped width controller (trained on the GPU by the engine's own
    kernels) evaluated by the float64 oracle forward pass on its committed
    training decisions: the accuracy recorded at training time, the logit
    sign per decision agreeing with the labels as often as the model claimed,
    and a training box that contains every training input."""
    import json
    from pathlib import Path
    models = Path(__file__).resolve().parents[1] / "amfls_solver_amd" / "models"
    with np.load(models / "width_mlp.npz", allow_pickle=False) as z:
        p, norm = z["params"].astype(np.float64), z["norm"].astype(np.float64)
        meta = json.loads(str(z["meta"]))
        lo, hi = z["domain_lo"], z["domain_hi"]
    with np.load(models / "width_training_decisions.npz", allow_pickle=False) as z:
        X, y = z["X"].astype(np.float64), z["y"].astype(np.float64)
    assert X.shape == (meta["decisions"], M.N_IN)
    assert abs(float(y.mean()) - meta["widen_fraction"]) < 1e-6
    logits, _ = M.forward(p, norm, X)
    acc = float(np.mean((logits > 0) == (y > 0.5)))
    # the GPU forward pass is FP32: a handful of logits within rounding of 0 may flip
    assert abs(acc - meta["train_accuracy"]) <= 2.0 / len(y), (acc, meta["train_accuracy"])
    assert np.all(X >= lo) and np.all(X <= hi)

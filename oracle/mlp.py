"""Width-controller MLP (oracle).

TEST INFRASTRUCTURE ONLY: the CPU restatement, in numpy float64, of the
12 -> 64 -> 64 -> 1 ReLU network that the HIP engine trains and evaluates on
the GPU (amfls_solver_amd/csrc/device/mlp.hip) as its learned replacement of
the reference's widening predicate (src/algorithms/amfls/amfls.cpp:322-358).
Parameter layout: W1 [64 x 12], b1 [64], W2 [64 x 64], b2 [64], W3 [64], b3,
flattened in that order; inputs are normalized as (x - mean) * inv_std;
loss = mean binary cross-entropy on the logits.

Pinned by tests/test_oracle_mlp.py: the analytic gradient against central
finite differences of the loss, the loss against a direct evaluation of
-log(sigmoid) terms, and the forward pass against an explicit per-sample
loop.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID = 12, 64
N_PARAMS = N_HID * N_IN + N_HID + N_HID * N_HID + N_HID + N_HID + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + N_HID * N_IN].reshape(N_HID, N_IN); o += N_HID * N_IN
    b1 = p[o:o + N_HID]; o += N_HID
    W2 = p[o:o + N_HID * N_HID].reshape(N_HID, N_HID); o += N_HID * N_HID
    b2 = p[o:o + N_HID]; o += N_HID
    W3 = p[o:o + N_HID]; o += N_HID
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def normalize(X, norm):
    norm = np.asarray(norm, dtype=np.float64)
    return (np.asarray(X, dtype=np.float64) - norm[:N_IN]) * norm[N_IN:]


def forward(p, norm, X):
    """Logits and the hidden activations (for the backward pass)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = normalize(X, norm)
    a1 = x @ W1.T + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0.0)
    z = h2 @ W3 + b3
    return z, (x, h1, h2)


def loss(p, norm, X, y):
    z, _ = forward(p, norm, X)
    y = np.asarray(y, dtype=np.float64)
    return float(np.mean(np.maximum(z, 0.0) - y * z + np.log1p(np.exp(-np.abs(z)))))


def gradient(p, norm, X, y):
    """d(mean BCE)/d(params), flattened in the parameter layout."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    z, (x, h1, h2) = forward(p, norm, X)
    n = z.shape[0]
    dz = (1.0 / (1.0 + np.exp(-z)) - np.asarray(y, dtype=np.float64)) / n
    gW3 = h2.T @ dz
    gb3 = dz.sum()
    d2 = np.outer(dz, W3) * (h2 > 0)
    gW2 = d2.T @ h1
    gb2 = d2.sum(axis=0)
    d1 = (d2 @ W2) * (h1 > 0)
    gW1 = d1.T @ x
    gb1 = d1.sum(axis=0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])


def init_params(seed: int = 0):
    """He-normal weights, zero biases (float32 values)."""
    r = np.random.default_rng(seed)
    W1 = r.standard_normal((N_HID, N_IN)) * np.sqrt(2.0 / N_IN)
    W2 = r.standard_normal((N_HID, N_HID)) * np.sqrt(2.0 / N_HID)
    W3 = r.standard_normal(N_HID) * np.sqrt(1.0 / N_HID)
    p = np.concatenate([W1.ravel(), np.zeros(N_HID), W2.ravel(), np.zeros(N_HID), W3, [0.0]])
    return p.astype(np.float32)

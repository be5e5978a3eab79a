"""ORACLE -- test infrastructure only.  CPU restatement (numpy, float64) of the planner's MLP cost
model: 24 -> 64 -> 64 -> 1 with ReLU, mean-squared-error loss, and the gradient of that loss
with respect to every weight (weights packed as W1[64][24], b1, W2[64][64], b2, w3, b3).
The MLP has no counterpart in the reference; this restatement is pinned by central finite
differences in tests/test_mlp.py."""

import numpy as np

N_IN, N_HID = 24, 64
SLICES = {}
_o = 0
for _name, _shape in (("W1", (N_HID, N_IN)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)), ("b2", (N_HID,)), ("w3", (N_HID,)), ("b3", (1,))):
    SLICES[_name] = (_o, _shape)
    _o += int(np.prod(_shape))
N_WEIGHTS = _o


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    return {k: w[o : o + int(np.prod(s))].reshape(s) for k, (o, s) in SLICES.items()}


def forward(w, X):
    p = unpack(w)
    X = np.asarray(X, dtype=np.float64)
    a1 = X @ p["W1"].T + p["b1"]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ p["W2"].T + p["b2"]
    h2 = np.maximum(a2, 0.0)
    y = h2 @ p["w3"] + p["b3"][0]
    return y, (X, a1, h1, a2, h2)


def loss(w, X, T):
    y, _ = forward(w, X)
    return float(np.mean((y - np.asarray(T, dtype=np.float64)) ** 2))


def loss_grad(w, X, T):
    p = unpack(w)
    y, (X, a1, h1, a2, h2) = forward(w, X)
    T = np.asarray(T, dtype=np.float64)
    n = len(T)
    dy = 2.0 * (y - T) / n
    g = {}
    g["w3"] = h2.T @ dy
    g["b3"] = np.array([dy.sum()])
    d2 = np.outer(dy, p["w3"]) * (a2 > 0)
    g["W2"] = d2.T @ h1
    g["b2"] = d2.sum(axis=0)
    d1 = (d2 @ p["W2"]) * (a1 > 0)
    g["W1"] = d1.T @ X
    g["b1"] = d1.sum(axis=0)
    flat = np.zeros(N_WEIGHTS)
    for k, (o, s) in SLICES.items():
        flat[o : o + int(np.prod(s))] = g[k].reshape(-1)
    return float(np.mean((y - T) ** 2)), flat

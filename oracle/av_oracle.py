"""numpy restatements for the shock-capturing workload.

TEST INFRASTRUCTURE ONLY (like the rest of oracle/): used by tests/ and the
CPU leg of bench.py as the checker. The product package never imports it.

* ``mlp_forward`` / ``mlp_grad``: the sensor network nin -> 64 -> 64 -> 1
  (tanh, tanh, sigmoid) and the gradient of the mean squared error, written
  from the chain rule. ``bf16_path=True`` rounds what the device's
  matrix-core path rounds (round-to-nearest-even to bfloat16): inputs, W1,
  W2 and the first hidden layer in the forward pass, and the operands of the
  backward products in the gradient.
* ``burgers_exact``: the entropy solution of u_t + div(u^2/2 (1,1,1)) = 0 on
  the periodic unit cube. On every line x0 + tau (1,1,1) the equation is the
  1D Burgers equation v_t + v v_tau = 0, so past shock formation
  v(0, t) = -y*/t, y* = argmin_y [V0(y) + y^2/(2t)], V0(y) = int_0^y v0
  (the Hopf-Lax / Lax-Oleinik formula; Evans, PDE, §3.4).
"""
from __future__ import annotations

import numpy as np

H = 64


def layout(nin: int) -> dict:
    o = {"W1": (0, (H, nin))}
    o["b1"] = (H * nin, (H,))
    o["W2"] = (H * nin + H, (H, H))
    o["b2"] = (H * nin + H + H * H, (H,))
    o["w3"] = (H * nin + 2 * H + H * H, (H,))
    o["b3"] = (H * nin + 3 * H + H * H, (1,))
    return o


def unpack(P: np.ndarray, nin: int) -> dict:
    out = {}
    for k, (o, shp) in layout(nin).items():
        out[k] = np.asarray(P[o:o + int(np.prod(shp))], dtype=np.float64).reshape(shp)
    return out


def bf16(x) -> np.ndarray:
    """Round float32 values to the nearest bfloat16 (ties to even), as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    b = a.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def mlp_forward(P: np.ndarray, x: np.ndarray, nin: int, bf16_path: bool = False) -> np.ndarray:
    w = unpack(P, nin)
    x = np.asarray(x, dtype=np.float64)
    W1, W2 = w["W1"], w["W2"]
    if bf16_path:
        x, W1, W2 = bf16(x), bf16(W1), bf16(W2)
    h1 = np.tanh(x @ W1.T + w["b1"])
    if bf16_path:
        h1 = bf16(h1)
    h2 = np.tanh(h1 @ W2.T + w["b2"])
    return 1.0 / (1.0 + np.exp(-(h2 @ w["w3"] + w["b3"][0])))


def mlp_grad(P: np.ndarray, x: np.ndarray, t: np.ndarray, nin: int, bf16_path: bool = False,
             rnd=None):
    """(gradient of mean (y - t)^2 as a flat array, loss), float64.

    ``bf16_path`` restates the device's matrix-core gradient: the forward
    rounds x, W1, W2 and h1 to bfloat16 (as ``mlp_forward``), the backward
    rounds dz2 and W2 (operands of W2^T dz2) and dz2, h1, dz1, x (operands of
    the weight-gradient products) to bfloat16, and takes the derivative of a
    rounding as 1 (tanh' from the unrounded tanh); the bias, w3 and b3
    gradients sum the unrounded values. ``rnd`` replaces the rounding (the
    identity gives the exact gradient: a check of the rounding points)."""
    w = unpack(P, nin)
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    q = rnd if rnd is not None else (bf16 if bf16_path else (lambda v: np.asarray(v, np.float64)))
    B = x.shape[0]
    xq, W1q, W2q = q(x), q(w["W1"]), q(w["W2"])
    h1 = np.tanh(xq @ W1q.T + w["b1"])
    h1q = q(h1)
    h2 = np.tanh(h1q @ W2q.T + w["b2"])
    y = 1.0 / (1.0 + np.exp(-(h2 @ w["w3"] + w["b3"][0])))
    err = y - t
    dy = 2.0 * err / B * y * (1.0 - y)
    dz2 = np.outer(dy, w["w3"]) * (1.0 - h2 * h2)
    dz2q = q(dz2)
    dz1 = (dz2q @ W2q) * (1.0 - h1 * h1)
    g = {"W1": q(dz1).T @ xq, "b1": dz1.sum(0), "W2": dz2q.T @ h1q, "b2": dz2.sum(0),
         "w3": h2.T @ dy, "b3": np.array([dy.sum()])}
    flat = np.zeros(sum(int(np.prod(s)) for _, s in layout(nin).values()))
    for k, (o, shp) in layout(nin).items():
        flat[o:o + int(np.prod(shp))] = g[k].reshape(-1)
    return flat, float(np.mean(err * err))


def burgers_exact(u0, X, Y, Z, t: float, umax: float, n_grid: int = 4001,
                  chunk: int = 4096) -> np.ndarray:
    """Entropy solution at points (X, Y, Z) and time t > 0 for the initial
    field u0(x, y, z) (vectorised, 1-periodic) with |u0| <= umax."""
    X, Y, Z = (np.asarray(a, dtype=np.float64).reshape(-1) for a in (X, Y, Z))
    L = umax * t * 1.05 + 1e-3
    ys = np.linspace(-L, L, n_grid)
    dy = ys[1] - ys[0]
    i0 = n_grid // 2  # ys[i0] == 0
    out = np.empty_like(X)
    for a in range(0, X.size, chunk):
        b = min(X.size, a + chunk)
        v0 = u0(X[a:b, None] + ys[None, :], Y[a:b, None] + ys[None, :], Z[a:b, None] + ys[None, :])
        V = np.concatenate([np.zeros((b - a, 1)), np.cumsum(0.5 * (v0[:, 1:] + v0[:, :-1]) * dy,
                                                            axis=1)], axis=1)
        V -= V[:, i0:i0 + 1]
        G = V + ys[None, :] ** 2 / (2.0 * t)
        j = np.clip(np.argmin(G, axis=1), 1, n_grid - 2)
        rows = np.arange(b - a)
        gm, g0, gp = G[rows, j - 1], G[rows, j], G[rows, j + 1]
        den = gm - 2.0 * g0 + gp
        shift = np.where(den > 0, 0.5 * (gm - gp) / np.where(den > 0, den, 1.0), 0.0)
        ystar = ys[j] + np.clip(shift, -1.0, 1.0) * dy
        out[a:b] = -ystar / t
    return out

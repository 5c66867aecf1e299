"""CPU restatement of the learned stage guess -- TEST INFRASTRUCTURE ONLY.

The network, its training step and its feature/target/guess maps have no upstream counterpart
(the upstream stage predictor is the linear extrapolation of core/solver.py:420,432-435).  This
module restates what the HIP kernels compute, rounding to BF16 at the same points, so that the
device forward pass, gradients and Adam step can be checked on the CPU.

Network: x (32) -> relu(W1 x + b1) (128) -> relu(W2 . + b2) -> relu(W3 . + b3) -> W4 . + b4 (16);
loss = mean over rows and the first 6 outputs of (out - (target - mean_t) / std_t)^2.
BF16 rounding points: normalised inputs, every weight, every hidden activation, the output
gradient and every hidden-layer gradient.  Products accumulate in float64 here (FP32 on the
matrix cores), so device and oracle differ by accumulation rounding only.

Pinning: ``bf16`` against torch's CPU bfloat16 conversion; the float64 variant of ``backward``
against central finite differences of ``loss``; ``forward`` against torch.nn.functional.linear
on the CPU; ``adam`` against torch.optim.Adam; ``features``/``targets``/``apply_guess`` by the
round trip apply(targets) = z_converged and by closed-form special cases.
"""
from __future__ import annotations

import numpy as np

MI, MH, MO, MO_USED = 32, 128, 16, 6
SHAPES = [("W1", (MH, MI)), ("b1", (MH,)), ("W2", (MH, MH)), ("b2", (MH,)), ("W3", (MH, MH)), ("b3", (MH,)),
          ("W4", (MO, MH)), ("b4", (MO,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)
NNORM = 2 * MI + 2 * MO


def unflatten(p):
    out, o = {}, 0
    for name, shp in SHAPES:
        k = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + k]).reshape(shp)
        o += k
    return out


def flatten(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def init_params(seed=0):
    """He-normal hidden weights, zero output layer and biases (the product's initialisation)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shp in SHAPES:
        if name.startswith("W") and name != "W4":
            d[name] = rng.standard_normal(shp) * np.sqrt(2.0 / shp[1])
        else:
            d[name] = np.zeros(shp)
    return flatten(d).astype(np.float32)


def _round(x, rounding):
    return bf16(x).astype(np.float64) if rounding else np.asarray(x, dtype=np.float64)


def forward(params, feat, norm, rounding=True):
    """Returns (out [rows][16], cache).  feat [rows][32] float32, norm [96]."""
    P = unflatten(np.asarray(params, dtype=np.float32).astype(np.float64))
    norm = np.asarray(norm, dtype=np.float32)
    xin = (np.asarray(feat, dtype=np.float32) - norm[:MI]) * norm[MI:2 * MI]     # float32 like the kernel
    x = _round(xin, rounding)
    W = {k: _round(np.asarray(P[k], dtype=np.float32), rounding) for k in ("W1", "W2", "W3", "W4")}
    hs = [x]
    h = x
    for k, b in (("W1", "b1"), ("W2", "b2"), ("W3", "b3")):
        h = _round(np.maximum(h @ W[k].T + P[b], 0.0), rounding)
        hs.append(h)
    out = h @ W["W4"].T + P["b4"]
    return out, {"h": hs, "W": W}


def loss(params, feat, target, norm, rounding=True):
    out, _ = forward(params, feat, norm, rounding)
    norm = np.asarray(norm, dtype=np.float64)
    tn = (np.asarray(target, dtype=np.float64)[:, :MO_USED] - norm[2 * MI:2 * MI + MO_USED]) / \
        norm[2 * MI + MO:2 * MI + MO + MO_USED]
    d = out[:, :MO_USED] - tn
    return float(np.mean(d * d))


def backward(params, feat, target, norm, rounding=True):
    """Gradient of ``loss`` in the flat parameter layout (float64)."""
    out, c = forward(params, feat, norm, rounding)
    rows = out.shape[0]
    norm = np.asarray(norm, dtype=np.float64)
    tn = (np.asarray(target, dtype=np.float64)[:, :MO_USED] - norm[2 * MI:2 * MI + MO_USED]) / \
        norm[2 * MI + MO:2 * MI + MO + MO_USED]
    g = np.zeros_like(out)
    g[:, :MO_USED] = (out[:, :MO_USED] - tn) * (2.0 / (rows * MO_USED))
    g = _round(g, rounding)
    hs, W = c["h"], c["W"]
    grads = {"W4": g.T @ hs[3], "b4": g.sum(0)}
    for L in (3, 2, 1):
        dh = g @ W[f"W{L + 1}"]
        g = _round(dh * (hs[L] > 0.0), rounding)
        grads[f"W{L}"] = g.T @ hs[L - 1]
        grads[f"b{L}"] = g.sum(0)
    return flatten(grads)


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (bias-corrected), float32 like the device; returns (p, m, v)."""
    p, g, m, v = (np.asarray(a, dtype=np.float32) for a in (p, g, m, v))
    m = np.float32(b1) * m + np.float32(1.0 - b1) * g
    v = np.float32(b2) * v + np.float32(1.0 - b2) * g * g
    c1 = np.float32(1.0 / (1.0 - b1 ** step))
    c2 = np.float32(1.0 / (1.0 - b2 ** step))
    p = p - np.float32(lr) * (m * c1) / (np.sqrt(v * c2) + np.float32(eps))
    return p, m, v


def features(y, z_pred, s, R, n, Nz, log_h, stage):
    """Feature rows [n][32] of one stage (cell variables first in the variable-major state);
    R is the residual term the caller supplies (the integrator passes M^-1 R(z_pred))."""
    Y = np.asarray(y[:6 * n]).reshape(6, n)
    k = np.arange(n) % Nz
    km = np.where(k > 0, np.arange(n) - 1, np.arange(n))
    kp = np.where(k < Nz - 1, np.arange(n) + 1, np.arange(n))
    S = np.asarray(s[:6 * n]).reshape(6, n)
    F = np.zeros((n, MI))
    F[:, 0:6] = Y.T
    F[:, 6:12] = Y[:, km].T
    F[:, 12:18] = Y[:, kp].T
    F[:, 18:24] = ((np.asarray(z_pred[:6 * n]).reshape(6, n) - Y) / S).T
    F[:, 24:30] = (np.asarray(R[:6 * n]).reshape(6, n) / S).T
    F[:, 30] = log_h
    F[:, 31] = stage
    return F.astype(np.float32)


def targets(z_conv, z_pred, s, n, r=None):
    """(z_conv - z_base)/s on the cell variables, z_base = z_pred - r (z_pred when r is None)."""
    base = np.asarray(z_pred, dtype=np.float64) - (0.0 if r is None else np.asarray(r, dtype=np.float64))
    T = np.zeros((n, 8), dtype=np.float32)
    T[:, :6] = ((np.asarray(z_conv[:6 * n]) - base[:6 * n]) / np.asarray(s[:6 * n])).reshape(6, n).T
    return T


def apply_guess(out, norm, z_pred, s, n, r=None):
    norm = np.asarray(norm, dtype=np.float32).astype(np.float64)
    zp = np.asarray(z_pred, dtype=np.float64)
    z = zp - (0.0 if r is None else np.asarray(r, dtype=np.float64))
    d = np.asarray(out, dtype=np.float32).astype(np.float64)[:, :6] * norm[2 * MI + MO:2 * MI + MO + 6] + \
        norm[2 * MI:2 * MI + 6]
    z[:6 * n] = z[:6 * n] + np.asarray(s[:6 * n]) * d.T.ravel()
    lo = 0.1 * np.maximum(zp[:5 * n], 0.0)          # positivity truncation of c and q
    z[:5 * n] = np.where(z[:5 * n] > lo, z[:5 * n], lo)
    return z

"""Learned stage guess (no upstream counterpart).

The upstream integrator starts every implicit stage from a linear extrapolation
(core/solver.py:420,432-435) and lets Newton-Krylov correct it.  Here the start is the
predictor after one preconditioned Newton step, z_pred - M^-1 R(z_pred), plus a correction
that a small per-cell network predicts; the network is trained on the device from the
simulator's own converged stages.  The Newton-Krylov solve still runs to the usual tolerance
from the guess, and the guess is used only when its preconditioned stage residual is smaller
than the predictor's (otherwise the predictor is kept).  Everything runs in libvtsa: the network's forward and backward passes are
BF16 matrix-core kernels with FP32 accumulation, the optimiser is Adam in FP32.

* ``StageNet``     -- the network (parameters, normalisation, forward, one training step).
* ``StageDataset`` -- device-resident (features, targets) rows of recorded stages of one grid.
* ``StageGuess``   -- what the integrator calls: features -> network -> guess.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _abi
from . import device as dv


def sizes():
    v = [C.c_int32() for _ in range(5)]
    _abi.call("vtsa_mlp_sizes", *[C.byref(x) for x in v])
    return {k: x.value for k, x in zip(("n_param", "n_in", "n_hidden", "n_out", "n_norm"), v)}


SHAPES = (("W1", (128, 32)), ("b1", (128,)), ("W2", (128, 128)), ("b2", (128,)), ("W3", (128, 128)),
          ("b3", (128,)), ("W4", (16, 128)), ("b4", (16,)))
N_USED = 6


def init_params(seed=0):
    """He-normal hidden weights, zero output layer and biases (an untrained network returns the
    target mean, so the guess starts as the preconditioned step)."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shp in SHAPES:
        if name.startswith("W") and name != "W4":
            parts.append((rng.standard_normal(shp) * np.sqrt(2.0 / shp[1])).ravel())
        else:
            parts.append(np.zeros(int(np.prod(shp))))
    return np.concatenate(parts).astype(np.float32)


def identity_norm():
    z = sizes()
    mi, mo = z["n_in"], z["n_out"]
    return np.concatenate([np.zeros(mi), np.ones(mi), np.zeros(mo), np.ones(mo)]).astype(np.float32)


class StageNet:
    """The per-cell network on the device; rows (cells) per call are at most ``max_rows``."""

    def __init__(self, max_rows: int, params=None, norm=None, seed: int = 0):
        z = sizes()
        self.n_param, self.n_in, self.n_out = z["n_param"], z["n_in"], z["n_out"]
        self.max_rows = int(max_rows)
        h = C.c_void_p()
        _abi.call("vtsa_mlp_create", self.max_rows, C.byref(h))
        self.handle = h
        self.set_params(init_params(seed) if params is None else params)
        self.set_norm(identity_norm() if norm is None else norm)

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.n_param:
            raise ValueError(f"expected {self.n_param} parameters, got {p.size}")
        _abi.call("vtsa_mlp_set_params", self.handle, p.ctypes.data_as(C.c_void_p), dv.stream())
        self._params = p.copy()

    def get_params(self):
        p = np.empty(self.n_param, np.float32)
        _abi.call("vtsa_mlp_get_params", self.handle, p.ctypes.data_as(C.c_void_p), dv.stream())
        return p

    def set_norm(self, norm):
        norm = np.ascontiguousarray(norm, dtype=np.float32)
        _abi.call("vtsa_mlp_set_norm", self.handle, norm.ctypes.data_as(C.c_void_p), dv.stream())
        self.norm = norm.copy()

    def forward(self, feat, rows, out):
        _abi.call("vtsa_mlp_forward", self.handle, dv.ptr(feat), int(rows), dv.ptr(out), dv.stream())
        return out

    def train_step(self, feat, target, rows, lr, want_loss=False, want_grad=False):
        """One Adam step on one minibatch; returns (loss or None, grad or None)."""
        loss = C.c_float() if want_loss else None
        grad = np.empty(self.n_param, np.float32) if want_grad else None
        _abi.call("vtsa_mlp_train_step", self.handle, dv.ptr(feat), dv.ptr(target), int(rows), C.c_float(lr),
                  C.byref(loss) if loss is not None else None,
                  grad.ctypes.data_as(C.c_void_p) if grad is not None else None, dv.stream())
        return (loss.value if loss is not None else None), grad

    def save(self, path):
        np.savez(path, params=self.get_params(), norm=self.norm)

    @classmethod
    def load(cls, path, max_rows):
        z = np.load(path)
        return cls(max_rows, params=z["params"], norm=z["norm"])

    def close(self):
        if self.handle is not None and self.handle.value:
            _abi.lib().vtsa_mlp_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class StageDataset:
    """Device-resident training rows of recorded stages on one grid (n cells per stage)."""

    def __init__(self, n: int, capacity: int, n_in: int = 32, stride: int = 1):
        """capacity stages of n cells; every stride-th stage offered to ``begin`` is recorded."""
        self.n, self.capacity, self.n_in = int(n), int(capacity), n_in
        self.stride, self.offered = max(1, int(stride)), 0
        self.feat = dv.empty(self.capacity * self.n * n_in, np.float32)
        self.targ = dv.empty(self.capacity * self.n * 8, np.float32)
        self.count = 0

    def slot(self, k):
        return (self.feat.view(k * self.n * self.n_in, self.n * self.n_in), self.targ.view(k * self.n * 8, self.n * 8))

    def begin(self, handle, y, z_pred, s, r, h, stage):
        """Features of the next stage (slot kept only if ``commit`` follows); r is the
        preconditioned stage residual M^-1 R(z_pred)."""
        self.offered += 1
        if self.count >= self.capacity or (self.offered - 1) % self.stride:
            return False
        f, _ = self.slot(self.count)
        _abi.call("vtsa_mlp_features", handle, dv.ptr(y), dv.ptr(z_pred), dv.ptr(s), dv.ptr(r), float(np.log(h)),
                  float(stage), dv.ptr(f), dv.stream())
        return True

    def commit(self, handle, z_conv, z_pred, r, s):
        """Targets of the stage begun last: (z_conv - (z_pred - r)) / s."""
        _, t = self.slot(self.count)
        _abi.call("vtsa_mlp_targets", handle, dv.ptr(z_conv), dv.ptr(z_pred), dv.ptr(r), dv.ptr(s), dv.ptr(t),
                  dv.stream())
        self.count += 1

    def norm(self, n_out=16, max_stages=64, seed=0):
        """Per-column mean and 1/std of the features, mean and std of the targets, from up to
        ``max_stages`` recorded stages (columns with no spread get std 1)."""
        if self.count == 0:
            raise RuntimeError("empty dataset")
        rng = np.random.default_rng(seed)
        pick = np.sort(rng.choice(self.count, min(self.count, max_stages), replace=False))
        F = np.concatenate([self.slot(k)[0].numpy().reshape(self.n, self.n_in) for k in pick]).astype(np.float64)
        T = np.concatenate([self.slot(k)[1].numpy().reshape(self.n, 8) for k in pick]).astype(np.float64)
        fm, fs = F.mean(0), F.std(0)
        fs = np.where(fs > 1e-12 * np.maximum(1.0, np.abs(fm)), fs, 1.0)
        tm, ts = np.zeros(n_out), np.ones(n_out)
        tm[:N_USED], ts[:N_USED] = T[:, :N_USED].mean(0), T[:, :N_USED].std(0)
        ts = np.where(ts > 1e-12, ts, 1.0)
        return np.concatenate([fm, 1.0 / fs, tm, ts]).astype(np.float32)

    def train(self, net: StageNet, epochs=20, lr=1e-3, lr_final=1e-4, seed=0, holdout=10, log=None,
              calibrate=True):
        """Adam over whole stages as minibatches, cosine learning-rate decay.  Every
        ``holdout``-th recorded stage is kept out for validation and, with ``calibrate``, for
        the per-variable shrinkage of the output (``calibrate``).  Returns a dict of per-epoch
        training and validation losses, the held-out stage ids, the gains and the wall time."""
        rng = np.random.default_rng(seed)
        ids = np.arange(self.count)
        held = (ids % holdout) == holdout - 1 if (holdout and self.count >= 2 * holdout) else np.zeros(self.count, bool)
        train_ids, val_ids = ids[~held], ids[held]
        total = epochs * len(train_ids)
        k = 0
        hist = {"train": [], "val": []}
        t0 = time.time()
        for ep in range(epochs):
            order = rng.permutation(train_ids)
            losses = []
            for j, sid in enumerate(order):
                a = lr_final + 0.5 * (lr - lr_final) * (1.0 + np.cos(np.pi * k / max(1, total)))
                f, t = self.slot(sid)
                want = j % 16 == 0
                loss, _ = net.train_step(f, t, self.n, a, want_loss=want)
                if want:
                    losses.append(loss)
                k += 1
            hist["train"].append(float(np.mean(losses)))
            if len(val_ids):
                hist["val"].append(self.evaluate(net, val_ids))
            if log:
                log(f"epoch {ep + 1}/{epochs}: train {hist['train'][-1]:.4g}"
                    + (f"  val {hist['val'][-1]:.4g}" if hist["val"] else ""))
        hist["gain"] = self.calibrate(net, val_ids).tolist() if (calibrate and len(val_ids)) else None
        dv.synchronize()
        hist["wall_s"] = time.time() - t0
        hist["steps"] = k
        hist["val_ids"] = val_ids
        return hist

    def errors(self, net: StageNet, ids):
        """Per-stage rms errors (tolerance units, per cell variable) against the converged stage
        of three starts: the predictor, the predictor after one preconditioned Newton step (the
        base of the targets) and the guess.  Returns three arrays [len(ids)][6]."""
        out = dv.empty(self.n * net.n_out, np.float32)
        tm, ts = net.norm[64:64 + N_USED].astype(np.float64), net.norm[80:80 + N_USED].astype(np.float64)
        ep, eb, eg = [], [], []
        for sid in ids:
            f, t = self.slot(int(sid))
            net.forward(f, self.n, out)
            O = out.numpy().reshape(self.n, net.n_out)[:, :N_USED].astype(np.float64) * ts + tm
            T = t.numpy().reshape(self.n, 8)[:, :N_USED].astype(np.float64)
            P = f.numpy().reshape(self.n, self.n_in)[:, 24:30].astype(np.float64)
            ep.append(np.sqrt(np.mean((P - T) ** 2, 0)))
            eb.append(np.sqrt(np.mean(T ** 2, 0)))
            eg.append(np.sqrt(np.mean((O - T) ** 2, 0)))
        return np.array(ep), np.array(eb), np.array(eg)

    def calibrate(self, net: StageNet, ids):
        """Shrink the network's correction per variable by the gain in [0, 1] that minimises the
        squared error on the given (held-out) stages, folded into the output layer.  Returns
        the gains."""
        out = dv.empty(self.n * net.n_out, np.float32)
        tm, ts = net.norm[64:64 + N_USED].astype(np.float64), net.norm[80:80 + N_USED].astype(np.float64)
        num, den = np.zeros(N_USED), np.zeros(N_USED)
        for sid in ids:
            f, t = self.slot(int(sid))
            net.forward(f, self.n, out)
            O = out.numpy().reshape(self.n, net.n_out)[:, :N_USED].astype(np.float64) * ts + tm
            T = t.numpy().reshape(self.n, 8)[:, :N_USED].astype(np.float64)
            num += np.sum(O * T, 0)
            den += np.sum(O * O, 0)
        gain = np.clip(np.where(den > 0.0, num / np.maximum(den, 1e-300), 0.0), 0.0, 1.0)
        # gain * (out std + mean) = (gain W4 h + gain b4 + (gain - 1) mean / std) std + mean
        p = net.get_params().astype(np.float64)
        o_w4 = sum(int(np.prod(s)) for name, s in SHAPES[:6])
        o_b4 = o_w4 + 16 * 128
        W4 = p[o_w4:o_b4].reshape(16, 128)
        W4[:N_USED] *= gain[:, None]
        p[o_w4:o_b4] = W4.ravel()
        p[o_b4:o_b4 + N_USED] = gain * p[o_b4:o_b4 + N_USED] + (gain - 1.0) * tm / ts
        net.set_params(p.astype(np.float32))
        return gain

    def evaluate(self, net: StageNet, ids):
        """Mean loss over the given stages without updating the network (gradient-only steps)."""
        vals = []
        for sid in ids:
            f, t = self.slot(int(sid))
            loss, _ = net.train_step(f, t, self.n, 0.0, want_loss=True)
            vals.append(loss)
        return float(np.mean(vals))


class StageGuess:
    """Features -> network -> guess for the stages of one model handle."""

    def __init__(self, net: StageNet, n: int, N: int):
        if n > net.max_rows:
            raise ValueError("the network's row capacity is smaller than the grid")
        self.net, self.n, self.N = net, n, N
        self.feat = dv.empty(n * net.n_in, np.float32)
        self.out = dv.empty(n * net.n_out, np.float32)

    def predict(self, handle, y, z_pred, s, r, h, stage, z_out):
        _abi.call("vtsa_mlp_features", handle, dv.ptr(y), dv.ptr(z_pred), dv.ptr(s), dv.ptr(r), float(np.log(h)),
                  float(stage), dv.ptr(self.feat), dv.stream())
        self.net.forward(self.feat, self.n, self.out)
        _abi.call("vtsa_mlp_apply", self.net.handle, handle, dv.ptr(self.out), dv.ptr(z_pred), dv.ptr(r), dv.ptr(s),
                  dv.ptr(z_out), dv.stream())
        return z_out

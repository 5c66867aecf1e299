"""Finite-difference Jacobian on the GPU (drop-in for upstream core/jacobian.py and
numerics/sparsity.py).

The device evaluates every column of the exact sparsity pattern in one launch and stores the
values in 6x6 cell blocks (``vtsa_fd_jacobian``); the integrator uses them in place.  The
upstream API -- ``FDJacobian(model)(t, y, f0) -> scipy CSC`` on the upstream pattern,
``theta_block``, ``P``, ``color``, ``ngroups``, ``nevals``, ``typ`` -- is provided on top by a
gather from the block storage into the CSC order of the pattern.
"""
from __future__ import annotations

from functools import lru_cache

import numpy as np
import scipy.sparse as sp

from .. import _abi
from .. import device as dv


def build_pattern(Nr, Nth, Nz, ng, na, wall):
    """Exact pattern of df/dy (numerics/sparsity.py:29-85): flux rows couple all variables of
    the cell, z+-1, z+-2, r+-1, theta+-1 (and the wall for the outer ring), loading rows the
    own cell, wall rows their wall neighbours and the outer cell."""
    n = Nr * Nth * Nz
    nv = ng + na + 1
    idx = np.arange(n).reshape(Nr, Nth, Nz)
    rc, cc = [idx.ravel()], [idx.ravel()]
    for dk in (1, 2):
        if Nz > dk:
            rc += [idx[:, :, dk:].ravel(), idx[:, :, :-dk].ravel()]
            cc += [idx[:, :, :-dk].ravel(), idx[:, :, dk:].ravel()]
    if Nr > 1:
        rc += [idx[1:].ravel(), idx[:-1].ravel()]
        cc += [idx[:-1].ravel(), idx[1:].ravel()]
    if Nth > 1:
        nb = np.roll(idx, -1, axis=1)
        rc += [idx.ravel(), nb.ravel()]
        cc += [nb.ravel(), idx.ravel()]
    rc, cc = np.concatenate(rc), np.concatenate(cc)
    flux = list(range(ng)) + [ng + na]
    rows, cols = [], []
    for vr in flux:
        for vc in range(nv):
            rows.append(vr * n + rc)
            cols.append(vc * n + cc)
    own = idx.ravel()
    for vr in range(ng, ng + na):
        for vc in range(nv):
            rows.append(vr * n + own)
            cols.append(vc * n + own)
    if wall:
        base = nv * n
        widx = np.arange(Nth * Nz).reshape(Nth, Nz)
        outer = idx[-1].ravel()
        wr, wc = [widx.ravel()], [widx.ravel()]
        if Nz > 1:
            wr += [widx[:, 1:].ravel(), widx[:, :-1].ravel()]
            wc += [widx[:, :-1].ravel(), widx[:, 1:].ravel()]
        if Nth > 1:
            nbw = np.roll(widx, -1, axis=0)
            wr += [widx.ravel(), nbw.ravel()]
            wc += [nbw.ravel(), widx.ravel()]
        rows.append(base + np.concatenate(wr))
        cols.append(base + np.concatenate(wc))
        for vc in range(nv):
            rows.append(base + widx.ravel())
            cols.append(vc * n + outer)
        for vr in flux:
            rows.append(vr * n + outer)
            cols.append(base + widx.ravel())
    r, c = np.concatenate(rows), np.concatenate(cols)
    N = nv * n + (Nth * Nz if wall else 0)
    P = sp.coo_matrix((np.ones(len(r), dtype=bool), (r, c)), shape=(N, N)).tocsc()
    P.sum_duplicates()
    P.data[:] = True
    return P


def color_columns(P):
    """Greedy largest-degree-first distance-2 colouring (numerics/sparsity.py:88-107).  Only
    used for the ``color``/``ngroups`` attributes: the device evaluates all columns at once."""
    P = P.astype(np.int8).tocsc()
    G = (P.T @ P).tocsr()
    color = -np.ones(P.shape[1], dtype=np.int64)
    for j in np.argsort(-np.diff(G.indptr), kind="stable"):
        used = color[G.indices[G.indptr[j]:G.indptr[j + 1]]]
        used = used[used >= 0]
        if used.size == 0:
            color[j] = 0
            continue
        mark = np.zeros(used.max() + 2, dtype=bool)
        mark[used] = True
        color[j] = int(np.argmin(mark))
    return color


def verify_coloring(P, color) -> bool:
    P = P.tocsr()
    for r in range(P.shape[0]):
        cc = color[P.indices[P.indptr[r]:P.indptr[r + 1]]]
        if len(np.unique(cc)) != len(cc):
            return False
    return True


@lru_cache(maxsize=16)
def _pattern(Nr, Nth, Nz, wall):
    return build_pattern(Nr, Nth, Nz, 3, 2, wall)


@lru_cache(maxsize=16)
def _gather(Nr, Nth, Nz, wall):
    """Index of every CSC entry of the pattern into the flat block storage [Jc | Jcw | Jw | Jwc]."""
    P = _pattern(Nr, Nth, Nz, wall)
    n = Nr * Nth * Nz
    nwall = Nth * Nz
    col = np.repeat(np.arange(P.shape[1]), np.diff(P.indptr))
    row = P.indices.astype(np.int64)
    out = np.full(len(row), -1, dtype=np.int64)
    o_cw = n * 9 * 36
    o_w = o_cw + nwall * 4
    o_wc = o_w + nwall * 5

    def ijk(cell):
        return cell // (Nth * Nz), (cell // Nz) % Nth, cell % Nz

    cr, cc = row < 6 * n, col < 6 * n
    m = cr & cc
    X, A = row[m] % n, col[m] % n
    vr, vc = row[m] // n, col[m] // n
    (i, j, k), (i2, j2, k2) = ijk(X), ijk(A)
    slot = np.full(X.shape, -1)
    same_col = (i == i2) & (j == j2)
    dk = k2 - k
    for s, d in ((0, 0), (1, -1), (2, 1), (3, -2), (4, 2)):
        slot[same_col & (dk == d)] = s
    rline = (j == j2) & (k == k2)
    slot[rline & (i2 == i - 1)] = 5
    slot[rline & (i2 == i + 1)] = 6
    if Nth > 1:
        tline = (i == i2) & (k == k2)
        dj = (j2 - j) % Nth
        slot[tline & (dj == Nth - 1)] = 7
        slot[tline & (dj == 1)] = 8
    assert np.all(slot >= 0)
    out[m] = ((X * 9 + slot) * 6 + vr) * 6 + vc
    if wall:
        m = cr & ~cc
        W = col[m] - 6 * n
        vr = row[m] // n
        out[m] = o_cw + W * 4 + np.where(vr == 5, 3, vr)
        m = ~cr & ~cc
        W, W2 = row[m] - 6 * n, col[m] - 6 * n
        jw, kw, j2, k2 = W // Nz, W % Nz, W2 // Nz, W2 % Nz
        ws = np.full(W.shape, -1)
        ws[(jw == j2) & (kw == k2)] = 0
        ws[(jw == j2) & (k2 == kw - 1)] = 1
        ws[(jw == j2) & (k2 == kw + 1)] = 2
        if Nth > 1:
            ws[(kw == k2) & ((j2 - jw) % Nth == Nth - 1)] = 3
            ws[(kw == k2) & ((j2 - jw) % Nth == 1)] = 4
        assert np.all(ws >= 0)
        out[m] = o_w + W * 5 + ws
        m = ~cr & cc
        W, vc = row[m] - 6 * n, col[m] // n
        out[m] = o_wc + W * 6 + vc
    assert np.all(out >= 0)
    return out


class FDJacobian:
    """Coloured-FD-equivalent Jacobian of f (upstream core/jacobian.py:21-91)."""

    def __init__(self, model):
        g = model.grid
        self.model = model
        self.P = _pattern(g.Nr, g.Nth, g.Nz, model.n_w > 0)
        assert self.P.shape[0] == model.n_core, (self.P.shape, model.n_core)
        self.n = self.P.shape[0]
        self.indptr, self.indices = self.P.indptr, self.P.indices
        self.typ = np.concatenate([np.full(g.ncell, v) for v in model.typ_var[:6]] +
                                  ([np.ones(model.n_w)] if model.n_w else []))
        self.nevals = 0
        self._color = None
        self._theta = None

    def _theta_pattern(self):
        """(kept entries, row indices, column pointers) of the theta-block pattern (upstream
        core/jacobian.py:71-91), formed on first use: the device path never needs it."""
        if self._theta is not None:
            return self._theta
        model, g = self.model, self.model.grid
        nv = model.ng + model.na + 1
        n = g.ncell
        idx = np.arange(self.n)
        cell = idx < nv * n
        jsec = np.empty(self.n, dtype=np.int64)
        jsec[cell] = ((idx[cell] % n) // g.Nz) % g.Nth
        jsec[~cell] = (idx[~cell] - nv * n) // g.Nz
        ring = np.empty(self.n, dtype=np.int64)
        ring[cell] = (idx[cell] % n) // (g.Nth * g.Nz)
        ring[~cell] = g.Nr - 1
        ecol = np.repeat(np.arange(self.n), np.diff(self.indptr))
        erow = self.indices
        cut = model.cfg.solver.precond_theta_rings - 1
        keep = (jsec[erow] == jsec[ecol]) | ((ring[erow] <= cut) & (ring[ecol] <= cut))
        kept = np.where(keep)[0]
        self._theta = (kept, erow[kept], np.concatenate([[0], np.cumsum(np.bincount(ecol[kept], minlength=self.n))]))
        return self._theta

    @property
    def color(self):
        if self._color is None:
            self._color = color_columns(self.P)
        return self._color

    @property
    def ngroups(self):
        return int(self.color.max()) + 1

    def evaluate(self, t, y_dev, f0_dev):
        """Device Jacobian blocks at (t, y) given f0 = f(t, y); stays in the model's handle."""
        _abi.call("vtsa_fd_jacobian", self.model.handle, float(t), dv.ptr(y_dev), dv.ptr(f0_dev), dv.stream())
        self.nevals += 1

    def blocks(self):
        """Host copies (Jc, Jcw, Jw, Jwc) of the block storage."""
        g = self.model.grid
        n, nwall = g.ncell, g.Nth * g.Nz
        Jc, Jcw, Jw, Jwc = np.zeros(n * 9 * 36), np.zeros(nwall * 4), np.zeros(nwall * 5), np.zeros(nwall * 6)
        ptr = lambda a: a.ctypes.data_as(_abi.P)  # noqa: E731
        _abi.call("vtsa_jacobian_get", self.model.handle, ptr(Jc), ptr(Jcw), ptr(Jw), ptr(Jwc))
        return Jc, Jcw, Jw, Jwc

    def __call__(self, t, y, f0=None) -> sp.csc_matrix:
        m = self.model
        yd = dv.to_dev(y) if isinstance(y, np.ndarray) else y
        if f0 is None:
            f0d = dv.empty(m.n_core)
            m.rhs_dev(t, yd, f0d)
        else:
            f0d = dv.to_dev(f0) if isinstance(f0, np.ndarray) else f0
        self.evaluate(t, yd, f0d)
        g = m.grid
        flat = np.concatenate(self.blocks())
        data = flat[_gather(g.Nr, g.Nth, g.Nz, m.n_w > 0)]
        return sp.csc_matrix((data, self.indices, self.indptr), shape=(self.n, self.n))

    def theta_block(self, J: sp.csc_matrix) -> sp.csc_matrix:
        kept, red_indices, red_indptr = self._theta_pattern()
        return sp.csc_matrix((J.data[kept], red_indices, red_indptr), shape=(self.n, self.n))

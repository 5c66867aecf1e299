"""Oracle restatement of the Jacobian sparsity, colouring and coloured FD Jacobian.

numerics/sparsity.py:29-123 and core/jacobian.py:21-91 of the upstream code.
"""
from __future__ import annotations

from functools import lru_cache

import numpy as np
import scipy.sparse as sp


def build_pattern(Nr, Nth, Nz, ng, na, wall):
    """Boolean CSC pattern of df/dy (numerics/sparsity.py:29-85)."""
    n = Nr * Nth * Nz
    nv = ng + na + 1
    idx = np.arange(n).reshape(Nr, Nth, Nz)
    pairs = [(idx.ravel(), idx.ravel())]
    for dk in (1, 2):
        if Nz > dk:
            pairs.append((idx[:, :, dk:].ravel(), idx[:, :, :-dk].ravel()))
            pairs.append((idx[:, :, :-dk].ravel(), idx[:, :, dk:].ravel()))
    if Nr > 1:
        pairs.append((idx[1:].ravel(), idx[:-1].ravel()))
        pairs.append((idx[:-1].ravel(), idx[1:].ravel()))
    if Nth > 1:
        nb = np.roll(idx, -1, axis=1)
        pairs.append((idx.ravel(), nb.ravel()))
        pairs.append((nb.ravel(), idx.ravel()))
    rc = np.concatenate([p[0] for p in pairs])
    cc = np.concatenate([p[1] for p in pairs])
    rows, cols = [], []
    flux_vars = list(range(ng)) + [ng + na]
    for vr in flux_vars:
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
        outer = idx[-1]
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
            cols.append(vc * n + outer.ravel())
        for vr in flux_vars:
            rows.append(vr * n + outer.ravel())
            cols.append(base + widx.ravel())
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    N = nv * n + (Nth * Nz if wall else 0)
    P = sp.coo_matrix((np.ones(len(r), dtype=bool), (r, c)), shape=(N, N)).tocsc()
    P.sum_duplicates()
    P.data[:] = True
    return P


def color_columns(P):
    """Greedy largest-degree-first distance-2 colouring (numerics/sparsity.py:88-107)."""
    P = P.astype(np.int8).tocsc()
    G = (P.T @ P).tocsr()
    n = P.shape[1]
    color = -np.ones(n, dtype=np.int64)
    indptr, indices = G.indptr, G.indices
    for j in np.argsort(-np.diff(indptr), kind="stable"):
        used = color[indices[indptr[j]:indptr[j + 1]]]
        used = used[used >= 0]
        if used.size == 0:
            color[j] = 0
            continue
        mark = np.zeros(used.max() + 2, dtype=bool)
        mark[used] = True
        color[j] = int(np.argmin(mark))
    return color


def verify_coloring(P, color):
    """True iff no two columns of one colour share a row (numerics/sparsity.py:110-117)."""
    P = P.tocsr()
    for r in range(P.shape[0]):
        cc = color[P.indices[P.indptr[r]:P.indptr[r + 1]]]
        if len(np.unique(cc)) != len(cc):
            return False
    return True


@lru_cache(maxsize=16)
def pattern_and_colors(Nr, Nth, Nz, ng, na, wall):
    P = build_pattern(Nr, Nth, Nz, ng, na, wall)
    return P, color_columns(P)


class OracleFDJacobian:
    """Coloured forward-difference Jacobian on the exact pattern (core/jacobian.py:21-91)."""

    def __init__(self, model):
        g = model.grid
        self.model = model
        self.P, self.color = pattern_and_colors(g.Nr, g.Nth, g.Nz, model.ng, model.na, model.n_w > 0)
        P = self.P.tocsc()
        self.indptr, self.indices = P.indptr, P.indices
        self.n = P.shape[0]
        self.ngroups = int(self.color.max()) + 1
        self.entry_col = np.repeat(np.arange(self.n), np.diff(self.indptr))
        self.entry_row = self.indices
        self.groups = [np.where(self.color == k)[0] for k in range(self.ngroups)]
        ent_group = self.color[self.entry_col]
        order = np.argsort(ent_group, kind="stable")
        bounds = np.searchsorted(ent_group[order], np.arange(self.ngroups + 1))
        self.group_entries = [order[bounds[k]:bounds[k + 1]] for k in range(self.ngroups)]
        n = g.ncell
        typ = [np.full(n, {"CO2": 1e-3, "H2O": 1e-2, "N2": 1.0}[s]) for s in model.species]
        typ += [np.full(n, 1e-3)] * model.na
        typ.append(np.full(n, model.rho_b * model.cfg.sorbent.cp_solid))
        if model.n_w:
            typ.append(np.ones(model.n_w))
        self.typ = np.concatenate(typ)
        self.nevals = 0
        nv = model.ng + model.na + 1
        idx = np.arange(self.n)
        cell = idx < nv * n
        jsec = np.empty(self.n, dtype=np.int64)
        jsec[cell] = ((idx[cell] % n) // g.Nz) % g.Nth
        jsec[~cell] = (idx[~cell] - nv * n) // g.Nz
        ring = np.empty(self.n, dtype=np.int64)
        ring[cell] = (idx[cell] % n) // (g.Nth * g.Nz)
        ring[~cell] = g.Nr - 1
        cut = model.cfg.solver.precond_theta_rings - 1
        keep = (jsec[self.entry_row] == jsec[self.entry_col]) | \
               ((ring[self.entry_row] <= cut) & (ring[self.entry_col] <= cut))
        self.keep_theta = np.where(keep)[0]
        self.red_indices = self.entry_row[self.keep_theta]
        self.red_indptr = np.concatenate([[0], np.cumsum(np.bincount(self.entry_col[self.keep_theta],
                                                                     minlength=self.n))])

    def steps(self, y):
        """Exactly representable FD steps h (core/jacobian.py:75-76)."""
        h = 1.5e-8 * np.maximum(np.abs(y), self.typ)
        return (y + h) - y

    def __call__(self, t, y, f0=None):
        m = self.model
        if f0 is None:
            f0, _ = m.rhs(t, y, want_acc=False)
        h = self.steps(y)
        data = np.empty(len(self.entry_row))
        for k in range(self.ngroups):
            cols = self.groups[k]
            yp = y.copy()
            yp[cols] += h[cols]
            fp, _ = m.rhs(t, yp, want_acc=False)
            ent = self.group_entries[k]
            data[ent] = (fp[self.entry_row[ent]] - f0[self.entry_row[ent]]) / h[self.entry_col[ent]]
        self.nevals += 1
        return sp.csc_matrix((data, self.indices, self.indptr), shape=(self.n, self.n))

    def theta_block(self, J):
        return sp.csc_matrix((J.data[self.keep_theta], self.red_indices, self.red_indptr),
                             shape=(self.n, self.n))

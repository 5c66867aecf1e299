"""Low-rank correction of the GMRES right preconditioner (no upstream counterpart).

CPU: lowrank_correction on Arnoldi data whose answer the mathematics fixes -- an invariant Krylov
space, where the harmonic Ritz pairs are exact eigenpairs and B Q must act as the identity on the
chosen subspace, and a complex pair, which must enter as one real plane.
GPU: the device apply of Q against the same formula on the host, and device GMRES cycles with and
without the correction on one Newton system.
"""
from __future__ import annotations

import numpy as np
import pytest

from dac_vtsa_amd.core.solver import lowrank_correction


def _arnoldi(B, b, m):
    n = len(b)
    V = np.zeros((n, m + 1))
    H = np.zeros((m + 1, m))
    V[:, 0] = b / np.linalg.norm(b)
    for j in range(m):
        w = B @ V[:, j]
        for _ in range(2):
            hj = V[:, :j + 1].T @ w
            w -= V[:, :j + 1] @ hj
            H[:j + 1, j] += hj
        H[j + 1, j] = np.linalg.norm(w)
        if H[j + 1, j] > 0:
            V[:, j + 1] = w / H[j + 1, j]
    return V, H


def test_invariant_space_gives_exact_eigenvectors():
    rng = np.random.default_rng(3)
    n = 24
    Qo, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.concatenate([[0.01, 0.03, 0.07], np.linspace(0.6, 2.0, n - 3)])
    B = Qo @ np.diag(lam) @ Qo.T
    V, H = _arnoldi(B, rng.standard_normal(n), n)
    G, D = lowrank_correction(H, n, 3)
    assert G.shape == (n, 3) and np.allclose(G.T @ G, np.eye(3), atol=1e-12)
    U = V[:, :n] @ G
    # U spans the eigenvectors of the three smallest eigenvalues
    assert np.allclose(Qo[:, :3] @ (Qo[:, :3].T @ U), U, atol=1e-8)
    Q = np.eye(n) + U @ D @ U.T
    assert np.allclose(B @ Q @ U, U, atol=1e-8)
    ev = np.sort(np.linalg.eigvals(B @ Q).real)
    assert np.allclose(ev, np.sort(np.concatenate([np.ones(3), lam[3:]])), atol=1e-8)


def test_complex_pair_enters_as_one_real_plane():
    rng = np.random.default_rng(5)
    n = 20
    A = np.diag(np.linspace(1.0, 3.0, n))
    A[:2, :2] = [[0.05, 0.04], [-0.04, 0.05]]      # eigenvalues 0.05 +- 0.04i
    X = np.linalg.qr(rng.standard_normal((n, n)))[0]
    B = X @ A @ X.T
    V, H = _arnoldi(B, rng.standard_normal(n), n)
    G, D = lowrank_correction(H, n, 1)
    assert G.shape[1] == 2                           # the pair is kept whole
    U = V[:, :n] @ G
    assert np.allclose(X[:, :2] @ (X[:, :2].T @ U), U, atol=1e-8)
    Q = np.eye(n) + U @ D @ U.T
    assert np.allclose(B @ Q @ U, U, atol=1e-8)


def test_singular_or_tiny_cycle_gives_none():
    assert lowrank_correction(np.zeros((3, 2)), 2, 1) is None
    assert lowrank_correction(np.ones((2, 1)), 1, 1) is None


DH = 0.3 * (1.0 - 1.0 / np.sqrt(2.0))


def _cfg1_system(dh=DH):
    from dac_vtsa_amd import device as dv
    from dac_vtsa_amd.config import load_config
    from dac_vtsa_amd.core.grid import Grid
    from dac_vtsa_amd.core.jacobian import FDJacobian
    from dac_vtsa_amd.core.residual import TVSAModel
    from dac_vtsa_amd.core.solver import DeviceOps
    from dac_vtsa_amd.properties import get_backend
    from tests._helpers import CFG1, golden, state_from
    z = golden("integ_cfg1.npz")
    y = z["sdirk2/snap_y"][2]
    t = float(z["sdirk2/snap_t"][2])
    cfg = load_config(override_dict=CFG1)
    m = TVSAModel(cfg, Grid.from_config(cfg), get_backend(cfg), cfg.cycle[0], state_from(z))
    ops = DeviceOps(m)
    yd = dv.to_dev(y)
    f = dv.empty(m.n_core)
    m.rhs_dev(t, yd, f)
    FDJacobian(m).evaluate(t, yd, f)
    s = ops.scale(yd, 1e-4)
    ops.precond_setup(dh, s, 2, 0, 2)
    b = dv.to_dev(dh * dv.to_host(f) / dv.to_host(s))
    return m, ops, t, yd, f, s, b


@pytest.mark.gpu
def test_device_apply_matches_host_formula():
    """Q v = v + u d (u . v) with u = b/|b|, the first basis vector of the last cycle (m = 1)."""
    from dac_vtsa_amd import device as dv
    m, ops, t, yd, f, s, b = _cfg1_system()
    w = dv.empty(m.n_core)
    ops.gmres(t, yd, f, s, DH, None, b, w, 1e-2, 0.0, 30)
    bh = dv.to_host(b)
    u = bh / np.linalg.norm(bh)
    ops.lowrank_set(np.array([[1.0]]), np.array([[0.75]]))
    v = np.random.default_rng(0).standard_normal(m.n_core)
    out = dv.to_host(ops.lowrank_apply(dv.to_dev(v), dv.empty(m.n_core)))
    ref = v + u * 0.75 * (u @ v)
    assert np.max(np.abs(out - ref)) < 1e-13 * np.max(np.abs(ref))
    # in place, and removal
    vd = dv.to_dev(v)
    ops.lowrank_apply(vd, vd)
    assert np.max(np.abs(dv.to_host(vd) - ref)) < 1e-13 * np.max(np.abs(ref))
    ops.lowrank_clear()
    assert np.array_equal(dv.to_host(ops.lowrank_apply(dv.to_dev(v), dv.empty(m.n_core))), v)
    m.close()


@pytest.mark.gpu
def test_gmres_with_the_correction_takes_fewer_iterations():
    """One Newton system (h = 30 s) solved to 1e-5 relative twice: plainly, then with the rank-8
    correction built from the first cycle.  GMRES's residual estimate meets the target both times
    and the corrected cycle takes fewer iterations (measured 8 -> 6).  The true residual after the
    cycle stays above the target on this state whatever the preconditioner: the finite-difference
    matvec is only linear to ~1e-2 here (limiter and upwind switches), so info is 1 for both.  A new
    factorisation removes the correction: the plain cycle is then reproduced bit for bit."""
    from dac_vtsa_amd import device as dv
    dh = DH * 100.0
    m, ops, t, yd, f, s, b = _cfg1_system(dh)
    w1, w2 = dv.empty(m.n_core), dv.empty(m.n_core)
    i1 = ops.gmres(t, yd, f, s, dh, None, b, w1, 1e-5, 0.0, 60)
    assert i1.presid <= i1.atol_eff and i1.iters >= 6, (i1.presid, i1.atol_eff, i1.iters)
    assert ops.lowrank_from_last_cycle(8)
    i2 = ops.gmres(t, yd, f, s, dh, None, b, w2, 1e-5, 0.0, 60)
    assert i2.presid <= i2.atol_eff and i2.iters < i1.iters, (i1.iters, i2.iters, i2.presid)
    # the corrected solution is as good through the same operator
    assert i2.rnorm < 1.5 * i1.rnorm, (i1.rnorm, i2.rnorm)
    ops.precond_setup(dh, s, 2, 0, 2)
    assert not ops.lowrank_on
    w3 = dv.empty(m.n_core)
    i3 = ops.gmres(t, yd, f, s, dh, None, b, w3, 1e-5, 0.0, 60)
    assert (i3.info, i3.iters) == (i1.info, i1.iters)
    assert np.array_equal(dv.to_host(w3), dv.to_host(w1))
    m.close()

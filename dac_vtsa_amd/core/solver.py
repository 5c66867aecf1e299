"""Implicit two-stage integrator with Newton-Krylov stages on the GPU (drop-in for upstream
core/solver.py).

The control flow -- SDIRK2 (default) / TR-BDF2 stages, inexact Newton with Armijo line search,
conservation polish, Hairer-Wanner early exit, M^-1-filtered error estimate, Newton-aware step
control, accumulator quadrature -- restates core/solver.py:47-511 decision by decision.  What
changes is where the work runs: every vector (states, residuals, Krylov basis) stays in HBM and
every vector operation is a HIP kernel of libvtsa; the host only sees the scalars the decisions
need.  The upstream SuperLU factorisation of I - D h J is replaced by the device z-line (or
block-Jacobi) preconditioner, in FP64 or FP32, and ``lu.solve`` by its application.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np

from .. import _abi
from .. import device as dv
from .jacobian import FDJacobian

GAMMA = 2.0 - np.sqrt(2.0)
D = GAMMA / 2.0
C1 = 1.0 / (GAMMA * (2.0 - GAMMA))
C2 = (1.0 - GAMMA) ** 2 / (GAMMA * (2.0 - GAMMA))
W1 = (3.0 * GAMMA - 1.0) / (6.0 * GAMMA)
W2 = 1.0 / (6.0 * GAMMA * (1.0 - GAMMA))
W3 = (2.0 - 3.0 * GAMMA) / (6.0 * (1.0 - GAMMA))
SQRT_EPS = np.sqrt(np.finfo(float).eps)
G_SD = 1.0 - 1.0 / np.sqrt(2.0)
SW1 = (3.0 * G_SD - 1.0) / (6.0 * G_SD)
SW2 = 1.0 / (6.0 * G_SD * (1.0 - G_SD))
SW3 = (2.0 - 3.0 * G_SD) / (6.0 * (1.0 - G_SD))

PRECONDITIONERS = {"block_jacobi": 0, "zline": 1, "zline_gs": 2}
PRECISIONS = {"fp64": 0, "fp32": 1}
MAX_RESTART = 64      # columns of the device Hessenberg (vtsa_gmres_hessenberg)
GUESS_MISS_LIMIT = 32  # learned guess: stages in a row lost to the predictor before it is tried less often
GUESS_RETRY_EVERY = 8  # ... and then tried on every 8th stage
LOWRANK_MAX = 16      # largest rank of the device low-rank correction


def lowrank_correction(H, m, k, theta_min=0.0, theta_max=np.inf):
    """Low-rank right-preconditioner correction from one Arnoldi cycle (no upstream counterpart).

    H: the (m+1) x m Arnoldi Hessenberg of the right-preconditioned operator B on the basis V_m.
    The harmonic Ritz pairs of B on span V_m solve (H_m + h^2 H_m^-T e_m e_m^T) g = theta g with
    H_m the square part and h = H[m, m-1] (Morgan, SIAM J. Matrix Anal. Appl. 16 (1995) 1154).
    The k values of smallest modulus within [theta_min, theta_max] (a complex pair is taken
    whole) give a real orthonormal G (m x kk, kk <= k + 1) spanning their vectors; with
    U = V_m G the correction is Q = I + U D U^T, D = T^-1 - I, T = G^T H_m G = U^T B U, so that B Q has eigenvalues near 1
    on span U.  Q stretches span U by up to 1/theta_min.  Returns (G, D), or None when no value
    qualifies or H_m or T is singular to working precision."""
    if m < 2 or k < 1:
        return None
    Hm = np.asarray(H[:m, :m], dtype=float)
    hl = float(H[m, m - 1])
    em = np.zeros(m)
    em[-1] = 1.0
    try:
        F = Hm + hl * hl * np.outer(np.linalg.solve(Hm.T, em), em)
        theta, vecs = np.linalg.eig(F)
    except np.linalg.LinAlgError:
        return None
    if not np.all(np.isfinite(theta)):
        return None
    mod = np.abs(theta)
    order = [i for i in np.argsort(mod, kind="stable") if theta_min <= mod[i] <= theta_max]
    if not order:
        return None
    cols, taken = [], set()
    for i in order:
        if len(cols) >= k:
            break
        if i in taken:
            continue
        taken.add(i)
        v = vecs[:, i]
        if abs(theta[i].imag) > 1e-12 * abs(theta[i]):
            # its conjugate partner spans the same real plane
            j = min((j for j in order if j not in taken), key=lambda j: abs(theta[j] - np.conj(theta[i])),
                    default=None)
            if j is not None:
                taken.add(j)
            cols += [v.real, v.imag]
        else:
            cols.append(v.real)
    if len(cols) > LOWRANK_MAX:
        cols = cols[:LOWRANK_MAX]
    G, R = np.linalg.qr(np.stack(cols, axis=1))
    sv = np.abs(np.diag(R))
    G = G[:, sv > 1e-10 * sv.max()]
    T = G.T @ Hm @ G
    if np.linalg.cond(T) > 1e12:
        return None
    return G, np.linalg.inv(T) - np.eye(G.shape[1])


class SolverError(RuntimeError):
    pass


@dataclass
class IntegrationResult:
    y: np.ndarray
    acc: np.ndarray
    t: float
    n_steps: int = 0
    n_reject: int = 0
    n_newton_fail: int = 0
    n_krylov: int = 0
    n_jac: int = 0
    n_lu: int = 0
    nfev: int = 0
    wall_time: float = 0.0
    snapshots: List = field(default_factory=list)       # (t, y, acc)
    stopped_early: bool = False
    stop_value: Optional[float] = None
    steps_taken: List = field(default_factory=list)     # accepted (t, h)


class DeviceOps:
    """Thin wrappers of the libvtsa Newton entry points for one model handle."""

    def __init__(self, model):
        self.m = model
        self.h = model.handle
        self.N = model.n_core
        # low-rank correction: rank asked for (0 off), the smallest cycle it is built from, and
        # whether the device holds one for the current factorisation
        self.lowrank = 0
        self.lowrank_window = (0.0, np.inf)
        self.lowrank_min_iters = 8
        self.lowrank_on = False
        self.lowrank_held = False     # dropped until the next factorisation
        self.lowrank_built = 0
        self.lowrank_dropped = 0
        self.gmres_iters_corrected = 0   # GMRES iterations of cycles run with / without the correction
        self.gmres_iters_plain = 0

    def st(self):
        return dv.stream()

    def scale(self, z, rtol, out=None):
        out = dv.empty(self.N) if out is None else out
        _abi.call("vtsa_scale", self.h, dv.ptr(z), float(rtol), dv.ptr(out), self.st())
        return out

    def rms(self, a, s=None):
        r = C.c_double()
        _abi.call("vtsa_rms", self.h, dv.ptr(a), dv.ptr(s), self.N, C.byref(r), self.st())
        return r.value

    def vec(self, op, out, x=None, y=None, z=None, w=None, a=0.0, b=0.0, c=0.0, e=0.0):
        _abi.call("vtsa_vec_op", self.h, op, float(a), float(b), float(c), float(e), dv.ptr(x), dv.ptr(y),
                  dv.ptr(z), dv.ptr(w), dv.ptr(out), self.st())
        return out

    def newton_eval(self, t, z, dh, rk, s, f, R):
        rn, pn = C.c_double(), C.c_double()
        _abi.call("vtsa_newton_eval", self.h, float(t), dv.ptr(z), float(dh), dv.ptr(rk), dv.ptr(s), dv.ptr(f),
                  dv.ptr(R), C.byref(rn), C.byref(pn), self.st())
        self.m.nfev += 1
        return rn.value, pn.value

    def newton_residual(self, t, z, dh, rk, s, f, R, norm=True):
        """f and R at z as newton_eval, without the preconditioned norm; rn (waited for) only when
        norm is true, else None and nothing is waited for."""
        rn = C.c_double()
        _abi.call("vtsa_newton_residual", self.h, float(t), dv.ptr(z), float(dh), dv.ptr(rk), dv.ptr(s),
                  dv.ptr(f), dv.ptr(R), C.byref(rn) if norm else None, self.st())
        self.m.nfev += 1
        return rn.value if norm else None

    def newton_try(self, t, z, dz, lam, dh, rk, s, zt, ft, Rt):
        rn, pn, step = C.c_double(), C.c_double(), C.c_double()
        _abi.call("vtsa_newton_try", self.h, float(t), dv.ptr(z), dv.ptr(dz), float(lam), float(dh), dv.ptr(rk),
                  dv.ptr(s), dv.ptr(zt), dv.ptr(ft), dv.ptr(Rt), C.byref(rn), C.byref(pn), C.byref(step), self.st())
        self.m.nfev += 1
        return rn.value, pn.value, step.value

    def gmres(self, t, z, f, s, dh, zs_norm, b, w, rtol, atol, restart, pw=None):
        """One GMRES cycle; zs_norm None: max(1, rms(z/s)) formed on the device.  ``pw`` (optional
        device vector): receives M^-1 (w*s), the step upstream forms with one more solve
        (core/solver.py:240-241), from the apply of the cycle's closing residual."""
        zs_norm = -1.0 if zs_norm is None else zs_norm
        info = _abi.GmresInfo()
        _abi.call("vtsa_gmres", self.h, float(t), dv.ptr(z), dv.ptr(f), dv.ptr(s), float(dh), float(zs_norm),
                  dv.ptr(b), dv.ptr(w), dv.ptr(pw) if pw is not None else None, float(rtol), float(atol),
                  int(restart), C.byref(info), self.st())
        self.m.nfev += info.iters + 1
        if self.lowrank_on:
            self.gmres_iters_corrected += info.iters
        else:
            self.gmres_iters_plain += info.iters
        if self.lowrank and not (self.lowrank_on or self.lowrank_held) and info.iters >= self.lowrank_min_iters:
            self.lowrank_from_last_cycle(self.lowrank)
        return info

    def hessenberg(self):
        """(H, m): the Arnoldi Hessenberg ((m+1) x m) of the last GMRES cycle."""
        buf = (C.c_double * ((MAX_RESTART + 1) * MAX_RESTART))()
        m = C.c_int32()
        _abi.call("vtsa_gmres_hessenberg", self.h, buf, C.byref(m))
        H = np.frombuffer(buf, dtype=float).reshape(MAX_RESTART + 1, MAX_RESTART)
        return H[:m.value + 1, :m.value].copy(), m.value

    def lowrank_set(self, G, Dm):
        """Install Q = I + (V_m G) Dm (V_m G)^T, V_m the basis of the last GMRES cycle."""
        m, k = G.shape
        Gc = (C.c_double * (m * k))(*np.ascontiguousarray(G, dtype=float).ravel())
        Dc = (C.c_double * (k * k))(*np.ascontiguousarray(Dm, dtype=float).ravel())
        _abi.call("vtsa_defl_set", self.h, int(k), int(m), Gc, Dc, self.st())
        self.lowrank_on = True
        self.lowrank_built += 1

    def lowrank_clear(self):
        _abi.call("vtsa_defl_set", self.h, 0, 0, None, None, self.st())
        self.lowrank_on = False

    def lowrank_drop(self):
        """Remove the correction and build none until the next factorisation."""
        self.lowrank_clear()
        self.lowrank_held = True
        self.lowrank_dropped += 1

    def lowrank_from_last_cycle(self, k):
        """Build the rank-k correction from the last GMRES cycle's Hessenberg and basis; returns
        whether one was installed."""
        H, m = self.hessenberg()
        gd = lowrank_correction(H, m, k, *self.lowrank_window)
        if gd is None:
            return False
        self.lowrank_set(*gd)
        return True

    def lowrank_apply(self, v, out):
        _abi.call("vtsa_defl_apply", self.h, dv.ptr(v), dv.ptr(out), self.st())
        return out

    def prec_apply(self, r, out):
        _abi.call("vtsa_precond_apply", self.h, dv.ptr(r), dv.ptr(out), self.st())
        return out

    def cons(self, R, s):
        wr, ws = (C.c_double * 4)(), (C.c_double * 4)()
        _abi.call("vtsa_cons", self.h, dv.ptr(R), dv.ptr(s), wr, ws, self.st())
        return np.array(wr[:]), np.array(ws[:])

    def cons_stash(self, R, s, slot):
        """W R and W s as cons, kept on the device in slot 0..3 (not waited for)."""
        _abi.call("vtsa_cons_stash", self.h, dv.ptr(R), dv.ptr(s), int(slot), self.st())

    def cons_fetch(self, n):
        """The stashed (W R, W s) of slots 0..n-1, as two (n, 4) arrays."""
        wr, ws = (C.c_double * (4 * n))(), (C.c_double * (4 * n))()
        _abi.call("vtsa_cons_fetch", self.h, int(n), wr, ws, self.st())
        return np.array(wr[:]).reshape(n, 4), np.array(ws[:]).reshape(n, 4)

    def cons_norm2(self, s):
        out = (C.c_double * 4)()
        _abi.call("vtsa_cons_norm2", self.h, dv.ptr(s), out, self.st())
        return np.array(out[:])

    def cons_shift(self, z, s, lam, out, truncate=False, report=True):
        """out = z - S^2 W^T lam.  Returns True when a concentration or loading would fall more
        than one tolerance unit below zero (with truncate=True those keep 0.1 of their value;
        truncate="zero": every one that would fall below zero does).  report=False: the plain
        shift, not waited for; returns False."""
        lam = (C.c_double * 4)(*[float(v) for v in lam])
        mode = (_abi.SHIFT_TRUNCATE if truncate else 0) | (_abi.SHIFT_AT_ZERO if truncate == "zero" else 0) | \
            (0 if report else _abi.SHIFT_NO_REPORT)
        rc = _abi.call("vtsa_cons_shift", self.h, dv.ptr(z), dv.ptr(s), lam, mode, dv.ptr(out), self.st())
        return rc == 1

    def error_norm(self, y, z2, y3, rtol):
        en = C.c_double()
        _abi.call("vtsa_error_norm", self.h, dv.ptr(y), dv.ptr(z2), dv.ptr(y3), float(rtol), C.byref(en), self.st())
        return en.value

    def precond_setup(self, dh, scale, kind, precision, sweeps=1):
        """New preconditioner factors; removes the low-rank correction of the previous ones."""
        _abi.call("vtsa_precond_setup", self.h, float(dh), dv.ptr(scale), int(kind), int(precision), int(sweeps),
                  self.st())
        self.lowrank_on = self.lowrank_held = False


def _finite(x):
    return math.isfinite(x)


def cons_ok(ops, cons_tol, R, s):
    """|W R| <= tol W s for every conserved quantity (core/solver.py:88-96)."""
    if cons_tol <= 0.0:
        return True
    wr, ws = ops.cons(R, s)
    return bool(np.all(np.abs(wr) <= cons_tol * ws))


class Factorisation:
    """Token for the handle's current preconditioner factorisation (the upstream ``lu``)."""

    def __init__(self, h):
        self.h = h


def project_conservation(ops, t, z, f, R, s, dh, rhs_known, cons_tol, iters=3, max_shift=1.0, strict=False):
    """Shift a converged stage value along S^2 W^T so that its residual conserves the four totals.

    No upstream counterpart (``conserve="project"``).  The shift z(lambda) = z - S^2 W^T lambda is
    the minimum-norm change in the s-scaled variables that moves the conserved totals W z; the
    four multipliers solve W R(z(lambda)) = 0 by Newton's method on this 4x4 system, its matrix
    B = d(W R)/d(lambda) taken by finite differences along the four directions (four residual
    evaluations per iteration; rows with no entries are left out).  When the shift would take a
    concentration or loading more than one tolerance unit below zero, those keep 0.1 of their
    value instead (the integrator's truncation rule).  The projection is
    accepted only when the totals close and the shift stays within max_shift tolerance units
    (rms of the change over s); otherwise None is returned and the caller falls back to the
    upstream polish.  strict: no shifted concentration or loading may go below zero (those keep 0.1
    of their value), at some cost in closure."""
    d2 = ops.cons_norm2(s)
    rows = np.nonzero(d2 > 0.0)[0]
    zc, fc, Rc = z, f, R
    zt, ft, Rt = dv.empty(ops.N), dv.empty(ops.N), dv.empty(ops.N)
    for _ in range(iters):
        wr, ws = ops.cons(Rc, s)
        if np.all(np.abs(wr) <= cons_tol * ws):
            return zc, fc, Rc
        B = np.zeros((4, 4))
        eps = 1e-2 * ws / np.where(d2 > 0.0, d2, 1.0)
        for q, l in enumerate(rows):     # the columns' residuals queued back to back, one wait
            lam = np.zeros(4)
            lam[l] = eps[l]
            ops.cons_shift(zc, s, lam, zt, report=False)
            ops.newton_residual(t, zt, dh, rhs_known, s, ft, Rt, norm=False)
            ops.cons_stash(Rt, s, q)
        wls, _ = ops.cons_fetch(len(rows))
        if not np.all(np.isfinite(wls)):
            return None
        for q, l in enumerate(rows):
            B[:, l] = (wls[q] - wr) / eps[l]
        lam = np.zeros(4)
        try:
            lam[rows] = np.linalg.solve(B[np.ix_(rows, rows)], -wr[rows])
        except np.linalg.LinAlgError:
            return None
        zn, fn, Rn = dv.empty(ops.N), dv.empty(ops.N), dv.empty(ops.N)
        # entries that would go negative are truncated as a Newton update does (an entry that
        # would not is left as the plain shift leaves it, so no report is needed)
        ops.cons_shift(zc, s, lam, zn, "zero" if strict else True, report=False)
        rn = ops.newton_residual(t, zn, dh, rhs_known, s, fn, Rn)
        if not _finite(rn):
            return None
        zc, fc, Rc = zn, fn, Rn
        if _shift(ops, z, zc, s, zt) > max_shift:
            return None
    wr, ws = ops.cons(Rc, s)
    return (zc, fc, Rc) if np.all(np.abs(wr) <= cons_tol * ws) else None


def _shift(ops, z0, z1, s, tmp):
    """rms((z1 - z0) / s): the size of a change in tolerance units."""
    ops.vec(7, tmp, x=z0, y=z1)
    return ops.rms(tmp, s)


def _unclosed(stats, z0, best, it, rate, new_lu, reject):
    """A converged stage whose balance stayed open: accepted as upstream (balance_unclosed), or,
    with reject, reported as a failed solve so that the step is retried with a smaller h
    (unclosed_rejected; no upstream counterpart)."""
    key = "unclosed_rejected" if reject else "balance_unclosed"
    if stats is not None:
        stats[key] = stats.get(key, 0) + 1
    return (z0, False, it + 1, rate, new_lu) if reject else (best[1], True, it + 1, rate, new_lu)


def newton_krylov_dev(ops, t, z0, rhs_known, dh, lu, rtol, max_iter, kappa, krylov_dim=30, eta=1e-2, counter=None,
                      cons_tol=0.0, stats=None, trace=None, refresh=None, project=False, strict=False,
                      reject_unclosed=False):
    """Inexact Newton-Krylov stage solve on the device (core/solver.py:154-294).

    ``project`` (no upstream counterpart): when the converged value misses the conservation
    test, try project_conservation before the upstream polish iterations.

    z0, rhs_known: device vectors.  Returns (z, ok, iterations, rate, new_lu)."""
    N = ops.N
    z = z0.clone()
    f, R, s = dv.empty(N), dv.empty(N), ops.scale(z, rtol)
    rn, pn = ops.newton_eval(t, z, dh, rhs_known, s, f, R)
    if not _finite(rn):
        return z0, False, 0, np.inf, None
    b, w, dz = dv.empty(N), dv.empty(N), dv.empty(N)
    zt, ft, Rt = dv.empty(N), dv.empty(N), dv.empty(N)
    step_prev = None
    polish = 0
    best = None
    new_lu = None
    rate = 0.0
    it = -1
    while it + 1 < max_iter + polish:
        it += 1
        ops.scale(z, rtol, s)
        ops.vec(8, b, x=R, y=s)
        zs_norm = None   # max(1, rms(z/s)), formed on the device by the GMRES cycle
        floor = (1e-4 if polish else 0.1 * kappa) * np.sqrt(N)
        info = ops.gmres(t, z, f, s, dh, zs_norm, b, w, (1e-3 if polish else eta), floor, krylov_dim, pw=dz)
        if polish and stats is not None:
            stats["gmres_polish"] = stats.get("gmres_polish", 0) + info.iters
        elif counter is not None:
            counter[0] += info.iters
        if info.info != 0 and refresh is not None and new_lu is None and not polish:
            lu = new_lu = refresh(z)
            rn, pn = ops.newton_eval(t, z, dh, rhs_known, s, f, R)
            if stats is not None:
                stats["jac_refresh_in_newton"] = stats.get("jac_refresh_in_newton", 0) + 1
            info = ops.gmres(t, z, f, s, dh, zs_norm, b, w, eta, floor, krylov_dim, pw=dz)
            if counter is not None:
                counter[0] += info.iters
        # dz = M^-1 (w*s) (core/solver.py:240-241) came with the GMRES cycle
        lam = 1.0
        ls = 0
        while ls < 4:
            rn_try, pn_try, step = ops.newton_try(t, z, dz, lam, dh, rhs_known, s, zt, ft, Rt)
            if math.isnan(step):
                # dz holds a non-finite entry (core/solver.py:241-242); that test evaluates no residual
                ops.m.nfev -= 1
                return z0, False, it + 1, np.inf, new_lu
            if not _finite(rn_try):
                pn_try = np.inf
            if _finite(pn_try) and (pn_try <= (1.0 - 1e-4 * lam) * pn or ls == 3):
                break
            if ls == 0 and ops.lowrank_on and not polish:
                # (no upstream counterpart) the full step of a direction found with the low-rank
                # correction fails the sufficient-decrease test: the correction is dropped until
                # the next factorisation and the direction is found again without it
                ops.lowrank_drop()
                info = ops.gmres(t, z, f, s, dh, zs_norm, b, w, eta, floor, krylov_dim, pw=dz)
                if counter is not None:
                    counter[0] += info.iters
                continue
            lam *= 0.5
            ls += 1
        if not _finite(pn_try):
            return z0, False, it + 1, np.inf, new_lu
        if trace is not None:
            trace.append({"it": it, "polish": polish, "gmres_info": info.info, "gmres_iters": info.iters,
                          "lam": lam, "step": step, "res_before": rn, "res_after": rn_try, "pres_before": pn,
                          "pres_after": pn_try})
        z, zt = zt, z
        f, ft = ft, f
        R, Rt = Rt, R
        rn, pn = rn_try, pn_try
        rate = 0.0 if step_prev is None else step / max(step_prev, 1e-300)
        converged = ((lam == 1.0 and step <= kappa and (info.info == 0 or step_prev is not None))
                     or (step_prev is not None and lam == 1.0 and info.info == 0 and rate < 1.0
                         and rate / (1.0 - rate) * step <= kappa)
                     or pn <= kappa)
        if converged:
            if best is None or rn < best[0]:
                best = (rn, z.clone())
            if cons_ok(ops, cons_tol, R, ops.scale(z, rtol, s)):
                return z, True, it + 1, rate, new_lu
            if project and cons_tol > 0.0:
                pr = project_conservation(ops, t, z, f, R, s, dh, rhs_known, cons_tol, strict=strict)
                if pr is not None:
                    if stats is not None:
                        stats["cons_projected"] = stats.get("cons_projected", 0) + 1
                    return pr[0], True, it + 1, rate, new_lu
            if polish < 2:
                polish += 1
                step_prev = None
                continue
            return _unclosed(stats, z0, best, it, rate, new_lu, reject_unclosed)
        if best is not None:
            return _unclosed(stats, z0, best, it, rate, new_lu, reject_unclosed)
        if step_prev is not None and lam == 1.0 and info.info == 0:
            left = max_iter - it - 1
            if it >= 2 and (rate >= 1.0 or rate ** left / (1.0 - rate) * step > kappa * 1e3):
                return z, False, it + 1, rate, new_lu
        step_prev = step
    return z, False, it + 1, 1.0, new_lu


def newton_modified_dev(ops, t, z0, rhs_known, dh, rtol, max_iter, kappa, fail_rate=0.9, cons_tol=0.0):
    """Modified Newton on the device with the handle's fixed preconditioner in the place of the
    upstream stored LU (core/solver.py:111-151): z <- trunc(z - M^-1 R(z)), the step norm
    rms((z_new - z_old) / s(z_new)), no acceptance on the first iteration, a contraction test
    against fail_rate, convergence when rate/(1 - rate) |dz| <= kappa, and, with a conservation
    tolerance, one more residual to test the balances before returning.

    z0, rhs_known: device vectors.  Returns (z, ok, iterations, rate)."""
    N = ops.N
    z = z0.clone()
    f, R, s, r = dv.empty(N), dv.empty(N), dv.empty(N), dv.empty(N)
    zn, fn, Rn, diff, sn = dv.empty(N), dv.empty(N), dv.empty(N), dv.empty(N), dv.empty(N)
    ops.scale(z, rtol, s)
    rn = ops.newton_residual(t, z, dh, rhs_known, s, f, R)
    nrm_prev = None
    rate = 0.0
    done = False
    for it in range(max_iter + 2):
        if not _finite(rn):
            return z0, False, it, np.inf
        if done:
            if cons_ok(ops, cons_tol, R, ops.scale(z, rtol, s)) or it >= max_iter + 1:
                return z, True, it, rate
        elif it >= max_iter:
            break
        ops.prec_apply(R, r)                                # M^-1 R; the update is -M^-1 R
        rn_new, _, _ = ops.newton_try(t, z, r, -1.0, dh, rhs_known, s, zn, fn, Rn)
        ops.vec(7, diff, x=z, y=zn)
        ops.scale(zn, rtol, sn)
        nrm = ops.rms(diff, sn)
        z, zn = zn, z
        f, fn = fn, f
        R, Rn = Rn, R
        s, sn = sn, s
        # the trial point's residual was formed with the previous scale; the next test uses R
        # itself, and its norm only for the finiteness test of the next iteration
        rn = rn_new
        if not _finite(nrm):
            return z0, False, it + 1, np.inf
        if done:
            continue
        if nrm_prev is not None:
            rate = nrm / max(nrm_prev, 1e-300)
            if rate >= fail_rate:
                return z, False, it + 1, rate
            if rate / (1.0 - rate) * nrm <= kappa:
                if cons_tol <= 0.0:
                    return z, True, it + 1, rate
                done = True
        elif nrm < 1e-8:
            if cons_tol <= 0.0:
                return z, True, it + 1, 0.0
            done = True
        nrm_prev = nrm
    return z, False, max_iter, rate


def _upstream_args(model, lu, scale_fn, pos, cons):
    """DeviceOps and rtol for the upstream-signature stage solvers, checking that what the caller
    passes is what the device computes: scale_fn(y) = atol + rtol |y| with the model's atol, the
    positivity set pos = all concentrations and loadings, cons = the model's conservation matrix,
    and lu the handle's current device factorisation."""
    if not isinstance(lu, Factorisation):
        raise TypeError("lu must be the Factorisation of this model's device preconditioner (TRBDF2._factor)")
    N = model.n_core
    ones = np.ones(N)
    s0, s1, s2 = (np.asarray(scale_fn(a * ones), dtype=float) for a in (0.0, 1.0, 2.0))
    owner = getattr(scale_fn, "__self__", None)      # an integrator's bound ``scale`` carries rtol
    k = int(np.argmin(model.atol))
    rtol = float(owner.rtol) if hasattr(owner, "rtol") else float(s1[k] - s0[k])
    # (atol + rtol) - atol loses digits where atol is large, hence the 1e-6 tolerances
    if not (np.array_equal(s0, model.atol) and rtol > 0.0
            and np.allclose(s1 - s0, rtol, rtol=1e-6, atol=0.0) and np.allclose(s2 - s1, rtol, rtol=1e-6, atol=0.0)):
        raise ValueError("scale_fn must be atol + rtol |y| with the model's atol (the device forms it)")
    if pos is not None and not np.array_equal(np.asarray(pos), np.arange(model.off_e)):
        raise ValueError("the device truncates exactly the concentrations and loadings (pos = arange(off_e))")
    if cons is not None:
        W = model.conservation_matrix()
        if cons.shape != W.shape or abs(cons - W).max() > 0.0:
            raise ValueError("cons must be the model's conservation_matrix() (the device applies its own)")
    return DeviceOps(model), rtol


def newton_krylov(model, t, z0, rhs_known, dh, lu, scale_fn, max_iter, kappa, krylov_dim=30, eta=1e-2,
                  counter=None, pos=None, cons=None, cons_tol=0.0, stats=None, trace=None, refresh=None):
    """Upstream signature (core/solver.py:154-156): NumPy vectors in and out, the stage solved on the
    device by newton_krylov_dev.  ``refresh(z)`` receives the iterate as a device vector and
    returns the new Factorisation.  Returns (z, ok, iterations, rate, new_lu_or_None)."""
    ops, rtol = _upstream_args(model, lu, scale_fn, pos, cons)
    z, ok, it, rate, new_lu = newton_krylov_dev(ops, t, dv.to_dev(np.asarray(z0, dtype=float)),
                                                dv.to_dev(np.asarray(rhs_known, dtype=float)), dh, lu, rtol, max_iter,
                                                kappa, krylov_dim=krylov_dim, eta=eta, counter=counter,
                                                cons_tol=cons_tol if cons is not None else 0.0, stats=stats,
                                                trace=trace, refresh=refresh)
    return dv.to_host(z).copy(), ok, it, rate, new_lu


def newton_modified(model, t, z0, rhs_known, dh, lu, scale_fn, max_iter, kappa, fail_rate=0.9, pos=None, cons=None,
                    cons_tol=0.0):
    """Upstream signature (core/solver.py:111-112) over newton_modified_dev.  Returns (z, ok,
    iterations, rate)."""
    ops, rtol = _upstream_args(model, lu, scale_fn, pos, cons)
    z, ok, it, rate = newton_modified_dev(ops, t, dv.to_dev(np.asarray(z0, dtype=float)),
                                          dv.to_dev(np.asarray(rhs_known, dtype=float)), dh, rtol, max_iter, kappa,
                                          fail_rate=fail_rate, cons_tol=cons_tol if cons is not None else 0.0)
    return dv.to_host(z).copy(), ok, it, rate


class TRBDF2:
    """Two-stage L-stable implicit integrator (core/solver.py:297-511) on the GPU.

    Extra keyword arguments (no upstream counterpart): ``precond`` ("zline_gs" | "zline" |
    "block_jacobi"), ``gs_sweeps`` (symmetric line Gauss-Seidel sweeps of "zline_gs"),
    ``gs_order`` ("symmetric": colours 0 1 0 1 0 for two sweeps; "forward": 0 1 0 1),
    ``precision`` ("fp64" | "fp32") of the preconditioner factors, ``krylov_dim``, ``conserve``
    ("polish": the upstream conservation polish; "project": project_conservation first;
    "project_strict": the same with no shifted entry below zero; with either projection a stage
    whose balance stays open fails its step, which is retried with a smaller h),
    ``guess`` (a learned.StageNet: each stage starts from the network's guess when its
    preconditioned residual is below the predictor's), ``record`` (a learned.StageDataset
    that receives the features and converged corrections of every solved stage), and the parity
    mode: ``h_sequence`` (the accepted step end times of another run, e.g. the upstream
    integrator's: each step ends at the next of them, every step that converges is accepted, a
    stage that fails with a fresh Jacobian raises SolverError) and ``error_filter`` ("precond":
    the local error filtered by the preconditioner, standing in for the upstream LU solve;
    "exact": by a GMRES(64) solve of (I - D h J) e = z2 - y3 to 1e-10 relative); ``kappa``
    overrides solver.newton_kappa (whose configured minimum is 1e-4) for such comparisons.
    ``lowrank`` (k > 0): after each new factorisation, the first GMRES cycle of at least 8
    iterations yields a rank-k correction of the preconditioner (lowrank_correction) from its
    harmonic Ritz values of modulus within ``lowrank_window``, used by every later cycle until the
    next factorisation.  The lower bound keeps the correction's stretch of the Krylov vectors
    (up to 1/theta) moderate: the finite-difference matvec of a vector stretched by 1e4 or more
    loses the true-residual test and forces Jacobian refreshes.  The
    preconditioner stands in for the upstream SuperLU factor of the theta-block matrix wherever
    that factor is used: GMRES, the Newton residual norm and the error-estimate filter."""

    def __init__(self, model, cfg, h_max: float, h_init: Optional[float] = None, logger=None,
                 precond: str = "zline_gs", precision: str = "fp64", krylov_dim: int = 30,
                 gs_sweeps: int = 2, conserve: str = "polish", guess=None, record=None, h_sequence=None,
                 error_filter: str = "precond", kappa: Optional[float] = None, lowrank: int = 0,
                 lowrank_window=(1e-3, 0.6), gs_order: str = "symmetric"):
        self.m = model
        s = cfg.solver
        self.rtol = s.rtol
        self.atol = model.atol
        self.h_max = h_max
        self.h_init = h_init if h_init is not None else s.h_initial
        self.max_iter = s.max_newton_iter
        self.kappa = s.newton_kappa if kappa is None else float(kappa)
        self.jac_max_age = s.jacobian_max_age
        self.max_steps = s.max_steps_per_step
        self.jac = FDJacobian(model)
        self.log = logger or (lambda *a, **k: None)
        self.stats = {"modified_fail": 0, "krylov_ok": 0, "krylov_fail": 0}
        self._gm = [0]
        self.scheme = s.scheme
        if s.newton_mode not in ("krylov", "hybrid", "modified"):
            raise ValueError(f"solver.newton_mode = {s.newton_mode!r}")
        self.mode = s.newton_mode
        if self.mode == "modified":
            # the device preconditioner is never an exact factorisation of I - D h J, so the upstream
            # rule for an approximate (theta-block) factor applies: modified Newton with the
            # Krylov fallback (core/solver.py:328-329)
            self.mode = "hybrid"
        self.gmres_refresh = s.gmres_refresh
        self.log_refactor = np.log(s.refactor_ratio)
        self.cons_tol = float(s.conservation_tol or 0.0)
        self.newton_refresh = True
        self.precond_kind = PRECONDITIONERS[precond]
        self.precision = PRECISIONS[precision]
        self.krylov_dim = krylov_dim
        self.gs_sweeps = int(gs_sweeps)
        if gs_order not in ("symmetric", "forward"):
            raise ValueError(f"gs_order={gs_order!r}")
        _abi.call("vtsa_set_option", model.handle, _abi.OPT_GS_FORWARD, int(gs_order == "forward"))
        if conserve not in ("polish", "project", "project_strict"):
            raise ValueError(f"conserve={conserve!r}")
        self.conserve = conserve
        self.cap_after_fail = 0.5
        self.cap_relax = 1.15
        self.ops = DeviceOps(model)
        if not 0 <= int(lowrank) <= LOWRANK_MAX:
            raise ValueError(f"lowrank={lowrank!r}: 0..{LOWRANK_MAX}")
        self.ops.lowrank = int(lowrank)
        self.ops.lowrank_window = (float(lowrank_window[0]), float(lowrank_window[1]))
        self.record = record
        self.guess = None
        if guess is not None:
            from ..learned import StageGuess
            self.guess = StageGuess(guess, model.n_c // model.ng, self.ops.N)
        if (guess is not None or record is not None) and (model.ng != 3 or model.na != 2):
            raise ValueError("the learned stage guess needs the three-gas, two-adsorbate state layout")
        self._guess_miss = self._guess_wait = 0
        if guess is not None or record is not None:
            self.stats.update({"guess_used": 0, "guess_kept_predictor": 0, "guess_skipped": 0})
            self._s_ref, self._gf, self._gR, self._gP, self._zg = (dv.empty(self.ops.N) for _ in range(5))
        # parity mode (no upstream counterpart): replay the accepted step end times of another run
        self.h_sequence = None if h_sequence is None else np.asarray(h_sequence, dtype=float)
        if error_filter not in ("precond", "exact"):
            raise ValueError(f"error_filter={error_filter!r}")
        self.error_filter = error_filter
        self.last_err = None
        self._new_lu = None
        self._new_J = None
        self._fresh = False

    def scale(self, y):
        return self.ops.scale(y, self.rtol)

    def _error_norm(self, t, y, z2, f2, y3, dh):
        """RMS of the filtered local error (core/solver.py:453-463): M^-1 (z2 - y3) scaled by
        atol + rtol max(|y|, |z2|), M the preconditioner ("precond") or I - dh J(z2) solved by
        GMRES ("exact")."""
        ops = self.ops
        if self.error_filter == "precond":
            return ops.error_norm(y, z2, y3, self.rtol)
        N = ops.N
        d, s, b, w, e = (dv.empty(N) for _ in range(5))
        ops.vec(7, d, x=y3, y=z2)                       # z2 - y3
        ops.scale(z2, self.rtol, s)
        ops.vec(8, b, x=d, y=s)                         # -(z2 - y3)/s
        info = ops.gmres(t, z2, f2, s, dh, None, b, w, 1e-10, 0.0, 64, pw=e)   # e = -M^-1 (z2 - y3)
        self.stats["exact_filter_unconverged"] = self.stats.get("exact_filter_unconverged", 0) + int(info.info != 0)
        eh, yh, zh = dv.to_host(e), dv.to_host(y), dv.to_host(z2)
        sc = self.atol + self.rtol * np.maximum(np.abs(yh), np.abs(zh))
        return float(np.sqrt(np.mean((eh / sc) ** 2)))

    def _factor(self, y_ref, h):
        self.ops.precond_setup(D * h, self.ops.scale(y_ref, self.rtol), self.precond_kind, self.precision,
                               self.gs_sweeps)
        return Factorisation(h)

    def _start(self, t, z0, rhs_known, dh, y_ref, stage):
        """Starting value of a stage: the predictor z0, or the learned guess when both its
        preconditioned and its plain stage residual are smaller; records the stage's features.
        After GUESS_MISS_LIMIT stages in a row whose guess lost to the predictor, the guess is
        tried on every GUESS_RETRY_EVERY-th stage only, until it wins again (no upstream
        counterpart: the trial costs two residuals and a preconditioner application)."""
        if self.guess is None and self.record is None:
            return z0
        if self.record is None and self._guess_miss >= GUESS_MISS_LIMIT:
            self._guess_wait += 1
            if self._guess_wait % GUESS_RETRY_EVERY:
                self.stats["guess_skipped"] += 1
                return z0
        h = dh / D
        ops = self.ops
        s = ops.scale(y_ref, self.rtol, self._s_ref)
        rn0, pn0 = ops.newton_eval(t, z0, dh, rhs_known, s, self._gf, self._gR)
        if not np.isfinite(pn0):
            return z0
        ops.prec_apply(self._gR, self._gP)          # M^-1 R(z0): minus the first Newton step
        self._recording = self.record is not None and self.record.begin(self.m.handle, y_ref, z0, s, self._gP, h,
                                                                        stage)
        if self.guess is None:
            return z0
        zg = self.guess.predict(self.m.handle, y_ref, z0, s, self._gP, h, stage, self._zg)
        rn1, pn1 = ops.newton_eval(t, zg, dh, rhs_known, s, self._gf, self._gR)
        if np.isfinite(pn1) and pn0 > 0.0:
            self.stats["guess_log_pn_ratio"] = self.stats.get("guess_log_pn_ratio", 0.0) + float(np.log(pn1 / pn0))
        # both the preconditioned and the plain residual must improve: a guess that only the
        # preconditioned norm favours can leave conservation errors Newton does not remove
        if np.isfinite(pn1) and np.isfinite(rn1) and pn1 < pn0 and rn1 <= rn0:
            self.stats["guess_used"] += 1
            self._guess_miss = 0
            return zg
        self.stats["guess_kept_predictor"] += 1
        self._guess_miss += 1
        return z0

    def _solve_stage(self, t, z0, rhs_known, dh, lu, res, y_ref, stage=0):
        h = dh / D
        z_pred = z0
        self._recording = False
        z0 = self._start(t, z0, rhs_known, dh, y_ref, stage)
        if self.mode == "hybrid":
            # modified Newton first (4 iterations, contraction limit 0.5), Krylov on failure
            # (core/solver.py:340-347)
            z, ok, it, rate = newton_modified_dev(self.ops, t, z0, rhs_known, dh, self.rtol, 4, self.kappa,
                                                  fail_rate=0.5, cons_tol=self.cons_tol)
            if ok:
                if self._recording:
                    self.record.commit(self.m.handle, z, z_pred, self._gP, self._s_ref)
                return z, ok, rate
            self.stats["modified_fail"] += 1

        def refresh(zc):
            fc = dv.empty(self.ops.N)
            self.m.rhs_dev(t, zc, fc)
            self.jac.evaluate(t, zc, fc)
            res.n_jac += 1
            res.n_lu += 1
            self._new_J = (True, h)
            return self._factor(zc, h)

        self._new_J = None
        z, ok, it, rate, new_lu = newton_krylov_dev(self.ops, t, z0, rhs_known, dh, lu, self.rtol, self.max_iter,
                                                self.kappa, krylov_dim=self.krylov_dim, counter=self._gm,
                                                cons_tol=self.cons_tol, stats=self.stats,
                                                refresh=refresh if (self.newton_refresh and not self._fresh) else None,
                                                project=self.conserve.startswith("project"),
                                                strict=self.conserve == "project_strict",
                                                reject_unclosed=self.conserve.startswith("project"))
        if new_lu is not None:
            self._new_lu = new_lu
        if ok and self._recording:
            self.record.commit(self.m.handle, z, z_pred, self._gP, self._s_ref)
        res.n_krylov += 1
        self.stats["krylov_ok" if ok else "krylov_fail"] += 1
        return z, ok, (0.0 if ok else rate)

    def integrate(self, y0, t_end: float, t_out, on_step: Optional[Callable] = None,
                  stop_fn: Optional[Callable] = None, stop_value: Optional[float] = None,
                  stop_direction: int = 1, device_result: bool = False) -> IntegrationResult:
        m, ops = self.m, self.ops
        N = ops.N
        wall0 = time.time()
        t = 0.0
        y = dv.to_dev(y0) if isinstance(y0, np.ndarray) else y0.clone()
        acc = np.zeros(m.n_acc)
        f_n, gbuf = dv.empty(N), dv.empty(m.n_acc)
        m.rhs_dev(t, y, f_n, gbuf)
        g_n = dv.to_host(gbuf).copy()
        res = IntegrationResult(y=None, acc=acc, t=t)
        t_out = np.asarray(sorted(set(np.clip(t_out, 0.0, t_end)) | {t_end}))
        k_out = 0
        if t_out[0] <= 0.0:
            res.snapshots.append((0.0, dv.to_host(y).copy(), acc.copy()))
            k_out = 1
        dev_cb = bool(getattr(on_step, "wants_device", False))
        h = min(self.h_init, self.h_max, t_end)
        have_J = False
        lu = None
        h_lu = None
        jac_age = 0
        fresh = False
        slope = dv.empty(N)
        have_slope = False
        rhs1, zp, rhs2, z2p, y3 = (dv.empty(N) for _ in range(5))
        f1, f2 = dv.empty(N), dv.empty(N)
        h_cap = np.inf
        self._gm = [0]
        sd = self.scheme == "sdirk2"
        c_mid = G_SD if sd else GAMMA
        replay = self.h_sequence is not None
        k_rep = 0
        while t < t_end * (1.0 - 1e-12):
            gm0 = self._gm[0]
            if res.n_steps + res.n_reject + res.n_newton_fail > self.max_steps:
                raise SolverError(f"too many steps in '{m.step['name']}' at t={t:.4g}s")
            t_target = t_out[k_out] if k_out < len(t_out) else t_end
            h_desired = h
            if replay:
                if k_rep >= len(self.h_sequence):
                    raise SolverError(f"replay: the step sequence ends before t_end (t={t:.6g}s)")
                h = float(self.h_sequence[k_rep]) - t
                truncated = False
            else:
                h = min(h, self.h_max, t_target - t)
                truncated = h < h_desired * (1.0 - 1e-9)
            if h < 1e-10 * max(1.0, t_end):
                raise SolverError(f"step size underflow (h={h:.3e}) in '{m.step['name']}' at t={t:.6g}s")
            if not have_J or jac_age >= self.jac_max_age:
                self.jac.evaluate(t, y, f_n)
                res.n_jac += 1
                have_J = True
                jac_age = 0
                fresh = True
                lu = None
            if lu is None or abs(np.log(h / h_lu)) > self.log_refactor:
                lu = self._factor(y, h)
                h_lu = h
                res.n_lu += 1
            dh = D * h
            if sd:
                r1 = y
            else:
                r1 = ops.vec(0, rhs1, x=y, y=f_n, a=dh)
            z_pred = y if not have_slope else ops.vec(1, zp, x=y, y=slope, a=c_mid * h)
            self._new_lu = None
            self._fresh = fresh
            z1, ok, r1_rate = self._solve_stage(t + c_mid * h, z_pred, r1, dh, lu, res, y, 0)
            if self._new_lu is not None:
                lu, h_lu = self._new_lu, self._new_J[1]
                jac_age, fresh, self._new_lu = 0, True, None
            if ok:
                if sd:
                    ops.vec(3, rhs2, x=y, y=z1, a=(1.0 - G_SD) * h, b=dh)
                    ops.vec(2, z2p, x=y, y=z1, a=G_SD)
                else:
                    ops.vec(4, rhs2, x=y, y=z1, a=C1, b=C2)
                    ops.vec(2, z2p, x=y, y=z1, a=GAMMA)
                z2, ok, r2_rate = self._solve_stage(t + h, z2p, rhs2, dh, lu, res, y, 1)
                if self._new_lu is not None:
                    lu, h_lu = self._new_lu, self._new_J[1]
                    jac_age, fresh, self._new_lu = 0, True, None
            if not ok:
                res.n_newton_fail += 1
                if not fresh:
                    have_J = False
                elif replay:
                    raise SolverError(f"replay: stage failed at t={t:.6g}s, h={h:.3e} with a fresh Jacobian")
                else:
                    h_cap = min(h_cap, self.cap_after_fail * h)
                    h *= 0.3
                    lu = None
                continue
            m.rhs_dev(t + c_mid * h, z1, f1, gbuf)
            g1 = dv.to_host(gbuf).copy()
            m.rhs_dev(t + h, z2, f2, gbuf)
            g2 = dv.to_host(gbuf).copy()
            if sd:
                ops.vec(5, y3, x=y, y=f_n, z=f1, w=f2, a=h, b=SW1, c=SW2, e=SW3)
            else:
                ops.vec(5, y3, x=y, y=f_n, z=f1, w=f2, a=h, b=W1, c=W2, e=W3)
            en = self._error_norm(t + h, y, z2, f2, y3, dh)
            self.last_err = en
            if not np.isfinite(en):
                en = 1e6
            if en > 1.0 and not replay:
                res.n_reject += 1
                h *= max(0.2, 0.9 * en ** (-1.0 / 3.0))
                continue
            if sd:
                acc = acc + h * ((1.0 - G_SD) * g1 + G_SD * g2)
            else:
                acc_g = acc + D * h * (g_n + g1)
                acc = C1 * acc_g - C2 * acc + D * h * g2
            ops.vec(6, slope, x=y, y=z2, a=h)
            have_slope = True
            t_new = t + h
            k_rep += 1
            res.steps_taken.append((t, h))
            if stop_fn is not None and stop_value is not None:
                sv = stop_fn(t_new, z2 if dev_cb else dv.to_host(z2))
                if sv is not None and (sv - stop_value) * stop_direction >= 0.0:
                    t, y, g_n = t_new, z2, g2
                    f_n, f2 = f2, f_n
                    res.stopped_early = True
                    res.stop_value = sv
                    res.n_steps += 1
                    if on_step:
                        on_step(t, y if dev_cb else dv.to_host(y), acc)
                    res.snapshots.append((t, dv.to_host(y).copy(), acc.copy()))
                    break
            t, y, g_n = t_new, z2, g2
            f_n, f2 = f2, f_n
            res.n_steps += 1
            jac_age += 1
            fresh = False
            if on_step:
                on_step(t, y if dev_cb else dv.to_host(y), acc)
            if k_out < len(t_out) and t >= t_out[k_out] * (1.0 - 1e-12):
                res.snapshots.append((t, dv.to_host(y).copy(), acc.copy()))
                k_out += 1
            gm_step = self._gm[0] - gm0
            if gm_step > self.gmres_refresh:
                jac_age = self.jac_max_age
            fac = 4.0 if en < 1e-8 else min(4.0, max(0.2, 0.9 * en ** (-1.0 / 3.0)))
            if 1.0 <= fac < 1.25:
                fac = 1.0
            h = max(h * fac, h_desired) if truncated else h * fac
            h_cap *= self.cap_relax
            h = min(h, h_cap)
        res.y = y if device_result else dv.to_host(y)
        res.acc, res.t = acc, t
        self.stats["gmres_iters"] = self._gm[0]
        if ops.lowrank:
            self.stats["lowrank_built"] = ops.lowrank_built
            self.stats["lowrank_dropped"] = ops.lowrank_dropped
            self.stats["gmres_iters_corrected"] = ops.gmres_iters_corrected
            self.stats["gmres_iters_uncorrected"] = ops.gmres_iters_plain
        res.nfev = m.nfev
        res.wall_time = time.time() - wall0
        return res

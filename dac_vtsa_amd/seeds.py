"""Seeded synthetic bed states for the benchmark workloads.

One generator, ``numpy.random.default_rng(seed)`` (PCG64), per-cell i.i.d. draws in the fixed
order T -> P -> y_CO2 -> y_H2O -> loading fractions (or loadings).  Workloads:

* ``small``  : T ~ U(288, 293) K, P = 1.013e5 Pa, y_CO2 = 4e-4, y_H2O ~ U(0.005, 0.015),
               y_N2 the rest; q_j = U(0, 0.2) x local equilibrium (state-basis water for CO2).
* ``midcycle``: a spatially uniform bed whose values are drawn once per seed:
               T ~ U(288, 373) K, P ~ U(1.5e4, 1.013e5) Pa, y_CO2 ~ U(0, 0.9), y_H2O ~ U(0, 0.05),
               y_N2 the rest, q_CO2 ~ U(0, 1.6), q_H2O ~ U(0, 6) mol/kg.
* ``midcycle_cells``: the same distributions drawn independently per cell.  The upstream Newton
               controller cannot start from such a field (every stage fails at t = 0 down to
               step-size underflow, with or without per-cell pressures), so it serves residual and
               Jacobian parity tests only.
The wall starts at the temperature of the outer bed ring.

The equilibrium loading is supplied by the caller (``qstar_fn(p_co2, p_h2o, T, q_h2o)`` returning
``(q_co2*, q_h2o*)``), so the same generator serves the HIP path and the CPU oracle.
"""
from __future__ import annotations

import numpy as np

from .core.state import CycleState, R_GAS


def seeded_state(grid, kind: str, seed: int, qstar_fn=None) -> CycleState:
    rng = np.random.default_rng(seed)
    shp = grid.shape
    if kind == "small":
        T = rng.uniform(288.0, 293.0, shp)
        P = np.full(shp, 1.013e5)
        y_co2 = np.full(shp, 4e-4)
        y_h2o = rng.uniform(0.005, 0.015, shp)
    elif kind in ("midcycle", "midcycle_cells"):
        draw = (lambda lo, hi: rng.uniform(lo, hi, shp)) if kind == "midcycle_cells" else \
            (lambda lo, hi: np.full(shp, rng.uniform(lo, hi)))
        T = draw(288.0, 373.0)
        P = draw(1.5e4, 1.013e5)
        y_co2 = draw(0.0, 0.9)
        y_h2o = draw(0.0, 0.05)
    else:
        raise ValueError(kind)
    y_n2 = 1.0 - y_co2 - y_h2o
    C = P / (R_GAS * T)
    c = np.stack([C * y_co2, C * y_h2o, C * y_n2])
    if kind == "small":
        if qstar_fn is None:
            raise ValueError("the 'small' workload needs qstar_fn for its loadings")
        f_w = rng.uniform(0.0, 0.2, shp)
        f_c = rng.uniform(0.0, 0.2, shp)
        _, qw_eq = qstar_fn(c[0] * R_GAS * T, c[1] * R_GAS * T, T, np.zeros(shp))
        q_w = f_w * qw_eq
        qc_eq, _ = qstar_fn(c[0] * R_GAS * T, c[1] * R_GAS * T, T, q_w)
        q_c = f_c * qc_eq
    else:
        q_c = draw(0.0, 1.6)
        q_w = draw(0.0, 6.0)
    q = np.stack([q_c, q_w])
    Tw = T[-1].copy()
    return CycleState(c=c, q=q, T=T, Tw=Tw, t=0.0, cycle=1, step="seeded")


WORKLOADS = {
    # name: (Nr, Ntheta, Nz, state kind, step index in the default cycle, duration override)
    "small_adsorption": (4, 4, 20, "small", 0, 200.0),
    "midcycle_heating": (10, 16, 50, "midcycle", 2, 22400.0),
    "large_cycle": (20, 32, 100, "small", None, None),
}


def device_qstar(cfg, grid, props):
    """``qstar_fn`` evaluated by the device isotherm kernels of a throwaway model of the grid."""
    from .core.residual import TVSAModel
    shp = grid.shape
    T = np.full(shp, cfg.feed.T)
    C = cfg.feed.P / (R_GAS * T)
    c = np.stack([C * 4e-4, C * 1e-2, C * (1.0 - 4e-4 - 1e-2)])
    st = CycleState(c=c, q=np.zeros((2,) + shp), T=T, Tw=T[-1].copy())
    model = TVSAModel(cfg, grid, props, cfg.cycle[0], st)

    def fn(pc, pw, T, qw):
        return model.equilibrium(pc, pw, T, qw)
    fn.model = model
    return fn

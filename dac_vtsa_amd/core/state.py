"""State container handed from step to step (upstream core/state.py:22-50).

Fields are cell averages on the (Nr, Ntheta, Nz) grid:
    c  (ng, Nr, Nth, Nz)  gas concentrations [CO2, (H2O), N2]   mol/m3
    q  (na, Nr, Nth, Nz)  loadings [CO2, (H2O)]                 mol/kg
    T  (Nr, Nth, Nz)      bed temperature                        K
    Tw (Nth, Nz)          wall temperature                       K
    P  derived: sum(c) R T                                       Pa
"""
from __future__ import annotations

import copy
from dataclasses import dataclass, field
from typing import Dict, Tuple

import numpy as np

R_GAS = 8.314462618


@dataclass
class CycleState:
    c: np.ndarray
    q: np.ndarray
    T: np.ndarray
    Tw: np.ndarray
    t: float = 0.0
    cycle: int = 1
    step: str = "initial"
    species: Tuple[str, ...] = ("CO2", "H2O", "N2")
    adsorbed: Tuple[str, ...] = ("CO2", "H2O")
    meta: Dict = field(default_factory=dict)

    @property
    def P(self) -> np.ndarray:
        return self.c.sum(axis=0) * R_GAS * self.T

    @property
    def y(self) -> np.ndarray:
        return self.c / np.maximum(self.c.sum(axis=0), 1e-30)

    def gas(self, name: str) -> np.ndarray:
        return self.c[self.species.index(name)]

    def loading(self, name: str) -> np.ndarray:
        return self.q[self.adsorbed.index(name)]

    def copy(self) -> "CycleState":
        # a step runner may attach the bed's device copy (_y_device, no upstream counterpart);
        # a copy is a host state of its own and does not carry it
        dev = self.__dict__.pop("_y_device", None)
        try:
            return copy.deepcopy(self)
        finally:
            if dev is not None:
                self._y_device = dev


def initial_state(cfg, grid, props) -> CycleState:
    """Bed state at the start of a run, chosen by ``run.initial_state`` (upstream
    core/state.py:62-93): ``clean`` is feed gas at feed T and P over a bare sorbent, ``feed_equilibrium``
    the same gas over sorbent at equilibrium with it (the water loading first, then CO2 on the
    state-basis water), ``csv`` a state file in the hand-off format.  The wall starts at feed T."""
    from ..valves import humid_air_composition
    water = bool(cfg.components.water)
    species = tuple(props.species)
    adsorbed = ("CO2", "H2O") if water else ("CO2",)
    mode = cfg.run.initial_state
    if mode == "csv":
        from ..simio.checkpoint import load_state_csv
        st = load_state_csv(cfg.run.initial_state_csv, grid, list(species), list(adsorbed))
        st.step, st.cycle = "initial", 1
        return st
    T0, P0 = cfg.feed.T, cfg.feed.P
    y = humid_air_composition(list(species), P0, T0, cfg.feed.RH if water else 0.0, cfg.feed.y_co2_dry,
                              float(props.p_sat(T0)))
    T = np.full(grid.shape, T0)
    c = np.stack([np.full(grid.shape, P0 / (R_GAS * T0) * yk) for yk in y])
    q = np.zeros((len(adsorbed),) + grid.shape)
    if mode == "feed_equilibrium":
        if not water:
            raise NotImplementedError("the device isotherms cover humid air (components.water = true)")
        from ..seeds import device_qstar
        qs = device_qstar(cfg, grid, props)
        p = c * R_GAS * T
        _, qw = qs(p[0], p[1], T, np.zeros(grid.shape))
        qc, qw2 = qs(p[0], p[1], T, qw)
        qs.model.close()
        q = np.stack([qc, qw2])
    elif mode != "clean":
        raise ValueError(f"run.initial_state = {mode!r}")
    return CycleState(c=c, q=q, T=T, Tw=np.full(grid.wall_shape, T0), t=0.0, cycle=1, step="initial",
                      species=species, adsorbed=adsorbed)

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
        return copy.deepcopy(self)

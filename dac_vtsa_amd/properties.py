"""Ideal-gas property constants for the device model (upstream properties/ideal_gas_backend.py).

Only constants are evaluated here, once per model: the pure-component tables, cp at
``properties.cp_reference_T``, the Wilke mass-ratio factors and the Fuller pair constants that
the kernel's mixture rules use.  ``p_sat`` is offered on the host for boundary-condition setup
(feed and ambient humidity); every per-cell evaluation happens on the GPU.
"""
from __future__ import annotations

import numpy as np

R_GAS = 8.314462618
# pure-component data (Young 2021 ESI S6/S7, Incropera A.4); the inert "N2" is air
CP_POLY = {"N2": (1.06569e-5, -6.4057779e-3, 30.1318), "CO2": (-3.23753e-5, 6.17517420e-2, 21.9165),
           "H2O": (1.72619108e-3, -1.22971184026, 250.9868), "O2": (9.08259e-6, 1.00528795e-3, 28.3000)}
MU_LIN = {"N2": (3.96929e-8, 6.15036e-6), "CO2": (4.41697e-8, 1.90440e-6),
          "H2O": (3.55595e-8, -9.930767e-7), "O2": (4.75391e-8, 6.53871e-6)}
K_POW = {"N2": (0.02598, 0.81), "O2": (0.02676, 0.84), "CO2": (0.01655, 1.335), "H2O": (0.0186, 1.18)}
MOLAR_MASS = {"CO2": 0.04401, "H2O": 0.018015, "N2": 0.028960}
AIR = (("N2", 0.79), ("O2", 0.21))
DIFFUSION_VOLUME = {"CO2": 26.7, "H2O": 13.1, "N2": 18.5, "O2": 16.3, "AIR": 19.7}
MOLAR_MASS_G = {"CO2": 44.01, "H2O": 18.015, "N2": 28.013, "O2": 31.999, "AIR": 28.96}


def arden_buck(T):
    Tc = np.asarray(T, dtype=float) - 273.15
    return 611.21 * np.exp((18.678 - Tc / 234.5) * (Tc / (257.14 + Tc)))


def species_for(cfg):
    return ["CO2", "H2O", "N2"] if cfg.components.water else ["CO2", "N2"]


def viscosity_terms(species):
    """Per species: (f1, a1, b1, f2, a2, b2, two) with mu = f1 (a1 T + b1) [+ f2 (a2 T + b2)]."""
    out = []
    for s in species:
        if s == "N2":
            (c1, f1), (c2, f2) = AIR
            out.append((f1, *MU_LIN[c1], f2, *MU_LIN[c2], 1))
        else:
            out.append((1.0, *MU_LIN[s], 0.0, 0.0, 0.0, 0))
    return out


def conductivity_terms(species):
    """Per species: (c1, n1, c2, n2) with k = c1 (T/300)^n1 [+ c2 (T/300)^n2]."""
    out = []
    for s in species:
        if s == "N2":
            (a, fa), (b, fb) = AIR
            out.append((fa * K_POW[a][0], K_POW[a][1], fb * K_POW[b][0], K_POW[b][1]))
        else:
            out.append((K_POW[s][0], K_POW[s][1], 0.0, 0.0))
    return out


def wilke_tables(M):
    """(sqrt(8 (1 + M_i/M_j)), (M_j/M_i)^(1/4)) of Wilke's mixing rule from the molar masses."""
    M = np.asarray(M, dtype=float)
    Mi, Mj = M[:, None], M[None, :]
    return np.sqrt(8.0 * (1.0 + Mi / Mj)), (Mj / Mi) ** 0.25


def fuller_pair(a, b):
    """(sqrt(M_ab), Sv_a^(1/3) + Sv_b^(1/3)) of the Fuller correlation."""
    Mab = 2.0 / (1.0 / MOLAR_MASS_G[a] + 1.0 / MOLAR_MASS_G[b])
    return float(np.sqrt(Mab)), DIFFUSION_VOLUME[a] ** (1 / 3) + DIFFUSION_VOLUME[b] ** (1 / 3)


class GasProps:
    """Constant tables of the humid-air mixture ``[CO2, H2O, N2]``."""

    name = "ideal_gas"

    def __init__(self, species, cp_reference_T=330.0):
        self.species = list(species)
        self.ng = len(self.species)
        if self.species != ["CO2", "H2O", "N2"]:
            raise NotImplementedError("the device model covers the humid-air species set CO2/H2O/N2")
        self.M = np.array([MOLAR_MASS[s] for s in self.species])
        self.cp = np.array([self._cp1(s, cp_reference_T) for s in self.species])
        self.i_co2, self.i_h2o, self.i_inert = 0, 1, 2
        self.wilke_den, self.mass_q = wilke_tables(self.M)

    @staticmethod
    def _cp1(s, T):
        if s == "N2":
            return sum(f * np.polyval(CP_POLY[c], T) for c, f in AIR)
        return float(np.polyval(CP_POLY[s], T))

    def p_sat(self, T):
        return arden_buck(T)

    def viscosity_terms(self):
        return viscosity_terms(self.species)

    def conductivity_terms(self):
        return conductivity_terms(self.species)

    fuller_pair = staticmethod(fuller_pair)


def get_backend(cfg, species=None):
    """Property backend for a configuration (upstream properties/__init__.py:9-14)."""
    if cfg.properties.backend != "ideal_gas":
        raise NotImplementedError("properties.backend must be ideal_gas on the device path")
    return GasProps(species or species_for(cfg), cfg.properties.cp_reference_T)

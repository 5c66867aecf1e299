"""Shared test helpers: fixture loading and oracle construction."""
import os

import numpy as np

from dac_vtsa_amd.config import load_config
from dac_vtsa_amd.core.state import CycleState

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

SMALL = {"grid.Nr": 3, "grid.Ntheta": 4, "grid.Nz": 6, "jacket.theta_coverage_deg": [0.0, 180.0],
         "feed.profile": "cosine_theta", "feed.profile_amplitude": 0.3}
CFG1 = {"grid.Nr": 4, "grid.Ntheta": 4, "grid.Nz": 20}
MID = {"grid.Nr": 4, "grid.Ntheta": 4, "grid.Nz": 10, "operation.P_vacuum": 1.5e4}


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def state_from(z, prefix=""):
    return CycleState(c=z[prefix + "c"], q=z[prefix + "q"], T=z[prefix + "T"], Tw=z[prefix + "Tw"])


def rel(a, b, floor=0.0):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), floor, 1e-300))


def oracle_model(over, step_index, state, T_jacket_prev=None):
    from oracle.runner import make
    from oracle.model import OracleModel
    cfg = load_config(override_dict=over)
    g, gas = make(cfg)
    return cfg, OracleModel(cfg, g, gas, cfg.cycle[step_index], state, T_jacket_prev)


def oracle_qstar(cfg):
    from oracle.model import Equilibrium, arden_buck
    eq = Equilibrium(cfg)

    def fn(pc, pw, T, qw):
        qs = eq.equilibrium(pc, pw, T, qw, arden_buck(T))
        return qs[0], qs[1]
    return fn

# option variants of the golden rhs fixture (tests/golden/make_golden.py writes them)
VARIANTS = {
    "base": {},
    "van_leer": {"solver.limiter": "van_leer"},
    "minmod": {"solver.limiter": "minmod"},
    "vatvd": {"solver.limiter": "van_albada_tvd"},
    "upwind": {"solver.limiter": "upwind"},
    "mech": {"isotherm.coadsorption.model": "mechanistic"},
    "sb": {"isotherm.coadsorption.model": "stampi_bombelli"},
    "none": {"isotherm.coadsorption.model": "none"},
    "eqbasis": {"isotherm.coadsorption.water_loading_basis": "equilibrium"},
    "young": {"transport.dispersion_model": "young2021"},
    "wakao": {"transport.dispersion_model": "wakao_kagei"},
    "parallel": {"transport.bed_conductivity_model": "parallel"},
    "constlam": {"transport.bed_conductivity_model": "constant", "transport.wall_htc_model": "constant"},
    "lumped": {"transport.heat_transfer_mode": "lumped"},
    "nowall": {"wall.enabled": False},
    "hx": {"internal_hx.enabled": True, "jacket.ramp_time": 30.0},
    "ramp": {"operation.vent_model": "ramp", "operation.vacuum_model": "ramp"},
    "nocheck": {"operation.vacuum_check_valve": False},
    "ea": {"kinetics.Ea_co2": 20000.0, "kinetics.Ea_h2o": 10000.0},
    "krish": {"energy.vacuum_efficiency_model": "krishnamurthy", "energy.pump_handles_water": True},
    "eqarea": {"grid.radial_spacing": "equal_area"},
    "linr": {"feed.profile": "linear_r"},
    "disc": {"feed.profile": "central_disc"},
    "sector": {"feed.profile": "sector", "feed.profile_theta0_deg": 30.0},
    "nth1": {"grid.Ntheta": 1},
    "nr1": {"grid.Nr": 1},
}

"""End-of-column boundary conditions of a cycle step (upstream valves/valve_logic.py:40-148).

A step opens or closes the valves V1..V3; each valve sits at one end of the column
(``cfg.ports``).  Per end, the open valve there decides the boundary-condition kind handed to
the device model:

* no open valve -> ``closed`` (no flow);
* ``OPEN`` (the feed blower) -> ``feed``: fixed molar inflow of feed gas;
* ``OPEN_ATMOSPHERIC`` -> ``valve`` (linear conductance to a header) or ``pressure`` (ramped
  Dirichlet), towards the feed header for V1 and the ambient for the others;
* ``OPEN_VACUUM`` -> ``pump`` (finite pump speed, optional check valve) or ``pressure``.

Evaluated once per step on the host; the kinds and their parameters go into the device
descriptor.  The feed profile gives the inlet distribution over the end face.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, Optional

import numpy as np

from .properties import R_GAS

KIND_CODE = {"closed": 0, "feed": 1, "pump": 2, "valve": 3, "pressure": 4}
VALVES = ("V1", "V2", "V3")


@dataclass
class EndCondition:
    """Boundary condition of one end: kind, the external state it connects to, pump/valve data."""
    end: str
    valve: Optional[str]
    state: str
    kind: str
    P_target: float = 101325.0
    y_ext: Optional[np.ndarray] = None
    T_ext: float = 298.15
    check_valve: bool = False
    is_product: bool = False
    is_vacuum: bool = False
    is_feed: bool = False
    speed: float = 0.0


def humid_air_composition(species, P, T, RH, y_co2_dry, p_sat):
    """Mole fractions of humid air in the order of ``species``: water from the relative humidity,
    CO2 from its dry-basis fraction, the inert gas the rest."""
    y_w = RH * p_sat / P if "H2O" in species else 0.0
    y_c = y_co2_dry * (1.0 - y_w)
    table = {"CO2": y_c, "H2O": y_w, "N2": 1.0 - y_c - y_w}
    return np.array([table[s] for s in species])


def _open_valve(cfg, step, end):
    """The valve that is open at ``end`` during the step (the last of V1..V3 when several are)."""
    found = None
    for v in VALVES:
        if cfg.ports[v] == end and step["valves"][v] != "CLOSED":
            found = v
    return found


def end_conditions(cfg, step: Dict, species, props) -> Dict[str, EndCondition]:
    """{"top": EndCondition, "bottom": EndCondition} of one step."""
    water = "H2O" in species
    feed, amb, op = cfg.feed, cfg.ambient, cfg.operation

    def air(src):
        return humid_air_composition(species, src.P, src.T, src.RH if water else 0.0, src.y_co2_dry,
                                     float(props.p_sat(src.T)))

    feed_gas = (feed.P, air(feed), feed.T)
    ambient_gas = (amb.P, air(amb), amb.T)
    product = step.get("product_port", "V3")
    conds = {}
    for end in ("top", "bottom"):
        v = _open_valve(cfg, step, end)
        if v is None:
            conds[end] = EndCondition(end, None, "CLOSED", "closed")
            continue
        state = step["valves"][v]
        if state == "OPEN":
            P, y, T = feed_gas
            conds[end] = EndCondition(end, v, state, "feed", P, y, T, is_feed=True)
        elif state == "OPEN_ATMOSPHERIC":
            P, y, T = feed_gas if v == "V1" else ambient_gas
            conds[end] = EndCondition(end, v, state, "valve" if op.vent_model == "valve" else "pressure", P, y, T,
                                      is_product=product == v, is_feed=v == "V1", speed=op.vent_conductance)
        else:
            _, y, T = ambient_gas
            conds[end] = EndCondition(end, v, state, "pump" if op.vacuum_model == "pump" else "pressure",
                                      op.P_vacuum, y, T, check_valve=bool(op.vacuum_check_valve),
                                      is_product=product == v, is_vacuum=True, speed=op.pump_speed_per_area)
    return conds


def feed_molar_flux(cfg):
    """Molar flux of feed gas per unit cross-section, u_feed P_feed / (R T_feed)."""
    return cfg.feed.velocity * cfg.feed.P / (R_GAS * cfg.feed.T)


def _sector_fraction(grid, centre_deg, width_deg):
    """Fraction of each theta sector's angle inside the sector [c - w/2, c + w/2] (periodic)."""
    half = 0.5 * np.radians(width_deg)
    c = np.radians(centre_deg)
    frac = np.zeros(grid.Nth)
    for j in range(grid.Nth):
        a, b = grid.th_f[j] - c, grid.th_f[j + 1] - c
        inside = 0.0
        for wrap in (-2.0 * np.pi, 0.0, 2.0 * np.pi):
            inside += max(0.0, min(b + wrap, half) - max(a + wrap, -half))
        frac[j] = inside / grid.dth
    return frac


def feed_profile(cfg, grid) -> np.ndarray:
    """Inlet distribution over the end face, shape (Nr, Nth), normalised to an area-weighted mean
    of 1 (``feed.profile``: uniform, cosine_theta, linear_r, central_disc, sector)."""
    fd = cfg.feed
    shape = (grid.Nr, grid.Nth)
    r = np.broadcast_to(np.asarray(grid.r_c, dtype=float)[:, None], shape)
    th = np.broadcast_to(np.asarray(grid.th_c, dtype=float)[None, :], shape)
    kind = fd.profile
    if kind == "uniform":
        f = np.ones(shape)
    elif kind == "cosine_theta":
        f = 1.0 + fd.profile_amplitude * np.cos(th - np.radians(fd.profile_theta0_deg))
    elif kind == "linear_r":
        f = 1.0 + fd.profile_amplitude * (2.0 * r / grid.R - 1.0)
    elif kind == "central_disc":
        # share of each ring's area inside the disc r <= r_in
        r_in = fd.profile_radius_fraction * grid.R
        inner, outer = np.asarray(grid.r_f[:-1]), np.asarray(grid.r_f[1:])
        share = (np.minimum(outer, r_in) ** 2 - inner ** 2) / (outer ** 2 - inner ** 2)
        f = np.broadcast_to(np.clip(share, 0.0, 1.0)[:, None], shape).copy()
    elif kind == "sector":
        f = np.broadcast_to(_sector_fraction(grid, fd.profile_theta0_deg, fd.profile_sector_deg)[None, :],
                            shape).copy()
    else:
        raise ValueError(kind)
    area = grid.Az * np.ones(shape)
    return f * area.sum() / (f * area).sum()

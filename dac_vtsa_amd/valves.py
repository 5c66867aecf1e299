"""Valve states -> end-of-column boundary conditions (upstream valves/valve_logic.py:40-148).

Evaluated once per step on the host; the resulting kinds and parameters go into the device
model descriptor.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, Optional

import numpy as np

from .properties import R_GAS

KIND_CODE = {"closed": 0, "feed": 1, "pump": 2, "valve": 3, "pressure": 4}


@dataclass
class EndCondition:
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
    y_w = RH * p_sat / P if "H2O" in species else 0.0
    y = {"H2O": y_w, "CO2": y_co2_dry * (1.0 - y_w)}
    y["N2"] = 1.0 - y["CO2"] - y_w
    return np.array([y[s] for s in species])


def end_conditions(cfg, step: Dict, species, props) -> Dict[str, EndCondition]:
    fd, amb, op = cfg.feed, cfg.ambient, cfg.operation
    water = "H2O" in species
    y_feed = humid_air_composition(species, fd.P, fd.T, fd.RH if water else 0.0, fd.y_co2_dry, float(props.p_sat(fd.T)))
    y_amb = humid_air_composition(species, amb.P, amb.T, amb.RH if water else 0.0, amb.y_co2_dry,
                                  float(props.p_sat(amb.T)))
    out = {}
    for end in ("top", "bottom"):
        valve = None
        for v in ("V1", "V2", "V3"):
            if cfg.ports[v] == end and step["valves"][v] != "CLOSED":
                valve = v
        if valve is None:
            out[end] = EndCondition(end, None, "CLOSED", "closed")
            continue
        st = step["valves"][valve]
        is_prod = step.get("product_port", "V3") == valve
        if st == "OPEN":
            out[end] = EndCondition(end, valve, st, "feed", fd.P, y_feed, fd.T, is_feed=True)
        elif st == "OPEN_ATMOSPHERIC":
            kind = "valve" if op.vent_model == "valve" else "pressure"
            if valve == "V1":
                out[end] = EndCondition(end, valve, st, kind, fd.P, y_feed, fd.T, is_product=is_prod, is_feed=True,
                                        speed=op.vent_conductance)
            else:
                out[end] = EndCondition(end, valve, st, kind, amb.P, y_amb, amb.T, is_product=is_prod,
                                        speed=op.vent_conductance)
        else:
            kind = "pump" if op.vacuum_model == "pump" else "pressure"
            out[end] = EndCondition(end, valve, st, kind, op.P_vacuum, y_amb, amb.T,
                                    check_valve=bool(op.vacuum_check_valve), is_product=is_prod, is_vacuum=True,
                                    speed=op.pump_speed_per_area)
    return out


def feed_molar_flux(cfg):
    return cfg.feed.velocity * cfg.feed.P / (R_GAS * cfg.feed.T)


def feed_profile(cfg, grid) -> np.ndarray:
    """Inlet distribution (Nr, Nth) with area-weighted mean 1."""
    fd = cfg.feed
    A = fd.profile_amplitude
    rr, th = np.meshgrid(grid.r_c, grid.th_c, indexing="ij")
    if fd.profile == "uniform":
        f = np.ones_like(rr)
    elif fd.profile == "cosine_theta":
        f = 1.0 + A * np.cos(th - np.radians(fd.profile_theta0_deg))
    elif fd.profile == "linear_r":
        f = 1.0 + A * (2.0 * rr / grid.R - 1.0)
    elif fd.profile == "central_disc":
        r_in = fd.profile_radius_fraction * grid.R
        lo, hi = grid.r_f[:-1], grid.r_f[1:]
        frac = np.clip((np.minimum(hi, r_in) ** 2 - lo ** 2) / (hi ** 2 - lo ** 2), 0.0, 1.0)
        f = np.repeat(frac[:, None], grid.Nth, axis=1)
    elif fd.profile == "sector":
        half = 0.5 * np.radians(fd.profile_sector_deg)
        t0 = np.radians(fd.profile_theta0_deg)
        lo, hi = grid.th_f[:-1] - t0, grid.th_f[1:] - t0
        ov = np.zeros(grid.Nth)
        for sh in (-2 * np.pi, 0.0, 2 * np.pi):
            ov += np.clip(np.minimum(hi + sh, half) - np.maximum(lo + sh, -half), 0.0, None)
        f = np.repeat((ov / grid.dth)[None, :], grid.Nr, axis=0)
    else:
        raise ValueError(fd.profile)
    w = grid.Az * np.ones_like(f)
    return f * w.sum() / (f * w).sum()

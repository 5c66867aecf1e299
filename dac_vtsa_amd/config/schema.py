"""Configuration loading and validation for the TVSA step engine.

Mirrors the public behaviour of the upstream loader (config/schema.py:529-594 and
config/parameters.py:25-44): defaults are deep-merged with an optional user YAML, a dict of
dotted overrides and ``key=value`` strings; string references such as ``operation.T_heat`` in
cycle steps are resolved; every key is type/range checked; unknown keys, ``Ntheta == 2`` and
inconsistent cycle definitions raise :class:`ConfigError`; derived quantities are added; the
result is a read-only attribute view.

The key table below gives, per dotted key, ``(type, lo, hi, choices, nullable)``.  Ranges are
the upstream ones (config/schema.py:58-281) so that the same inputs are accepted and rejected.
"""
from __future__ import annotations

import copy
import math
import os
from typing import Any, Dict, List, Optional, Sequence, Tuple

import yaml

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULTS_PATH = os.path.join(HERE, "defaults.yaml")

F, I, S, B, L, A = "float", "int", "str", "bool", "list", "any"
INF = None


def _f(lo=None, hi=None):
    return (F, lo, hi, None, False)


def _i(lo=None, hi=None):
    return (I, lo, hi, None, False)


def _c(*choices):
    return (S, None, None, tuple(choices), False)


_STR = (S, None, None, None, False)
_BOOL = (B, None, None, None, False)
_LIST = (L, None, None, None, False)
_ANY = (A, None, None, None, False)

KEYS: Dict[str, Tuple] = {
    "run.name": _STR, "run.results_dir": _STR, "run.max_cycles": _i(1, 500),
    "run.min_cycles": _i(1, 500), "run.css_tolerance": _f(1e-8, 1.0),
    "run.css_mb_tolerance": _f(1e-8, 1.0), "run.handoff_via_csv": _BOOL,
    "run.snapshots_per_step": _i(2, 10000), "run.write_fields_every_cycle": _BOOL,
    "run.initial_state": _c("clean", "feed_equilibrium", "csv"),
    "run.initial_state_csv": (S, None, None, None, True), "run.make_plots": _BOOL,
    "components.water": _BOOL,
    "geometry.bed_length": _f(1e-4, 50.0), "geometry.bed_radius": _f(1e-4, 10.0),
    "geometry.wall_thickness": _f(1e-5, 0.5),
    "grid.Nr": _i(1, 200), "grid.Ntheta": _i(1, 360), "grid.Nz": _i(2, 2000),
    "grid.radial_spacing": _c("uniform", "equal_area"),
    "sorbent.name": _STR, "sorbent.particle_diameter": _f(1e-6, 0.05),
    "sorbent.particle_density": _f(50, 5000), "sorbent.bed_voidage": _f(0.2, 0.95),
    "sorbent.total_voidage": _f(0.2, 0.99), "sorbent.cp_solid": _f(100, 5000),
    "sorbent.particle_conductivity": _f(0.01, 50), "sorbent.max_temperature": _f(300, 700),
    "isotherm.co2.q_inf0": _f(0.01, 50), "isotherm.co2.chi": _f(-50, 50),
    "isotherm.co2.b0": _f(0, 1e10), "isotherm.co2.dH0": _f(-5e5, 0),
    "isotherm.co2.tau0": _f(0.01, 5), "isotherm.co2.alpha": _f(-10, 10),
    "isotherm.co2.T0": _f(200, 500), "isotherm.co2.p_regularization": _f(0.0, 10.0),
    "isotherm.h2o.q_m": _f(0.01, 100), "isotherm.h2o.C": _f(0, 2e5), "isotherm.h2o.D": _f(0, 1),
    "isotherm.h2o.F": _f(0, 2e5), "isotherm.h2o.G": _f(-1000, 1000),
    "isotherm.h2o.E10_slope": _f(-1000, 1000), "isotherm.h2o.E10_intercept": _f(0, 2e5),
    "isotherm.h2o.rh_cap": _f(0.5, 0.999),
    "isotherm.coadsorption.model": _c("none", "wadst", "mechanistic", "stampi_bombelli"),
    "isotherm.coadsorption.water_loading_basis": _c("state", "equilibrium"),
    "isotherm.coadsorption.wadst.b0_wet": _f(0, 1e10),
    "isotherm.coadsorption.wadst.q_inf0_wet": _f(0, 100),
    "isotherm.coadsorption.wadst.tau0_wet": _f(0.001, 5),
    "isotherm.coadsorption.wadst.chi_wet": _f(-50, 50),
    "isotherm.coadsorption.wadst.alpha_wet": _f(-10, 10),
    "isotherm.coadsorption.wadst.dH_wet": _f(-5e5, 0),
    "isotherm.coadsorption.wadst.A": _f(1e-6, 100),
    "isotherm.coadsorption.mechanistic.f_blocked_max": _f(0, 1),
    "isotherm.coadsorption.mechanistic.k": _f(0, 100),
    "isotherm.coadsorption.mechanistic.phi_dry": _f(0.01, 2),
    "isotherm.coadsorption.mechanistic.A": _f(1e-6, 100),
    "isotherm.coadsorption.mechanistic.dH_wet": _f(-5e5, 0),
    "isotherm.coadsorption.mechanistic.n": _f(0.1, 10),
    "isotherm.coadsorption.stampi_bombelli.gamma": _f(-10, 10),
    "isotherm.coadsorption.stampi_bombelli.beta": _f(-10, 100),
    "kinetics.k_ldf_co2": _f(1e-7, 100), "kinetics.k_ldf_h2o": _f(1e-7, 100),
    "kinetics.Ea_co2": _f(0, 2e5), "kinetics.Ea_h2o": _f(0, 2e5), "kinetics.T_ref": _f(200, 500),
    "thermo.dH_ads_co2": _f(-3e5, 0), "thermo.dH_ads_h2o": _f(-3e5, 0),
    "thermo.cp_ads_co2": _f(0, 500), "thermo.cp_ads_h2o": _f(0, 500), "thermo.T_ref": _f(200, 500),
    "properties.backend": _c("ideal_gas", "coolprop"), "properties.cp_reference_T": _f(200, 600),
    "transport.dispersion_model": _c("edwards_richardson", "young2021", "wakao_kagei"),
    "transport.er_gamma1": _f(0, 5), "transport.er_gamma2": _f(0, 5),
    "transport.er_beta": _f(0, 100), "transport.Pe_mass_radial_inf": _f(1, 100),
    "transport.Pe_heat_radial_inf": _f(1, 100), "transport.Pe_heat_axial_inf": _f(0.1, 100),
    "transport.bed_conductivity_model": _c("zehner_schlunder", "parallel", "constant"),
    "transport.bed_conductivity_constant": _f(1e-3, 50), "transport.zs_shape_factor": _f(0.5, 5),
    "transport.wall_htc_model": _c("martin_nilles", "constant"),
    "transport.wall_htc_constant": _f(0.1, 1e5), "transport.ergun_A": _f(0, 1e4),
    "transport.ergun_B": _f(0, 100), "transport.heat_transfer_mode": _c("resolved", "lumped"),
    "transport.lumped_U": _f(0.01, 1e5),
    "wall.enabled": _BOOL, "wall.density": _f(100, 30000), "wall.cp": _f(100, 5000),
    "wall.conductivity": _f(0.01, 500),
    "jacket.h_jacket": _f(0, 1e5), "jacket.theta_coverage_deg": _LIST, "jacket.z_coverage": _LIST,
    "jacket.h_ambient": _f(0, 1e4), "jacket.ramp_time": _f(0, 1e5),
    "internal_hx.enabled": _BOOL, "internal_hx.hA_per_volume": _f(0, 1e7),
    "internal_hx.follows_jacket": _BOOL,
    "feed.T": _f(200, 400), "feed.P": _f(1e4, 1e6), "feed.RH": _f(0, 1), "feed.y_co2_dry": _f(0, 1),
    "feed.velocity": _f(1e-6, 10),
    "feed.profile": _c("uniform", "cosine_theta", "linear_r", "central_disc", "sector"),
    "feed.profile_amplitude": _f(0, 0.99), "feed.profile_theta0_deg": _f(0, 360),
    "feed.profile_radius_fraction": _f(0.05, 1.0), "feed.profile_sector_deg": _f(1.0, 360.0),
    "ambient.T": _f(200, 400), "ambient.P": _f(1e4, 1e6), "ambient.RH": _f(0, 1),
    "ambient.y_co2_dry": _f(0, 1),
    "operation.P_vacuum": _f(100, 1e5), "operation.T_heat": _f(273, 600),
    "operation.T_cool": _f(200, 600), "operation.T_adsorption_jacket": _f(200, 600),
    "operation.valve_time_constant": _f(1e-3, 1e4), "operation.vent_model": _c("valve", "ramp"),
    "operation.vent_conductance": _f(1e-6, 1e4), "operation.vacuum_model": _c("pump", "ramp"),
    "operation.pump_speed_per_area": _f(1e-7, 100.0), "operation.vacuum_check_valve": _BOOL,
    "ports.V1": _c("top", "bottom"), "ports.V2": _c("top", "bottom"), "ports.V3": _c("top", "bottom"),
    "energy.eta_vacuum": _f(0.01, 1), "energy.vacuum_efficiency_model": _c("constant", "krishnamurthy"),
    "energy.eta_blower": _f(0.01, 1), "energy.gamma_gas": _f(1.0, 1.7),
    "energy.pump_handles_water": _BOOL, "energy.eta_turbine": _f(0, 1),
    "energy.T_low_equiv": _f(200, 400),
    "solver.method": _c("trbdf2", "scipy_bdf", "rk4"), "solver.scheme": _c("sdirk2", "trbdf2"),
    "solver.rtol": _f(1e-10, 1e-1), "solver.atol_c_co2": _f(1e-14, 1),
    "solver.atol_c_h2o": _f(1e-14, 1), "solver.atol_c_inert": _f(1e-14, 10),
    "solver.atol_q": _f(1e-14, 1), "solver.atol_T": _f(1e-8, 10), "solver.h_initial": _f(1e-10, 100),
    "solver.h_max_fraction": _f(1e-5, 1), "solver.max_newton_iter": _i(2, 50),
    "solver.preconditioner": _c("theta_block", "full"), "solver.precond_theta_rings": _i(0, 1000),
    "solver.refactor_ratio": _f(1.05, 100.0), "solver.gmres_refresh": _i(1, 100000),
    "solver.newton_mode": _c("hybrid", "krylov", "modified"), "solver.newton_kappa": _f(1e-4, 1),
    "solver.conservation_tol": _f(0.0, 1.0), "solver.jacobian_max_age": _i(1, 10000),
    "solver.limiter": _c("van_albada", "van_albada_tvd", "van_leer", "minmod", "upwind"),
    "solver.max_steps_per_step": _i(10, 10 ** 8), "solver.outlet_record_every": _i(1, 10 ** 6),
    "sensitivity.relative_step": _f(1e-4, 0.9), "sensitivity.cycles": _i(1, 200),
    "sensitivity.grid": _ANY, "sensitivity.processes": _i(1, 256), "sensitivity.parameters": _LIST,
    "sensitivity.kpis": _LIST,
}

STEP_TYPES = ("adsorption", "blowdown", "heating", "cooling", "pressurization")
VALVES = ("V1", "V2", "V3")
VALVE_STATES = ("CLOSED", "OPEN", "OPEN_ATMOSPHERIC", "OPEN_VACUUM")
STOP_CRITERIA = ("outlet_ratio", "end_saturation", "min_temperature", "max_temperature")


class ConfigError(ValueError):
    """Invalid configuration (unknown key, wrong type, out of range, inconsistent cycle)."""


class Config:
    """Read-only nested attribute view over a validated configuration dict."""

    __slots__ = ("_data",)

    def __init__(self, data: Dict[str, Any]):
        object.__setattr__(self, "_data", data)

    def __getattr__(self, name):
        try:
            v = self._data[name]
        except KeyError as exc:
            raise AttributeError(name) from exc
        return Config(v) if isinstance(v, dict) else v

    def __setattr__(self, key, value):
        raise AttributeError("Config is read-only; use load_config(overrides=...) or with_overrides")

    def __getitem__(self, name):
        return self.__getattr__(name)

    def get(self, dotted: str, default=None):
        node: Any = self._data
        for part in dotted.split("."):
            if not (isinstance(node, dict) and part in node):
                return default
            node = node[part]
        return node

    def to_dict(self) -> Dict[str, Any]:
        return copy.deepcopy(self._data)

    def __repr__(self):
        return f"Config({list(self._data)})"


def deep_merge(base: Dict, upd: Optional[Dict]) -> Dict:
    out = copy.deepcopy(base)
    for k, v in (upd or {}).items():
        out[k] = deep_merge(out[k], v) if isinstance(v, dict) and isinstance(out.get(k), dict) \
            else copy.deepcopy(v)
    return out


def set_dotted(d: Dict, dotted: str, value: Any) -> None:
    *head, last = dotted.split(".")
    for p in head:
        if not isinstance(d.get(p), dict):
            d[p] = {}
        d = d[p]
    d[last] = value


def get_dotted(d: Dict, dotted: str):
    for p in dotted.split("."):
        d = d[p]
    return d


def _leaves(d: Dict, prefix: str = "") -> Dict[str, Any]:
    out = {}
    for k, v in d.items():
        key = f"{prefix}.{k}" if prefix else str(k)
        if isinstance(v, dict) and key not in KEYS:
            out.update(_leaves(v, key))
        else:
            out[key] = v
    return out


def parse_override(text: str) -> Tuple[str, Any]:
    if "=" not in text:
        raise ConfigError(f"override '{text}' must look like key=value")
    k, v = text.split("=", 1)
    return k.strip(), yaml.safe_load(v)


def _coerce(key: str, v: Any) -> Any:
    typ, lo, hi, choices, nullable = KEYS[key]
    if v is None:
        if nullable:
            return None
        raise ConfigError(f"{key}: value is required")
    if typ in (F, I) and isinstance(v, str):
        try:                      # YAML 1.1 reads '2.0e4' as a string
            v = float(v)
        except ValueError:
            raise ConfigError(f"{key}: expected a number, got {v!r}")
    if typ == F:
        if isinstance(v, bool) or not isinstance(v, (int, float)):
            raise ConfigError(f"{key}: expected a number, got {v!r}")
        v = float(v)
        if not math.isfinite(v):
            raise ConfigError(f"{key}: must be finite")
    elif typ == I:
        if isinstance(v, bool) or not isinstance(v, (int, float)) or int(v) != v:
            raise ConfigError(f"{key}: expected an integer, got {v!r}")
        v = int(v)
    elif typ == B and not isinstance(v, bool):
        raise ConfigError(f"{key}: expected true/false, got {v!r}")
    elif typ == S and not isinstance(v, str):
        raise ConfigError(f"{key}: expected a string, got {v!r}")
    elif typ == L:
        if not isinstance(v, (list, tuple)):
            raise ConfigError(f"{key}: expected a list, got {v!r}")
        v = list(v)
    if choices is not None and v not in choices:
        raise ConfigError(f"{key}: {v!r} not in {choices}")
    if typ in (F, I):
        if lo is not None and v < lo:
            raise ConfigError(f"{key} = {v} below allowed minimum {lo}")
        if hi is not None and v > hi:
            raise ConfigError(f"{key} = {v} above allowed maximum {hi}")
    return v


def _ref(raw: Dict, value: Any) -> Any:
    if isinstance(value, str) and "." in value:
        try:
            return float(get_dotted(raw, value))
        except (KeyError, TypeError, ValueError):
            raise ConfigError(f"cannot resolve reference '{value}'")
    return value


def _is_number(x: Any) -> bool:
    try:
        float(x)
        return True
    except (TypeError, ValueError):
        return False


def _cycle(raw: Dict) -> List[Dict[str, Any]]:
    steps = raw.get("cycle")
    if not isinstance(steps, list) or not steps:
        raise ConfigError("cycle: must be a non-empty list of steps")
    allowed = {"name", "type", "duration", "valves", "jacket", "stop_on", "product_port", "notes"}
    out, names = [], set()
    for i, st in enumerate(steps):
        if not isinstance(st, dict):
            raise ConfigError(f"cycle[{i}] must be a mapping")
        st = copy.deepcopy(st)
        if set(st) - allowed:
            raise ConfigError(f"cycle[{i}]: unknown keys {sorted(set(st) - allowed)}")
        name = st.get("name") or st.get("type")
        if name in names:
            raise ConfigError(f"cycle: duplicate step name '{name}'")
        names.add(name)
        st["name"] = name
        if st.get("type") not in STEP_TYPES:
            raise ConfigError(f"cycle[{i}].type must be one of {STEP_TYPES}")
        dur = _ref(raw, st.get("duration"))
        if not isinstance(dur, (int, float)) or isinstance(dur, bool) or dur <= 0:
            raise ConfigError(f"cycle[{i}].duration must be > 0 s")
        st["duration"] = float(dur)
        valves = st.get("valves") or {}
        if set(valves) - set(VALVES):
            raise ConfigError(f"cycle[{i}].valves: unknown valves {sorted(set(valves) - set(VALVES))}")
        vv = {}
        for v in VALVES:
            s = valves.get(v, "CLOSED")
            if s not in VALVE_STATES:
                raise ConfigError(f"cycle[{i}].valves.{v} = {s!r} not in {VALVE_STATES}")
            vv[v] = s
        st["valves"] = vv
        jk = st.get("jacket") or {"mode": "adiabatic"}
        mode = jk.get("mode", "temperature")
        if mode not in ("temperature", "adiabatic"):
            raise ConfigError(f"cycle[{i}].jacket.mode must be temperature|adiabatic")
        jout = {"mode": mode}
        if mode == "temperature":
            src = jk.get("T_source", jk.get("T"))
            T = _ref(raw, src)
            if not isinstance(T, (int, float)) or not (150 < T < 900):
                raise ConfigError(f"cycle[{i}].jacket.T invalid ({T!r})")
            jout["T"] = float(T)
            if isinstance(src, str) and "." in src and not _is_number(src):
                jout["T_source"] = src
        st["jacket"] = jout
        stop = st.get("stop_on")
        if stop is not None:
            crit = stop.get("criterion")
            if crit not in STOP_CRITERIA:
                raise ConfigError(f"cycle[{i}].stop_on.criterion invalid: {crit!r}")
            vsrc = stop.get("value_source", stop.get("value"))
            new = {"criterion": crit, "value": float(_ref(raw, vsrc))}
            if isinstance(vsrc, str) and "." in vsrc and not _is_number(vsrc):
                new["value_source"] = vsrc
            stop = new
        st["stop_on"] = stop
        pp = st.get("product_port") or "V3"
        if pp not in VALVES + ("none",):
            raise ConfigError(f"cycle[{i}].product_port must be V1/V2/V3/none")
        st["product_port"] = pp
        out.append(st)
    ports = raw["ports"]
    for st in out:
        per_end: Dict[str, List[str]] = {"top": [], "bottom": []}
        for v, s in st["valves"].items():
            if s != "CLOSED":
                per_end[ports[v]].append(v)
            if s == "OPEN" and v != "V1":
                raise ConfigError(f"step '{st['name']}': state OPEN (feed blower) is only valid for V1")
        for end, lst in per_end.items():
            if len(lst) > 1:
                raise ConfigError(f"step '{st['name']}': valves {lst} are open at the same end ({end})")
    return out


def validate(raw: Dict[str, Any]) -> Config:
    raw = copy.deepcopy(raw)
    raw.pop("derived", None)
    unknown = [k for k in _leaves({k: v for k, v in raw.items() if k != "cycle"}) if k not in KEYS]
    if unknown:
        raise ConfigError(f"unknown configuration keys (typo?): {unknown}")
    for key in KEYS:
        try:
            v = get_dotted(raw, key)
        except (KeyError, TypeError):
            raise ConfigError(f"missing configuration key: {key}")
        set_dotted(raw, key, _coerce(key, v))
    s = raw["sorbent"]
    if s["total_voidage"] < s["bed_voidage"]:
        raise ConfigError("sorbent.total_voidage must be >= sorbent.bed_voidage")
    if raw["run"]["min_cycles"] > raw["run"]["max_cycles"]:
        raise ConfigError("run.min_cycles > run.max_cycles")
    th, zc = raw["jacket"]["theta_coverage_deg"], raw["jacket"]["z_coverage"]
    if len(th) != 2 or not (0 <= th[0] < th[1] <= 360):
        raise ConfigError("jacket.theta_coverage_deg must be [start, end] with 0 <= start < end <= 360")
    if len(zc) != 2 or not (0 <= zc[0] < zc[1] <= 1):
        raise ConfigError("jacket.z_coverage must be [start, end] with 0 <= start < end <= 1")
    if raw["operation"]["P_vacuum"] >= raw["ambient"]["P"]:
        raise ConfigError("operation.P_vacuum must be below ambient.P")
    if raw["grid"]["Ntheta"] == 2:
        raise ConfigError("grid.Ntheta = 2 is degenerate; use 1 (axisymmetric) or >= 3")
    raw["cycle"] = _cycle(raw)
    eb, et = s["bed_voidage"], s["total_voidage"]
    R, Lb = raw["geometry"]["bed_radius"], raw["geometry"]["bed_length"]
    d = {"particle_porosity": (et - eb) / (1.0 - eb),
         "bulk_density": s["particle_density"] * (1.0 - eb),
         "bed_volume": math.pi * R ** 2 * Lb,
         "cross_section": math.pi * R ** 2}
    d["sorbent_mass"] = d["bulk_density"] * d["bed_volume"]
    raw["derived"] = d
    return Config(raw)


def load_raw_defaults() -> Dict[str, Any]:
    with open(DEFAULTS_PATH) as fh:
        return yaml.safe_load(fh)


def load(user_yaml: Optional[str] = None, overrides: Optional[Sequence[str]] = None,
         override_dict: Optional[Dict[str, Any]] = None) -> Config:
    raw = load_raw_defaults()
    if user_yaml:
        with open(user_yaml) as fh:
            user = yaml.safe_load(fh) or {}
        user.pop("derived", None)
        if "cycle" in user:
            raw["cycle"] = user.pop("cycle")
        raw = deep_merge(raw, user)
    for k, v in (override_dict or {}).items():
        set_dotted(raw, k, v)
    for ov in overrides or []:
        set_dotted(raw, *parse_override(ov))
    return validate(raw)

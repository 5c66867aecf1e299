"""Configuration loading and validation for the TVSA step engine.

Accepts and rejects the same inputs as the upstream loader (config/schema.py:529-594 and
config/parameters.py:25-44): the defaults are merged with an optional user YAML (a user ``cycle``
replaces the default one), a dict of dotted overrides and ``key=value`` strings; every leaf key
must be known and is checked against its spec (type, range, choices); cycle steps are normalised
(valve states, jacket, stop criterion, product port; string references such as
``operation.T_heat`` are resolved against the configuration and remembered, so that a later
override of the referenced key reaches the step); cross-field rules and derived quantities follow;
the result is a read-only attribute view.

KEYS gives, per dotted key, ``(type, lo, hi, choices, nullable)``; the ranges are the upstream ones
(config/schema.py:58-281).
"""
from __future__ import annotations

import copy
import math
import os
from dataclasses import dataclass
from typing import Any, Callable, Dict, Iterator, List, Optional, Sequence, Tuple

import yaml

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULTS_PATH = os.path.join(HERE, "defaults.yaml")

F, I, S, B, L, A = "float", "int", "str", "bool", "list", "any"


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
STEP_FIELDS = frozenset({"name", "type", "duration", "valves", "jacket", "stop_on", "product_port", "notes"})


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
        try:
            return get_dotted(self._data, dotted)
        except (KeyError, TypeError):
            return default

    def to_dict(self) -> Dict[str, Any]:
        return copy.deepcopy(self._data)

    def __repr__(self):
        return f"Config({list(self._data)})"


# ------------------------------------------------------------------ nested-dict helpers
def get_dotted(tree: Dict, dotted: str):
    node = tree
    for part in dotted.split("."):
        node = node[part]
    return node


def set_dotted(tree: Dict, dotted: str, value: Any) -> None:
    """Set tree[a][b]...[z] = value for dotted = "a.b...z", creating (or replacing non-dict)
    intermediate levels."""
    parts = dotted.split(".")
    node = tree
    for part in parts[:-1]:
        child = node.get(part)
        if not isinstance(child, dict):
            child = node[part] = {}
        node = child
    node[parts[-1]] = value


def deep_merge(base: Dict, update: Optional[Dict]) -> Dict:
    """A copy of base with update merged in: sub-dicts merge recursively, other values replace."""
    merged = copy.deepcopy(base)
    stack = [(merged, update or {})]
    while stack:
        dst, src = stack.pop()
        for key, value in src.items():
            if isinstance(value, dict) and isinstance(dst.get(key), dict):
                stack.append((dst[key], value))
            else:
                dst[key] = copy.deepcopy(value)
    return merged


def _leaf_keys(tree: Dict, prefix: str = "") -> Iterator[str]:
    """Dotted paths of the leaves (a spec'd key whose value is a dict counts as a leaf)."""
    for key, value in tree.items():
        path = f"{prefix}.{key}" if prefix else str(key)
        if isinstance(value, dict) and path not in SPECS:
            yield from _leaf_keys(value, path)
        else:
            yield path


def parse_override(text: str) -> Tuple[str, Any]:
    """``"a.b=value"`` -> ("a.b", YAML value)."""
    key, sep, value = text.partition("=")
    if not sep:
        raise ConfigError(f"override {text!r} is not of the form key=value")
    return key.strip(), yaml.safe_load(value)


# ------------------------------------------------------------------ per-key specs
@dataclass(frozen=True)
class Spec:
    kind: str
    lo: Optional[float] = None
    hi: Optional[float] = None
    choices: Optional[Tuple] = None
    nullable: bool = False

    def check(self, key: str, value: Any) -> Any:
        """The value converted to the spec's type, or ConfigError."""
        if value is None:
            if self.nullable:
                return None
            raise ConfigError(f"{key} must be set")
        value = self._typed(key, value)
        if self.choices is not None and value not in self.choices:
            raise ConfigError(f"{key} = {value!r}: expected one of {list(self.choices)}")
        if self.kind in (F, I):
            if self.lo is not None and value < self.lo:
                raise ConfigError(f"{key} = {value} is below its minimum {self.lo}")
            if self.hi is not None and value > self.hi:
                raise ConfigError(f"{key} = {value} is above its maximum {self.hi}")
        return value

    def _typed(self, key, value):
        kind = self.kind
        if kind in (F, I):
            if isinstance(value, str):
                # YAML 1.1 leaves e.g. 2.0e4 (no sign in the exponent) a string
                try:
                    value = float(value)
                except ValueError:
                    raise ConfigError(f"{key} = {value!r} is not a number") from None
            if isinstance(value, bool) or not isinstance(value, (int, float)):
                raise ConfigError(f"{key} = {value!r} is not a number")
            if kind == F:
                if not math.isfinite(value):
                    raise ConfigError(f"{key} = {value!r} is not finite")
                return float(value)
            if int(value) != value:
                raise ConfigError(f"{key} = {value!r} is not an integer")
            return int(value)
        if kind == B and not isinstance(value, bool):
            raise ConfigError(f"{key} = {value!r}: expected true or false")
        if kind == S and not isinstance(value, str):
            raise ConfigError(f"{key} = {value!r}: expected a string")
        if kind == L:
            if not isinstance(value, (list, tuple)):
                raise ConfigError(f"{key} = {value!r}: expected a list")
            return list(value)
        return value


SPECS: Dict[str, Spec] = {key: Spec(*row) for key, row in KEYS.items()}


# ------------------------------------------------------------------ cycle steps
def _is_reference(value) -> bool:
    """A string naming another configuration key (it holds a dot and is not a number)."""
    if not isinstance(value, str) or "." not in value:
        return False
    try:
        float(value)
    except ValueError:
        return True
    return False


class _StepReader:
    """Normalises one cycle step of the raw configuration."""

    def __init__(self, raw: Dict, index: int):
        self.raw, self.where = raw, f"cycle[{index}]"

    def fail(self, what: str):
        raise ConfigError(f"{self.where}: {what}")

    def number(self, value, what: str) -> float:
        """A number, or the value of the key a reference string names."""
        if _is_reference(value):
            try:
                value = get_dotted(self.raw, value)
            except (KeyError, TypeError):
                self.fail(f"{what} refers to {value!r}, which is not a configuration key")
        if isinstance(value, str):
            try:
                value = float(value)
            except ValueError:
                self.fail(f"{what} = {value!r} is not a number")
        if isinstance(value, bool) or not isinstance(value, (int, float)):
            self.fail(f"{what} = {value!r} is not a number")
        return float(value)

    def read(self, step: Any) -> Dict[str, Any]:
        if not isinstance(step, dict):
            self.fail("a step must be a mapping")
        extra = sorted(set(step) - STEP_FIELDS)
        if extra:
            self.fail(f"unknown step fields {extra}")
        out = copy.deepcopy(step)
        out["name"] = step.get("name") or step.get("type")
        if step.get("type") not in STEP_TYPES:
            self.fail(f"type must be one of {list(STEP_TYPES)}")
        if step.get("duration") is None:
            self.fail("a duration is required")
        out["duration"] = self.number(step.get("duration"), "duration")
        if out["duration"] <= 0.0:
            self.fail("duration must be positive")
        out["valves"] = self.valves(step.get("valves") or {})
        out["jacket"] = self.jacket(step.get("jacket") or {"mode": "adiabatic"})
        out["stop_on"] = self.stop(step.get("stop_on"))
        port = step.get("product_port") or "V3"
        if port not in VALVES + ("none",):
            self.fail("product_port must be V1, V2, V3 or none")
        out["product_port"] = port
        return out

    def valves(self, given: Dict) -> Dict[str, str]:
        extra = sorted(set(given) - set(VALVES))
        if extra:
            self.fail(f"unknown valves {extra}")
        states = {}
        for v in VALVES:
            states[v] = given.get(v, "CLOSED")
            if states[v] not in VALVE_STATES:
                self.fail(f"valve {v} state {states[v]!r} is not one of {list(VALVE_STATES)}")
        return states

    def jacket(self, given: Dict) -> Dict[str, Any]:
        mode = given.get("mode", "temperature")
        if mode == "adiabatic":
            return {"mode": mode}
        if mode != "temperature":
            self.fail("jacket mode must be temperature or adiabatic")
        source = given.get("T_source", given.get("T"))
        if source is None:
            self.fail("a temperature jacket needs T")
        T = self.number(source, "jacket T")
        if not 150.0 < T < 900.0:
            self.fail(f"jacket T = {T} K is out of range")
        out = {"mode": mode, "T": T}
        if _is_reference(source):
            out["T_source"] = source
        return out

    def stop(self, given: Optional[Dict]) -> Optional[Dict[str, Any]]:
        if given is None:
            return None
        if given.get("criterion") not in STOP_CRITERIA:
            self.fail(f"stop_on criterion must be one of {list(STOP_CRITERIA)}")
        source = given.get("value_source", given.get("value"))
        out = {"criterion": given["criterion"], "value": self.number(source, "stop_on value")}
        if _is_reference(source):
            out["value_source"] = source
        return out


def _read_cycle(raw: Dict) -> List[Dict[str, Any]]:
    steps = raw.get("cycle")
    if not isinstance(steps, list) or not steps:
        raise ConfigError("cycle must be a non-empty list of steps")
    cycle = [_StepReader(raw, i).read(st) for i, st in enumerate(steps)]
    seen = set()
    for st in cycle:
        if st["name"] in seen:
            raise ConfigError(f"cycle: two steps are named {st['name']!r}")
        seen.add(st["name"])
        open_at = {"top": [], "bottom": []}
        for v, state in st["valves"].items():
            if state == "OPEN" and v != "V1":
                raise ConfigError(f"step {st['name']!r}: only V1 (the feed blower) can be OPEN")
            if state != "CLOSED":
                open_at[raw["ports"][v]].append(v)
        for end, vs in open_at.items():
            if len(vs) > 1:
                raise ConfigError(f"step {st['name']!r}: valves {vs} are open at the same end ({end})")
    return cycle


# ------------------------------------------------------------------ whole configuration
def _interval(x, lo, hi) -> bool:
    return len(x) == 2 and lo <= x[0] < x[1] <= hi


RULES: List[Tuple[Callable[[Dict], bool], str]] = [
    (lambda r: r["sorbent"]["total_voidage"] >= r["sorbent"]["bed_voidage"],
     "sorbent.total_voidage must not be below sorbent.bed_voidage"),
    (lambda r: r["run"]["min_cycles"] <= r["run"]["max_cycles"], "run.min_cycles exceeds run.max_cycles"),
    (lambda r: _interval(r["jacket"]["theta_coverage_deg"], 0, 360),
     "jacket.theta_coverage_deg must be [start, end] with 0 <= start < end <= 360"),
    (lambda r: _interval(r["jacket"]["z_coverage"], 0, 1),
     "jacket.z_coverage must be [start, end] with 0 <= start < end <= 1"),
    (lambda r: r["operation"]["P_vacuum"] < r["ambient"]["P"], "operation.P_vacuum must be below ambient.P"),
    (lambda r: r["grid"]["Ntheta"] != 2, "grid.Ntheta = 2 is degenerate: use 1 (axisymmetric) or at least 3"),
]


def _derived(raw: Dict) -> Dict[str, float]:
    sb, geo = raw["sorbent"], raw["geometry"]
    area = math.pi * geo["bed_radius"] ** 2
    volume = area * geo["bed_length"]
    rho_b = sb["particle_density"] * (1.0 - sb["bed_voidage"])
    return {"particle_porosity": (sb["total_voidage"] - sb["bed_voidage"]) / (1.0 - sb["bed_voidage"]),
            "bulk_density": rho_b, "bed_volume": volume, "cross_section": area, "sorbent_mass": rho_b * volume}


def validate(raw: Dict[str, Any]) -> Config:
    raw = copy.deepcopy(raw)
    raw.pop("derived", None)
    unknown = [k for k in _leaf_keys({k: v for k, v in raw.items() if k != "cycle"}) if k not in SPECS]
    if unknown:
        raise ConfigError(f"unknown configuration keys {unknown}")
    for key, spec in SPECS.items():
        try:
            value = get_dotted(raw, key)
        except (KeyError, TypeError):
            raise ConfigError(f"configuration key {key} is missing") from None
        set_dotted(raw, key, spec.check(key, value))
    for ok, message in RULES:
        if not ok(raw):
            raise ConfigError(message)
    raw["cycle"] = _read_cycle(raw)
    raw["derived"] = _derived(raw)
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
        cycle = user.pop("cycle", None)
        raw = deep_merge(raw, user)
        if cycle is not None:
            raw["cycle"] = cycle
    for key, value in (override_dict or {}).items():
        set_dotted(raw, key, value)
    for text in overrides or []:
        set_dotted(raw, *parse_override(text))
    return validate(raw)

"""One-at-a-time sensitivity of the cycle KPIs on one GPU (upstream analysis/sensitivity.py:77-138).

For every parameter p (a dotted configuration key in ``sensitivity.parameters``) the cycle is run
to cyclic steady state at p+ and p- -- p (1 +- sensitivity.relative_step), or p +- an absolute step
for temperatures and humidities -- and each KPI K gets the normalised sensitivity

    S = (K+ - K-) / (2 K0)  /  ((p+ - p-) / (2 p0)).

As upstream: the runs use ``sensitivity.grid`` and at most ``sensitivity.cycles`` cycles; the
perturbed runs start from the base run's cyclic steady state (in memory here, not through a CSV
file) with at most cycles // 2 + 1 cycles (at least 2).  What changes is where the cases run: all
perturbed cases at once, one HIP stream per worker thread (dac_vtsa_amd.batch), instead of a
process pool.
"""
from __future__ import annotations

import time
from typing import Dict, List, Optional

import numpy as np

# absolute steps of the parameters that are perturbed by an amount, not a fraction
ABS_STEP = {"operation.T_heat": 5.0, "operation.T_cool": 5.0, "operation.T_adsorption_jacket": 2.0,
            "feed.T": 2.0, "feed.RH": 0.05, "ambient.RH": 0.05}


def _param(entry):
    """(key, options) of a ``sensitivity.parameters`` entry: a key, or {key: ..., options}."""
    if isinstance(entry, str):
        return entry, {}
    if isinstance(entry, dict) and "key" in entry:
        return entry["key"], {k: v for k, v in entry.items() if k != "key"}
    raise ValueError(f"sensitivity parameter {entry!r}")


def _flat_overrides(cfg) -> Dict:
    """Every leaf of a configuration as dotted overrides (the cycle as one value)."""
    out = {}
    stack = [("", {k: v for k, v in cfg.to_dict().items() if k not in ("cycle", "derived")})]
    while stack:
        prefix, node = stack.pop()
        for k, v in node.items():
            key = f"{prefix}.{k}" if prefix else k
            if isinstance(v, dict) and key != "sensitivity.grid":
                stack.append((key, v))
            else:
                out[key] = v
    out["cycle"] = cfg.to_dict()["cycle"]
    return out


def perturbations(cfg) -> List[Dict]:
    """[{key, p0, p_plus, p_minus}] for ``sensitivity.parameters``."""
    sc = cfg.sensitivity
    out = []
    for entry in sc.parameters:
        key, opts = _param(entry)
        p0 = float(cfg.get(key))
        if "delta_abs" in opts or key in ABS_STEP:
            d = float(opts.get("delta_abs", ABS_STEP.get(key)))
            pp, pm = p0 + d, p0 - d
        else:
            d = float(opts.get("relative_step", sc.relative_step))
            pp, pm = p0 * (1.0 + d), p0 * (1.0 - d)
        out.append({"key": key, "p0": p0, "p_plus": pp, "p_minus": pm})
    return out


def run_sensitivity(cfg, workers: int = 16, integrator_kw: Optional[Dict] = None, log=None) -> Dict:
    """Base case, then every +- perturbation concurrently on the GPU; returns {"rows": one row per
    (parameter, KPI) as upstream, "base": the base case's KPIs, "cases": per-case summaries, wall
    times, "base_overrides" and "base_state": the base case's flat configuration and end state}."""
    from .batch import run_batch, run_case_to_css
    log = log or (lambda *a, **k: None)
    sc = cfg.sensitivity
    flat = _flat_overrides(cfg)
    grid = dict(cfg.get("sensitivity.grid"))
    base_over = {**flat, "run.max_cycles": int(sc.cycles), "run.min_cycles": min(2, int(sc.cycles)),
                 "run.make_plots": False, "run.snapshots_per_step": 3}
    base_over.update({f"grid.{k}": v for k, v in grid.items()})
    t0 = time.time()
    base = run_case_to_css({"id": "base", "overrides": base_over}, integrator_kw=integrator_kw)
    if not base["ok"]:
        raise RuntimeError(f"sensitivity base case failed: {base['error']}")
    t_base = time.time() - t0
    log(f"sensitivity: base case {base['cycles']} cycles in {t_base:.1f} s, recovery "
        f"{base['kpis']['recovery_pct']:.2f} %")
    warm_cycles = max(2, int(sc.cycles) // 2 + 1)
    pert = perturbations(cfg)
    cases = []
    for p in pert:
        for tag in ("plus", "minus"):
            over = {**base_over, "run.max_cycles": warm_cycles, p["key"]: p["p_" + tag]}
            cases.append({"id": f"{p['key']}|{tag}", "overrides": over})

    def warm(case, **kw):
        return run_case_to_css(case, state=base["state"], **kw)
    t1 = time.time()
    res = run_batch(cases, workers=workers, runner=warm, integrator_kw=integrator_kw,
                    log=lambda m: log("  " + m))
    t_pert = time.time() - t1
    by_id = {r["id"]: r for r in res["cases"]}
    failed = [r["id"] for r in res["cases"] if not r["ok"]]
    rows = []
    for p in pert:
        kp, km = by_id[f"{p['key']}|plus"], by_id[f"{p['key']}|minus"]
        for kpi in sc.kpis:
            K0 = base["kpis"][kpi]
            Kp = kp["kpis"][kpi] if kp["ok"] else np.nan
            Km = km["kpis"][kpi] if km["ok"] else np.nan
            dlnp = (p["p_plus"] - p["p_minus"]) / (2.0 * p["p0"]) if p["p0"] != 0 else np.nan
            S = (Kp - Km) / (2.0 * K0) / dlnp if (K0 not in (0, None) and dlnp) else np.nan
            rows.append({"parameter": p["key"], "p_base": p["p0"], "p_plus": p["p_plus"], "p_minus": p["p_minus"],
                         "kpi": kpi, "K_base": K0, "K_plus": Kp, "K_minus": Km, "normalized_sensitivity": S,
                         "change_pct_plus": 100 * (Kp - K0) / K0 if K0 else np.nan,
                         "change_pct_minus": 100 * (Km - K0) / K0 if K0 else np.nan})
    return {"rows": rows, "base": base["kpis"], "base_cycles": base["cycles"], "failed": failed,
            "cases": [{k: r.get(k) for k in ("id", "ok", "cycles", "css", "wall_s", "sim_s")} for r in res["cases"]],
            "wall_base_s": t_base, "wall_perturbed_s": t_pert, "base_overrides": base_over,
            "base_state": base["state"]}

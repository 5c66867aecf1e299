"""One-at-a-time sensitivity driver (analysis/sensitivity.py:77-138).

CPU: the perturbation table for the default configuration (relative steps and the absolute steps
of temperatures and humidities), and the flat override set that hands the whole configuration to
every case.  GPU: a reduced study -- two parameters, two cycles, the 6x1x10 sensitivity grid --
whose perturbed cases run concurrently on one GPU, warm-started from the base case's end state;
its sensitivities follow the central-difference formula from the case KPIs, and a concurrently
run case reproduces the same case run alone bit for bit.
"""
from __future__ import annotations

import math

import numpy as np
import pytest


def test_perturbation_table_of_the_default_configuration():
    from dac_vtsa_amd.config import load_config
    from dac_vtsa_amd.sensitivity import perturbations
    cfg = load_config()
    rows = {p["key"]: p for p in perturbations(cfg)}
    assert list(rows) == [e if isinstance(e, str) else e["key"] for e in cfg.sensitivity.parameters]
    d = float(cfg.sensitivity.relative_step)
    k = rows["kinetics.k_ldf_co2"]
    assert k["p0"] == float(cfg.kinetics.k_ldf_co2)
    assert k["p_plus"] == pytest.approx(k["p0"] * (1 + d), rel=1e-15)
    assert k["p_minus"] == pytest.approx(k["p0"] * (1 - d), rel=1e-15)
    t = rows["operation.T_heat"]
    assert (t["p_plus"] - t["p0"], t["p0"] - t["p_minus"]) == (5.0, 5.0)
    if "feed.RH" in rows:
        assert rows["feed.RH"]["p_plus"] == pytest.approx(rows["feed.RH"]["p0"] + 0.05)


def test_flat_overrides_rebuild_the_same_configuration():
    from dac_vtsa_amd.config import load_config
    from dac_vtsa_amd.sensitivity import _flat_overrides
    cfg = load_config(override_dict={"grid.Nr": 5, "operation.T_heat": 371.0})
    again = load_config(override_dict=_flat_overrides(cfg))
    a, b = cfg.to_dict(), again.to_dict()
    a.pop("derived", None)
    b.pop("derived", None)
    assert a == b


@pytest.mark.gpu
def test_reduced_sensitivity_study_on_the_gpu():
    from dac_vtsa_amd.batch import run_case_to_css
    from dac_vtsa_amd.config import load_config
    from dac_vtsa_amd.sensitivity import run_sensitivity
    kw = {"gs_sweeps": 2, "conserve": "project"}
    cfg = load_config(override_dict={"sensitivity.cycles": 2, "solver.conservation_tol": 1e-5,
                                     "sensitivity.parameters": ["kinetics.k_ldf_co2", "operation.T_heat"]})
    out = run_sensitivity(cfg, workers=4, integrator_kw=kw)
    assert not out["failed"], out["failed"]
    assert len(out["cases"]) == 4 and all(c["cycles"] >= 1 for c in out["cases"])
    kpis = list(cfg.sensitivity.kpis)
    assert len(out["rows"]) == 2 * len(kpis)
    for r in out["rows"]:
        dlnp = (r["p_plus"] - r["p_minus"]) / (2 * r["p_base"])
        S = (r["K_plus"] - r["K_minus"]) / (2 * r["K_base"]) / dlnp
        assert math.isfinite(r["normalized_sensitivity"]) and r["normalized_sensitivity"] == pytest.approx(S)
    # heating hotter recovers more CO2
    rec = next(r for r in out["rows"] if r["parameter"] == "operation.T_heat" and r["kpi"] == "recovery_pct")
    assert rec["K_plus"] > rec["K_minus"]
    # one perturbed case alone, from the same warm start, gives the same KPIs bit for bit
    flat = out["base_overrides"]
    case = {"id": "alone", "overrides": {**flat, "run.max_cycles": 2, "operation.T_heat": rec["p_plus"]}}
    alone = run_case_to_css(case, integrator_kw=kw, state=out["base_state"])
    assert alone["ok"], alone.get("error")
    assert alone["kpis"]["recovery_pct"] == rec["K_plus"]
    assert np.isfinite(out["base"]["recovery_pct"])

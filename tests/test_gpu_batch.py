"""Concurrent independent cases on one GPU (needs an MI355X): one HIP stream per worker thread.

A case's arithmetic must not depend on what runs beside it, so the concurrent batch has to
reproduce the one-after-another run bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_batch_matches_sequential_runs_exactly():
    from dac_vtsa_amd.batch import draw_cases, run_batch
    cases = draw_cases(4, seed=11)
    for c in cases:              # keep the test short: at most 12x6x40 cells
        o = c["overrides"]
        o["grid.Nr"], o["grid.Nz"] = min(o["grid.Nr"], 12), min(o["grid.Nz"], 40)
        o["grid.Ntheta"] = min(o["grid.Ntheta"], 6) if o["grid.Ntheta"] != 2 else 3
    kw = dict(mode="cycle", state_kind="small", t_end=5.0,
              base_overrides={"solver.conservation_tol": 1e-5},
              integrator_kw={"gs_sweeps": 2, "conserve": "project"})
    seq = run_batch(cases, workers=1, **kw)
    par = run_batch(cases, workers=4, **kw)
    assert not seq["failed"] and not par["failed"], (seq["failed"], par["failed"])
    for a, b in zip(seq["cases"], par["cases"]):
        assert a["sim_s"] == b["sim_s"] and a["grid"] == b["grid"]
        for f in ("c", "q", "T", "Tw"):
            assert np.array_equal(getattr(a["state"], f), getattr(b["state"], f)), (a["id"], f)
        # 5 s steps: the balance rows divide by small step throughputs, so the gate here is loose
        assert a["max_mass_rel_err"] <= 1e-4 and a["max_energy_rel_err"] <= 1e-4
        assert set(a["kpis"]) >= {"recovery_pct", "purity_dry_pct"}
    print(f"sequential {seq['wall_s']:.2f} s, 4 streams {par['wall_s']:.2f} s")


def test_uneven_cases_run_to_cyclic_steady_state():
    """The batched-sensitivity workload reduced to three drawn cases (grids clipped to at most
    6x5x30 cells to keep the test short), each run cycle after cycle from the clean bed to the CSS
    test of validation/cyclic_steady_state.py:22-44 (css 1e-3, CO2 balance 5e-3, 2..12 cycles),
    concurrently on three streams: every case reaches CSS with its recovery, purity and specific
    energy (postprocessing/reporting.py:52-110) in range and every step's balances closed."""
    from dac_vtsa_amd.batch import draw_cases, run_batch, run_case_to_css
    cases = draw_cases(3, seed=23)
    for c in cases:
        o = c["overrides"]
        o["grid.Nr"], o["grid.Nz"] = min(o["grid.Nr"], 6), min(o["grid.Nz"], 30)
        o["grid.Ntheta"] = min(o["grid.Ntheta"], 5) if o["grid.Ntheta"] != 2 else 3
    out = run_batch(cases, workers=3, runner=run_case_to_css, base_overrides={"solver.conservation_tol": 1e-5},
                    integrator_kw={"gs_sweeps": 2, "conserve": "project", "precision": "fp32"})
    assert not out["failed"], [c.get("error") for c in out["cases"]]
    for c in out["cases"]:
        assert c["css"], (c["id"], c["cycles"])
        assert 2 <= c["cycles"] <= 12
        k = c["kpis"]
        assert 0.0 < k["recovery_pct"] <= 100.0 and 0.0 < k["purity_dry_pct"] <= 100.0
        assert k["specific_total_kJ_per_kg"] > 0.0
        assert c["max_mass_rel_err"] <= 1e-4 and c["max_energy_rel_err"] <= 1e-4
    print([(c["grid"], c["cycles"], round(c["wall_s"], 1)) for c in out["cases"]], f"{out['wall_s']:.1f} s")

"""Cycle steps: run_<type>(state_in, cfg, grid, props, step=None, **kw) -> (state_out, StepResult)
(upstream steps/__init__.py:8-9)."""
from ._runner import _make_runner, run_step  # noqa: F401

run_adsorption = _make_runner("adsorption")
run_blowdown = _make_runner("blowdown")
run_heating = _make_runner("heating")
run_cooling = _make_runner("cooling")
run_pressurization = _make_runner("pressurization")

RUNNERS = {"adsorption": run_adsorption, "blowdown": run_blowdown, "heating": run_heating,
           "cooling": run_cooling, "pressurization": run_pressurization}

"""Outlet rows reduced on the device (vtsa_outlet_row into a device ring, one host copy per
ring) against the rows the upstream runner forms on the host from the full diagnostic fields
(steps/_runner.py:104-121), on the same trajectory: 20 s of adsorption (feed end and a valve
end with its outflow part) and of heating (vacuum pump end), from the golden cfg1 bed.  The sums
run in another order than NumPy's, so the rows agree to rounding, not bit for bit."""
import numpy as np
import pytest

from dac_vtsa_amd.config import load_config
from tests._helpers import CFG1, golden, state_from

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("step_index", [0, 2])
def test_device_outlet_rows_match_the_host_rows(step_index):
    from dac_vtsa_amd.core.grid import Grid
    from dac_vtsa_amd.properties import get_backend
    from dac_vtsa_amd.steps import run_step
    st = state_from(golden("runner_cfg1.npz"))
    cfg = load_config(override_dict=CFG1)
    g, props = Grid.from_config(cfg), get_backend(cfg)
    step = dict(cfg.cycle[step_index])
    step["duration"] = 20.0
    step.pop("stop_on", None)
    _, host = run_step(st, cfg, g, props, step, record_outlet="host")
    _, dev = run_step(st, cfg, g, props, step, record_outlet=True)
    assert len(dev.outlet) == len(host.outlet) > 2
    for a, b in zip(dev.outlet, host.outlet):
        assert list(a) == list(b)
        for k, vb in b.items():
            va = a[k]
            if isinstance(vb, str) or k in ("cycle", "step"):
                assert va == vb, k
            else:
                assert va == pytest.approx(vb, rel=1e-11, abs=1e-13 * max(1.0, abs(vb))), (k, va, vb)

"""The cycle harness keeps the bed in device memory between steps and re-initialises one device
model per grid in place (vtsa_reset) instead of building one per step (upstream
vtsa_cycle.py:80-86 rebuilds the model and hands the state on through host memory).  Neither
may change a number: a cycle run with both, and one run with a fresh model per step and the
state uploaded from its host arrays, must agree bit for bit (needs an MI355X)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GRID = {"grid.Nr": 4, "grid.Ntheta": 3, "grid.Nz": 20, "solver.conservation_tol": 1e-5}


def _cycle(reuse):
    from dac_vtsa_amd import device as dv
    from dac_vtsa_amd.config import load_config
    from dac_vtsa_amd.core import residual
    from dac_vtsa_amd.core.grid import Grid
    from dac_vtsa_amd.core.state import initial_state
    from dac_vtsa_amd.cycle import run_cycle
    from dac_vtsa_amd.properties import get_backend
    cfg = load_config(override_dict=GRID)
    for st in cfg.cycle:                       # a short cycle: every step type, 30 s each
        st["duration"] = min(float(st["duration"]), 30.0)
    grid = Grid.from_config(cfg)
    props = get_backend(cfg)
    keep = residual.HANDLE_POOL_MAX
    residual.release_handles()
    residual.HANDLE_POOL_MAX = keep if reuse else 0
    try:
        state = initial_state(cfg, grid, props)
        ends = []
        for n in (1, 2):
            # reuse: the second cycle also starts from the device copy of the first one's end
            state, steps, _ = run_cycle(state, cfg, grid, props, n, device_handoff=reuse,
                                        integrator_kw={"gs_sweeps": 2, "conserve": "project"})
            if not reuse:
                state = state.copy()           # a copy carries no device copy: uploaded again
            ends.append((state, steps))
        dv.synchronize()
    finally:
        residual.HANDLE_POOL_MAX = keep
    return ends


def test_device_handoff_and_model_reuse_change_nothing():
    from dac_vtsa_amd.core import residual
    a = _cycle(True)
    assert len(residual._handle_pool) >= 1     # the models of the run left their handle for reuse
    b = _cycle(False)
    for (sa, ra), (sb, rb) in zip(a, b):
        for f in ("c", "q", "T", "Tw"):
            assert np.array_equal(getattr(sa, f), getattr(sb, f)), f
        for x, y in zip(ra, rb):
            assert x.acc == y.acc and x.inv_end == y.inv_end and x.duration == y.duration
            assert [r["t_step_s"] for r in x.outlet] == [r["t_step_s"] for r in y.outlet]
    residual.release_handles()
    assert not residual._handle_pool


def test_device_energy_round_trip_equals_the_host_state_round_trip():
    """vtsa_state_roundtrip(y) == pack(to_state(y)) bit for bit on a seeded mid-cycle bed."""
    from dac_vtsa_amd import _abi
    from dac_vtsa_amd import device as dv
    from dac_vtsa_amd.config import load_config
    from dac_vtsa_amd.core.grid import Grid
    from dac_vtsa_amd.core.residual import TVSAModel
    from dac_vtsa_amd.properties import get_backend
    from dac_vtsa_amd.seeds import seeded_state
    cfg = load_config(override_dict=GRID)
    grid = Grid.from_config(cfg)
    st = seeded_state(grid, "midcycle", 3)
    m = TVSAModel(cfg, grid, get_backend(cfg), cfg.cycle[2], st)
    y = m.pack(st)
    y[m.off_e:m.off_w] *= 1.0 + 1e-3 * np.sin(np.arange(grid.ncell))   # e not of the form energy(T)
    want = m.pack(m.to_state(y, st, 0.0, "x"))
    yd = dv.to_dev(y)
    _abi.call("vtsa_state_roundtrip", m.handle, dv.ptr(yd), dv.stream())
    got = dv.to_host(yd)
    m.close()
    assert not np.array_equal(want[m.off_e:m.off_w], y[m.off_e:m.off_w])   # the round trip moves bits
    assert np.array_equal(got, want)

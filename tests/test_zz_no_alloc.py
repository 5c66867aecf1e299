"""The allocation-free entry point vtsa_rms_noalloc against vtsa_rms (needs an MI355X).

Both run the same reduction kernel on the same grid, so the results must be bit-identical: two
calls in a row, the second on a smaller input, so the shared scratch array's reuse (its counter
left at zero by the first call) is exercised."""
import ctypes as C

import numpy as np
import pytest

from dac_vtsa_amd.config import load_config

pytestmark = pytest.mark.gpu


def test_rms_noalloc_matches_rms_on_two_calls():
    from dac_vtsa_amd import _abi
    from dac_vtsa_amd import device as dv
    from dac_vtsa_amd.core.grid import Grid
    from dac_vtsa_amd.core.residual import TVSAModel
    from dac_vtsa_amd.core.state import initial_state
    from dac_vtsa_amd.properties import get_backend
    cfg = load_config(override_dict={"grid.Nr": 3, "grid.Ntheta": 4, "grid.Nz": 6})
    g = Grid.from_config(cfg)
    props = get_backend(cfg)
    m = TVSAModel(cfg, g, props, cfg.cycle[0], initial_state(cfg, g, props))
    L = _abi.lib()
    rng = np.random.default_rng(7)
    out = dv.zeros(1)
    for n, with_s in ((300_001, True), (1_234, True), (77, False)):
        a = rng.standard_normal(n) * 10.0 ** rng.uniform(-3, 3, n)
        s = rng.uniform(0.5, 2.0, n)
        ad, sd = dv.to_dev(a), dv.to_dev(s) if with_s else None
        ref = C.c_double()
        _abi.call("vtsa_rms", m.handle, dv.ptr(ad), dv.ptr(sd), n, C.byref(ref), dv.stream())
        rc = L.vtsa_rms_noalloc(dv.ptr(ad), dv.ptr(sd), n, dv.ptr(out), dv.stream())
        assert rc == 0, L.vtsa_last_error()
        got = float(dv.to_host(out)[0])
        assert got == ref.value, (n, got, ref.value)
        host = np.sqrt(np.mean((a / s) ** 2)) if with_s else np.sqrt(np.mean(a * a))
        assert abs(got - host) <= 1e-12 * host
    assert L.vtsa_rms_noalloc(dv.ptr(out), None, 0, dv.ptr(out), None) == -2
    assert L.vtsa_rms_noalloc(dv.ptr(out), None, (1 << 30) + 1, dv.ptr(out), None) == -2
    assert L.vtsa_rms_noalloc(None, None, 10, dv.ptr(out), None) == -1
    m.close()

"""Learned stage guess on the device vs its CPU restatement (needs an MI355X).

The network runs in BF16 on the matrix cores with FP32 accumulation; the oracle rounds to BF16
at the same points and accumulates in float64, so the two differ by accumulation order only
(occasionally one BF16 ulp of an activation).  Tolerances below are set for that: outputs 1e-2
of their maximum, per-layer gradients 2e-2 in relative Frobenius norm, Adam steps 1e-5.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(rows, seed=3):
    from oracle import mlp
    rng = np.random.default_rng(seed)
    p = mlp.init_params(seed) + (0.05 * rng.standard_normal(mlp.NPARAM)).astype(np.float32)
    feat = (rng.standard_normal((rows, 32)) * 3.0 + 1.0).astype(np.float32)
    norm = np.concatenate([np.full(32, 1.0), np.full(32, 1.0 / 3.0), rng.standard_normal(16),
                           0.5 + rng.random(16)]).astype(np.float32)
    target = rng.standard_normal((rows, 8)).astype(np.float32)
    return mlp, p, feat, norm, target


def _net(rows, p, norm):
    from dac_vtsa_amd.learned import StageNet
    return StageNet(rows, params=p, norm=norm)


@pytest.mark.parametrize("rows", [1, 37, 1000, 4099])
def test_forward_matches_oracle(rows):
    from dac_vtsa_amd import device as dv
    mlp, p, feat, norm, _ = _case(rows)
    net = _net(rows, p, norm)
    f = dv.to_dev(feat.ravel(), np.float32)
    out = dv.empty(rows * 16, np.float32)
    net.forward(f, rows, out)
    got = out.numpy().reshape(rows, 16)
    ref, _ = mlp.forward(p, feat, norm, rounding=True)
    assert np.max(np.abs(got - ref)) <= 1e-2 * np.abs(ref).max(), np.max(np.abs(got - ref))
    assert np.linalg.norm(got - ref) <= 3e-3 * np.linalg.norm(ref)
    net.close()


@pytest.mark.parametrize("rows", [37, 1000, 8000])
def test_gradient_matches_oracle(rows):
    """Loss and every layer's gradient (forward, loss, backward data pass, weight products)."""
    from dac_vtsa_amd import device as dv
    mlp, p, feat, norm, target = _case(rows, seed=rows)
    net = _net(rows, p, norm)
    f = dv.to_dev(feat.ravel(), np.float32)
    t = dv.to_dev(target.ravel(), np.float32)
    loss, g = net.train_step(f, t, rows, 0.0, want_loss=True, want_grad=True)
    assert abs(loss - mlp.loss(p, feat, target, norm)) <= 1e-3 * mlp.loss(p, feat, target, norm)
    ref = mlp.backward(p, feat, target, norm, rounding=True)
    G, R = mlp.unflatten(g.astype(np.float64)), mlp.unflatten(ref)
    for k in R:
        err = np.linalg.norm(G[k] - R[k]) / max(np.linalg.norm(R[k]), 1e-30)
        assert err < 2e-2, (k, err)
    assert np.array_equal(net.get_params(), p)      # lr = 0 leaves the parameters alone
    net.close()


def test_adam_steps_match_oracle():
    from dac_vtsa_amd import device as dv
    mlp, p, feat, norm, target = _case(500, seed=11)
    net = _net(500, p, norm)
    f = dv.to_dev(feat.ravel(), np.float32)
    t = dv.to_dev(target.ravel(), np.float32)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    q = p.copy()
    for k in (1, 2, 3):
        _, g = net.train_step(f, t, 500, 1e-3, want_grad=True)
        q, m, v = mlp.adam(q, g, m, v, k, 1e-3)
        np.testing.assert_allclose(net.get_params(), q, rtol=1e-5, atol=1e-7)
    net.close()


def test_training_fits_a_smooth_map():
    """Device training reduces the loss of a smooth nonlinear regression well below its start."""
    from dac_vtsa_amd import device as dv
    from dac_vtsa_amd.learned import StageNet, identity_norm
    rng = np.random.default_rng(7)
    rows, nb = 2048, 4
    X = rng.standard_normal((nb, rows, 32)).astype(np.float32)
    Y = np.zeros((nb, rows, 8), np.float32)
    Y[..., 0] = np.tanh(X[..., 0] + 0.5 * X[..., 1])
    Y[..., 1] = X[..., 2] * X[..., 3] * 0.5
    Y[..., 2] = np.maximum(X[..., 4], 0.0) - 0.4
    Y[..., 3] = np.sin(X[..., 5])
    Y[..., 4] = 0.3 * X[..., 6] - 0.2 * X[..., 7]
    Y[..., 5] = X[..., 8] ** 2 - 1.0
    net = StageNet(rows, norm=identity_norm(), seed=1)
    fs = [dv.to_dev(X[b].ravel(), np.float32) for b in range(nb)]
    ts = [dv.to_dev(Y[b].ravel(), np.float32) for b in range(nb)]
    first = np.mean([net.train_step(fs[b], ts[b], rows, 0.0, want_loss=True)[0] for b in range(nb)])
    for it in range(400):
        net.train_step(fs[it % nb], ts[it % nb], rows, 3e-3)
    last = np.mean([net.train_step(fs[b], ts[b], rows, 0.0, want_loss=True)[0] for b in range(nb)])
    assert last < 0.1 * first, (first, last)
    net.close()


def test_feature_target_guess_kernels_match_oracle():
    """vtsa_mlp_features / _targets / _apply on a 3x4x6 model handle, bit for bit."""
    import ctypes as C
    from dac_vtsa_amd import _abi
    from dac_vtsa_amd import device as dv
    from oracle import mlp
    from tests._helpers import SMALL, golden, state_from
    from tests.test_gpu_parity import gpu_model
    z = golden("rhs_small.npz")
    st = state_from(z, "base/")
    cfg, m = gpu_model(SMALL, 0, st)
    n, N = m.n_c // m.ng, m.n_core
    rng = np.random.default_rng(5)
    y = m.pack(st)
    s = 1e-6 + 1e-4 * np.abs(y)
    zp = y + s * rng.standard_normal(N)
    zc = zp + s * rng.standard_normal(N)
    R = rng.standard_normal(N) * s
    Y, ZP, ZC, S, RR = (dv.to_dev(a) for a in (y, zp, zc, s, R))
    feat = dv.empty(n * 32, np.float32)
    targ = dv.empty(n * 8, np.float32)
    _abi.call("vtsa_mlp_features", m.handle, Y.ptr, ZP.ptr, S.ptr, RR.ptr, float(np.log(0.37)), 1.0, feat.ptr, None)
    _abi.call("vtsa_mlp_targets", m.handle, ZC.ptr, ZP.ptr, RR.ptr, S.ptr, targ.ptr, None)
    F = feat.numpy().reshape(n, 32)
    T = targ.numpy().reshape(n, 8)
    assert np.array_equal(F, mlp.features(y, zp, s, R, n, cfg.grid.Nz, np.float32(np.log(0.37)), 1.0))
    assert np.array_equal(T, mlp.targets(zc, zp, s, n, R))
    _abi.call("vtsa_mlp_targets", m.handle, ZC.ptr, ZP.ptr, None, S.ptr, targ.ptr, None)
    assert np.array_equal(targ.numpy().reshape(n, 8), mlp.targets(zc, zp, s, n))
    from dac_vtsa_amd.learned import StageNet
    norm = np.concatenate([np.zeros(32), np.ones(32), rng.standard_normal(16), 0.5 + rng.random(16)]).astype(np.float32)
    net = StageNet(n, norm=norm)
    out = (rng.standard_normal((n, 16))).astype(np.float32)
    O = dv.to_dev(out.ravel(), np.float32)
    ZG = dv.empty(N)
    _abi.call("vtsa_mlp_apply", net.handle, m.handle, O.ptr, ZP.ptr, RR.ptr, S.ptr, ZG.ptr, None)
    assert np.array_equal(ZG.numpy(), mlp.apply_guess(out, norm, zp, s, n, R))
    _abi.call("vtsa_mlp_apply", net.handle, m.handle, O.ptr, ZP.ptr, None, S.ptr, ZG.ptr, None)
    assert np.array_equal(ZG.numpy(), mlp.apply_guess(out, norm, zp, s, n))
    net.close()
    m.close()


def test_learned_guess_on_a_heating_step():
    """End to end on the 10x16x50 heating step: a plain run from seed 1 records its stages, the
    device trains the network, and the seed-0 run starts its stages from the guess (FP32
    preconditioner).  Shows what the network gains on held-out stages, in tolerance units
    against the converged stage: the predictor, the predictor after one preconditioned Newton
    step (the network's base), and the guess.  The integrated run must keep the balance gate
    and agree with the plain FP64 run to the integration tolerance."""
    from dac_vtsa_amd.config import load_config
    from dac_vtsa_amd.core.grid import Grid
    from dac_vtsa_amd.learned import StageDataset, StageNet
    from dac_vtsa_amd.properties import get_backend
    from dac_vtsa_amd.seeds import seeded_state
    from dac_vtsa_amd.steps import run_step
    from dac_vtsa_amd import balances
    cfg = load_config(override_dict={"grid.Nr": 10, "grid.Ntheta": 16, "grid.Nz": 50, "operation.P_vacuum": 1.5e4,
                                     "solver.conservation_tol": 1e-5})
    grid, props = Grid.from_config(cfg), get_backend(cfg)
    step = cfg.cycle[2]
    base = {"gs_sweeps": 2, "conserve": "project"}
    ds = StageDataset(grid.ncell, 700)
    run_step(seeded_state(grid, "midcycle", 1), cfg, grid, props, step, integrator_kw={**base, "record": ds})
    assert ds.count > 50
    net = StageNet(grid.ncell, norm=ds.norm(), seed=0)
    hist = ds.train(net, epochs=10)
    ep, eb, eg = ds.errors(net, hist["val_ids"])
    print("held-out rms error (tolerance units, c_CO2 c_H2O c_N2 q_CO2 q_H2O e):\n"
          f"  predictor            {np.round(ep.mean(0), 3)}\n  preconditioned step  {np.round(eb.mean(0), 3)}\n"
          f"  learned guess        {np.round(eg.mean(0), 3)}   gain {np.round(hist['gain'], 3)}")
    # the calibrated guess is never worse than its base on the held-out stages (sum of squares),
    # and it is closer to the converged stages than the predictor every solve starts from
    assert np.sum(eg ** 2) <= np.sum(eb ** 2) * (1.0 + 1e-6)
    assert np.sum(eg ** 2) < np.sum(ep ** 2)
    st0 = seeded_state(grid, "midcycle", 0)
    st_plain, sr_plain = run_step(st0, cfg, grid, props, step, integrator_kw={**base, "precision": "fp64"})
    st, sr = run_step(st0, cfg, grid, props, step, integrator_kw={**base, "precision": "fp32", "guess": net,
                                                                  "lowrank": 4})
    s = sr.solver
    assert s["guess_used"] + s["guess_kept_predictor"] > 0
    # iterations of the learned form (guess, rank-4 preconditioner correction, FP32 factors) and of
    # the plain FP64 method on this step, for the record: they take different step sequences
    print(f"GMRES iterations: learned form {s['gmres_iters']}, plain FP64 {sr_plain.solver['gmres_iters']}; "
          f"guess used {s['guess_used']}, kept the predictor {s['guess_kept_predictor']}, "
          f"skipped {s['guess_skipped']}; corrections built {s['lowrank_built']}")
    mb, eb_, _ = balances.check_cycle([sr], props.species)
    assert max(abs(r["relative_error"]) for r in mb if r["step"] != "CYCLE") <= 1e-6
    assert max(abs(r["relative_error"]) for r in eb_ if r["step"] != "CYCLE") <= 1e-6
    # end fields agree with the plain FP64 run to the integration tolerance: the runs take
    # different Newton paths, each accepted step carries up to one tolerance unit of local error
    s_ = cfg.solver
    for a, b, atol in ((st.c[0], st_plain.c[0], s_.atol_c_co2), (st.c[1], st_plain.c[1], s_.atol_c_h2o),
                       (st.c[2], st_plain.c[2], s_.atol_c_inert), (st.q, st_plain.q, s_.atol_q),
                       (st.T, st_plain.T, s_.atol_T), (st.Tw, st_plain.Tw, s_.atol_T)):
        err = np.max(np.abs(a - b) / (atol + s_.rtol * np.abs(b)))
        print(f"end-field difference {err:.3g} tolerance units")
        assert err < 3.0, err
    net.close()


def test_rows_beyond_capacity_are_rejected():
    from dac_vtsa_amd import device as dv
    from dac_vtsa_amd.learned import StageNet
    net = StageNet(64)
    f = dv.empty(128 * 32, np.float32)
    o = dv.empty(128 * 16, np.float32)
    with pytest.raises(RuntimeError, match="rows out of range"):
        net.forward(f, 65, o)
    net.close()

"""Write golden input/output vectors from the upstream simulator (run in the build container only).

The upstream code is imported from /root/reference (read-only); nothing of it is copied.  Each
fixture stores the seeded inputs and the upstream outputs as plain arrays (``.npz``, no pickle),
for the functions that no closed form, invariant or known answer pins on its own:

* ``rhs_small.npz``   : f(t, y), accumulator rates and diag fields on a 3x4x6 grid for every step
                        type, with a partial jacket and a cosine feed profile, plus option variants
                        (limiters, co-adsorption, dispersion, conductivity, wall/lumped/hx);
* ``rhs_cfg1.npz``    : seeded 4x4x20 'small' state, adsorption step: f, acc, stage residual;
* ``rhs_mid.npz``     : seeded 'midcycle' state (4x4x10 grid) on the heating step at P_vac=1.5e4;
* ``jac_small.npz``   : coloured FD Jacobian (CSC values) on 3x4x6 and its theta-block subset;
* ``jac_cfg1.npz``    : the same on the 4x4x20 seeded state;
* ``integ_cfg1.npz``  : 20 s of the adsorption step from the seeded 4x4x20 state, sdirk2 and
                        trbdf2: accepted (t, h), counters, end state and accumulators;
* ``integ_cfg1_hybrid.npz`` : 10 s of the same step with solver.newton_mode = hybrid (modified
                        Newton on the stored factor, Krylov on failure);
* ``closures.npz``    : isotherm / p_sat / property values over a grid of inputs;
* ``runner_cfg1.npz`` : on the seeded 4x4x20 state, per step type: inventories, the four stop
                        metrics and q_feed_eq; the step runner over 20 s of adsorption:
                        accepted times, outlet rows, StepResult accumulators and inventories,
                        snapshots and the end state; the clean and feed-equilibrium initial beds;
* ``integ_cfg1_200s.npz`` : the whole 200 s adsorption step of the seeded 4x4x20 state with a
                        tight Newton tolerance (newton_kappa 1e-4, the smallest the configuration allows), sdirk2 and trbdf2: accepted
                        times, counters, snapshots, end state and accumulators;
* ``dry_air.json``    : the upstream dry-air case (components.water = false) on 3x4x6, stage by
                        stage (initial bed, model, residual): which complete, and the exception the
                        first failing one raises with its upstream file:line;
* ``state_ref.csv``   : a state file written by the upstream save_state_csv (the seeded 3x4x6
                        state), read back by this package's loader.

Usage (from the repository root):  python -m tests.golden.make_golden [fixture stems ...]
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, REF)

from config.parameters import load_config  # noqa: E402  (upstream)
from core.grid import Grid  # noqa: E402
from core.residual import TVSAModel  # noqa: E402
from core.jacobian import FDJacobian  # noqa: E402
from core.solver import TRBDF2  # noqa: E402
from core.state import initial_state, CycleState  # noqa: E402
from physics.isotherms import EquilibriumModel  # noqa: E402
from properties import get_backend, arden_buck  # noqa: E402
from steps._runner import run_step  # noqa: E402
from simio.checkpoint import save_state_csv  # noqa: E402

from dac_vtsa_amd.seeds import seeded_state  # noqa: E402  (this repository)

from tests._helpers import SMALL, VARIANTS  # noqa: E402


def _state_arrays(prefix, st, out):
    out[prefix + "c"], out[prefix + "q"], out[prefix + "T"], out[prefix + "Tw"] = st.c, st.q, st.T, st.Tw


def _diag(prefix, d, out):
    for k in ("T", "P", "y", "qstar", "uz", "ur", "uth", "Q_jac", "Q_wb", "Q_amb", "Q_hx", "R_bed"):
        v = d[k]
        if v is not None:
            out[prefix + "d_" + k] = np.asarray(v, dtype=float)
    for end in ("top", "bottom"):
        for k, v in d["bnd"][end].items():
            out[prefix + f"d_bnd_{end}_{k}"] = np.asarray(v, dtype=float)


def small_state(cfg, g, props, rng_seed=1):
    st = initial_state(cfg, g, props)
    rng = np.random.default_rng(rng_seed)
    st.T = st.T + 5.0 * rng.random(st.T.shape)
    st.q = st.q + 0.3 * rng.random(st.q.shape)
    st.Tw = st.Tw + 2.0 * rng.random(st.Tw.shape)
    return st


def qstar_fn_for(cfg):
    eq = EquilibriumModel(cfg)

    def fn(pc, pw, T, qw):
        qs = eq.equilibrium(pc, pw, T, qw, arden_buck(T))
        return qs[0], qs[1]
    return fn


def make_rhs_small(path):
    out = {"variants": np.array(list(VARIANTS))}
    for name, over in VARIANTS.items():
        cfg = load_config(override_dict={**SMALL, **over})
        g = Grid.from_config(cfg)
        props = get_backend(cfg)
        st = small_state(cfg, g, props)
        _state_arrays(f"{name}/", st, out)
        for si in range(len(cfg.cycle)):
            m = TVSAModel(cfg, g, props, cfg.cycle[si], st, T_jacket_prev=300.0)
            y = m.pack(st)
            f, acc = m.rhs(12.5, y)
            p = f"{name}/s{si}/"
            out[p + "y"], out[p + "f"], out[p + "acc"] = y, f, acc
            if name == "base":
                _diag(p, m.rhs(12.5, y, diag=True), out)
    np.savez_compressed(path, **out)


def _seeded(cfg, kind, seed):
    g = Grid.from_config(cfg)
    return g, seeded_state(g, kind, seed, qstar_fn_for(cfg))


def make_rhs_cfg1(path):
    cfg = load_config(override_dict={"grid.Nr": 4, "grid.Ntheta": 4, "grid.Nz": 20})
    g, st = _seeded(cfg, "small", 0)
    props = get_backend(cfg)
    out = {}
    _state_arrays("", st, out)
    m = TVSAModel(cfg, g, props, cfg.cycle[0], st)
    y = m.pack(st)
    f, acc = m.rhs(5.0, y)
    rng = np.random.default_rng(7)
    rhs_known = y + 1e-3 * rng.standard_normal(y.size) * np.abs(y)
    out.update(y=y, f=f, acc=acc, rhs_known=rhs_known, R=m.stage_residual(5.0, y, 0.37, rhs_known),
               atol=m.atol)
    _diag("", m.rhs(5.0, y, diag=True), out)
    np.savez_compressed(path, **out)


def make_rhs_mid(path):
    cfg = load_config(override_dict={"grid.Nr": 4, "grid.Ntheta": 4, "grid.Nz": 10, "operation.P_vacuum": 1.5e4})
    g, st = _seeded(cfg, "midcycle_cells", 0)
    props = get_backend(cfg)
    out = {}
    _state_arrays("", st, out)
    for si in range(4):
        m = TVSAModel(cfg, g, props, cfg.cycle[si], st)
        y = m.pack(st)
        f, acc = m.rhs(3.0, y)
        out[f"s{si}/y"], out[f"s{si}/f"], out[f"s{si}/acc"] = y, f, acc
    np.savez_compressed(path, **out)


def make_jac(path, cfg, st, t):
    g = Grid.from_config(cfg)
    props = get_backend(cfg)
    m = TVSAModel(cfg, g, props, cfg.cycle[0], st)
    y = m.pack(st)
    jac = FDJacobian(m)
    J = jac(t, y)
    Jt = jac.theta_block(J)
    np.savez_compressed(path, y=y, data=J.data, indices=J.indices, indptr=J.indptr, color=jac.color,
                        theta_data=Jt.data, theta_indices=Jt.indices, theta_indptr=Jt.indptr,
                        ngroups=jac.ngroups, typ=jac.typ, **{"c": st.c, "q": st.q, "T": st.T, "Tw": st.Tw})


def make_integ(path, t_end=20.0):
    out = {}
    for scheme in ("sdirk2", "trbdf2"):
        cfg = load_config(override_dict={"grid.Nr": 4, "grid.Ntheta": 4, "grid.Nz": 20, "solver.scheme": scheme})
        g, st = _seeded(cfg, "small", 0)
        if scheme == "sdirk2":
            _state_arrays("", st, out)
        props = get_backend(cfg)
        m = TVSAModel(cfg, g, props, cfg.cycle[0], st)
        y0 = m.pack(st)
        integ = TRBDF2(m, cfg, h_max=0.02 * 200.0)
        times = []
        t0 = time.time()
        res = integ.integrate(y0, t_end, np.linspace(0, t_end, 5), on_step=lambda t, y, a: times.append(t))
        p = f"{scheme}/"
        out[p + "wall"] = time.time() - t0
        out[p + "y0"], out[p + "y"], out[p + "acc"], out[p + "t_acc"] = y0, res.y, res.acc, np.array(times)
        out[p + "counters"] = np.array([res.n_steps, res.n_reject, res.n_newton_fail, res.n_krylov,
                                        res.n_jac, res.n_lu, res.nfev])
        out[p + "snap_t"] = np.array([s[0] for s in res.snapshots])
        out[p + "snap_y"] = np.array([s[1] for s in res.snapshots])
        out[p + "gmres_iters"] = integ.stats["gmres_iters"]
        print(scheme, "steps", res.n_steps, "wall", out[p + "wall"])
    np.savez_compressed(path, **out)


def make_integ_hybrid(path, t_end=10.0):
    """10 s of the 4x4x20 adsorption step with solver.newton_mode = hybrid (modified Newton with
    the factor first, Krylov on failure; core/solver.py:111-151)."""
    out = {}
    cfg = load_config(override_dict={"grid.Nr": 4, "grid.Ntheta": 4, "grid.Nz": 20, "solver.newton_mode": "hybrid"})
    g, st = _seeded(cfg, "small", 0)
    props = get_backend(cfg)
    m = TVSAModel(cfg, g, props, cfg.cycle[0], st)
    y0 = m.pack(st)
    integ = TRBDF2(m, cfg, h_max=0.02 * 200.0)
    times = []
    res = integ.integrate(y0, t_end, np.linspace(0, t_end, 3), on_step=lambda t, y, a: times.append(t))
    out["y0"], out["y"], out["acc"], out["t_acc"] = y0, res.y, res.acc, np.array(times)
    out["counters"] = np.array([res.n_steps, res.n_reject, res.n_newton_fail, res.n_krylov,
                                res.n_jac, res.n_lu, res.nfev])
    out["modified_fail"] = integ.stats["modified_fail"]
    print("hybrid steps", res.n_steps, "modified_fail", integ.stats["modified_fail"])
    np.savez_compressed(path, **out)


def make_closures(path):
    cfg = load_config()
    props = get_backend(cfg)
    out = {}
    T = np.linspace(250.0, 420.0, 35)
    out["T"] = T
    out["psat"] = arden_buck(T)
    TT, PP = np.meshgrid(np.linspace(260, 400, 8), np.logspace(-3, 5.3, 12), indexing="ij")
    QW = np.linspace(0.0, 7.0, TT.size).reshape(TT.shape)
    out["grid_T"], out["grid_p"], out["grid_qw"] = TT, PP, QW
    for model in ("none", "wadst", "mechanistic", "stampi_bombelli"):
        from config.parameters import with_overrides
        eq = EquilibriumModel(with_overrides(cfg, {"isotherm.coadsorption.model": model}))
        out[f"qco2_{model}"] = eq.q_co2(PP, TT, QW)
    eq = EquilibriumModel(cfg)
    ps = arden_buck(TT)
    out["qh2o"] = eq.q_h2o(PP * 1e-2, TT, ps)
    yy = np.random.default_rng(3).dirichlet([1, 1, 1], size=TT.shape).transpose(2, 0, 1)
    out["y"] = yy
    mu, k = props.mu_k_mix(TT, yy)
    out["mu"], out["k"] = mu, k
    out["D"] = props.D_co2_mix(TT, PP, yy)
    out["cp"] = props.cp
    np.savez_compressed(path, **out)


INV_KEYS = ("gas_CO2", "ads_CO2", "tot_CO2", "gas_H2O", "ads_H2O", "tot_H2O", "gas_N2", "ads_N2", "tot_N2",
            "E_bed", "E_wall", "E_total")
STOP_CRITERIA = (("min_temperature", 300.0), ("max_temperature", 350.0), ("end_saturation", 0.5),
                 ("outlet_ratio", 0.5))


def make_runner(path):
    cfg = load_config(override_dict={"grid.Nr": 4, "grid.Ntheta": 4, "grid.Nz": 20})
    g, st = _seeded(cfg, "small", 0)
    props = get_backend(cfg)
    out = {"inv_keys": np.array(INV_KEYS), "stop_criteria": np.array([c for c, _ in STOP_CRITERIA])}
    _state_arrays("", st, out)
    for si, step in enumerate(cfg.cycle):
        stops = []
        for crit, val in STOP_CRITERIA:
            stp = dict(step)
            stp["stop_on"] = {"criterion": crit, "value": val}
            m = TVSAModel(cfg, g, props, stp, st)
            stops.append(m.stop_metric(7.0, m.pack(st)))
        out[f"s{si}/stop"] = np.array(stops)
        inv = m.inventories(m.pack(st))
        out[f"s{si}/inv"] = np.array([inv[k] for k in INV_KEYS])
        out[f"s{si}/q_feed_eq"] = m.q_feed_eq()
    step = dict(cfg.cycle[0])
    step["duration"] = 20.0
    times = []
    st_out, sr = run_step(st, cfg, g, props, step)
    rows = sr.outlet
    num = [k for k, v in rows[0].items() if isinstance(v, (int, float, np.floating)) and not isinstance(v, bool)]
    out["outlet_keys"] = np.array(num)
    out["outlet_str_keys"] = np.array([k for k in rows[0] if k not in num])
    out["outlet_str"] = np.array([[str(r[k]) for k in out["outlet_str_keys"]] for r in rows])
    out["outlet"] = np.array([[float(r[k]) for k in num] for r in rows])
    m = TVSAModel(cfg, g, props, step, st)
    out["acc_names"] = np.array(m.acc_names)
    out["acc"] = np.array([sr.acc[k] for k in m.acc_names])
    out["inv_start"] = np.array([sr.inv_start[k] for k in INV_KEYS])
    out["inv_end"] = np.array([sr.inv_end[k] for k in INV_KEYS])
    out["duration"], out["T_jacket"] = sr.duration, sr.T_jacket
    out["counters"] = np.array([sr.solver[k] for k in ("steps", "rejected", "newton_fail", "krylov", "jacobians",
                                                       "lu", "nfev")])
    out["snap_t"] = np.array([sn["t_step"] for sn in sr.snapshots])
    for k in ("c", "q", "T", "Tw"):
        out["snap_" + k] = np.array([getattr(sn["state"], k) for sn in sr.snapshots])
        out["end_" + k] = getattr(st_out, k)
    out["t_acc"] = np.array([r["t_step_s"] for r in rows[1:]])
    for mode in ("clean", "feed_equilibrium"):
        cm = load_config(override_dict={"grid.Nr": 4, "grid.Ntheta": 4, "grid.Nz": 20, "run.initial_state": mode})
        si = initial_state(cm, Grid.from_config(cm), get_backend(cm))
        _state_arrays(f"init_{mode}/", si, out)
    np.savez_compressed(path, **out)


def make_integ_200(path, kappa=1e-4):
    out = {}
    for scheme in ("sdirk2", "trbdf2"):
        cfg = load_config(override_dict={"grid.Nr": 4, "grid.Ntheta": 4, "grid.Nz": 20, "solver.scheme": scheme,
                                         "solver.newton_kappa": kappa})
        g, st = _seeded(cfg, "small", 0)
        if scheme == "sdirk2":
            _state_arrays("", st, out)
        props = get_backend(cfg)
        m = TVSAModel(cfg, g, props, cfg.cycle[0], st)
        y0 = m.pack(st)
        integ = TRBDF2(m, cfg, h_max=0.02 * 200.0)
        times = []
        t0 = time.time()
        res = integ.integrate(y0, 200.0, np.linspace(0, 200.0, 12), on_step=lambda t, y, a: times.append(t))
        p = f"{scheme}/"
        out[p + "wall"] = time.time() - t0
        out[p + "y0"], out[p + "y"], out[p + "acc"], out[p + "t_acc"] = y0, res.y, res.acc, np.array(times)
        out[p + "counters"] = np.array([res.n_steps, res.n_reject, res.n_newton_fail, res.n_krylov,
                                        res.n_jac, res.n_lu, res.nfev])
        out[p + "snap_t"] = np.array([s[0] for s in res.snapshots])
        out[p + "snap_y"] = np.array([s[1] for s in res.snapshots])
        out[p + "inv0"] = np.array([m.inventories(y0)[k] for k in INV_KEYS])
        out[p + "inv1"] = np.array([m.inventories(res.y)[k] for k in INV_KEYS])
        out[p + "stats"] = np.array([integ.stats.get(k, 0) for k in ("gmres_iters", "balance_unclosed")])
        print(scheme, "200 s steps", res.n_steps, "wall", out[p + "wall"], flush=True)
    out["kappa"] = kappa
    np.savez_compressed(path, **out)


def make_state_csv(path):
    cfg = load_config(override_dict=SMALL)
    g = Grid.from_config(cfg)
    save_state_csv(small_state(cfg, g, get_backend(cfg)), g, path)


def make_dry_air(path):
    """The upstream dry-air case (config/cases/dry_air.yaml: components.water = false) on a 3x4x6
    grid, stage by stage: which stages complete and, for the first that does not, the exception
    type, message and the upstream frame (file:line) it was raised in."""
    import json
    import traceback
    cfg = load_config(os.path.join(REF, "config", "cases", "dry_air.yaml"), override_dict=SMALL)
    g = Grid.from_config(cfg)
    props = get_backend(cfg)
    out = {"case": "config/cases/dry_air.yaml", "grid": [g.Nr, g.Nth, g.Nz], "species": list(props.species),
           "stages": {}}
    st = m = None
    for stage in ("initial_state", "model", "rhs"):
        try:
            if stage == "initial_state":
                st = initial_state(cfg, g, props)
            elif stage == "model":
                m = TVSAModel(cfg, g, props, cfg.cycle[0], st)
            else:
                m.rhs(1.0, m.pack(st))
            out["stages"][stage] = "ok"
        except Exception as e:  # the outcome is the fixture
            fr = traceback.extract_tb(e.__traceback__)[-1]
            out["stages"][stage] = {"raised": type(e).__name__, "message": str(e),
                                    "at": f"{os.path.relpath(fr.filename, REF)}:{fr.lineno}"}
            break
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)


MAKERS = {"closures": make_closures, "rhs_small": make_rhs_small, "rhs_cfg1": make_rhs_cfg1, "rhs_mid": make_rhs_mid,
          "runner_cfg1": make_runner, "integ_cfg1_200s": make_integ_200}


if __name__ == "__main__" and len(sys.argv) > 1:
    for stem in sys.argv[1:]:
        t0 = time.time()
        if stem == "state_ref":
            make_state_csv(os.path.join(HERE, "state_ref.csv"))
        elif stem == "dry_air":
            make_dry_air(os.path.join(HERE, "dry_air.json"))
        else:
            MAKERS[stem](os.path.join(HERE, stem + ".npz"))
        print(stem, "done in %.1f s" % (time.time() - t0), flush=True)
elif __name__ == "__main__":
    t0 = time.time()
    make_closures(os.path.join(HERE, "closures.npz"))
    make_rhs_small(os.path.join(HERE, "rhs_small.npz"))
    make_rhs_cfg1(os.path.join(HERE, "rhs_cfg1.npz"))
    make_rhs_mid(os.path.join(HERE, "rhs_mid.npz"))
    cfg = load_config(override_dict=SMALL)
    g = Grid.from_config(cfg)
    make_jac(os.path.join(HERE, "jac_small.npz"), cfg, small_state(cfg, g, get_backend(cfg)), 5.0)
    cfg = load_config(override_dict={"grid.Nr": 4, "grid.Ntheta": 4, "grid.Nz": 20})
    g, st = _seeded(cfg, "small", 0)
    make_jac(os.path.join(HERE, "jac_cfg1.npz"), cfg, st, 5.0)
    make_integ(os.path.join(HERE, "integ_cfg1.npz"))
    make_integ_hybrid(os.path.join(HERE, "integ_cfg1_hybrid.npz"))
    print("done in %.1f s" % (time.time() - t0))

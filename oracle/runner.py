"""Oracle restatement of the step runner (test infrastructure only).

``run_step`` follows steps/_runner.py:87-188 (with ``record``/``_end_summary`` :50-84,104-121);
pinned by tests/golden/runner_cfg1.npz (outlet rows, accumulators, inventories, snapshots).
"""
from __future__ import annotations

import time
from typing import List, Optional

import numpy as np

from dac_vtsa_amd.core.state import CycleState
from dac_vtsa_amd.steps.result import StepResult
from dac_vtsa_amd import balances

from .model import Equilibrium, Gas, Grid, OracleModel, R_GAS, humid_air
from .solver import OracleTRBDF2


def species_for(cfg):
    return ["CO2", "H2O", "N2"] if cfg.components.water else ["CO2", "N2"]


def initial_state(cfg, grid, gas) -> CycleState:
    """Clean or feed-equilibrium initial bed (core/state.py:62-93); csv mode is host I/O."""
    species = tuple(gas.species)
    adsorbed = ("CO2", "H2O") if cfg.components.water else ("CO2",)
    T0, P0 = cfg.feed.T, cfg.feed.P
    y = humid_air(species, P0, T0, cfg.feed.RH if cfg.components.water else 0.0, cfg.feed.y_co2_dry,
                  float(gas.p_sat(T0)))
    C = P0 / (R_GAS * T0)
    c = np.empty((len(species),) + grid.shape)
    for k in range(len(species)):
        c[k] = C * y[k]
    T = np.full(grid.shape, T0)
    q = np.zeros((len(adsorbed),) + grid.shape)
    if cfg.run.initial_state == "feed_equilibrium":
        eq = Equilibrium(cfg)
        p = c * R_GAS * T
        iw = species.index("H2O") if "H2O" in species else None
        psat = gas.p_sat(T)
        qw = eq.q_h2o(p[iw], T, psat) if iw is not None else None
        q = np.array(eq.equilibrium(p[0], p[iw] if iw is not None else None, T, qw, psat))
    return CycleState(c=c, q=q, T=T, Tw=np.full(grid.wall_shape, T0), t=0.0, cycle=1, step="initial",
                      species=species, adsorbed=adsorbed)


def end_summary(model, d, end, c_feed_co2):
    """Outlet summary of one end (steps/_runner.py:50-84)."""
    b = d["bnd"][end]
    Fout = -b["F_in"]
    tot = Fout.sum(axis=0)
    Ftot = float(tot.sum())
    k_lay = 0 if end == "top" else -1
    T_lay = d["T"][..., k_lay]
    A = model.grid.Az * np.ones_like(T_lay)
    P_face = float(np.sum(b["P_face"] * A) / np.sum(A))
    out_faces = tot > 0.0
    if "F_outpart" in b:
        Fout = b["F_outpart"]
        tot = Fout.sum(axis=0)
        out_faces = tot > 0.0
    F_leaving = float(tot[out_faces].sum()) if np.any(out_faces) else 0.0
    if F_leaving > 1e-14:
        ys = [float(Fout[k][out_faces].sum()) / F_leaving for k in range(model.ng)]
        T_g = float(np.sum(T_lay[out_faces] * tot[out_faces]) / F_leaving)
    else:
        yl = d["y"][..., k_lay]
        ys = [float(np.sum(yl[k] * A) / np.sum(A)) for k in range(model.ng)]
        T_g = float(np.sum(T_lay * A) / np.sum(A))
    out = {f"{end}_valve": model.conds[end].valve or "closed",
           f"{end}_valve_state": model.conds[end].state,
           f"{end}_flow_out_mol_s": Ftot, f"{end}_P_Pa": P_face, f"{end}_T_K": T_g}
    for k, s in enumerate(model.species):
        out[f"{end}_y_{s}"] = ys[k]
    c_co2 = ys[model.i_co2] * P_face / (R_GAS * T_g)
    out[f"{end}_c_CO2_mol_m3"] = c_co2
    out[f"{end}_C_over_Cfeed"] = c_co2 / c_feed_co2 if c_feed_co2 > 0 else np.nan
    return out


def to_state(model, y, template, t_abs, step_name):
    c, q, e, Tw = model.unpack(y)
    T = model.temperature(e, c, q)
    st = template.copy()
    st.c, st.q, st.T = c.copy(), q.copy(), T.copy()
    st.Tw = Tw.copy() if Tw is not None else T[-1].copy()
    st.t = t_abs
    st.step = step_name
    return st


def run_step(state_in, cfg, grid, gas, step, cycle_idx=1, t_cycle0=0.0, T_jacket_prev=None,
             logger=None, snapshots=None, record_outlet=True, integrator_kw=None):
    """One cycle step on the CPU (steps/_runner.py:87-188)."""
    model = OracleModel(cfg, grid, gas, step, state_in, T_jacket_prev)
    y0 = model.pack(state_in)
    dur = float(step["duration"])
    n_snap = snapshots if snapshots is not None else cfg.run.snapshots_per_step
    t_out = np.linspace(0.0, dur, max(2, n_snap))
    yf = humid_air(model.species, cfg.feed.P, cfg.feed.T, cfg.feed.RH if model.water else 0.0,
                   cfg.feed.y_co2_dry, float(gas.p_sat(cfg.feed.T)))
    c_feed_co2 = yf[model.i_co2] * cfg.feed.P / (R_GAS * cfg.feed.T)
    outlet: List[dict] = []
    every = cfg.solver.outlet_record_every
    counter = [0]
    V = np.broadcast_to(grid.V, grid.shape)
    Vt = float(V.sum())

    def record(t, y, acc):
        d = model.rhs(t, y, diag=True)
        row = {"cycle": cycle_idx, "step": step["name"], "t_step_s": t, "t_cycle_s": t_cycle0 + t,
               "t_total_s": state_in.t + t}
        for end in ("top", "bottom"):
            row.update(end_summary(model, d, end, c_feed_co2))
        row["bed_T_mean_K"] = float(np.sum(d["T"] * V) / Vt)
        row["bed_P_mean_Pa"] = float(np.sum(d["P"] * V) / Vt)
        c, q, e, Tw = model.unpack(y)
        for n, s in enumerate(model.adsorbed):
            row[f"bed_q_{s}_mean"] = float(np.sum(q[n] * V) / Vt)
        ai = model.acc_index
        for v in ("V1", "V2", "V3"):
            row[f"cum_CO2_out_{v}_mol"] = acc[ai[f"n_out_{v}_CO2"]]
        row["cum_Q_jacket_J"] = acc[ai["Q_jac_pos"]] + acc[ai["Q_jac_neg"]]
        outlet.append(row)

    def on_step(t, y, acc):
        counter[0] += 1
        if record_outlet and counter[0] % every == 0:
            record(t, y, acc)

    if record_outlet:
        record(0.0, y0, np.zeros(model.n_acc))
    stop = step.get("stop_on")
    stop_fn = stop_val = None
    direction = 1
    if stop:
        stop_fn, stop_val = model.stop_metric, stop["value"]
        direction = -1 if stop["criterion"] == "min_temperature" else 1
    wall0 = time.time()
    integ = OracleTRBDF2(model, cfg, h_max=cfg.solver.h_max_fraction * dur, **(integrator_kw or {}))
    res = integ.integrate(y0, dur, t_out, on_step=on_step, stop_fn=stop_fn, stop_value=stop_val,
                          stop_direction=direction)
    stats = {"method": "trbdf2", "steps": res.n_steps, "rejected": res.n_reject,
             "newton_fail": res.n_newton_fail, "krylov": res.n_krylov, "jacobians": res.n_jac,
             "lu": res.n_lu, "nfev": res.nfev, "groups": integ.jac.ngroups, **integ.stats}
    if record_outlet and outlet[-1]["t_step_s"] < res.t:
        record(res.t, res.y, res.acc)
    wall = time.time() - wall0
    state_out = to_state(model, res.y, state_in, state_in.t + res.t, step["name"])
    state_out.cycle = cycle_idx
    snaps = []
    for (ts, ys, _a) in res.snapshots:
        st = to_state(model, ys, state_in, state_in.t + ts, step["name"])
        st.cycle = cycle_idx
        snaps.append({"cycle": cycle_idx, "step": step["name"], "t_step": ts, "t_cycle": t_cycle0 + ts,
                      "t_total": state_in.t + ts, "state": st})
    inv0, inv1 = model.inventories(y0), model.inventories(res.y)
    T_jac = model.T_jacket(res.t) if model.jacket_on else cfg.ambient.T
    sr = StepResult(name=step["name"], type=step["type"], cycle=cycle_idx, t_start_cycle=t_cycle0,
                    duration_planned=dur, duration=res.t, stopped_early=res.stopped_early,
                    acc={nm: float(res.acc[i]) for i, nm in enumerate(model.acc_names)},
                    inv_start=inv0, inv_end=inv1, outlet=outlet, snapshots=snaps, solver=stats,
                    wall_time=wall, T_jacket=float(T_jac), jacket_on=model.jacket_on,
                    product_port=step.get("product_port", "V3"), valves=dict(step["valves"]),
                    ads_start={s: inv0[f"ads_{s}"] for s in model.adsorbed},
                    ads_end={s: inv1[f"ads_{s}"] for s in model.adsorbed})
    sr.steps_taken = res.steps_taken
    return state_out, sr


def make(cfg):
    """(grid, gas) for a config."""
    return Grid.from_config(cfg), Gas(species_for(cfg), cfg.properties.cp_reference_T)

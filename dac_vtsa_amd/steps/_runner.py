"""Step runner (drop-in for upstream steps/_runner.py:26-188).

``run_step`` builds the device model of the step, integrates it on the GPU, records one
outlet row per accepted step from the device diagnostics, and returns ``(CycleState,
StepResult)`` exactly as upstream; the state stays in HBM for the whole step.
"""
from __future__ import annotations

import time
from typing import Dict, List, Optional

import numpy as np

from .. import _abi
from .. import device as dv
from ..core.residual import TVSAModel
from ..core.solver import TRBDF2, SolverError
from ..properties import R_GAS
from ..valves import humid_air_composition
from .result import StepResult


def _end_summary(model, d, end, c_feed_co2):
    """Outlet composition/flow summary of one end (steps/_runner.py:50-84)."""
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
    out = {f"{end}_valve": model.conds[end].valve or "closed", f"{end}_valve_state": model.conds[end].state,
           f"{end}_flow_out_mol_s": Ftot, f"{end}_P_Pa": P_face, f"{end}_T_K": T_g}
    for k, s in enumerate(model.species):
        out[f"{end}_y_{s}"] = ys[k]
    c_co2 = ys[model.i_co2] * P_face / (R_GAS * T_g)
    out[f"{end}_c_CO2_mol_m3"] = c_co2
    out[f"{end}_C_over_Cfeed"] = c_co2 / c_feed_co2 if c_feed_co2 > 0 else np.nan
    return out


OUTLET_ROW = 32      # doubles per device outlet row (vtsa_outlet_row)
OUTLET_RING = 2048   # rows held on the device before they are copied to the host


def _end_from_sums(model, r, e, end, c_feed_co2, A_tot):
    """The outlet summary of one end (the keys and formulas of _end_summary) from the device
    sums of vtsa_outlet_row, r = that row, e = 0 (top) or 1 (bottom)."""
    b = 11 * e
    Ftot, PA, F_leaving = float(r[b]), float(r[b + 1]), float(r[b + 2])
    P_face = PA / A_tot
    if F_leaving > 1e-14:
        ys = [float(r[b + 3 + k]) / F_leaving for k in range(model.ng)]
        T_g = float(r[b + 6]) / F_leaving
    else:
        ys = [float(r[b + 7 + k]) / A_tot for k in range(model.ng)]
        T_g = float(r[b + 10]) / A_tot
    out = {f"{end}_valve": model.conds[end].valve or "closed", f"{end}_valve_state": model.conds[end].state,
           f"{end}_flow_out_mol_s": Ftot, f"{end}_P_Pa": P_face, f"{end}_T_K": T_g}
    for k, s in enumerate(model.species):
        out[f"{end}_y_{s}"] = ys[k]
    c_co2 = ys[model.i_co2] * P_face / (R_GAS * T_g)
    out[f"{end}_c_CO2_mol_m3"] = c_co2
    out[f"{end}_C_over_Cfeed"] = c_co2 / c_feed_co2 if c_feed_co2 > 0 else np.nan
    return out


class _OutletRing:
    """Outlet rows recorded on the device (no upstream counterpart: steps/_runner.py:104-121 copies
    the diagnostic fields to the host at every recorded step).  Each record queues one
    vtsa_outlet_row into a device ring; the host keeps the row's times and accumulator values and
    forms the rows when the ring is full and at the end of the step, one copy for many rows."""

    def __init__(self, model, head, c_feed_co2, out_rows):
        self.m = model
        self.head = head
        self.c_feed = c_feed_co2
        self.rows = out_rows
        self.buf = dv.empty(OUTLET_RING * OUTLET_ROW)
        self.pending = []
        self.mask = sum(1 << e for e, end in enumerate(("top", "bottom")) if model.conds[end].kind == "valve")
        g = model.grid
        self.A_tot = float(np.sum(g.Az * np.ones((g.Nr, g.Nth))))
        self.Vt = float(g.volume_full().sum())

    def record(self, t, y_dev, acc):
        if len(self.pending) == OUTLET_RING:
            self.flush()
        slot = len(self.pending)
        _abi.call("vtsa_outlet_row", self.m.handle, float(t), dv.ptr(y_dev), self.mask,
                  dv.ptr(self.buf.view(slot * OUTLET_ROW, OUTLET_ROW)), dv.stream())
        ai = self.m.acc_index
        extra = [float(acc[ai[f"n_out_{v}_CO2"]]) for v in ("V1", "V2", "V3")]
        extra.append(float(acc[ai["Q_jac_pos"]] + acc[ai["Q_jac_neg"]]))
        self.pending.append((float(t), extra))

    def flush(self):
        if not self.pending:
            return
        sums = dv.to_host(self.buf.view(0, len(self.pending) * OUTLET_ROW)).reshape(-1, OUTLET_ROW)
        m = self.m
        for (t, extra), r in zip(self.pending, sums):
            row = self.head(t)
            for e, end in enumerate(("top", "bottom")):
                row.update(_end_from_sums(m, r, e, end, self.c_feed, self.A_tot))
            row["bed_T_mean_K"] = float(r[22]) / self.Vt
            row["bed_P_mean_Pa"] = float(r[23]) / self.Vt
            for k, s in enumerate(m.adsorbed):
                row[f"bed_q_{s}_mean"] = float(r[24 + 3 + k]) / m.rho_b / self.Vt
            for v, val in zip(("V1", "V2", "V3"), extra[:3]):
                row[f"cum_CO2_out_{v}_mol"] = val
            row["cum_Q_jacket_J"] = extra[3]
            self.rows.append(row)
        self.pending.clear()

    def last_t(self):
        return self.pending[-1][0] if self.pending else (self.rows[-1]["t_step_s"] if self.rows else -np.inf)


def run_step(state_in, cfg, grid, props, step: Dict, cycle_idx: int = 1, t_cycle0: float = 0.0,
             T_jacket_prev: Optional[float] = None, logger=None, snapshots: Optional[int] = None,
             record_outlet: bool = True, integrator_kw: Optional[Dict] = None, progress_s: float = 30.0,
             stop_at: Optional[float] = None, device_handoff: bool = False):
    """One cycle step on the GPU -> (state_out, StepResult) (steps/_runner.py:87-188).

    ``record_outlet``, ``integrator_kw`` (preconditioner kind/precision), ``stop_at`` (end the
    step at the first accepted time step reaching this step time, as a stop criterion does; only
    for steps without one of their own) and ``device_handoff`` have no upstream counterpart; the
    defaults reproduce the upstream behaviour.  The outlet rows are reduced on the device
    (record_outlet=True) or, with record_outlet="host", from diagnostic fields copied to the host
    as upstream does.  The returned state carries the bed's device copy; with device_handoff=True
    a state returned by an earlier step on this stream (and not modified since) starts this step
    from that copy, so the bed stays in HBM from step to step.  The host arrays are filled either
    way."""
    log = logger or (lambda *a, **k: None)
    model = TVSAModel(cfg, grid, props, step, state_in, T_jacket_prev)
    y_in = getattr(state_in, "_y_device", None) if device_handoff else None
    if y_in is not None and (y_in.n != model.n_core or y_in._key[0] != dv._stream_key()):
        y_in = None     # another layout, or made on another stream: uploaded from the host arrays
    if y_in is not None:
        # the host arrays hold T, from which pack re-forms e; the same round trip on the device
        # keeps this start bit-identical to one from the host state
        y0 = y_in.clone()
        _abi.call("vtsa_state_roundtrip", model.handle, dv.ptr(y0), dv.stream())
    else:
        y0 = dv.to_dev(model.pack(state_in))
    dur = float(step["duration"])
    n_snap = snapshots if snapshots is not None else cfg.run.snapshots_per_step
    t_out = np.linspace(0.0, dur, max(2, n_snap))
    yf = humid_air_composition(model.species, cfg.feed.P, cfg.feed.T, cfg.feed.RH, cfg.feed.y_co2_dry,
                               float(props.p_sat(cfg.feed.T)))
    c_feed_co2 = yf[model.i_co2] * cfg.feed.P / (R_GAS * cfg.feed.T)
    outlet: List[Dict] = []
    every = cfg.solver.outlet_record_every
    counter = [0]
    V = grid.volume_full()
    Vt = float(V.sum())
    Vr = np.ascontiguousarray(V)

    def head(t):
        return {"cycle": cycle_idx, "step": step["name"], "t_step_s": t, "t_cycle_s": t_cycle0 + t,
                "t_total_s": state_in.t + t}

    ring = _OutletRing(model, head, c_feed_co2, outlet) if record_outlet and record_outlet != "host" else None

    def record(t, y, acc):
        if ring is not None:
            ring.record(t, y if not isinstance(y, np.ndarray) else dv.to_dev(y), acc)
            return
        d = model.diag(t, y if isinstance(y, np.ndarray) else dv.to_host(y))
        row = head(t)
        for end in ("top", "bottom"):
            row.update(_end_summary(model, d, end, c_feed_co2))
        row["bed_T_mean_K"] = float(np.sum(d["T"] * Vr) / Vt)
        row["bed_P_mean_Pa"] = float(np.sum(d["P"] * Vr) / Vt)
        q = model.inventories(y)
        for s in model.adsorbed:
            row[f"bed_q_{s}_mean"] = q[f"ads_{s}"] / model.rho_b / Vt
        ai = model.acc_index
        for v in ("V1", "V2", "V3"):
            row[f"cum_CO2_out_{v}_mol"] = acc[ai[f"n_out_{v}_CO2"]]
        row["cum_Q_jacket_J"] = acc[ai["Q_jac_pos"]] + acc[ai["Q_jac_neg"]]
        outlet.append(row)

    last_print = [time.time()]

    def on_step(t, y, acc):
        counter[0] += 1
        if record_outlet and counter[0] % every == 0:
            record(t, y, acc)
        if time.time() - last_print[0] > progress_s:
            last_print[0] = time.time()
            log(f"      ... {step['name']}: t = {t:.0f}/{dur:.0f} s ({100 * t / dur:.0f} %), {counter[0]} steps")

    on_step.wants_device = True
    if record_outlet:
        record(0.0, y0, np.zeros(model.n_acc))
    stop = step.get("stop_on")
    stop_fn = stop_val = None
    direction = 1
    if stop:
        stop_fn, stop_val = model.stop_metric, stop["value"]
        direction = -1 if stop["criterion"] == "min_temperature" else 1
    elif stop_at is not None:
        stop_fn, stop_val = (lambda t, y: t), float(stop_at)
    wall0 = time.time()
    if cfg.solver.method != "trbdf2":
        raise RuntimeError(f"solver.method = {cfg.solver.method}: only the implicit trbdf2 path runs on the GPU")
    kw = dict(integrator_kw or {})
    integ = TRBDF2(model, cfg, h_max=cfg.solver.h_max_fraction * dur, **kw)
    try:
        res = integ.integrate(y0, dur, t_out, on_step=on_step, stop_fn=stop_fn, stop_value=stop_val,
                              stop_direction=direction, device_result=True)
    except SolverError as e:
        if kw.get("conserve") != "project":
            raise
        # the conservation projection can leave states (slightly negative concentrations) the
        # step controller cannot follow; the step is run again with the projection kept at or
        # above zero, and if that fails too, with the upstream polish
        fallback = 0
        for level, mode in ((1, "project_strict"), (2, "polish")):
            log(f"    {step['name']}: {e}; running the step again with conserve={mode}")
            outlet.clear()
            if ring is not None:
                ring.pending.clear()
            counter[0] = 0
            if record_outlet:
                record(0.0, y0, np.zeros(model.n_acc))
            kw["conserve"] = mode
            integ = TRBDF2(model, cfg, h_max=cfg.solver.h_max_fraction * dur, **kw)
            try:
                res = integ.integrate(y0, dur, t_out, on_step=on_step, stop_fn=stop_fn, stop_value=stop_val,
                                      stop_direction=direction, device_result=True)
            except SolverError as e2:
                if level == 2:
                    raise
                e = e2
                continue
            fallback = level
            break
        integ.stats["projection_fallback"] = fallback
    solver_stats = {"method": "trbdf2", "steps": res.n_steps, "rejected": res.n_reject,
                    "newton_fail": res.n_newton_fail, "krylov": res.n_krylov, "jacobians": res.n_jac,
                    "lu": res.n_lu, "nfev": res.nfev, "groups": "all-columns", **integ.stats}
    if record_outlet:
        last = ring.last_t() if ring is not None else outlet[-1]["t_step_s"]
        if last < res.t:
            record(res.t, res.y, res.acc)
        if ring is not None:
            ring.flush()
    wall = time.time() - wall0
    state_out = model.to_state(dv.to_host(res.y), state_in, state_in.t + res.t, step["name"])
    state_out.cycle = cycle_idx
    state_out._y_device = res.y
    snaps = []
    for (ts, ys, _a) in res.snapshots:
        st = model.to_state(ys, state_in, state_in.t + ts, step["name"])
        st.cycle = cycle_idx
        snaps.append({"cycle": cycle_idx, "step": step["name"], "t_step": ts, "t_cycle": t_cycle0 + ts,
                      "t_total": state_in.t + ts, "state": st})
    inv0, inv1 = model.inventories_dev(y0), model.inventories_dev(res.y)
    T_jac = model.T_jacket(res.t) if model.jacket_on else cfg.ambient.T
    sr = StepResult(name=step["name"], type=step["type"], cycle=cycle_idx, t_start_cycle=t_cycle0,
                    duration_planned=dur, duration=res.t, stopped_early=res.stopped_early,
                    acc={nm: float(res.acc[i]) for i, nm in enumerate(model.acc_names)}, inv_start=inv0,
                    inv_end=inv1, outlet=outlet, snapshots=snaps, solver=solver_stats, wall_time=wall,
                    T_jacket=float(T_jac), jacket_on=model.jacket_on, product_port=step.get("product_port", "V3"),
                    valves=dict(step["valves"]), ads_start={s: inv0[f"ads_{s}"] for s in model.adsorbed},
                    ads_end={s: inv1[f"ads_{s}"] for s in model.adsorbed})
    sr.steps_taken = res.steps_taken
    log(f"    {step['name']:<14s} {res.t:9.1f} s  steps={res.n_steps:>5}  wall={wall:6.1f}s")
    model.close()
    return state_out, sr


def _spec(cfg, step_type, name=None):
    for s in cfg.cycle:
        if s["type"] == step_type and (name is None or s["name"] == name):
            return s
    raise KeyError(f"no '{step_type}' step in cfg.cycle")


def _make_runner(step_type):
    def run(state_in, cfg, grid, props, step=None, **kw):
        return run_step(state_in, cfg, grid, props, step or _spec(cfg, step_type), **kw)
    run.__name__ = f"run_{step_type}"
    run.__doc__ = f"Run the {step_type} step (spec from cfg.cycle unless given) -> (state_out, result)."
    return run

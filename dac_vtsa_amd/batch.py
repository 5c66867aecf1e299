"""Many independent cases on one GPU (no upstream counterpart in the engine; the reference runs
sensitivity cases one per process, analysis/sensitivity.py:63-67,121).

Each worker thread owns a HIP stream and runs its share of the cases through the drop-in step
runners; every device call of a case goes to its worker's stream, so the kernels of different
cases overlap on the GPU while each case's own sequence stays ordered.  A case's arithmetic does
not depend on what runs beside it: results are identical to running the cases one after another.

Cases follow the batched-sensitivity workload: grid N_r in 4..12, N_theta in {1, 3..16},
N_z in 20..100, T_heat ~ U(363, 383) K, P_vacuum ~ U(5e3, 2e4) Pa, feed RH ~ U(0.2, 0.8),
feed y_CO2 = 400 ppm, drawn per case from one seed.
"""
from __future__ import annotations

import argparse
import json
import sys
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from typing import Dict, List, Optional

import numpy as np

from . import device as dv


def draw_cases(n_cases: int, seed: int = 0) -> List[Dict]:
    """Per-case grid and operating parameters (fixed draw order per case)."""
    rng = np.random.default_rng(seed)
    cases = []
    for k in range(n_cases):
        Nr = int(rng.integers(4, 13))
        Nth = int(rng.choice([1] + list(range(3, 17))))
        Nz = int(rng.integers(20, 101))
        cases.append({"id": k, "seed": int(rng.integers(0, 2 ** 31 - 1)),
                      "overrides": {"grid.Nr": Nr, "grid.Ntheta": Nth, "grid.Nz": Nz,
                                    "operation.T_heat": float(rng.uniform(363.0, 383.0)),
                                    "operation.P_vacuum": float(rng.uniform(5e3, 2e4)),
                                    "feed.RH": float(rng.uniform(0.2, 0.8)), "feed.y_co2_dry": 4e-4}})
    return cases


def run_case(case: Dict, mode: str = "cycle", state_kind: str = "small", base_overrides: Optional[Dict] = None,
             integrator_kw: Optional[Dict] = None, t_end: Optional[float] = None) -> Dict:
    """One case on the calling thread's current stream.  mode: "cycle" (one full cycle) or the
    name of one step of the cycle.  t_end (optional) shortens every step to that many seconds.
    Returns the case's simulated and wall seconds, balances, KPIs (cycle) and end state."""
    from .balances import check_cycle, cycle_kpis
    from .config import load_config
    from .core.grid import Grid
    from .cycle import run_cycle
    from .properties import get_backend
    from .seeds import seeded_state
    from .steps import run_step
    over = {**(base_overrides or {}), **case["overrides"]}
    cfg = load_config(override_dict=over)
    if t_end is not None:
        for st in cfg.cycle:
            st["duration"] = min(float(st["duration"]), float(t_end))
    grid = Grid.from_config(cfg)
    props = get_backend(cfg)
    qstar = None
    if state_kind == "small":
        # the seeded loadings are fractions of the local equilibrium: device isotherms of a
        # model built on a provisional state of the same grid
        from .core.residual import TVSAModel
        probe = TVSAModel(cfg, grid, props, cfg.cycle[0], seeded_state(grid, "midcycle", case["seed"]))
        qstar = probe.equilibrium
    state = seeded_state(grid, state_kind, case["seed"], qstar)
    if qstar is not None:
        probe.close()
    kw = {"integrator_kw": dict(integrator_kw or {})}
    out = {"id": case["id"], "cells": grid.ncell, "grid": [grid.Nr, grid.Nth, grid.Nz], "ok": True}
    t0 = time.time()
    try:
        if mode == "cycle":
            st, steps, _ = run_cycle(state, cfg, grid, props, 1, **kw)
        else:
            step = next(s for s in cfg.cycle if s["name"] == mode)
            st, sr = run_step(state, cfg, grid, props, step, **kw)
            steps = [sr]
        dv.synchronize()
    except Exception as e:                     # a failed case is reported, the batch goes on
        out.update(ok=False, error=f"{type(e).__name__}: {e}", wall_s=time.time() - t0, sim_s=0.0)
        return out
    out["wall_s"] = time.time() - t0
    out["sim_s"] = sum(s.duration for s in steps)
    mb, eb, _ = check_cycle(steps, props.species)
    out["max_mass_rel_err"] = max(abs(r["relative_error"]) for r in mb if r["step"] != "CYCLE")
    out["max_energy_rel_err"] = max(abs(r["relative_error"]) for r in eb if r["step"] != "CYCLE")
    # how far a step fell back after a failed projection run (0 none, 1 the projection kept at or
    # above zero, 2 the upstream polish), and stages the polish left unclosed
    out["projection_fallback"] = max(int(s.solver.get("projection_fallback", 0) or 0) for s in steps)
    out["balance_unclosed"] = sum(int(s.solver.get("balance_unclosed", 0) or 0) for s in steps)
    if mode == "cycle":
        k = cycle_kpis(steps, cfg, props.species)
        out["kpis"] = {q: k.get(q) for q in ("recovery_pct", "purity_dry_pct", "specific_total_kJ_per_kg",
                                             "productivity_kg_m3_day") if q in k}
    out["state"] = st
    return out


def run_case_to_css(case: Dict, base_overrides: Optional[Dict] = None, integrator_kw: Optional[Dict] = None,
                    state=None, max_cycles: Optional[int] = None, log=None) -> Dict:
    """One case run cycle after cycle to cyclic steady state (vtsa_cycle.py:93-138, the CSS test of
    validation/cyclic_steady_state.py:22-44) on the calling thread's current stream, from
    ``state`` (default: the configuration's initial bed).  Returns per-cycle balances, the last
    cycle's KPIs, whether CSS was reached, and the end state.  ``log`` (optional) receives the step
    runners' progress lines, prefixed with the case id."""
    from .balances import is_css
    from .config import load_config
    from .core.grid import Grid
    from .core.state import initial_state
    from .cycle import run_to_css
    from .properties import get_backend
    cfg = load_config(override_dict={**(base_overrides or {}), **case["overrides"]})
    grid = Grid.from_config(cfg)
    out = {"id": case["id"], "cells": grid.ncell, "grid": [grid.Nr, grid.Nth, grid.Nz], "ok": True}
    t0 = time.time()
    try:
        st0 = state if state is not None else initial_state(cfg, grid, get_backend(cfg))
        case_log = (lambda msg: log(f"[case {case['id']}] {msg}")) if log else None
        records, st = run_to_css(cfg, st0, log=case_log, max_cycles=max_cycles,
                                 integrator_kw=dict(integrator_kw or {}))
        dv.synchronize()
    except Exception as e:                     # a failed case is reported, the batch goes on
        out.update(ok=False, error=f"{type(e).__name__}: {e}", wall_s=time.time() - t0, sim_s=0.0)
        return out
    last = records[-1]
    out.update(wall_s=time.time() - t0, sim_s=sum(r.kpis["cycle_time_s"] for r in records), cycles=len(records),
               css=bool(is_css(last.change, last.kpis, cfg, len(records))),
               kpis=dict(last.kpis), state=st,
               max_mass_rel_err=max(abs(r["relative_error"]) for rec in records for r in rec.mb_rows),
               max_energy_rel_err=max(abs(r["relative_error"]) for rec in records for r in rec.eb_rows))
    return out


def run_batch(cases: List[Dict], workers: int = 16, log=None, runner=None, **kw) -> Dict:
    """Run the cases on ``workers`` threads, one HIP stream each; larger cases are started first.
    ``runner`` is run_case (default) or run_case_to_css, called with the case and ``kw``; ``log``
    (optional) receives one line per finished case.  Returns {"cases": [per-case dicts in input
    order], "wall_s", "sim_s", "value" (aggregate simulated seconds per wall second)}."""
    runner = runner or run_case
    order = sorted(range(len(cases)), key=lambda i: -np.prod(list(cases[i]["overrides"].get(g, 1) for g in
                                                                 ("grid.Nr", "grid.Ntheta", "grid.Nz"))))
    local = threading.local()
    streams, lock = [], threading.Lock()

    def work(i):
        s = getattr(local, "stream", None)
        if s is None:
            s = local.stream = dv.new_stream()
            with lock:
                streams.append(s)
        with dv.use_stream(s):
            r = runner(cases[i], **kw)
        if log:
            cyc = f", {r['cycles']} cycles{' (CSS)' if r.get('css') else ''}" if "cycles" in r else ""
            log(f"case {r['id']} {r['grid']}: {'ok' if r['ok'] else r['error']}, {r['sim_s']:.0f} s simulated "
                f"in {r['wall_s']:.2f} s{cyc}")
        return i, r

    results = [None] * len(cases)
    dv.synchronize()
    t0 = time.time()
    try:
        with ThreadPoolExecutor(max_workers=max(1, int(workers))) as ex:
            for i, r in ex.map(work, order):
                results[i] = r
        wall = time.time() - t0
    finally:
        from .core.residual import release_handles
        for s in streams:
            release_handles(s)
            dv.free_stream(s)
    sim = sum(r["sim_s"] for r in results)
    return {"cases": results, "wall_s": wall, "sim_s": sim, "value": sim / wall,
            "failed": [r["id"] for r in results if not r["ok"]]}


def main(argv=None) -> int:
    """Command line: drawn cases on this GPU, one cycle each or each to cyclic steady state; one
    progress line per finished step and per finished case, then one JSON line with the aggregate
    rate, per-case grids, cycles, KPIs and balances."""
    ap = argparse.ArgumentParser(prog="python -m dac_vtsa_amd.batch", description=main.__doc__)
    ap.add_argument("--cases", type=int, default=64, help="number of drawn cases")
    ap.add_argument("--seed", type=int, default=5, help="seed of the draw")
    ap.add_argument("--workers", type=int, default=16, help="worker threads, one HIP stream each")
    ap.add_argument("--css", action="store_true", help="run every case to cyclic steady state (else one cycle)")
    ap.add_argument("--max-cycles", type=int, default=None, help="cycle limit per case (default: run.max_cycles)")
    ap.add_argument("--precision", choices=("fp32", "fp64"), default="fp32", help="preconditioner factors")
    ap.add_argument("--lowrank", type=int, default=0, help="rank of the preconditioner correction (0 off)")
    ap.add_argument("--part", default="0/1", help="k/n: run only the drawn cases whose id is k mod n (one share "
                                                  "of a batch split across processes)")
    ap.add_argument("--json", default=None, help="also write the result line to this file")
    args = ap.parse_args(argv)
    k_part, n_part = (int(v) for v in args.part.split("/"))
    if not 0 <= k_part < n_part:
        ap.error("--part k/n needs 0 <= k < n")

    def say(msg):
        print(msg, flush=True)
    cases = [c for c in draw_cases(args.cases, seed=args.seed) if c["id"] % n_part == k_part]
    ikw = {"gs_sweeps": 2, "conserve": "project", "precision": args.precision}
    if args.lowrank:
        ikw["lowrank"] = args.lowrank
    out = run_batch(cases, workers=args.workers, log=say, runner=lambda case, **kw: run_case_to_css(case, log=say, **kw),
                    max_cycles=args.max_cycles if args.css else 1,
                    base_overrides={"solver.conservation_tol": 1e-5}, integrator_kw=ikw)
    keys = ("recovery_pct", "purity_dry_pct", "specific_total_kJ_per_kg")
    by_id = {c["id"]: c for c in cases}
    rows = []
    for c in out["cases"]:
        row = {"id": c["id"], "grid": c["grid"], "cells": c["cells"], "ok": c["ok"], "wall_s": round(c["wall_s"], 2)}
        if c["ok"]:
            row.update(cycles=c["cycles"], css=c["css"], sim_s=c["sim_s"],
                       max_mass_rel_err=c["max_mass_rel_err"], max_energy_rel_err=c["max_energy_rel_err"],
                       **{k: c["kpis"][k] for k in keys},
                       **{k: v for k, v in by_id[c["id"]]["overrides"].items() if k.startswith(("operation.", "feed.RH"))})
        else:
            row["error"] = c["error"]
        rows.append(row)
    ok = [r for r in rows if r["ok"]]
    line = {"workload": f"{len(cases)} of {args.cases} drawn cases (seed {args.seed}, part {args.part}), "
                        + ("each to cyclic steady state" if args.css else "one cycle each")
                        + f" from the clean bed, {args.precision} factors"
                        + (f", rank-{args.lowrank} correction" if args.lowrank else "") + f", {args.workers} HIP streams",
            "value": out["value"], "unit": "sim-s/wall-s (all cases together)", "wall_s": out["wall_s"],
            "sim_s": out["sim_s"], "failed": out["failed"],
            "css_reached": sum(1 for r in ok if r.get("css")), "cycles_total": sum(r["cycles"] for r in ok),
            "kpis_min_max": {k: [min(r[k] for r in ok), max(r[k] for r in ok)] for k in keys} if ok else {},
            "max_mass_rel_err": max((r["max_mass_rel_err"] for r in ok), default=None),
            "max_energy_rel_err": max((r["max_energy_rel_err"] for r in ok), default=None),
            "cases": rows}
    text = json.dumps(line)
    if args.json:
        with open(args.json, "w") as f:
            f.write(text + "\n")
    print(text, flush=True)
    return 0 if not out["failed"] else 1


if __name__ == "__main__":
    sys.exit(main())

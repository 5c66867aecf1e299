"""Benchmark: cycle simulated seconds per wall second of the implicit step engine on one MI355X.

Workload (the metric's configuration): grid Nr=20, Ntheta=32, Nz=100 (64,000 cells, 387,200
unknowns), the default cycle adsorption 8,000 s -> blowdown 60 s -> heating 22,400 s -> cooling
400 s (30,860 s), from the clean bed (feed gas at 288.15 K and 1 atm over bare sorbent, the
upstream default ``run.initial_state``).  FP64 residual, FD Jacobian, Newton and GMRES(60)
(Arnoldi orthogonalisation in FP64), line Gauss-Seidel preconditioner (2 forward sweeps over the
two colours of z-lines, each line solved by its block LU), conservation projection,
conservation_tol 1e-5 so that every step's mass and energy balance closes below 1e-6.  Every
form uses the same preconditioner order; they differ in the factors' precision and the learned
parts only.

One bench "step" is one window of 1,543 s of simulated time (1/20 of the cycle) of the cycle
schedule run continuously from the clean bed: ``--steps 20`` times exactly one full cycle.  The
runs that are timed, each over the same K windows from the same bed:

* value: the learned + FP32 form -- the learned stage guess (network trained on the device from a
  plain FP64 run of another initial bed, the feed-equilibrium one), the low-rank correction of
  the preconditioner (rank ``--lowrank``, rebuilt after every factorisation from the harmonic Ritz
  vectors of the first GMRES cycle) and FP32 preconditioner factors; Newton, residual, Jacobian
  and GMRES stay FP64 and decide every step;
* ``plain_fp64``: the plain method with FP64 factors, and ``vs_plain_fp64`` = value / its rate;
* ``fp32``: FP32 factors without the guess.

Beside them, ``batch``: the batched-sensitivity workload (``--batch-cases`` drawn cases of uneven
grids, one full cycle each, on 16 HIP streams at once), reported as its own aggregate rate.

Per-field errors (c_CO2, c_H2O, c_N2, q_CO2, q_H2O, T, Tw) of the FP32 and learned runs against
the plain FP64 run are taken at the end of the last cycle step all runs completed.  The roofline
is that of the dominant kernel (the line Gauss-Seidel half-sweep), from every 7th of its launches
(7 is prime to the 4 half-sweeps of an apply, so every position in the apply is sampled alike)
in the value's timed region bracketed by HIP events on its stream.  The CPU baseline is the
NumPy/SciPy oracle (the upstream algorithms) on one host thread: its unit costs measured on this
workload within a bounded budget, extrapolated with the operation counts of the plain FP64 run
(the extrapolation is labelled in the line).
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
GRID = (20, 32, 100)
CYCLE_S = 30860.0
WINDOW_S = CYCLE_S / 20.0
# HBM bytes per launch of the dominant kernel (k_gs_half<float, 4>, the value form's half-sweep) on
# this workload from two rocprofv3 counter passes (FETCH_SIZE x 1.9993 + WRITE_SIZE x 1.0, the
# factors of a 64 MiB device copy), over the first 1,543 s of the cycle.
TRAFFIC_BYTES_PER_LAUNCH = 56659456
# FP64 operations per launch on the kernel-probe workload (the heating step's model on the clean
# bed at t = 0) from the rocprofv3 counters SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 (wave
# instructions x 64 lanes, an FMA counted as 2); the FD Jacobian's figure sums every kernel its
# launch runs (k_props, k_faces, k_fd_props, k_fd_cells, k_fd_wall).
FP64_FLOP_PER_LAUNCH = {"fd_jacobian": 5127644800, "residual": 280426880}
METRIC = "cycle sim-s/wall-s, 1 GPU: learned+FP32 vs plain FP64 at MB/EB err <=1e-6"


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_cfg(extra=None):
    from dac_vtsa_amd.config import load_config
    nr, nth, nz = GRID
    over = {"grid.Nr": nr, "grid.Ntheta": nth, "grid.Nz": nz, "solver.conservation_tol": 1e-5}
    over.update(extra or {})
    return load_config(override_dict=over)


def setup(extra=None, initial="clean"):
    from dac_vtsa_amd.core.grid import Grid
    from dac_vtsa_amd.core.state import initial_state
    from dac_vtsa_amd.properties import get_backend
    cfg = make_cfg({**(extra or {}), "run.initial_state": initial})
    grid = Grid.from_config(cfg)
    props = get_backend(cfg)
    return cfg, grid, props, initial_state(cfg, grid, props)


def run_windows(cfg, grid, props, state, n_windows, integrator_kw, log=None):
    """The cycle schedule from ``state`` for n_windows x WINDOW_S simulated seconds, the bed handed
    from step to step in HBM-resident runs of ``run_step``; the last step ends at the first
    accepted time step reaching the window boundary.  Returns the simulated seconds advanced, the
    step results and the end state of every completed cycle step."""
    from dac_vtsa_amd.steps import RUNNERS
    t_goal = n_windows * WINDOW_S
    t_done, cycle, T_jac = 0.0, 1, None
    results, ends = [], []
    while t_done < t_goal * (1.0 - 1e-12):
        for step in cfg.cycle:
            left = t_goal - t_done
            dur = float(step["duration"])
            stop = left if left < dur * (1.0 - 1e-12) else None
            state, sr = RUNNERS[step["type"]](state, cfg, grid, props, step=step, cycle_idx=cycle,
                                              T_jacket_prev=T_jac, logger=log, integrator_kw=integrator_kw,
                                              stop_at=stop, progress_s=60.0, device_handoff=True)
            t_done += sr.duration
            T_jac = sr.T_jacket if sr.jacket_on else T_jac
            results.append(sr)
            if stop is None:
                ends.append((cycle, step["name"], state))
            if t_done >= t_goal * (1.0 - 1e-12):
                break
        cycle += 1
    return t_done, results, ends


def balance_errors(results, species):
    """Largest relative mass and energy balance error over every step (upstream definitions,
    postprocessing/reporting.py:113-140 with the validation/conservation_checks.py floors)."""
    from dac_vtsa_amd import balances
    mb_max = eb_max = 0.0
    by_cycle = {}
    for sr in results:
        by_cycle.setdefault(sr.cycle, []).append(sr)
    for steps in by_cycle.values():
        mb, eb, _ = balances.check_cycle(steps, species)
        mb_max = max([mb_max] + [abs(r["relative_error"]) for r in mb if r["step"] != "CYCLE"])
        eb_max = max([eb_max] + [abs(r["relative_error"]) for r in eb if r["step"] != "CYCLE"])
    return mb_max, eb_max


def solver_totals(results):
    keys = ("steps", "rejected", "newton_fail", "jacobians", "lu", "gmres_iters", "nfev", "cons_projected",
            "balance_unclosed", "unclosed_rejected", "guess_used", "guess_kept_predictor", "lowrank_built",
            "lowrank_dropped", "gmres_iters_corrected", "gmres_iters_uncorrected")
    return {k: sum(int(sr.solver.get(k) or 0) for sr in results) for k in keys}


def field_errors(st, ref, cfg):
    """Per field: the max relative difference where |ref| exceeds both 1e-3 of the field's maximum
    and its absolute tolerance, and the max difference in integration-tolerance units
    |a - b| / (atol + rtol |b|)."""
    s = cfg.solver
    fields = [("c_CO2", st.c[0], ref.c[0], s.atol_c_co2), ("c_H2O", st.c[1], ref.c[1], s.atol_c_h2o),
              ("c_N2", st.c[2], ref.c[2], s.atol_c_inert), ("q_CO2", st.q[0], ref.q[0], s.atol_q),
              ("q_H2O", st.q[1], ref.q[1], s.atol_q), ("T", st.T, ref.T, s.atol_T), ("Tw", st.Tw, ref.Tw, s.atol_T)]
    rel, tol = {}, {}
    for name, a, b, atol in fields:
        a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
        big = np.abs(b) > max(1e-3 * np.abs(b).max(), atol)
        rel[name] = float(np.max(np.abs(a - b)[big] / np.abs(b)[big])) if big.any() else 0.0
        tol[name] = float(np.max(np.abs(a - b) / (atol + s.rtol * np.abs(b))))
    return {"max_rel": rel, "max_tol_units": tol}


def timed_leg(name, cfg, grid, props, state, n_windows, kw):
    from dac_vtsa_amd import device as dv
    dv.synchronize()
    t0 = time.time()
    sim, results, ends = run_windows(cfg, grid, props, state, n_windows, kw)
    dv.synchronize()
    wall = time.time() - t0
    mb, eb = balance_errors(results, props.species)
    tot = solver_totals(results)
    _log(f"{name}: {sim:.0f} s simulated in {wall:.2f} s ({sim / wall:.1f} sim-s/wall-s), {tot['steps']} steps, "
         f"{tot['gmres_iters']} GMRES iterations, balances {mb:.2e} / {eb:.2e}")
    per_step = {}
    for sr in results:
        d = per_step.setdefault(sr.name, {"sim_s": 0.0, "wall_s": 0.0, "accepted": 0, "gmres_iters": 0})
        d["sim_s"] += sr.duration
        d["wall_s"] += sr.wall_time
        d["accepted"] += int(sr.solver["steps"])
        d["gmres_iters"] += int(sr.solver.get("gmres_iters") or 0)
    return {"value": sim / wall, "unit": "sim-s/wall-s", "sim_s": sim, "wall_s": wall,
            "balance": {"max_mass_rel_err": mb, "max_energy_rel_err": eb}, "solver": tot,
            "steps": per_step}, ends


def train_network(args, base_kw, log=_log):
    """Record the stages of a plain FP64 run of the cycle from the feed-equilibrium bed (every
    ``--record-stride``-th stage) and train the stage network on the device."""
    from dac_vtsa_amd import device as dv
    from dac_vtsa_amd.learned import StageDataset, StageNet
    cfg, grid, props, st = setup(initial="feed_equilibrium")
    ds = StageDataset(grid.ncell, args.record_stages, stride=args.record_stride)
    t0 = time.time()
    # untimed: the first --warmup windows of this run are the bench's warm-up of the hot path
    n_win = max(args.train_windows, args.warmup)
    run_windows(cfg, grid, props, st, n_win, {**base_kw, "precision": "fp64", "record": ds})
    dv.synchronize()
    t_rec = time.time() - t0
    net = StageNet(grid.ncell, norm=ds.norm(), seed=0)
    hist = ds.train(net, epochs=args.epochs, lr=1e-3, lr_final=1e-4)
    ep, eb, eg = ds.errors(net, hist["val_ids"])
    log(f"learned: {ds.count} stages recorded in {t_rec:.2f} s, trained in {hist['wall_s']:.2f} s "
        f"({hist['steps']} Adam steps), held-out val loss {hist['val'][-1]:.3g}")
    info = {"initial_bed": "feed_equilibrium", "windows": n_win, "stages": ds.count,
            "record_stride": args.record_stride, "rows": ds.count * grid.ncell, "record_s": t_rec,
            "train_s": hist["wall_s"], "adam_steps": hist["steps"], "epochs": args.epochs,
            "val_loss": hist["val"][-1],
            "heldout_rms_err_tol_units": {"predictor": ep.mean(0).tolist(), "preconditioned_step": eb.mean(0).tolist(),
                                          "learned_guess": eg.mean(0).tolist()}}
    return net, info


def kernel_probes(cfg, grid, props, state, n_rhs=200, n_jac=20):
    """Average duration of one residual evaluation f(t, y) (k_props + k_faces + k_flux) and of one
    coloured FD Jacobian (base-point properties and face table, then k_fd_props + k_fd_cells +
    k_fd_wall) on the heating step's model at the bed ``state``, each
    timed with HIP events recorded on the stream it is launched on; the residual's algorithmic
    traffic (read y, write f) and, where measured, both kernels' FP64 throughput."""
    import ctypes as C
    from dac_vtsa_amd import _abi
    from dac_vtsa_amd import device as dv
    from dac_vtsa_amd.core.residual import TVSAModel
    m = TVSAModel(cfg, grid, props, cfg.cycle[2], state)
    y = dv.to_dev(m.pack(state))
    f = dv.empty(m.n_core)
    m.rhs_dev(0.0, y, f)
    dv.synchronize()
    ms = C.c_double()
    _abi.call("vtsa_time_rhs", m.handle, 0.0, dv.ptr(y), dv.ptr(f), n_rhs, C.byref(ms), dv.stream())
    rhs_s = ms.value / 1e3 / n_rhs
    m.rhs_dev(0.0, y, f)
    _abi.call("vtsa_time_fdjac", m.handle, 0.0, dv.ptr(y), dv.ptr(f), n_jac, C.byref(ms), dv.stream())
    jac_s = ms.value / 1e3 / n_jac
    algo = 96 * grid.ncell + 16 * m.n_w
    m.close()
    out = {"residual": {"kernels": "k_props + k_faces + k_flux", "avg_launch_us": rhs_s * 1e6,
                        "algo_bytes_per_launch": algo, "achieved_GBs": algo / rhs_s / 1e9},
           "fd_jacobian": {"kernels": "k_props + k_faces + k_fd_props + k_fd_cells + k_fd_wall (all colours)",
                           "avg_launch_us": jac_s * 1e6}}
    for key, sec in (("residual", rhs_s), ("fd_jacobian", jac_s)):
        flop = FP64_FLOP_PER_LAUNCH[key]
        if flop:
            tf = flop / sec / 1e12
            out[key].update({"fp64_flop_per_launch": flop, "achieved_tflops": tf, "peak_tflops": FP64_PEAK_TFLOPS,
                             "fp64_frac": tf / FP64_PEAK_TFLOPS})
    return out


def batch_leg(n_cases, workers, seed=5):
    """The batched-sensitivity workload on this GPU: n_cases drawn cases (grid Nr 4..12, Ntheta in
    {1, 3..16}, Nz 20..100; T_heat, P_vacuum, feed RH per case), one full cycle each from the clean
    bed with FP32 factors, on `workers` HIP streams at once; aggregate simulated seconds per wall
    second and the per-case KPIs."""
    from dac_vtsa_amd.batch import draw_cases, run_batch, run_case_to_css
    cases = draw_cases(n_cases, seed=seed)
    out = run_batch(cases, workers=workers, runner=run_case_to_css, max_cycles=1,
                    base_overrides={"solver.conservation_tol": 1e-5},
                    integrator_kw={"gs_sweeps": 2, "conserve": "project", "precision": "fp32"})
    ok = [c for c in out["cases"] if c["ok"]]
    cells = [c["cells"] for c in out["cases"]]
    _log(f"batch: {n_cases} cases on {workers} streams, {out['sim_s']:.0f} s simulated in {out['wall_s']:.2f} s "
         f"({out['value']:.0f} sim-s/wall-s), failed {out['failed']}")
    kp = {k: [c["kpis"][k] for c in ok] for k in ("recovery_pct", "purity_dry_pct", "specific_total_kJ_per_kg")}
    return {"workload": f"{n_cases} drawn cases ({min(cells)}-{max(cells)} cells), one full cycle each from the "
                        f"clean bed, FP32 factors, {workers} HIP streams",
            "value": out["value"], "unit": "sim-s/wall-s (all cases together)", "wall_s": out["wall_s"],
            "sim_s": out["sim_s"], "failed": out["failed"],
            "max_mass_rel_err": max(c["max_mass_rel_err"] for c in ok) if ok else None,
            "max_energy_rel_err": max(c["max_energy_rel_err"] for c in ok) if ok else None,
            "kpis_min_max": {k: [min(v), max(v)] for k, v in kp.items() if v},
            "cases": [{"grid": c["grid"], "wall_s": c["wall_s"], **{k: c["kpis"][k] for k in kp}} for c in ok]}


def copy_probe(mib=64, repeat=5):
    """Calibration of the HBM traffic counters: `repeat` device-to-device copies of `mib` MiB."""
    from dac_vtsa_amd import device as dv
    n = mib * (1 << 20) // 8
    a, b = dv.zeros(n), dv.empty(n)
    for _ in range(repeat):
        b.copy_from(a)
    dv.synchronize()
    print(json.dumps({"copy_probe_bytes_per_dispatch": n * 8, "dispatches": repeat}), flush=True)


def _cpu_units(queue, budget_s):
    """Child process: the oracle's unit costs on the workload, one host thread."""
    from threadpoolctl import threadpool_limits
    with threadpool_limits(1):
        import scipy.sparse as sp
        import scipy.sparse.linalg as spla
        from oracle.jacobian import OracleFDJacobian
        from oracle.model import OracleModel
        from oracle.runner import make, initial_state
        cfg = make_cfg()
        g, gas = make(cfg)
        st = initial_state(cfg, g, gas)
        m = OracleModel(cfg, g, gas, cfg.cycle[2], st)
        y = m.pack(st)
        out = {}
        t0 = time.time()
        f, _ = m.rhs(0.0, y, want_acc=False)
        t0 = time.time()
        for _ in range(3):
            f, _ = m.rhs(0.0, y, want_acc=False)
        out["rhs_s"] = (time.time() - t0) / 3
        queue.put(dict(out))
        t0 = time.time()
        jac = OracleFDJacobian(m)
        out["pattern_s"] = time.time() - t0
        out["colours"] = jac.ngroups
        t0 = time.time()
        J = jac(0.0, y, f)
        out["jacobian_s"] = time.time() - t0
        queue.put(dict(out))
        D = 1.0 - 1.0 / np.sqrt(2.0)
        M = (sp.identity(J.shape[0], format="csc") - (D * 100.0) * jac.theta_block(J)).tocsc()
        t0 = time.time()
        out["splu_started"] = True
        queue.put(dict(out))
        lu = spla.splu(M, permc_spec="COLAMD")
        out["splu_s"] = time.time() - t0
        t0 = time.time()
        lu.solve(f)
        out["lu_solve_s"] = time.time() - t0
        queue.put(dict(out))


def cpu_baseline(budget_s, counts, sim_s):
    """Oracle unit costs (one host thread, bounded by budget_s; a factorisation still running at
    the budget is reported as at least that long) x the operation counts of the plain FP64 GPU
    run, as an extrapolated CPU time for the same simulated window."""
    ctx = mp.get_context("spawn")
    q = ctx.Queue()
    p = ctx.Process(target=_cpu_units, args=(q, budget_s), daemon=True)
    t0 = time.time()
    p.start()
    units = {}
    while time.time() - t0 < budget_s and (p.is_alive() or not q.empty()):
        try:
            units = q.get(timeout=1.0)
        except Exception:
            pass
        if "lu_solve_s" in units:
            break
    elapsed = time.time() - t0
    if p.is_alive():
        p.terminate()
        p.join(5)
    if "splu_started" in units and "splu_s" not in units:
        units["splu_s_lower_bound"] = elapsed - sum(units.get(k, 0.0) for k in ("pattern_s", "jacobian_s")) \
            - 4 * units.get("rhs_s", 0.0)
    rhs = units.get("rhs_s")
    if rhs is None:
        return None
    lu = units.get("splu_s", units.get("splu_s_lower_bound", 0.0))
    solve = units.get("lu_solve_s", 0.0)
    est = (counts["nfev"] * rhs + counts["jacobians"] * units.get("jacobian_s", 90 * rhs)
           + counts["lu"] * lu + (counts["gmres_iters"] + 2 * counts["steps"]) * solve + units.get("pattern_s", 0.0))
    try:
        cpu_model = next(line.split(":", 1)[1].strip() for line in open("/proc/cpuinfo") if line.startswith("model name"))
    except Exception:
        cpu_model = platform.processor()
    lower = "splu_s" not in units
    return {"value": sim_s / est, "unit": "sim-s/wall-s", "cores": 1, "kind": "port",
            "sample": (f"oracle (NumPy/SciPy restatement of the upstream algorithms) on 1 host thread of "
                       f"'{cpu_model}' (nproc {os.cpu_count()}), unit costs measured in {elapsed:.0f} s on this "
                       f"workload: rhs {rhs:.3f} s, coloured FD Jacobian {units.get('jacobian_s', float('nan')):.2f} s, "
                       f"theta-block SuperLU factor {'>=' if lower else ''}{lu:.1f} s, LU solve "
                       f"{solve:.3f} s; EXTRAPOLATED to {sim_s:.0f} simulated s with the plain FP64 GPU run's "
                       f"counts ({counts['nfev']} rhs, {counts['jacobians']} Jacobians, {counts['lu']} "
                       f"factorisations, {counts['gmres_iters']} GMRES iterations): {est:.0f} s"
                       + (" (a lower bound: the factorisation did not finish within the budget)" if lower else "")),
            "units": units}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="windows of 1,543 simulated s (20 = one cycle)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--sweeps", type=int, default=2, help="line Gauss-Seidel sweeps of the preconditioner")
    ap.add_argument("--gs-order", default="forward", choices=("forward", "symmetric"),
                    help="line Gauss-Seidel colour order (forward: 0 1 0 1; symmetric: 0 1 0 1 0)")
    ap.add_argument("--krylov", type=int, default=60, help="GMRES restart length")
    ap.add_argument("--lowrank", type=int, default=8, help="rank of the value leg's preconditioner correction (0 off)")
    ap.add_argument("--batch-cases", type=int, default=16, help="cases of the batch leg (0 skips it)")
    ap.add_argument("--cpu-budget-s", type=float, default=40.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 leg")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-only leg")
    ap.add_argument("--no-learned", action="store_true", help="value = FP32 factors without the learned guess")
    ap.add_argument("--train-windows", type=int, default=20)
    ap.add_argument("--record-stages", type=int, default=400)
    ap.add_argument("--record-stride", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--copy-probe", action="store_true", help="only run the traffic-counter calibration copy")
    ap.add_argument("--kernel-probe", action="store_true", help="only time the residual and FD Jacobian kernels")
    ap.add_argument("--profile", choices=("plain", "fp32", "value"),
                    help="only run --steps windows of one form (plain FP64; FP32 factors; FP32 factors with the "
                         "low-rank correction, without the learned guess), for profiling")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (independent cases would shard "
              "one group per GPU with no communication; no multi-GPU run is part of this build)", file=sys.stderr)
        return 2
    from dac_vtsa_amd.core import residual
    base_kw = {"gs_sweeps": args.sweeps, "gs_order": args.gs_order, "conserve": "project", "krylov_dim": args.krylov}
    cfg, grid, props, bed = setup()
    if args.kernel_probe:
        print(json.dumps(kernel_probes(cfg, grid, props, bed)), flush=True)
        return 0
    if args.profile:
        form = {"plain": {"precision": "fp64"}, "fp32": {"precision": "fp32"},
                "value": {"precision": "fp32", "lowrank": args.lowrank}}[args.profile]
        res, _ = timed_leg(args.profile, cfg, grid, props, bed, args.steps, {**base_kw, **form})
        print(json.dumps({"profile": args.profile, "windows": args.steps,
                          **{k: res[k] for k in ("value", "sim_s", "wall_s", "solver")}}), flush=True)
        return 0
    _log(f"workload {GRID[0]}x{GRID[1]}x{GRID[2]} cycle from the clean bed, {args.steps} windows of {WINDOW_S:.0f} s")
    net, train = (None, None)
    if args.no_learned:
        cfg_w, grid_w, props_w, bed_w = setup(initial="feed_equilibrium")
        t0 = time.time()
        run_windows(cfg_w, grid_w, props_w, bed_w, args.warmup, {**base_kw, "precision": "fp64"})
        _log(f"warm-up: {args.warmup} windows in {time.time() - t0:.2f} s")
    else:
        net, train = train_network(args, base_kw)
    main_kw = {**base_kw, "precision": "fp32", "lowrank": args.lowrank}
    if net is not None:
        main_kw["guess"] = net
    residual.reset_kernel_profile(7)
    main, ends_main = timed_leg("learned+fp32" if net is not None or args.lowrank else "fp32", cfg, grid, props,
                                bed, args.steps, main_kw)
    prof = dict(residual.KERNEL_PROFILE)
    residual.reset_kernel_profile(0)
    avg_s = prof["ms"] / 1e3 / max(prof["launches"], 1)
    bpl = prof["bytes"] / max(prof["launches"], 1)
    achieved = bpl / avg_s / 1e9 if avg_s > 0 else 0.0
    line = {
        "metric": METRIC, "value": main["value"], "unit": "sim-s/wall-s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * main["wall_s"] / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 (fp32 preconditioner factors)",
        "data": "synthetic: clean bed (feed gas over bare sorbent); network trained on the simulator's own stages",
        "config": {"workload": f"full TVSA cycle (adsorption 8000 s, blowdown 60 s, heating 22400 s, cooling 400 s) "
                               f"on grid {GRID[0]}x{GRID[1]}x{GRID[2]} (387200 unknowns), clean bed",
                   "step": f"{WINDOW_S:.0f} simulated s of the cycle schedule (20 steps = one cycle)",
                   "form": ("learned stage guess + " if net is not None else "") +
                           (f"rank-{args.lowrank} preconditioner correction + " if args.lowrank else "") +
                           "FP32 preconditioner factors",
                   "solver": f"sdirk2 Newton-GMRES({args.krylov}), line Gauss-Seidel ({args.sweeps} {args.gs_order} "
                             f"sweeps), projection",
                   "rtol": cfg.solver.rtol, "conservation_tol": cfg.solver.conservation_tol},
        "balance": main["balance"], "solver": main["solver"], "steps_detail": main["steps"],
        "roofline": {"bound": "hbm", "kernel": "k_gs_half (line Gauss-Seidel half-sweep)", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": TRAFFIC_BYTES_PER_LAUNCH, "avg_launch_us": avg_s * 1e6, "algo_bytes_per_launch": bpl,
                     "sampled_launches": prof["launches"]},
    }
    if train is not None:
        line["training"] = train
    ends = {"main": ends_main}
    if not args.no_plain:
        plain, ends["plain"] = timed_leg("plain fp64", cfg, grid, props, bed, args.steps, {**base_kw, "precision": "fp64"})
        line["plain_fp64"] = plain
        line["vs_plain_fp64"] = main["value"] / plain["value"]
    learned = net is not None or args.lowrank > 0
    if not args.no_fp32 and learned:
        fp32, ends["fp32"] = timed_leg("fp32", cfg, grid, props, bed, args.steps, {**base_kw, "precision": "fp32"})
        line["fp32"] = fp32
    if "plain" in ends and ends["plain"]:
        key = min(len(e) for e in ends.values())
        if key:
            ref = ends["plain"][key - 1]
            line["field_err_vs_plain_fp64"] = {"at": f"cycle {ref[0]} end of {ref[1]}"}
            for nm in ("main", "fp32"):
                if nm in ends:
                    line["field_err_vs_plain_fp64"]["learned_fp32" if nm == "main" and learned else
                                                     ("fp32" if nm == "fp32" else "value_run")] = \
                        field_errors(ends[nm][key - 1][2], ref[2], cfg)
    line.update(kernel_probes(cfg, grid, props, bed))
    if args.batch_cases > 0:
        line["batch"] = batch_leg(args.batch_cases, workers=min(16, args.batch_cases))
    if not args.no_cpu and "plain_fp64" in line:
        cb = cpu_baseline(args.cpu_budget_s, line["plain_fp64"]["solver"], line["plain_fp64"]["sim_s"])
        if cb is not None:
            line["cpu_baseline"] = cb
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

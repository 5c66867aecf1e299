"""Benchmark: simulated seconds per wall second of the implicit step engine on one MI355X.

Workload (the metric's single-GPU configuration): grid Nr=10, Ntheta=16, Nz=50 (8,000 cells,
48,800 unknowns), seeded mid-cycle state (seed 0: T~U(288,373) K, P~U(1.5e4,1.013e5) Pa,
y_CO2~U(0,0.9), y_H2O~U(0,0.05), q_CO2~U(0,1.6), q_H2O~U(0,6) mol/kg), one full heating step of
22,400 s (jacket 373.15 K, product vacuum pump at 1.5e4 Pa), FP64 residual / Jacobian / Newton /
GMRES with the line Gauss-Seidel preconditioner (``--precision fp32`` stores its factors in FP32).

One bench "step" = one complete heating step through the drop-in ``run_step`` (outlet recording
and balances included, as upstream), from the same seeded state.  value = steps x 22,400 s /
wall seconds of the timed region.

Also reported: the per-step mass/energy balance errors; the roofline of the dominant kernel, the
line Gauss-Seidel half-sweep, from every 8th of its launches in the timed region bracketed by HIP
events on its stream (algorithmic bytes per launch / average launch time); the residual kernel
pair timed the same way as a secondary figure; a CPU baseline: the NumPy/SciPy oracle (same
algorithms as upstream, 1 thread) on the same state for a bounded wall-clock sample, with the GPU
wall time over the same simulated window; one full cycle (``cycle``: adsorption, blowdown,
heating, cooling, 30,860 s) on the same grid; and the learned form (``learned_fp32``): the same step
with the learned stage guess and FP32 preconditioner factors, the network trained on the device
from a plain run of another seeded state, its rate against the plain FP64 rate, its balances and
its end-field differences from the plain FP64 run.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per k_gs_half launch on this workload, from two rocprofv3 counter passes over one bench
# step (FETCH_SIZE x 1.9993 + WRITE_SIZE x 1.0000 KiB, the factors from a 64 MiB device copy):
# 5643.4 x 1.9993 + 373.1 KiB.  Re-measure when the kernel changes.
TRAFFIC_BYTES_PER_LAUNCH = (5643.4 * 1.9993 + 373.1) * 1024
FP64_PEAK_TFLOPS = 78.6
WORKLOAD = {"grid.Nr": 10, "grid.Ntheta": 16, "grid.Nz": 50, "operation.P_vacuum": 1.5e4}
STEP_INDEX = 2          # heating in the default cycle
SEED = 0


def setup(precision, overrides=None):
    from dac_vtsa_amd.config import load_config
    from dac_vtsa_amd.core.grid import Grid
    from dac_vtsa_amd.properties import get_backend
    from dac_vtsa_amd.seeds import seeded_state
    cfg = load_config(override_dict={**WORKLOAD, **(overrides or {})})
    grid = Grid.from_config(cfg)
    props = get_backend(cfg)
    state = seeded_state(grid, "midcycle", SEED)
    return cfg, grid, props, state


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def one_step(cfg, grid, props, state, precision, sweeps=2, conserve="polish"):
    from dac_vtsa_amd.steps import run_step
    return run_step(state, cfg, grid, props, cfg.cycle[STEP_INDEX], logger=_log, progress_s=20.0,
                    integrator_kw={"precision": precision, "gs_sweeps": sweeps, "conserve": conserve})


def copy_probe(mib=64, repeat=5):
    """Calibration for the HBM traffic counters: `repeat` device-to-device copies of `mib` MiB
    (the runtime's copy kernel reads and writes exactly that many bytes per dispatch)."""
    from dac_vtsa_amd import device as dv
    n = mib * (1 << 20) // 8
    a, b = dv.zeros(n), dv.empty(n)
    for _ in range(repeat):
        b.copy_from(a)
    dv.synchronize()
    print(json.dumps({"copy_probe_bytes_per_dispatch": n * 8, "dispatches": repeat}), flush=True)


def residual_roofline(cfg, grid, props, state, n_launch=200):
    """Average duration of one residual evaluation (k_props + k_flux, the kernel pair of every
    line-search point and error estimate), timed with HIP events recorded by libvtsa on the
    stream it launches on, and its algorithmic traffic (read y, write f)."""
    import ctypes as C
    from dac_vtsa_amd import _abi
    from dac_vtsa_amd import device as dv
    from dac_vtsa_amd.core.residual import TVSAModel
    m = TVSAModel(cfg, grid, props, cfg.cycle[STEP_INDEX], state)
    y = dv.to_dev(m.pack(state))
    f = dv.empty(m.n_core)
    m.rhs_dev(0.0, y, f)
    dv.synchronize()
    ms = C.c_double()
    _abi.call("vtsa_time_rhs", m.handle, 0.0, dv.ptr(y), dv.ptr(f), n_launch, C.byref(ms), dv.stream())
    avg_s = ms.value / 1e3 / n_launch
    n, nw = grid.ncell, m.n_w
    algo_bytes = 96 * n + 16 * nw          # read y, write f (FP64)
    m.close()
    return avg_s, algo_bytes


def gpu_window(cfg, grid, props, state, t_end, precision, sweeps=2, conserve="polish"):
    """GPU wall seconds for the first t_end simulated seconds of the step (same integrator)."""
    from dac_vtsa_amd import device as dv
    from dac_vtsa_amd.core.residual import TVSAModel
    from dac_vtsa_amd.core.solver import TRBDF2
    step = cfg.cycle[STEP_INDEX]
    m = TVSAModel(cfg, grid, props, step, state)
    integ = TRBDF2(m, cfg, h_max=cfg.solver.h_max_fraction * float(step["duration"]), precision=precision,
                   gs_sweeps=sweeps, conserve=conserve)
    dv.synchronize()
    t0 = time.time()
    integ.integrate(m.pack(state), t_end, np.array([t_end]))
    dv.synchronize()
    wall = time.time() - t0
    m.close()
    return wall


def cpu_baseline(cfg, state, budget_s, sr):
    """The oracle (NumPy/SciPy, the upstream algorithms) on 1 host thread, two windows of
    budget_s / 2 wall seconds each: from the start of the heating step, and from the GPU
    trajectory's snapshot nearest mid-step (with the GPU's step size there as the first step).
    Returns [(t_start, simulated seconds, wall seconds, accepted steps)] per window."""
    from oracle.model import OracleModel
    from oracle.runner import make
    from oracle.solver import OracleTRBDF2
    g, gas = make(cfg)
    step = cfg.cycle[STEP_INDEX]
    dur = float(step["duration"])
    mid = min(sr.snapshots, key=lambda sn: abs(sn["t_step"] - 0.5 * dur))
    t_mid = float(mid["t_step"])
    h_mid = min(sr.steps_taken, key=lambda th: abs(th[0] - t_mid))[1]
    out = []
    for t0, st, h0 in ((0.0, state, None), (t_mid, mid["state"], h_mid)):
        m = OracleModel(cfg, g, gas, step, state)     # the step's model is built from its initial state
        integ = OracleTRBDF2(m, cfg, h_max=cfg.solver.h_max_fraction * dur, h_init=h0)
        w0 = time.time()
        res = integ.integrate(m.pack(st), dur, np.array([dur]), stop_fn=lambda t, y: time.time() - w0,
                              stop_value=0.5 * budget_s, t0=t0)
        out.append((t0, res.t - t0, time.time() - w0, res.n_steps))
    return out


def field_errors(st, ref, cfg):
    """Per field (c_CO2, c_H2O, c_N2, q_CO2, q_H2O, T, Tw) against a reference end state: the max
    relative difference where |ref| is above both 1e-3 of the field's maximum and atol, and the max difference
    in integration-tolerance units |a - b| / (atol + rtol |b|) (atol of the configuration)."""
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


def cycle_leg(cfg, grid, props, state, args):
    """One full adsorption -> blowdown -> heating -> cooling cycle (30,860 s) on the same grid from
    the same seeded state, plain FP64, through the cycle driver (vtsa_cycle.py:82-84 order)."""
    from dac_vtsa_amd import device as dv
    from dac_vtsa_amd.cycle import run_cycle
    from dac_vtsa_amd import balances
    kw = {"integrator_kw": {"precision": "fp64", "gs_sweeps": args.sweeps, "conserve": args.conserve}}
    dv.synchronize()
    t0 = time.time()
    _, steps, _ = run_cycle(state, cfg, grid, props, 1, **kw)
    dv.synchronize()
    wall = time.time() - t0
    sim = sum(s.duration for s in steps)
    mb, eb, _ = balances.check_cycle(steps, props.species)
    _log(f"cycle: {sim:.0f} s simulated in {wall:.2f} s")
    return {"value": sim / wall, "unit": "sim-s/wall-s", "sim_s": sim, "wall_s": wall,
            "max_mass_rel_err": max(abs(r["relative_error"]) for r in mb if r["step"] != "CYCLE"),
            "max_energy_rel_err": max(abs(r["relative_error"]) for r in eb if r["step"] != "CYCLE"),
            "steps": {s.name: {"sim_s": s.duration, "wall_s": s.wall_time, "accepted": s.solver["steps"],
                               "gmres_iters": s.solver.get("gmres_iters")} for s in steps}}


def batch_leg(args):
    """Independent sensitivity cases (grid N_r 4-12, N_theta 1-16, N_z 20-100, T_heat, P_vacuum and
    feed RH drawn per case) run concurrently, one HIP stream per worker thread: one heating step
    of 22,400 s each from a seeded mid-cycle state.  value = total simulated seconds / wall."""
    from dac_vtsa_amd.batch import draw_cases, run_batch
    cases = draw_cases(args.batch, seed=0)
    r = run_batch(cases, workers=args.workers, log=_log, mode="heating", state_kind="midcycle",
                  base_overrides={"solver.conservation_tol": args.cons_tol},
                  integrator_kw={"gs_sweeps": args.sweeps, "conserve": args.conserve})
    ok = [c for c in r["cases"] if c["ok"]]
    return {"value": r["value"], "unit": "sim-s/wall-s (all cases)", "cases": len(cases), "workers": args.workers,
            "wall_s": r["wall_s"], "sim_s": r["sim_s"], "failed": r["failed"],
            "cells_range": [min(c["cells"] for c in r["cases"]), max(c["cells"] for c in r["cases"])],
            "max_mass_rel_err": max((c["max_mass_rel_err"] for c in ok), default=None),
            "max_energy_rel_err": max((c["max_energy_rel_err"] for c in ok), default=None),
            "projection_fallback_cases": sum(1 for c in ok if c.get("projection_fallback")),
            "unclosed_stages": sum(c.get("balance_unclosed", 0) for c in ok)}


def learned_leg(cfg, grid, props, state, args, st_plain, plain_value):
    """Learned stage guess + FP32 preconditioner on the same heating step.  The network is trained
    on the device from a plain FP64 run of another seeded state (``--train-seed``): that run
    records every converged stage, 20 epochs of Adam follow, the output is shrunk per variable on
    held-out stages; then the seed-0 step is timed with the guess and FP32 factors."""
    from dac_vtsa_amd import device as dv
    from dac_vtsa_amd.learned import StageDataset, StageNet
    from dac_vtsa_amd.seeds import seeded_state
    from dac_vtsa_amd.steps import run_step
    step = cfg.cycle[STEP_INDEX]
    dur = float(step["duration"])
    n = grid.ncell
    base = {"gs_sweeps": args.sweeps, "conserve": args.conserve}
    ds = StageDataset(n, 700)
    t0 = time.time()
    run_step(seeded_state(grid, "midcycle", args.train_seed), cfg, grid, props, step,
             integrator_kw={**base, "precision": "fp64", "record": ds})
    dv.synchronize()
    t_rec = time.time() - t0
    net = StageNet(n, norm=ds.norm(), seed=0)
    hist = ds.train(net, epochs=args.epochs, lr=1e-3, lr_final=1e-4)
    ep, eb, eg = ds.errors(net, hist["val_ids"])
    _log(f"learned: {ds.count} stages recorded in {t_rec:.2f} s, trained in {hist['wall_s']:.2f} s "
         f"({hist['steps']} Adam steps), held-out val loss {hist['val'][-1]:.3g}")
    kw = {**base, "precision": "fp32", "guess": net}
    for _ in range(args.warmup):
        run_step(state, cfg, grid, props, step, integrator_kw=kw)
    dv.synchronize()
    t0 = time.time()
    for _ in range(args.steps):
        st, sr = run_step(state, cfg, grid, props, step, integrator_kw=kw)
    dv.synchronize()
    rate = args.steps * dur / (time.time() - t0)
    mb, eb_ = balance_errors(sr, props.species)
    _log(f"learned+fp32: {rate:.0f} sim-s/wall-s ({rate / plain_value:.2f}x plain fp64)")
    return {
        "value": rate, "unit": "sim-s/wall-s", "vs_plain_fp64": rate / plain_value,
        "balance": {"max_mass_rel_err": mb, "max_energy_rel_err": eb_},
        "field_diff_vs_plain_fp64": field_errors(st, st_plain, cfg),
        "solver": {k: sr.solver.get(k) for k in ("steps", "rejected", "gmres_iters", "guess_used",
                                                 "guess_kept_predictor")},
        "training": {"train_seed": args.train_seed, "stages": ds.count, "rows": ds.count * n,
                     "record_s": t_rec, "train_s": hist["wall_s"], "adam_steps": hist["steps"],
                     "epochs": args.epochs, "val_loss": hist["val"][-1], "gain": hist["gain"],
                     "heldout_rms_err_tol_units": {"predictor": ep.mean(0).tolist(),
                                                   "preconditioned_step": eb.mean(0).tolist(),
                                                   "learned_guess": eg.mean(0).tolist()}},
    }


def balance_errors(sr, species):
    from dac_vtsa_amd import balances
    mb, eb, _ = balances.check_cycle([sr], species)
    return (max(abs(r["relative_error"]) for r in mb if r["step"] != "CYCLE"),
            max(abs(r["relative_error"]) for r in eb if r["step"] != "CYCLE"))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=("fp64", "fp32"), default="fp64")
    ap.add_argument("--cpu-sample-s", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sweeps", type=int, default=2, help="line Gauss-Seidel sweeps of the preconditioner")
    ap.add_argument("--cons-tol", type=float, default=1e-5,
                    help="solver.conservation_tol (upstream default 1e-3; the metric needs balances <= 1e-6)")
    ap.add_argument("--rtol", type=float, default=None, help="solver.rtol (default: config)")
    ap.add_argument("--conserve", choices=("polish", "project"), default="project",
                    help="stage conservation: upstream polish, or the conservation projection first")
    ap.add_argument("--copy-probe", action="store_true", help="only run the traffic-counter calibration copy")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-guess + FP32 preconditioner leg")
    ap.add_argument("--no-cycle", action="store_true", help="skip the full-cycle leg")
    ap.add_argument("--batch", type=int, default=0,
                    help="also run this many independent sensitivity cases concurrently (heating step each)")
    ap.add_argument("--workers", type=int, default=16, help="worker threads (HIP streams) of the batch leg")
    ap.add_argument("--train-seed", type=int, default=1, help="seed of the state whose run trains the network")
    ap.add_argument("--epochs", type=int, default=20)
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (independent cases would shard "
              "one group per GPU with no communication; no multi-GPU run is part of this build)", file=sys.stderr)
        return 2
    from dac_vtsa_amd import device as dv
    over = {}
    if args.cons_tol is not None:
        over["solver.conservation_tol"] = args.cons_tol
    if args.rtol is not None:
        over["solver.rtol"] = args.rtol
    cfg, grid, props, state = setup(args.precision, over)
    dur = float(cfg.cycle[STEP_INDEX]["duration"])
    print(f"[bench] workload 10x16x50 heating {dur:.0f} s, precision {args.precision}", file=sys.stderr, flush=True)
    for i in range(args.warmup):
        t0 = time.time()
        one_step(cfg, grid, props, state, args.precision, args.sweeps, args.conserve)
        print(f"[bench] warmup {i + 1}/{args.warmup}: {time.time() - t0:.2f} s", file=sys.stderr, flush=True)
    from dac_vtsa_amd.core import residual
    residual.reset_kernel_profile(8)
    dv.synchronize()
    t0 = time.time()
    results = []
    for i in range(args.steps):
        st_plain, sr = one_step(cfg, grid, props, state, args.precision, args.sweeps, args.conserve)
        results.append(sr)
        print(f"[bench] step {i + 1}/{args.steps}: {sr.solver['steps']} accepted steps, "
              f"{sr.solver['rejected']} rejected, {sr.solver.get('gmres_iters', 0)} GMRES iterations",
              file=sys.stderr, flush=True)
    dv.synchronize()
    elapsed = time.time() - t0
    prof = dict(residual.KERNEL_PROFILE)
    residual.reset_kernel_profile(0)
    value = args.steps * dur / elapsed
    mb, eb = balance_errors(results[-1], props.species)
    avg_s = prof["ms"] / 1e3 / max(prof["launches"], 1)
    bytes_per_launch = prof["bytes"] / max(prof["launches"], 1)
    achieved = bytes_per_launch / avg_s / 1e9
    res_s, res_bytes = residual_roofline(cfg, grid, props, state)
    line = {
        "metric": "cycle sim-s/wall-s, 1 GPU: learned+FP32 vs plain FP64 at MB/EB err <=1e-6",
        "value": value, "unit": "sim-s/wall-s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f64 (fp32 preconditioner)",
        "data": "synthetic seeded mid-cycle state (seed 0)",
        "config": {"workload": "heating step 22400 s, grid 10x16x50 (48800 unknowns), P_vac 1.5e4 Pa",
                   "preconditioner": f"zline_gs ({args.sweeps} sweep{'s' if args.sweeps > 1 else ''})", "precision": args.precision,
                   "solver": "sdirk2 Newton-GMRES(30)", "conserve": args.conserve, "rtol": cfg.solver.rtol,
                   "conservation_tol": cfg.solver.conservation_tol},
        "balance": {"max_mass_rel_err": mb, "max_energy_rel_err": eb},
        "solver": {k: results[-1].solver.get(k) for k in ("steps", "rejected", "newton_fail", "jacobians", "lu",
                                                          "gmres_iters", "gmres_polish", "balance_unclosed",
                                                          "cons_projected", "nfev")},
        "roofline": {"bound": "hbm", "kernel": "k_gs_half (line Gauss-Seidel half-sweep)", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": TRAFFIC_BYTES_PER_LAUNCH,
                     "avg_launch_us": avg_s * 1e6, "algo_bytes_per_launch": bytes_per_launch,
                     "sampled_launches": prof["launches"]},
        "residual_pair": {"kernels": "k_props + k_flux", "avg_launch_us": res_s * 1e6,
                          "algo_bytes_per_launch": res_bytes, "achieved_GBs": res_bytes / res_s / 1e9},
    }
    if not args.no_cycle:
        line["cycle"] = cycle_leg(cfg, grid, props, state, args)
    if args.batch > 0:
        line["batch"] = batch_leg(args)
    if not args.no_learned:
        line["learned_fp32"] = learned_leg(cfg, grid, props, state, args, st_plain, value)
    if not args.no_cpu:
        from threadpoolctl import threadpool_limits
        with threadpool_limits(1):
            wins = cpu_baseline(cfg, state, args.cpu_sample_s, results[-1])
        (_, s0, w0, n0), (tm, s1, w1, n1) = wins
        g_wall = gpu_window(cfg, grid, props, state, s0, args.precision, args.sweeps, args.conserve)
        line["cpu_baseline"] = {"value": s1 / w1, "unit": "sim-s/wall-s", "cores": 1, "kind": "port",
                                "sample": f"oracle on 1 host thread, two windows of the same heating step: "
                                          f"from t={tm:.0f} s (GPU snapshot, mid-step) {s1:.3g} s simulated in "
                                          f"{w1:.1f} s ({n1} steps, this rate is the value); from t=0 {s0:.3g} s "
                                          f"in {w0:.1f} s ({n0} steps, {s0 / w0:.3g} sim-s/wall-s; the GPU runs "
                                          f"that first window in {g_wall:.3f} s)"}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

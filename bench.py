"""Benchmark: simulated seconds per wall second of the implicit step engine on one MI355X.

Workload (the metric's single-GPU configuration): grid Nr=10, Ntheta=16, Nz=50 (8,000 cells,
48,800 unknowns), seeded mid-cycle state (seed 0: T~U(288,373) K, P~U(1.5e4,1.013e5) Pa,
y_CO2~U(0,0.9), y_H2O~U(0,0.05), q_CO2~U(0,1.6), q_H2O~U(0,6) mol/kg), one full heating step of
22,400 s (jacket 373.15 K, product vacuum pump at 1.5e4 Pa), FP64 residual / Jacobian / Newton /
GMRES with the z-line preconditioner (``--precision fp32`` stores the preconditioner in FP32).

One bench "step" = one complete heating step through the drop-in ``run_step`` (outlet recording
and balances included, as upstream), from the same seeded state.  value = steps x 22,400 s /
wall seconds of the timed region.

Also reported: the per-step mass/energy balance errors, the roofline of the dominant kernel
(the finite-difference matvec residual, timed with HIP events on the library's stream), and a
CPU baseline: the NumPy/SciPy oracle (same algorithms as upstream, 1 thread) on the same state
for a bounded wall-clock sample, with the GPU time over the same simulated window.
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
FP64_PEAK_TFLOPS = 78.6
WORKLOAD = {"grid.Nr": 10, "grid.Ntheta": 16, "grid.Nz": 50, "operation.P_vacuum": 1.5e4}
STEP_INDEX = 2          # heating in the default cycle
SEED = 0


def setup(precision):
    from dac_vtsa_amd.config import load_config
    from dac_vtsa_amd.core.grid import Grid
    from dac_vtsa_amd.properties import get_backend
    from dac_vtsa_amd.seeds import seeded_state
    cfg = load_config(override_dict=WORKLOAD)
    grid = Grid.from_config(cfg)
    props = get_backend(cfg)
    state = seeded_state(grid, "midcycle", SEED)
    return cfg, grid, props, state


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def one_step(cfg, grid, props, state, precision):
    from dac_vtsa_amd.steps import run_step
    return run_step(state, cfg, grid, props, cfg.cycle[STEP_INDEX], logger=_log, progress_s=20.0,
                    integrator_kw={"precision": precision})


def kernel_roofline(cfg, grid, props, state, n_launch=200):
    """Average duration of one residual evaluation (k_props + k_flux, the kernel pair every
    GMRES matvec, line-search point and error estimate launches), timed with HIP events
    recorded by libvtsa on the stream it launches on, and its algorithmic traffic."""
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


def cpu_baseline(cfg, state, budget_s):
    """Oracle integration of the same step on 1 host thread for a bounded wall-clock budget."""
    from oracle.model import OracleModel
    from oracle.runner import make
    from oracle.solver import OracleTRBDF2
    g, gas = make(cfg)
    step = cfg.cycle[STEP_INDEX]
    m = OracleModel(cfg, g, gas, step, state)
    y0 = m.pack(state)
    dur = float(step["duration"])
    integ = OracleTRBDF2(m, cfg, h_max=cfg.solver.h_max_fraction * dur)
    t0 = time.time()
    res = integ.integrate(y0, dur, np.array([dur]), stop_fn=lambda t, y: time.time() - t0, stop_value=budget_s)
    wall = time.time() - t0
    return res.t, wall, res.n_steps


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
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (independent cases would shard "
              "one group per GPU with no communication; no multi-GPU run is part of this build)", file=sys.stderr)
        return 2
    from dac_vtsa_amd import device as dv
    cfg, grid, props, state = setup(args.precision)
    dur = float(cfg.cycle[STEP_INDEX]["duration"])
    print(f"[bench] workload 10x16x50 heating {dur:.0f} s, precision {args.precision}", file=sys.stderr, flush=True)
    for i in range(args.warmup):
        t0 = time.time()
        one_step(cfg, grid, props, state, args.precision)
        print(f"[bench] warmup {i + 1}/{args.warmup}: {time.time() - t0:.2f} s", file=sys.stderr, flush=True)
    dv.synchronize()
    t0 = time.time()
    results = []
    for i in range(args.steps):
        _, sr = one_step(cfg, grid, props, state, args.precision)
        results.append(sr)
        print(f"[bench] step {i + 1}/{args.steps}: {sr.solver['steps']} accepted steps, "
              f"{sr.solver['rejected']} rejected, {sr.solver.get('gmres_iters', 0)} GMRES iterations",
              file=sys.stderr, flush=True)
    dv.synchronize()
    elapsed = time.time() - t0
    value = args.steps * dur / elapsed
    mb, eb = balance_errors(results[-1], props.species)
    avg_s, algo_bytes = kernel_roofline(cfg, grid, props, state)
    achieved = algo_bytes / avg_s / 1e9
    line = {
        "metric": "cycle sim-s/wall-s, 1 GPU: learned+FP32 vs plain FP64 at MB/EB err <=1e-6",
        "value": value, "unit": "sim-s/wall-s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f64 (fp32 preconditioner)",
        "data": "synthetic seeded mid-cycle state (seed 0)",
        "config": {"workload": "heating step 22400 s, grid 10x16x50 (48800 unknowns), P_vac 1.5e4 Pa",
                   "preconditioner": "zline", "precision": args.precision, "solver": "sdirk2 Newton-GMRES(30)"},
        "balance": {"max_mass_rel_err": mb, "max_energy_rel_err": eb},
        "solver": {k: results[-1].solver.get(k) for k in ("steps", "rejected", "newton_fail", "jacobians", "lu",
                                                          "gmres_iters", "nfev")},
        "roofline": {"bound": "hbm", "kernel": "residual (k_props + k_flux)", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "avg_launch_us": avg_s * 1e6, "algo_bytes_per_launch": algo_bytes},
    }
    if not args.no_cpu:
        from threadpoolctl import threadpool_limits
        with threadpool_limits(1):
            t_sim, wall, nsteps = cpu_baseline(cfg, state, args.cpu_sample_s)
        line["cpu_baseline"] = {"value": t_sim / wall, "unit": "sim-s/wall-s", "cores": 1, "kind": "port",
                                "sample": f"first {t_sim:.3g} s of the same heating step ({nsteps} accepted "
                                          f"steps) in {wall:.1f} s wall on 1 host thread"}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Cycle harness (upstream vtsa_cycle.py:72-138): steps chained with the bed kept in device memory
and each step's device model re-initialised in place (the CSV hand-off is optional), balances,
KPIs and the cyclic-steady-state test."""
from __future__ import annotations

import os
import time
from dataclasses import dataclass, field
from typing import Dict, List

from . import balances
from .core.grid import Grid
from .properties import get_backend
from .simio.checkpoint import load_state_csv, save_state_csv
from .steps import RUNNERS


@dataclass
class CycleRecord:
    cycle: int
    kpis: Dict[str, float]
    mb_rows: List[Dict]
    eb_rows: List[Dict]
    change: Dict[str, float]
    co2_balance: float
    balances_ok: bool
    wall_time: float


def run_cycle(state, cfg, grid, props, cycle_idx: int = 1, cycle_dir=None, log=None, T_jac_prev=None,
              handoff_csv: bool = False, device_handoff: bool = False, **kw):
    """The steps of cfg.cycle from ``state`` -> (end state, step results, last jacket set-point).
    Between steps the bed stays in device memory (a CSV hand-off, when asked for, goes through the
    file); ``device_handoff`` lets the first step start from the device copy ``state`` carries too
    (a state returned by a step runner and not modified since)."""
    steps_out, t_cycle = [], 0.0
    prev_csv = None
    if handoff_csv and cycle_dir:
        prev_csv = save_state_csv(state, grid, os.path.join(cycle_dir, "state_cycle_start.csv"))
    for step in cfg.cycle:
        if prev_csv:
            state = load_state_csv(prev_csv, grid, list(props.species), list(state.adsorbed))
        # the bed goes from step to step in device memory unless it went through a state file
        state, sr = RUNNERS[step["type"]](state, cfg, grid, props, step=step, cycle_idx=cycle_idx,
                                          t_cycle0=t_cycle, T_jacket_prev=T_jac_prev, logger=log,
                                          device_handoff=not prev_csv and (device_handoff or bool(steps_out)), **kw)
        if handoff_csv and cycle_dir:
            prev_csv = save_state_csv(state, grid, os.path.join(cycle_dir, f"state_after_{step['name']}.csv"))
        t_cycle += sr.duration
        T_jac_prev = sr.T_jacket if sr.jacket_on else T_jac_prev
        steps_out.append(sr)
    return state, steps_out, T_jac_prev


def run_to_css(cfg, state, log=None, max_cycles=None, **kw):
    """Cycles until the CSS test passes (vtsa_cycle.py:93-138); returns the records and state."""
    grid = Grid.from_config(cfg)
    props = get_backend(cfg)
    species = list(props.species)
    records = []
    T_jac_prev = None
    prev_end = state.copy()
    for n in range(1, (max_cycles or cfg.run.max_cycles) + 1):
        t0 = time.time()
        state, steps, T_jac_prev = run_cycle(state, cfg, grid, props, n, log=log, T_jac_prev=T_jac_prev,
                                             device_handoff=n > 1, **kw)
        mb, eb, ok = balances.check_cycle(steps, species)
        kp = balances.cycle_kpis(steps, cfg, species)
        ch = balances.state_change(prev_end, state, grid, cfg)
        rec = CycleRecord(n, kp, mb, eb, ch, balances.cycle_co2_balance(kp), ok, time.time() - t0)
        records.append(rec)
        if balances.is_css(ch, kp, cfg, n):
            break
        prev_end = state.copy()
    return records, state

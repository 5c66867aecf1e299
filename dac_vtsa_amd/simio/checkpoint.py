"""State CSV hand-off files (upstream simio/checkpoint.py:27-96), bit-exact round trip.

``<name>.csv`` holds one row per bed cell (indices, centre coordinates, T, P, c_k, y_k, q_j)
and ``<name>_wall.csv`` one row per wall cell; floats are written with 17 significant digits.
The column names and header line are the upstream ones, so files written by either code load
in the other.
"""
from __future__ import annotations

import os
from typing import Sequence

import numpy as np

from ..core.state import CycleState

FMT = "%.17g"


def save_state_csv(state: CycleState, grid, path: str) -> str:
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    r, th, z = grid.cell_centres()
    P, y = state.P, state.y
    cols = ["i_r", "i_theta", "i_z", "r_m", "theta_deg", "z_m", "T_K", "P_Pa"]
    cols += [f"c_{s}_mol_m3" for s in state.species] + [f"y_{s}" for s in state.species]
    cols += [f"q_{s}_mol_kg" for s in state.adsorbed]
    ii, jj, kk = np.meshgrid(np.arange(grid.Nr), np.arange(grid.Nth), np.arange(grid.Nz), indexing="ij")
    data = [ii.ravel(), jj.ravel(), kk.ravel(), r.ravel(), np.degrees(th).ravel(), z.ravel(), state.T.ravel(),
            P.ravel()]
    data += [state.c[k].ravel() for k in range(len(state.species))]
    data += [y[k].ravel() for k in range(len(state.species))]
    data += [state.q[k].ravel() for k in range(len(state.adsorbed))]
    arr = np.column_stack(data)
    header = (f"# t_s={float(state.t)!r} cycle={int(state.cycle)} step={state.step} "
              f"grid={grid.Nr}x{grid.Nth}x{grid.Nz}\n" + ",".join(cols))
    np.savetxt(path, arr, delimiter=",", header=header, comments="", fmt=["%d"] * 3 + [FMT] * (arr.shape[1] - 3))
    jw, kw = np.meshgrid(np.arange(grid.Nth), np.arange(grid.Nz), indexing="ij")
    warr = np.column_stack([jw.ravel(), kw.ravel(), np.degrees(grid.th_c)[jw.ravel()], grid.z_c[kw.ravel()],
                            state.Tw.ravel()])
    np.savetxt(path[:-4] + "_wall.csv", warr, delimiter=",", header="i_theta,i_z,theta_deg,z_m,T_wall_K",
               comments="", fmt=["%d", "%d", FMT, FMT, FMT])
    return path


def load_state_csv(path: str, grid, species: Sequence[str], adsorbed: Sequence[str]) -> CycleState:
    with open(path) as fh:
        first = fh.readline().strip()
        header = fh.readline().strip().split(",")
    meta = {}
    for tok in first.lstrip("# ").split():
        if "=" in tok:
            k, v = tok.split("=", 1)
            meta[k] = v
    arr = np.loadtxt(path, delimiter=",", skiprows=2, ndmin=2)
    col = {h: i for i, h in enumerate(header)}
    if arr.shape[0] != grid.Nr * grid.Nth * grid.Nz:
        raise ValueError(f"{path}: has {arr.shape[0]} cells but the grid has {grid.Nr * grid.Nth * grid.Nz} "
                         f"({grid.Nr}x{grid.Nth}x{grid.Nz}); use the same grid as the run that wrote it")
    i, j, k = (arr[:, col[c]].astype(int) for c in ("i_r", "i_theta", "i_z"))
    T = np.empty(grid.shape)
    T[i, j, k] = arr[:, col["T_K"]]
    c = np.empty((len(species),) + grid.shape)
    for n, s in enumerate(species):
        name = f"c_{s}_mol_m3"
        if name not in col:
            raise ValueError(f"{path}: missing column {name} (was the run dry/humid?)")
        c[n][i, j, k] = arr[:, col[name]]
    q = np.empty((len(adsorbed),) + grid.shape)
    for n, s in enumerate(adsorbed):
        q[n][i, j, k] = arr[:, col[f"q_{s}_mol_kg"]]
    wpath = path[:-4] + "_wall.csv"
    Tw = np.empty(grid.wall_shape)
    if os.path.exists(wpath):
        w = np.loadtxt(wpath, delimiter=",", skiprows=1, ndmin=2)
        Tw[w[:, 0].astype(int), w[:, 1].astype(int)] = w[:, 4]
    else:
        Tw[:] = T[-1]
    return CycleState(c=c, q=q, T=T, Tw=Tw, t=float(meta.get("t_s", 0.0)), cycle=int(meta.get("cycle", 1)),
                      step=meta.get("step", "loaded"), species=tuple(species), adsorbed=tuple(adsorbed))

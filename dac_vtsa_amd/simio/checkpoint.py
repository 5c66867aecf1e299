"""State files of the step-to-step hand-off (upstream format of simio/checkpoint.py:27-96).

Format (shared with the upstream simulator, so either code reads the other's files):

``<name>.csv``
    line 1: ``# t_s=<repr float> cycle=<int> step=<name> grid=<Nr>x<Nth>x<Nz>``
    line 2: column names
    then one row per bed cell in (i_r, i_theta, i_z) C order: the three indices as integers, then
    r_m, theta_deg, z_m, T_K, P_Pa, c_<gas>_mol_m3 and y_<gas> per gas, q_<sorbate>_mol_kg per
    sorbate, each float with 17 significant digits (so that reading it back is bit-exact);
``<name>_wall.csv``
    line 1: ``i_theta,i_z,theta_deg,z_m,T_wall_K``, then one row per wall cell.

The reader locates values by column name and cell indices, checks the cell count against the
grid, and without a wall file starts the wall at the outer ring's temperature.
"""
from __future__ import annotations

import os
from typing import Dict, List, Sequence, Tuple

import numpy as np

from ..core.state import CycleState

WALL_HEADER = ("i_theta", "i_z", "theta_deg", "z_m", "T_wall_K")


def _g17(v) -> str:
    return format(float(v), ".17g")


def _wall_path(path: str) -> str:
    root, ext = os.path.splitext(path)
    return root + "_wall" + ext


def _bed_columns(species: Sequence[str], adsorbed: Sequence[str]) -> List[str]:
    return (["i_r", "i_theta", "i_z", "r_m", "theta_deg", "z_m", "T_K", "P_Pa"]
            + [f"c_{s}_mol_m3" for s in species] + [f"y_{s}" for s in species]
            + [f"q_{s}_mol_kg" for s in adsorbed])


def _write_rows(path: str, lines: List[str]):
    with open(path, "w") as fh:
        fh.write("\n".join(lines))
        fh.write("\n")


def save_state_csv(state: CycleState, grid, path: str) -> str:
    """Write ``state`` (and its wall) to ``path`` in the hand-off format; returns path."""
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    species, adsorbed = list(state.species), list(state.adsorbed)
    r, th, z = grid.cell_centres()
    th_deg = np.degrees(th)
    P, y = state.P, state.y
    lines = [f"# t_s={float(state.t)!r} cycle={int(state.cycle)} step={state.step} "
             f"grid={grid.Nr}x{grid.Nth}x{grid.Nz}", ",".join(_bed_columns(species, adsorbed))]
    for i in range(grid.Nr):
        for j in range(grid.Nth):
            for k in range(grid.Nz):
                at = (i, j, k)
                vals = [r[at], th_deg[at], z[at], state.T[at], P[at]]
                vals += [state.c[n][at] for n in range(len(species))]
                vals += [y[n][at] for n in range(len(species))]
                vals += [state.q[n][at] for n in range(len(adsorbed))]
                lines.append(f"{i},{j},{k}," + ",".join(_g17(v) for v in vals))
    _write_rows(path, lines)
    th_w = np.degrees(grid.th_c)
    wall = [",".join(WALL_HEADER)]
    for j in range(grid.Nth):
        for k in range(grid.Nz):
            wall.append(f"{j},{k}," + ",".join(_g17(v) for v in (th_w[j], grid.z_c[k], state.Tw[j, k])))
    _write_rows(_wall_path(path), wall)
    return path


def _read_table(path: str, skip: int) -> Tuple[List[str], List[str], np.ndarray]:
    """(comment lines, column names, float rows) of a comma-separated file."""
    with open(path) as fh:
        head = [fh.readline().rstrip("\n") for _ in range(skip)]
        names = fh.readline().strip().split(",")
        rows = [ln.split(",") for ln in fh if ln.strip()]
    data = np.array(rows, dtype=float).reshape(len(rows), len(names))
    return head, names, data


def _metadata(comment: str) -> Dict[str, str]:
    out = {}
    for token in comment.lstrip("#").split():
        key, sep, value = token.partition("=")
        if sep:
            out[key] = value
    return out


def load_state_csv(path: str, grid, species: Sequence[str], adsorbed: Sequence[str]) -> CycleState:
    """Read a state file written in the hand-off format for ``grid``; raises ValueError when its
    cell count differs from the grid's or a gas column is missing."""
    (comment,), names, data = _read_table(path, 1)
    ncell = grid.Nr * grid.Nth * grid.Nz
    if data.shape[0] != ncell:
        raise ValueError(f"{path}: {data.shape[0]} cell rows, but grid {grid.Nr}x{grid.Nth}x{grid.Nz} has "
                         f"{ncell} cells")
    col = {name: data[:, n] for n, name in enumerate(names)}
    at = tuple(col[k].astype(int) for k in ("i_r", "i_theta", "i_z"))

    def field(name):
        if name not in col:
            raise ValueError(f"{path}: no column {name!r} (the file's gases/sorbates differ from the model's)")
        out = np.empty(grid.shape)
        out[at] = col[name]
        return out

    T = field("T_K")
    c = np.stack([field(f"c_{s}_mol_m3") for s in species])
    q = np.stack([field(f"q_{s}_mol_kg") for s in adsorbed]) if len(adsorbed) else np.empty((0,) + grid.shape)
    wall = _wall_path(path)
    if os.path.exists(wall):
        _, wnames, w = _read_table(wall, 0)
        wc = {name: w[:, n] for n, name in enumerate(wnames)}
        Tw = np.empty(grid.wall_shape)
        Tw[wc["i_theta"].astype(int), wc["i_z"].astype(int)] = wc["T_wall_K"]
    else:
        Tw = T[-1].copy()
    meta = _metadata(comment)
    return CycleState(c=c, q=q, T=T, Tw=Tw, t=float(meta.get("t_s", 0.0)), cycle=int(meta.get("cycle", 1)),
                      step=meta.get("step", "loaded"), species=tuple(species), adsorbed=tuple(adsorbed))

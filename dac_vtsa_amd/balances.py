"""Mass/energy balance rows, cycle KPIs and the cyclic-steady-state test (host arithmetic).

These are the metric's balance gate.  Definitions follow the upstream code:
postprocessing/reporting.py:32-140 (flows, KPIs, balance rows),
validation/conservation_checks.py:17-71 (tolerances, floors, cycle rows) and
validation/cyclic_steady_state.py:22-44 (state change, CO2 cycle balance, CSS).
"""
from __future__ import annotations

from typing import Dict, List, Tuple

import numpy as np

M_CO2 = 0.04401
M_H2O = 0.018015
VALVES = ("V1", "V2", "V3")
MASS_TOL = 1e-4
ENERGY_TOL = 1e-4
FLOOR = 1e-3


def _sum(acc, prefix, s):
    return sum(acc.get(f"{prefix}_{v}_{s}", 0.0) for v in VALVES)


def step_flows(sr, species):
    pp = sr.product_port
    out = {}
    for s in species:
        fed, allout = _sum(sr.acc, "n_in", s), _sum(sr.acc, "n_out", s)
        prod = (sr.acc.get(f"n_out_{pp}_{s}", 0.0)
                if (pp in VALVES and sr.valves.get(pp, "CLOSED") != "CLOSED") else 0.0)
        out[s] = {"fed": fed, "product": prod, "vented": allout - prod, "out": allout}
    return out


def is_heating_step(sr, cfg) -> bool:
    return sr.jacket_on and sr.T_jacket > cfg.ambient.T + 1.0


def cycle_kpis(steps: List, cfg, species) -> Dict[str, float]:
    tot = {s: {"fed": 0.0, "product": 0.0, "vented": 0.0} for s in species}
    for sr in steps:
        fl = step_flows(sr, species)
        for s in species:
            for k in ("fed", "product", "vented"):
                tot[s][k] += fl[s][k]
    t_cycle = sum(sr.duration for sr in steps)
    m_sorb, V_bed = cfg.derived.sorbent_mass, cfg.derived.bed_volume
    n_co2 = tot["CO2"]["product"]
    kg = n_co2 * M_CO2
    heat = [sr for sr in steps if is_heating_step(sr, cfg)]
    Q_th = sum(sr.acc["Q_jac_pos"] + sr.acc["Q_hx_pos"] for sr in heat)
    Q_cool = -sum(sr.acc["Q_jac_neg"] + sr.acc["Q_hx_neg"] for sr in steps)
    W_vac = sum(sr.acc.get(f"W_pump_{v}", 0.0) for sr in steps for v in VALVES)
    W_fan = sum(sr.acc.get(f"W_fan_{v}", 0.0) for sr in steps for v in VALVES)
    W_el = W_vac + W_fan
    T_H = max([sr.T_jacket for sr in heat] or [cfg.operation.T_heat])
    W_eq = W_el + cfg.energy.eta_turbine * (1.0 - cfg.energy.T_low_equiv / T_H) * Q_th
    prod_all = sum(tot[s]["product"] for s in species)
    nn = tot["N2"]["product"]
    safe = kg if n_co2 > 1e-6 * max(tot["CO2"]["fed"], 1e-30) else np.nan
    fed = tot["CO2"]["fed"]
    k = {"cycle_time_s": t_cycle, "CO2_fed_mol": fed, "CO2_product_mol": n_co2,
         "CO2_vented_mol": tot["CO2"]["vented"], "CO2_product_kg": kg,
         "recovery_pct": 100.0 * n_co2 / fed if fed > 0 else np.nan,
         "purity_dry_pct": 100.0 * n_co2 / (n_co2 + nn) if (n_co2 + nn) > 0 else np.nan,
         "purity_wet_pct": 100.0 * n_co2 / prod_all if prod_all > 0 else np.nan,
         "N2_product_mol": nn,
         "productivity_kg_m3_day": kg / V_bed / (t_cycle / 86400.0),
         "productivity_mol_kg_day": n_co2 / m_sorb / (t_cycle / 86400.0),
         "capture_rate_kg_day": kg / (t_cycle / 86400.0),
         "working_capacity_mol_kg": n_co2 / m_sorb,
         "Q_thermal_J": Q_th, "Q_cooling_removed_J": Q_cool, "W_vacuum_J": W_vac, "W_blower_J": W_fan,
         "specific_thermal_kJ_per_kg": Q_th / safe / 1e3,
         "specific_electric_kJ_per_kg": W_el / safe / 1e3,
         "specific_vacuum_kJ_per_kg": W_vac / safe / 1e3,
         "specific_blower_kJ_per_kg": W_fan / safe / 1e3,
         "specific_total_kJ_per_kg": (Q_th + W_el) / safe / 1e3,
         "work_equivalent_kJ_per_kg": W_eq / safe / 1e3,
         "T_heat_K": T_H, "sorbent_mass_kg": m_sorb, "bed_volume_m3": V_bed}
    if "H2O" in species:
        k["H2O_product_mol"] = tot["H2O"]["product"]
        k["H2O_per_CO2_kg_kg"] = tot["H2O"]["product"] * M_H2O / safe
        k["H2O_fed_mol"] = tot["H2O"]["fed"]
    return k


def mass_balance_rows(sr, species, floors=None) -> List[Dict]:
    rows, floors = [], floors or {}
    for s in species:
        a, b = sr.inv_start[f"tot_{s}"], sr.inv_end[f"tot_{s}"]
        fin, fout = _sum(sr.acc, "n_in", s), _sum(sr.acc, "n_out", s)
        res = (b - a) - (fin - fout)
        scale = max(abs(fin), abs(fout), abs(a), abs(b), floors.get(s, 1e-30))
        rows.append({"cycle": sr.cycle, "step": sr.name, "species": s, "inventory_start_mol": a,
                     "inventory_end_mol": b, "in_mol": fin, "out_mol": fout, "residual_mol": res,
                     "relative_error": res / scale})
    return rows


def energy_balance_row(sr, floor: float = 0.0) -> Dict:
    acc = sr.acc
    E0, E1 = sr.inv_start["E_total"], sr.inv_end["E_total"]
    H_in = sum(acc.get(f"H_in_{v}", 0.0) for v in VALVES)
    H_out = sum(acc.get(f"H_out_{v}", 0.0) for v in VALVES)
    Q_j = acc["Q_jac_pos"] + acc["Q_jac_neg"]
    Q_hx = acc["Q_hx_pos"] + acc["Q_hx_neg"]
    Q_amb = acc["Q_amb"]
    res = (E1 - E0) - (H_in - H_out + Q_j + Q_hx + Q_amb)
    scale = max(abs(E1 - E0), abs(H_in), abs(H_out), abs(Q_j), abs(Q_hx), abs(Q_amb), floor, 1e-30)
    return {"cycle": sr.cycle, "step": sr.name, "E_start_J": E0, "E_end_J": E1, "dE_J": E1 - E0,
            "H_in_J": H_in, "H_out_J": H_out, "Q_jacket_J": Q_j, "Q_internal_hx_J": Q_hx,
            "Q_ambient_J": Q_amb, "residual_J": res, "relative_error": res / scale}


def energy_breakdown_row(sr, cfg) -> Dict:
    acc = sr.acc
    dH = {"CO2": cfg.thermo.dH_ads_co2, "H2O": cfg.thermo.dH_ads_h2o}
    q_des = {s: dH[s] * (sr.ads_end[s] - sr.ads_start[s]) for s in sr.ads_start}
    dE_bed = sr.inv_end["E_bed"] - sr.inv_start["E_bed"]
    dE_wall = sr.inv_end["E_wall"] - sr.inv_start["E_wall"]
    H_net_out = sum(acc.get(f"H_out_{v}", 0.0) for v in VALVES) - sum(acc.get(f"H_in_{v}", 0.0) for v in VALVES)
    Q_in = acc["Q_jac_pos"] + acc["Q_jac_neg"] + acc["Q_hx_pos"] + acc["Q_hx_neg"]
    row = {"cycle": sr.cycle, "step": sr.name, "heat_input_J": Q_in,
           "desorption_CO2_J": q_des.get("CO2", 0.0), "desorption_H2O_J": q_des.get("H2O", 0.0),
           "bed_sensible_J": dE_bed - sum(q_des.values()), "wall_sensible_J": dE_wall,
           "gas_enthalpy_out_J": H_net_out, "ambient_loss_J": -acc["Q_amb"]}
    row["closure_J"] = Q_in - (row["desorption_CO2_J"] + row["desorption_H2O_J"] + row["bed_sensible_J"]
                               + row["wall_sensible_J"] + row["gas_enthalpy_out_J"] + row["ambient_loss_J"])
    return row


def check_step(sr, species, floors=None, e_floor: float = 0.0) -> Tuple[List[Dict], Dict, bool]:
    mb = mass_balance_rows(sr, species, floors)
    eb = energy_balance_row(sr, e_floor)
    ok = all(abs(r["relative_error"]) < MASS_TOL for r in mb) and abs(eb["relative_error"]) < ENERGY_TOL
    return mb, eb, ok


def check_cycle(step_results, species):
    mb_rows, eb_rows, all_ok = [], [], True
    floors = {}
    for s in species:
        thr = sum(sum(sr.acc.get(f"n_in_{v}_{s}", 0.0) + sr.acc.get(f"n_out_{v}_{s}", 0.0) for v in VALVES)
                  for sr in step_results)
        floors[s] = FLOOR * max(thr, 1e-30)
    e_thr = 0.0
    for sr in step_results:
        a = sr.acc
        e_thr += (sum(abs(a.get(f"H_in_{v}", 0.0)) + abs(a.get(f"H_out_{v}", 0.0)) for v in VALVES)
                  + abs(a["Q_jac_pos"]) + abs(a["Q_jac_neg"]) + abs(a["Q_hx_pos"]) + abs(a["Q_hx_neg"])
                  + abs(a["Q_amb"]))
    for sr in step_results:
        mb, eb, ok = check_step(sr, species, floors, FLOOR * e_thr)
        mb_rows += mb
        eb_rows.append(eb)
        all_ok &= ok
    for s in species:
        a, b = step_results[0].inv_start[f"tot_{s}"], step_results[-1].inv_end[f"tot_{s}"]
        fin = sum(r["in_mol"] for r in mb_rows if r["species"] == s)
        fout = sum(r["out_mol"] for r in mb_rows if r["species"] == s)
        res = (b - a) - (fin - fout)
        mb_rows.append({"cycle": step_results[0].cycle, "step": "CYCLE", "species": s,
                        "inventory_start_mol": a, "inventory_end_mol": b, "in_mol": fin, "out_mol": fout,
                        "residual_mol": res, "relative_error": res / max(abs(fin), abs(fout), abs(a), 1e-30)})
    E0, E1 = step_results[0].inv_start["E_total"], step_results[-1].inv_end["E_total"]
    keys = ("H_in_J", "H_out_J", "Q_jacket_J", "Q_internal_hx_J", "Q_ambient_J")
    tot = {k: sum(r[k] for r in eb_rows) for k in keys}
    res = (E1 - E0) - (tot["H_in_J"] - tot["H_out_J"] + tot["Q_jacket_J"] + tot["Q_internal_hx_J"]
                       + tot["Q_ambient_J"])
    scale = max(abs(E1 - E0), *[abs(v) for v in tot.values()], 1e-30)
    eb_rows.append({"cycle": step_results[0].cycle, "step": "CYCLE", "E_start_J": E0, "E_end_J": E1,
                    "dE_J": E1 - E0, **tot, "residual_J": res, "relative_error": res / scale})
    return mb_rows, eb_rows, all_ok


def _wrms(a, w):
    return float(np.sqrt(np.sum(w * a * a) / np.sum(w)))


def state_change(prev, now, grid, cfg) -> Dict[str, float]:
    w = np.broadcast_to(grid.V, grid.shape)
    out = {}
    for n, s in enumerate(now.adsorbed):
        out[f"q_{s}"] = _wrms(now.q[n] - prev.q[n], w) / max(_wrms(now.q[n], w), 1e-6)
    out["T"] = _wrms(now.T - prev.T, w) / max(cfg.operation.T_heat - cfg.feed.T, 1.0)
    Csc = max(_wrms(now.c.sum(axis=0), w), 1e-9)
    for k, s in enumerate(now.species):
        out[f"c_{s}"] = _wrms(now.c[k] - prev.c[k], w) / Csc
    out["max"] = max(out.values())
    return out


def cycle_co2_balance(kpis: Dict[str, float]) -> float:
    fed = kpis["CO2_fed_mol"]
    return abs(fed - kpis["CO2_product_mol"] - kpis["CO2_vented_mol"]) / fed if fed > 0 else np.inf


def is_css(change, kpis, cfg, cycle: int) -> bool:
    return (cycle >= cfg.run.min_cycles and change["max"] < cfg.run.css_tolerance
            and cycle_co2_balance(kpis) < cfg.run.css_mb_tolerance)


def worst_step_balance(step_results, species) -> Tuple[float, float]:
    """Largest |relative error| over the per-step mass rows and energy rows of a step list."""
    mb, eb, _ = check_cycle(step_results, species)
    return (max(abs(r["relative_error"]) for r in mb if r["step"] != "CYCLE"),
            max(abs(r["relative_error"]) for r in eb if r["step"] != "CYCLE"))

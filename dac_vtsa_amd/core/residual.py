"""Semi-discrete TVSA model dy/dt = f(t, y) on the GPU (drop-in for upstream core/residual.py).

``TVSAModel`` keeps the upstream constructor, layout, attributes and methods
(core/residual.py:40-482).  The host side only evaluates per-step constants (boundary
conditions, feed profile, jacket coverage, tolerances) and fills the device descriptor; every
evaluation of f, the stage residual, the diagnostics, the inventories and the Jacobian runs in
the HIP kernels of libvtsa.  The NumPy-facing methods copy to and from the device; the
integrator uses the ``*_dev`` variants, which keep vectors resident in HBM.
"""
from __future__ import annotations

import ctypes as C
import threading
from typing import Dict, List, Optional

import numpy as np

from .. import _abi
from .. import device as dv
from ..properties import R_GAS, conductivity_terms, fuller_pair, viscosity_terms, wilke_tables
from ..valves import KIND_CODE, end_conditions, feed_molar_flux, feed_profile, humid_air_composition

LIMITERS = {"van_albada": 0, "van_albada_tvd": 1, "van_leer": 2, "minmod": 3, "upwind": 4}
DISPERSION = {"edwards_richardson": 0, "young2021": 1, "wakao_kagei": 2}
COADS = {"none": 0, "wadst": 1, "mechanistic": 2, "stampi_bombelli": 3}
BED = {"zehner_schlunder": 0, "parallel": 1, "constant": 2}
WALLHTC = {"martin_nilles": 0, "constant": 1}
VALVE_INDEX = {"V1": 0, "V2": 1, "V3": 2}

# Sampled kernel timing (no upstream counterpart): when KERNEL_PROFILE["every"] > 0, every model
# times each every-th launch of its dominant kernel (the line Gauss-Seidel half-sweep) with HIP
# events on the launch stream and adds launches / milliseconds / algorithmic bytes here on close.
KERNEL_PROFILE = {"every": 0, "launches": 0, "ms": 0.0, "bytes": 0.0}


def reset_kernel_profile(every: int):
    KERNEL_PROFILE.update(every=int(every), launches=0, ms=0.0, bytes=0.0)


# Device handles kept for the next model of the same grid (no upstream counterpart: upstream
# builds a new model for every step, vtsa_cycle.py:80-86).  A closed model leaves its handle here,
# keyed by grid and stream; the next model of that key re-initialises it with vtsa_reset instead of
# allocating ~1 GB (20x32x100) anew, and no vtsa_destroy (a device-wide synchronisation) runs
# between the steps of a cycle.  At most HANDLE_POOL_MAX handles are kept, the oldest destroyed.
HANDLE_POOL_MAX = 4
_handle_pool: List = []            # [(key, handle)], most recently released last
_pool_lock = threading.RLock()     # reentrant: TVSAModel.__del__ releases into the pool


def _pool_take(key):
    with _pool_lock:
        for i in range(len(_handle_pool) - 1, -1, -1):
            if _handle_pool[i][0] == key:
                return _handle_pool.pop(i)[1]
    return None


def _pool_put(key, h):
    with _pool_lock:
        _handle_pool.append((key, h))
        evict = _handle_pool[:-HANDLE_POOL_MAX] if len(_handle_pool) > HANDLE_POOL_MAX else []
        del _handle_pool[:len(evict)]
    for _, old in evict:
        _abi.lib().vtsa_destroy(old)


def release_handles(stream=None):
    """Destroy the pooled handles (of the models built on ``stream`` when given)."""
    sk = None if stream is None else (stream.value if isinstance(stream, C.c_void_p) else stream)
    with _pool_lock:
        drop = [e for e in _handle_pool if stream is None or e[0][-1] == sk]
        _handle_pool[:] = [e for e in _handle_pool if not any(e is d for d in drop)]
    for _, h in drop:
        _abi.lib().vtsa_destroy(h)


class Thermo:
    """Conservative energy variable e <-> T (upstream physics/energy.py:31-70), host side."""

    def __init__(self, cfg, props):
        self.eps_t = cfg.sorbent.total_voidage
        self.rho_b = cfg.derived.bulk_density
        self.cp_s = cfg.sorbent.cp_solid
        self.cpg = np.asarray(props.cp, dtype=float)
        th = cfg.thermo
        self.cpa = np.array([th.cp_ads_co2, th.cp_ads_h2o])
        self.dH = np.array([th.dH_ads_co2, th.dH_ads_h2o])
        self.T_ref = th.T_ref

    @staticmethod
    def _wsum(w, a):
        out = w[0] * a[0]
        for k in range(1, len(w)):
            out = out + w[k] * a[k]
        return out

    def _terms(self, c, q):
        Cg = c[0] + c[1] + c[2]
        A = self.eps_t * self._wsum(self.cpg, c) + self.rho_b * (self.cp_s + self._wsum(self.cpa, q))
        B = self.eps_t * R_GAS * Cg
        Hq = self.rho_b * self._wsum(self.dH, q)
        return A, B, Hq

    def temperature(self, e, c, q):
        A, B, Hq = self._terms(c, q)
        return (e - Hq + A * self.T_ref) / (A - B)

    def energy(self, T, c, q):
        A, B, Hq = self._terms(c, q)
        return A * (T - self.T_ref) - B * T + Hq

    def heat_capacity(self, c, q):
        A, B, _ = self._terms(c, q)
        return A - B


class WallInfo:
    def __init__(self, cfg, grid):
        w = cfg.wall
        self.enabled = bool(w.enabled)
        self.rho, self.cp, self.lam = w.density, w.cp, w.conductivity
        self.C = self.rho * self.cp * grid.V_wall
        self.T_ref = cfg.thermo.T_ref

    def energy(self, Tw):
        return float(np.sum(self.C * (Tw - self.T_ref)))


class TVSAModel:
    """One cycle step of the bed model on the GPU (upstream core/residual.py:40-123)."""

    def __init__(self, cfg, grid, props, step: Dict, state0, T_jacket_prev: Optional[float] = None):
        if not cfg.components.water:
            # upstream builds the dry model (core/residual.py:46-50) but its first residual raises
            # TypeError in the equilibrium (physics/isotherms.py:118, np.maximum(None, 1e-12)), so
            # there is no dry-air behaviour to reproduce; refused here before any device work
            raise NotImplementedError("components.water = false (dry air): the upstream residual of this case "
                                      "raises TypeError at physics/isotherms.py:118; the device model covers "
                                      "humid air (components.water = true)")
        self.cfg, self.grid, self.props, self.step = cfg, grid, props, step
        g = grid
        self.species = list(props.species)
        self.ng = len(self.species)
        self.water = True
        self.adsorbed = ["CO2", "H2O"]
        self.na = 2
        self.i_co2, self.i_h2o, self.i_inert = 0, 1, 2
        self.ads_to_gas = [0, 1]
        self.thermo = Thermo(cfg, props)
        self.wall = WallInfo(cfg, g)
        self.eps_b = cfg.sorbent.bed_voidage
        self.eps_t = cfg.sorbent.total_voidage
        self.rho_b = cfg.derived.bulk_density
        self.cpg = np.asarray(props.cp)
        self.T_ref = cfg.thermo.T_ref
        self.limiter = cfg.solver.limiter
        self.lumped = cfg.transport.heat_transfer_mode == "lumped"
        self.conds = end_conditions(cfg, step, self.species, props)
        self.tau_valve = cfg.operation.valve_time_constant
        P0 = state0.P
        self.P_start = {"top": float(np.mean(P0[:, :, 0])), "bottom": float(np.mean(P0[:, :, -1]))}
        prof = feed_profile(cfg, g)
        self.feed_flux = feed_molar_flux(cfg) * prof * g.Az
        has_feed = any(c.kind == "feed" for c in self.conds.values())
        self.u_nom = cfg.feed.velocity if has_feed else 0.0
        self.N_nom = (feed_molar_flux(cfg) * prof)[..., None] if has_feed else np.zeros(g.Az.shape + (1,))
        jk = step["jacket"]
        self.jacket_on = jk["mode"] == "temperature"
        self.T_jac_set = jk.get("T", cfg.ambient.T)
        self.T_jac_prev = self.T_jac_set if T_jacket_prev is None else T_jacket_prev
        self.jac_tau = cfg.jacket.ramp_time
        self.h_jac, self.h_amb = cfg.jacket.h_jacket, cfg.jacket.h_ambient
        self.T_amb, self.P_amb = cfg.ambient.T, cfg.ambient.P
        self.cov = g.coverage_fraction(cfg.jacket.theta_coverage_deg, cfg.jacket.z_coverage)
        self.hx_on = bool(cfg.internal_hx.enabled) and self.jacket_on
        en = cfg.energy
        self.gam, self.eta_vac, self.vac_model = en.gamma_gas, en.eta_vacuum, en.vacuum_efficiency_model
        self.eta_bl, self.pump_water = en.eta_blower, bool(en.pump_handles_water)
        n = g.ncell
        self.n_c = self.ng * n
        self.n_q = self.na * n
        self.off_q = self.n_c
        self.off_e = self.n_c + self.n_q
        self.off_w = self.off_e + n
        self.n_w = g.Nth * g.Nz if self.wall.enabled else 0
        self.n_core = self.off_w + self.n_w
        self.acc_names = self._acc_names()
        self.n_acc = len(self.acc_names)
        self.acc_index = {nm: i for i, nm in enumerate(self.acc_names)}
        s = cfg.solver
        self.atol_var = np.array([s.atol_c_co2, s.atol_c_h2o, s.atol_c_inert, s.atol_q, s.atol_q,
                                  self.rho_b * cfg.sorbent.cp_solid * s.atol_T, s.atol_T])
        self.atol = np.concatenate([np.full(n, v) for v in self.atol_var[:6]] +
                                   ([np.full(self.n_w, self.atol_var[6])] if self.n_w else []))
        self.typ_var = np.array([1e-3, 1e-2, 1.0, 1e-3, 1e-3, self.rho_b * cfg.sorbent.cp_solid, 1.0])
        self.nfev = 0
        self.desc = self._descriptor()
        self._handle = None
        self._create()

    # ---------------------------------------------------------------- device model
    def _descriptor(self) -> _abi.ModelDesc:
        cfg, g, pr = self.cfg, self.grid, self.props
        d = _abi.ModelDesc()
        d.Nr, d.Nth, d.Nz = g.Nr, g.Nth, g.Nz
        d.wall_on, d.lumped, d.hx_on, d.jacket_on = int(self.wall.enabled), int(self.lumped), int(self.hx_on), \
            int(self.jacket_on)
        d.limiter = LIMITERS[self.limiter]
        tr = cfg.transport
        d.disp_model = DISPERSION[tr.dispersion_model]
        d.coads_model = COADS[cfg.isotherm.coadsorption.model]
        d.water_basis_eq = int(cfg.isotherm.coadsorption.water_loading_basis == "equilibrium")
        d.bed_model = BED[tr.bed_conductivity_model]
        d.wall_model = WALLHTC[tr.wall_htc_model]
        kn = cfg.kinetics
        d.arrhenius = int(not (kn.Ea_co2 == 0.0 and kn.Ea_h2o == 0.0))
        d.mn_dynamic = int(abs(self.u_nom) > 0.0)
        d.vac_krish = int(self.vac_model == "krishnamurthy")
        d.pump_water = int(self.pump_water)
        for e, end in enumerate(("top", "bottom")):
            c = self.conds[end]
            d.end_kind[e] = KIND_CODE[c.kind]
            d.end_check[e] = int(c.check_valve)
            d.end_valve[e] = VALVE_INDEX[c.valve] if c.valve is not None else -1
            d.end_is_vacuum[e] = int(c.is_vacuum)
            d.end_is_feed[e] = int(c.is_feed)
            d.end_P_target[e] = c.P_target
            d.end_T_ext[e] = c.T_ext
            d.end_speed[e] = c.speed
            d.end_P_start[e] = self.P_start[end]
            y = c.y_ext if c.y_ext is not None else np.zeros(3)
            for s in range(3):
                d.end_y_ext[e][s] = float(y[s])
        d.dz, d.dth = g.dz, g.dth
        d.A_wall_in, d.A_wall_out, d.A_wall_axial = g.A_wall_in, g.A_wall_out, g.A_wall_axial
        d.R_wall_mid, d.wall_thickness, d.half_dr_outer = g.R_wall_mid, g.wall_thickness, g.half_dr_outer
        d.C_wall = self.wall.C
        sb = cfg.sorbent
        d.R, d.eps_b, d.eps_t, d.rho_b, d.cp_s, d.T_ref = R_GAS, sb.bed_voidage, sb.total_voidage, self.rho_b, \
            sb.cp_solid, cfg.thermo.T_ref
        # only the attributes every upstream property backend has (species, cp, M) are read from
        # props; the pure-component correlations and mixing-rule tables are this package's own
        species = list(pr.species)
        cp, M = np.asarray(pr.cp, dtype=float), np.asarray(pr.M, dtype=float)
        visc, cond = viscosity_terms(species), conductivity_terms(species)
        wden, wmq = wilke_tables(M)
        for s in range(3):
            d.cpg[s], d.M[s] = cp[s], M[s]
            f1, a1, b1, f2, a2, b2, two = visc[s]
            d.mu_f1[s], d.mu_a1[s], d.mu_b1[s], d.mu_f2[s], d.mu_a2[s], d.mu_b2[s], d.mu_two[s] = \
                f1, a1, b1, f2, a2, b2, two
            c1, n1, c2, n2 = cond[s]
            d.k_c1[s], d.k_n1[s], d.k_c2[s], d.k_n2[s] = c1, n1, c2, n2
            for t in range(3):
                d.wilke_mq[s][t] = wmq[s, t]
                d.wilke_den[s][t] = wden[s, t]
        th = cfg.thermo
        d.cpa[0], d.cpa[1] = th.cp_ads_co2, th.cp_ads_h2o
        d.dHa[0], d.dHa[1] = th.dH_ads_co2, th.dH_ads_h2o
        d.fuller_sqrtM_air, d.fuller_s_air = fuller_pair("CO2", "AIR")
        d.fuller_sqrtM_h2o, d.fuller_s_h2o = fuller_pair("CO2", "H2O")
        eps, dp = sb.bed_voidage, sb.particle_diameter
        d.lam_const, d.lam_s = tr.bed_conductivity_constant, sb.particle_conductivity
        d.zs_B = tr.zs_shape_factor * ((1.0 - eps) / eps) ** (10.0 / 9.0)
        d.zs_s = float(np.sqrt(1.0 - eps))
        iso = cfg.isotherm
        c2 = iso.co2
        d.co2_q_inf0, d.co2_chi, d.co2_b0, d.co2_dH0 = c2.q_inf0, c2.chi, c2.b0, c2.dH0
        d.co2_tau0, d.co2_alpha, d.co2_T0, d.co2_p_reg = c2.tau0, c2.alpha, c2.T0, c2.p_regularization
        h2 = iso.h2o
        d.h2o_q_m, d.h2o_C, d.h2o_D, d.h2o_F, d.h2o_G = h2.q_m, h2.C, h2.D, h2.F, h2.G
        d.h2o_E10_slope, d.h2o_E10_intercept, d.h2o_rh_cap = h2.E10_slope, h2.E10_intercept, h2.rh_cap
        w = iso.coadsorption.wadst
        d.wadst_b0, d.wadst_q_inf0, d.wadst_tau0, d.wadst_chi = w.b0_wet, w.q_inf0_wet, w.tau0_wet, w.chi_wet
        d.wadst_alpha, d.wadst_dH, d.wadst_A = w.alpha_wet, w.dH_wet, w.A
        mm = iso.coadsorption.mechanistic
        d.mech_f_blocked_max, d.mech_k, d.mech_phi_dry = mm.f_blocked_max, mm.k, mm.phi_dry
        d.mech_A, d.mech_dH_wet, d.mech_n = mm.A, mm.dH_wet, mm.n
        d.sb_gamma, d.sb_beta = iso.coadsorption.stampi_bombelli.gamma, iso.coadsorption.stampi_bombelli.beta
        d.k_ldf[0], d.k_ldf[1], d.Ea[0], d.Ea[1], d.k_T_ref = kn.k_ldf_co2, kn.k_ldf_h2o, kn.Ea_co2, kn.Ea_h2o, \
            kn.T_ref
        d.ka = tr.ergun_A * (1.0 - eps) ** 2 / (eps ** 3 * dp ** 2)
        d.kb = tr.ergun_B * (1.0 - eps) / (eps ** 3 * dp)
        d.dp, d.er_g1, d.er_g2, d.er_beta, d.Pe_T = dp, tr.er_gamma1, tr.er_gamma2, tr.er_beta, tr.Pe_mass_radial_inf
        d.Pe_r, d.Pe_z = tr.Pe_heat_radial_inf, tr.Pe_heat_axial_inf
        d.N_ratio = 2.0 * cfg.geometry.bed_radius / dp
        d.alpha_const, d.U_lumped = tr.wall_htc_constant, tr.lumped_U
        d.u_nom, d.tau_valve = self.u_nom, self.tau_valve
        d.h_jac, d.h_amb, d.T_amb, d.P_amb = self.h_jac, self.h_amb, self.T_amb, self.P_amb
        d.T_jac_set, d.T_jac_prev, d.jac_tau = self.T_jac_set, self.T_jac_prev, self.jac_tau
        d.hA_v, d.lam_w = cfg.internal_hx.hA_per_volume, cfg.wall.conductivity
        d.gam, d.eta_vac, d.eta_bl = self.gam, self.eta_vac, self.eta_bl
        for v in range(7):
            d.typ[v] = self.typ_var[v]
            d.atol[v] = self.atol_var[v]
        return d

    def _device_arrays(self):
        """Host arrays of vtsa_create, broadcast to the sizes the kernels index."""
        g = self.grid
        Nr, Nth, Nz = g.Nr, g.Nth, g.Nz
        specs = [(g.V, (Nr,)), (g.Az, (Nr,)), (g.Ath, (Nr,)), (g.arc, (Nr,)),
                 (g.Ar if Nr > 1 else np.zeros(1), (max(Nr - 1, 1),)),
                 (g.drc if Nr > 1 else np.zeros(1), (max(Nr - 1, 1),)),
                 (np.broadcast_to(self.feed_flux, (Nr, Nth)), (Nr, Nth)),
                 (np.broadcast_to(self.N_nom.reshape(self.N_nom.shape[:2]), (Nr, Nth)), (Nr, Nth)),
                 (np.broadcast_to(self.cov, (Nth, Nz)), (Nth, Nz))]
        out = []
        for a, shp in specs:
            a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(shp)).ravel()
            assert a.size == int(np.prod(shp))
            out.append(a)
        return out

    def _create(self):
        arrs = self._device_arrays()
        self._host_arrays = arrs
        ptrs = [a.ctypes.data_as(C.c_void_p) for a in arrs]
        g = self.grid
        self._pool_key = (g.Nr, g.Nth, g.Nz, int(self.desc.wall_on), dv._stream_key())
        h = _pool_take(self._pool_key)
        if h is not None and _abi.lib().vtsa_reset(h, C.byref(self.desc), *ptrs, dv.stream()) != 0:
            _abi.lib().vtsa_destroy(h)
            h = None
        if h is None:
            h = C.c_void_p()
            _abi.call("vtsa_create", C.byref(self.desc), *ptrs, C.byref(h))
        self._handle = h
        if KERNEL_PROFILE["every"] > 0:
            _abi.call("vtsa_profile", h, int(KERNEL_PROFILE["every"]))
        self._y = dv.empty(self.n_core)
        self._f = dv.empty(self.n_core)
        self._acc = dv.empty(self.n_acc)

    def close(self):
        if self._handle is not None:
            if KERNEL_PROFILE["every"] > 0:
                n, ms, b = C.c_int64(), C.c_double(), C.c_double()
                _abi.call("vtsa_profile_read", self._handle, C.byref(n), C.byref(ms), C.byref(b))
                KERNEL_PROFILE["launches"] += n.value
                KERNEL_PROFILE["ms"] += ms.value
                KERNEL_PROFILE["bytes"] += b.value
            _pool_put(self._pool_key, self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._handle

    def _acc_names(self) -> List[str]:
        names = []
        for v in ("V1", "V2", "V3"):
            names += [f"n_in_{v}_{s}" for s in self.species]
            names += [f"n_out_{v}_{s}" for s in self.species]
            names += [f"H_in_{v}", f"H_out_{v}", f"W_pump_{v}", f"W_fan_{v}"]
        names += ["Q_jac_pos", "Q_jac_neg", "Q_amb", "Q_wb", "Q_hx_pos", "Q_hx_neg"]
        return names

    # ---------------------------------------------------------------- layout (core/residual.py:146-196)
    def conservation_matrix(self):
        import scipy.sparse as sp
        g = self.grid
        n = g.ncell
        Vc = np.broadcast_to(g.V, g.shape).ravel()
        rows, cols, vals = [], [], []
        for k, s in enumerate(self.species):
            rows.append(np.full(n, k)); cols.append(np.arange(k * n, (k + 1) * n)); vals.append(self.eps_t * Vc)
            if s in self.adsorbed:
                j = self.adsorbed.index(s)
                rows.append(np.full(n, k)); cols.append(self.off_q + np.arange(j * n, (j + 1) * n))
                vals.append(self.rho_b * Vc)
        rows.append(np.full(n, self.ng)); cols.append(self.off_e + np.arange(n)); vals.append(Vc)
        if self.n_w:
            rows.append(np.full(self.n_w, self.ng)); cols.append(self.off_w + np.arange(self.n_w))
            vals.append(np.ravel(self.wall.C) * np.ones(self.n_w))
        return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                             shape=(self.ng + 1, self.n_core))

    def pack(self, st) -> np.ndarray:
        e = self.thermo.energy(st.T, st.c, st.q)
        parts = [st.c.ravel(), st.q.ravel(), e.ravel()]
        if self.n_w:
            parts.append(st.Tw.ravel())
        return np.concatenate(parts)

    def unpack(self, y):
        g = self.grid
        c = y[:self.n_c].reshape((self.ng,) + g.shape)
        q = y[self.off_q:self.off_e].reshape((self.na,) + g.shape)
        e = y[self.off_e:self.off_w].reshape(g.shape)
        Tw = y[self.off_w:self.n_core].reshape(g.wall_shape) if self.n_w else None
        return c, q, e, Tw

    def to_state(self, y, template, t_abs: float, step_name: str):
        c, q, e, Tw = self.unpack(y)
        T = self.thermo.temperature(e, c, q)
        st = template.copy()
        st.c, st.q, st.T = c.copy(), q.copy(), T.copy()
        st.Tw = Tw.copy() if Tw is not None else T[-1].copy()
        st.t = t_abs
        st.step = step_name
        return st

    def T_jacket(self, t):
        if self.jac_tau <= 0:
            return self.T_jac_set
        return self.T_jac_set + (self.T_jac_prev - self.T_jac_set) * np.exp(-t / self.jac_tau)

    def eta_vacuum(self, P):
        if self.vac_model == "krishnamurthy":
            pb = P / 1e5
            return np.maximum(0.8 * 19.55 * pb / (1.0 + 19.55 * pb), 1e-3)
        return self.eta_vac

    # ---------------------------------------------------------------- device evaluations
    def rhs_dev(self, t, y_dev, f_dev, acc_dev=None):
        """f = f(t, y) with y, f (and acc, 36) resident on the device."""
        self.nfev += 1
        _abi.call("vtsa_rhs", self._handle, float(t), dv.ptr(y_dev), dv.ptr(f_dev), dv.ptr(acc_dev), dv.stream())

    def rhs(self, t: float, y, want_acc: bool = True, diag: bool = False):
        """Return (f, acc) -- or the diagnostics dict if diag=True (core/residual.py:211-417)."""
        if diag:
            return self.diag(t, y)
        self._y.upload(y)
        self.rhs_dev(t, self._y, self._f, self._acc if want_acc else None)
        f = dv.to_host(self._f)
        return f, (dv.to_host(self._acc).copy() if want_acc else None)

    def stage_residual(self, t, y, gamma_h, rhs_known):
        """R = y - gamma_h f(t, y) - rhs_known (core/residual.py:420-423)."""
        self.nfev += 1
        yd, rk, R = dv.to_dev(y), dv.to_dev(rhs_known), dv.empty(self.n_core)
        _abi.call("vtsa_stage_residual", self._handle, float(t), dv.ptr(yd), float(gamma_h), dv.ptr(rk), dv.ptr(R),
                  dv.stream())
        return dv.to_host(R)

    def diag(self, t, y):
        """Diagnostics dict of rhs(diag=True)."""
        g = self.grid
        n, Nr, Nth, Nz = g.ncell, g.Nr, g.Nth, g.Nz
        nf, nw2 = Nr * Nth, Nth * Nz
        yd = y if not isinstance(y, np.ndarray) else dv.to_dev(y)
        cell, uz = dv.zeros(8 * n), dv.zeros(max(1, Nr * Nth * (Nz - 1)))
        ur, uth = dv.zeros(max(1, (Nr - 1) * Nth * Nz)), dv.zeros(n)
        bnd, wall = dv.zeros(2 * 14 * nf), dv.zeros(4 * nw2)
        self.nfev += 1
        _abi.call("vtsa_diag", self._handle, float(t), dv.ptr(yd), dv.ptr(cell), dv.ptr(uz), dv.ptr(ur), dv.ptr(uth),
                  dv.ptr(bnd), dv.ptr(wall), dv.stream())
        cell, uz, ur, uth = dv.to_host(cell), dv.to_host(uz), dv.to_host(ur), dv.to_host(uth)
        bnd, wall = dv.to_host(bnd).reshape(2, 14, Nr, Nth), dv.to_host(wall).reshape(4, Nth, Nz)
        shp = g.shape
        out = {"T": cell[:n].reshape(shp), "P": cell[n:2 * n].reshape(shp),
               "y": cell[2 * n:5 * n].reshape((3,) + shp), "qstar": cell[5 * n:7 * n].reshape((2,) + shp),
               "uz": uz[:Nr * Nth * (Nz - 1)].reshape(Nr, Nth, Nz - 1),
               "ur": ur.reshape(Nr - 1, Nth, Nz) if Nr > 1 else None,
               "uth": uth.reshape(shp) if Nth > 1 else None}
        b = {}
        for e, end in enumerate(("top", "bottom")):
            bb = bnd[e]
            d = {"F_in": bb[0:3].copy(), "H_in": bb[3].copy(), "u_out": bb[4].copy(), "P_face": bb[5].copy()}
            if self.conds[end].kind == "valve":
                d.update(F_inpart=bb[6:9].copy(), F_outpart=bb[9:12].copy(), H_inpart=bb[12].copy(),
                         H_outpart=bb[13].copy())
            b[end] = d
        out["bnd"] = b
        out["Q_jac"], out["Q_wb"], out["Q_amb"] = wall[0], wall[1], wall[2]
        out["Q_hx"] = cell[7 * n:8 * n].reshape(shp) if self.hx_on else None
        out["R_bed"] = (1.0 / self.cfg.transport.lumped_U) if self.lumped else wall[3]
        return out

    def equilibrium(self, p_co2, p_h2o, T, q_h2o_state):
        """(q*_CO2, q*_H2O) at arbitrary points, evaluated by the device isotherm code."""
        shp = np.shape(p_co2)
        arrs = [dv.to_dev(np.broadcast_to(a, shp).ravel()) for a in (p_co2, p_h2o, T, q_h2o_state)]
        m = int(np.prod(shp)) if shp else 1
        qc, qw = dv.empty(m), dv.empty(m)
        _abi.call("vtsa_equilibrium", self._handle, m, *[dv.ptr(a) for a in arrs], dv.ptr(qc), dv.ptr(qw),
                  dv.stream())
        return dv.to_host(qc).reshape(shp), dv.to_host(qw).reshape(shp)

    def inventories_dev(self, y_dev) -> Dict[str, float]:
        out = (C.c_double * 7)()
        _abi.call("vtsa_inventories", self._handle, dv.ptr(y_dev), out, dv.stream())
        inv = {}
        for k, s in enumerate(self.species):
            ads = float(out[3 + k]) if k < 2 else 0.0
            inv[f"gas_{s}"], inv[f"ads_{s}"], inv[f"tot_{s}"] = float(out[k]), ads, float(out[k]) + ads
        inv["E_bed"] = float(out[5])
        inv["E_wall"] = float(out[6]) if self.n_w else 0.0
        inv["E_total"] = inv["E_bed"] + inv["E_wall"]
        return inv

    def inventories(self, y) -> Dict[str, float]:
        """Moles of each species (gas + adsorbed) and bed + wall energy (core/residual.py:426-443)."""
        return self.inventories_dev(y if not isinstance(y, np.ndarray) else dv.to_dev(y))

    def q_feed_eq(self) -> float:
        cfg = self.cfg
        T = cfg.feed.T
        psat = float(self.props.p_sat(T))
        yv = humid_air_composition(self.species, cfg.feed.P, T, cfg.feed.RH, cfg.feed.y_co2_dry, psat)
        p = yv * cfg.feed.P
        _, qw = self.equilibrium(np.array([p[0]]), np.array([p[1]]), np.array([T]), np.array([0.0]))
        qc, _ = self.equilibrium(np.array([p[0]]), np.array([p[1]]), np.array([T]), qw)
        return float(qc[0])

    def stop_metric(self, t, y) -> Optional[float]:
        """Step stop criteria (core/residual.py:445-469)."""
        stop = self.step.get("stop_on")
        if not stop:
            return None
        crit = stop["criterion"]
        g = self.grid
        y = y if isinstance(y, np.ndarray) else dv.to_host(y)
        c, q, e, Tw = self.unpack(y)
        if crit in ("min_temperature", "max_temperature"):
            T = self.thermo.temperature(e, c, q)
            return float(np.sum(T * g.V) / np.sum(g.V * np.ones_like(T)))
        if crit == "end_saturation":
            return float(np.mean(q[0][..., -1]) / self.q_feed_eq())
        if crit == "outlet_ratio":
            d = self.diag(t, y)
            b = d["bnd"]["bottom"] if self.conds["top"].kind == "feed" else d["bnd"]["top"]
            Fout = -b["F_in"]
            tot = Fout.sum(axis=0)
            m = tot > 0
            if not np.any(m):
                return 0.0
            y_out = Fout[self.i_co2][m].sum() / tot[m].sum()
            yf = self.conds["top"].y_ext[self.i_co2] if self.conds["top"].y_ext is not None else 4e-4
            return float(y_out / yf)
        return None

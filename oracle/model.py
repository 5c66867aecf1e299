"""Oracle restatement of the semi-discrete TVSA model (test infrastructure only).

Grid geometry, gas properties, closures, boundary fluxes and the right-hand side f(t, y) of
the finite-volume model, in NumPy float64 with the upstream evaluation order so that results
match the upstream code to round-off.  Upstream references are ``path:line`` in the upstream
repository.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, Optional

import numpy as np

R_GAS = 8.314462618                       # properties/ideal_gas_backend.py:23

# ------------------------------------------------------------------ pure-component data
# properties/ideal_gas_backend.py:25-44 (data tables, Young 2021 ESI S6/S7 + Incropera A.4)
_CP = {"N2": (1.06569e-5, -6.4057779e-3, 30.1318), "CO2": (-3.23753e-5, 6.17517420e-2, 21.9165),
       "H2O": (1.72619108e-3, -1.22971184026, 250.9868), "O2": (9.08259e-6, 1.00528795e-3, 28.3000)}
_MU = {"N2": (3.96929e-8, 6.15036e-6), "CO2": (4.41697e-8, 1.90440e-6),
       "H2O": (3.55595e-8, -9.930767e-7), "O2": (4.75391e-8, 6.53871e-6)}
_K = {"N2": (0.02598, 0.81), "O2": (0.02676, 0.84), "CO2": (0.01655, 1.335), "H2O": (0.0186, 1.18)}
_M = {"CO2": 0.04401, "H2O": 0.018015, "N2": 0.028960}
_AIR = (("N2", 0.79), ("O2", 0.21))
# utils/diffusion.py:14-15 (Poling Table 11-1)
_DV = {"CO2": 26.7, "H2O": 13.1, "N2": 18.5, "O2": 16.3, "AIR": 19.7}
_MG = {"CO2": 44.01, "H2O": 18.015, "N2": 28.013, "O2": 31.999, "AIR": 28.96}


def arden_buck(T):
    """Water saturation pressure [Pa] (properties/ideal_gas_backend.py:48-51)."""
    Tc = np.asarray(T, dtype=float) - 273.15
    return 611.21 * np.exp((18.678 - Tc / 234.5) * (Tc / (257.14 + Tc)))


def fuller(T, P, a, b):
    """Fuller binary diffusivity [m2/s] (utils/diffusion.py:19-26)."""
    Mab = 2.0 / (1.0 / _MG[a] + 1.0 / _MG[b])
    s = _DV[a] ** (1 / 3) + _DV[b] ** (1 / 3)
    T = np.asarray(T, dtype=float)
    P = np.maximum(np.asarray(P, dtype=float), 1.0)
    return 1.43e-7 * T ** 1.75 / ((P / 1e5) * np.sqrt(Mab) * s * s)


class Gas:
    """Ideal-gas mixture properties (properties/ideal_gas_backend.py:54-166)."""

    def __init__(self, species, cp_T=330.0):
        self.species = list(species)
        self.ng = len(self.species)
        self.M = np.array([_M[s] for s in self.species])
        self.cp = np.array([self._cp1(s, cp_T) for s in self.species])
        self.i_co2 = self.species.index("CO2")
        self.i_h2o = self.species.index("H2O") if "H2O" in self.species else None
        self.i_inert = self.species.index("N2")
        Mi, Mj = self.M[:, None], self.M[None, :]
        self.wden = np.sqrt(8.0 * (1.0 + Mi / Mj))
        self.mq = (Mj / Mi) ** 0.25

    @staticmethod
    def _cp1(s, T):
        if s == "N2":
            return sum(f * np.polyval(_CP[c], T) for c, f in _AIR)
        return float(np.polyval(_CP[s], T))

    @staticmethod
    def _mu1(s, T):
        if s == "N2":
            return sum(f * (_MU[c][0] * T + _MU[c][1]) for c, f in _AIR)
        return _MU[s][0] * T + _MU[s][1]

    @staticmethod
    def _k1(s, T):
        if s == "N2":
            return sum(f * _K[c][0] * (T / 300.0) ** _K[c][1] for c, f in _AIR)
        return _K[s][0] * (T / 300.0) ** _K[s][1]

    def p_sat(self, T):
        return arden_buck(T)

    def mu_k_mix(self, T, y):
        """Wilke viscosity / Mason-Saxena conductivity (:94-113)."""
        T = np.asarray(T, dtype=float)
        mu = np.stack([self._mu1(s, T) for s in self.species])
        k = np.stack([self._k1(s, T) for s in self.species])
        sq = np.sqrt(mu)
        mu_m = 0.0
        k_m = 0.0
        for i in range(self.ng):
            den = 0.0
            for j in range(self.ng):
                if i == j:
                    den = den + y[j]
                else:
                    den = den + y[j] * ((1.0 + (sq[i] / sq[j]) * self.mq[i, j]) ** 2 / self.wden[i, j])
            den = np.maximum(den, 1e-30)
            mu_m = mu_m + y[i] * mu[i] / den
            k_m = k_m + y[i] * k[i] / den
        return mu_m, k_m

    def _lin(self, w, y):
        out = w[0] * y[0]
        for k in range(1, self.ng):
            out = out + w[k] * y[k]
        return out

    def cp_mix(self, y):
        return self._lin(self.cp, y)

    def molar_mass_mix(self, y):
        return self._lin(self.M, y)

    def D_co2_mix(self, T, P, y):
        """Blanc's law over air and water (:127-137)."""
        D_ca = fuller(T, P, "CO2", "AIR")
        yi = y[self.i_inert]
        if self.i_h2o is None:
            return D_ca
        D_cw = fuller(T, P, "CO2", "H2O")
        yw = y[self.i_h2o]
        return (yi + yw + 1e-9) / (yi / D_ca + yw / D_cw + 1e-9 / D_ca)


# ------------------------------------------------------------------ geometry
class Grid:
    """Cylindrical FV grid (core/grid.py:26-104); z = 0 is the top (feed) end."""

    def __init__(self, R, L, Nr, Nth, Nz, wall_thickness=0.002, radial_spacing="uniform"):
        self.R, self.L, self.Nr, self.Nth, self.Nz = R, L, Nr, Nth, Nz
        self.wall_thickness = wall_thickness
        if radial_spacing == "equal_area":
            self.r_f = R * np.sqrt(np.linspace(0.0, 1.0, Nr + 1))
        else:
            self.r_f = np.linspace(0.0, R, Nr + 1)
        self.r_c = 0.5 * (self.r_f[1:] + self.r_f[:-1])
        self.dr = np.diff(self.r_f)
        self.dth = 2.0 * np.pi / Nth
        self.th_f = np.linspace(0.0, 2.0 * np.pi, Nth + 1)
        self.th_c = 0.5 * (self.th_f[1:] + self.th_f[:-1])
        self.dz = L / Nz
        self.z_f = np.linspace(0.0, L, Nz + 1)
        self.z_c = 0.5 * (self.z_f[1:] + self.z_f[:-1])
        ring = 0.5 * (self.r_f[1:] ** 2 - self.r_f[:-1] ** 2)
        self.Az = (ring * self.dth)[:, None]
        self.V = (ring * self.dth * self.dz)[:, None, None]
        self.Ar = (self.r_f[1:-1] * self.dth * self.dz)[:, None, None]
        self.Ath = (self.dr * self.dz)[:, None, None]
        self.drc = np.diff(self.r_c)[:, None, None]
        self.arc = (self.r_c * self.dth)[:, None, None]
        self.A_wall_in = R * self.dth * self.dz
        Ro = R + wall_thickness
        self.A_wall_out = Ro * self.dth * self.dz
        self.V_wall = 0.5 * (Ro ** 2 - R ** 2) * self.dth * self.dz
        self.A_wall_axial = 0.5 * (Ro ** 2 - R ** 2) * self.dth
        self.R_wall_mid = R + 0.5 * wall_thickness
        self.half_dr_outer = R - self.r_c[-1]
        self.ncell = Nr * Nth * Nz
        self.shape = (Nr, Nth, Nz)
        self.wall_shape = (Nth, Nz)

    @classmethod
    def from_config(cls, cfg):
        return cls(cfg.geometry.bed_radius, cfg.geometry.bed_length, cfg.grid.Nr, cfg.grid.Ntheta,
                   cfg.grid.Nz, cfg.geometry.wall_thickness, cfg.grid.radial_spacing)

    def cell_centres(self):
        return np.meshgrid(self.r_c, self.th_c, self.z_c, indexing="ij")

    def volume_full(self):
        return np.broadcast_to(self.V, self.shape)

    def coverage_fraction(self, theta_deg, z_frac):
        """Jacket coverage of each wall cell (core/grid.py:98-104)."""
        t0, t1 = np.radians(theta_deg[0]), np.radians(theta_deg[1])
        z0, z1 = z_frac[0] * self.L, z_frac[1] * self.L
        ov_t = np.clip(np.minimum(self.th_f[1:], t1) - np.maximum(self.th_f[:-1], t0), 0, None) / self.dth
        ov_z = np.clip(np.minimum(self.z_f[1:], z1) - np.maximum(self.z_f[:-1], z0), 0, None) / self.dz
        return ov_t[:, None] * ov_z[None, :]


# ------------------------------------------------------------------ closures
def toth(p, T, q_inf0, chi, b0, dH0, tau0, alpha, T0):
    """Temperature-dependent Toth isotherm (physics/isotherms.py:166-172)."""
    p = np.maximum(p, 0.0)
    qinf = q_inf0 * np.exp(chi * (1.0 - T / T0))
    b = b0 * np.exp(-dH0 / (R_GAS * T))
    tau = tau0 + alpha * (1.0 - T0 / T)
    bp = b * p
    return qinf * bp / (1.0 + bp ** tau) ** (1.0 / tau)


def gab(x, T, q_m, C, D, F, G, E10_slope, E10_intercept):
    """GAB water isotherm with Anderson T-dependence (physics/isotherms.py:175-183)."""
    E1 = C - np.exp(D * T)
    E29 = F + G * T
    E10 = E10_slope * T + E10_intercept
    RT = R_GAS * T
    c = np.exp((E1 - E10) / RT)
    k = np.exp((E29 - E10) / RT)
    kx = k * x
    return q_m * c * kx / ((1.0 - kx) * (1.0 + (c - 1.0) * kx))


def smooth_cap(x, cap, width=0.02):
    """C1 smoothed min(x, cap) (physics/isotherms.py:186-196)."""
    x = np.maximum(x, -0.2)
    lo = cap - width
    over = np.maximum(x - lo, 0.0)
    blend = lo + over - over ** 2 / (4.0 * width)
    return np.where(x <= lo, x, np.where(over >= 2.0 * width, cap, blend))


class Equilibrium:
    """q* for CO2 (+ co-adsorption) and H2O (physics/isotherms.py:199-269)."""

    def __init__(self, cfg):
        iso = cfg.isotherm
        c = iso.co2
        self.co2 = dict(q_inf0=c.q_inf0, chi=c.chi, b0=c.b0, dH0=c.dH0, tau0=c.tau0, alpha=c.alpha, T0=c.T0)
        h = iso.h2o
        self.h2o = dict(q_m=h.q_m, C=h.C, D=h.D, F=h.F, G=h.G, E10_slope=h.E10_slope,
                        E10_intercept=h.E10_intercept)
        self.rh_cap = h.rh_cap
        self.p_reg = c.p_regularization
        self.water = bool(cfg.components.water)
        self.model = iso.coadsorption.model if self.water else "none"
        self.basis = iso.coadsorption.water_loading_basis
        co = iso.coadsorption
        self.wadst = co.wadst.to_dict()
        self.mech = co.mechanistic.to_dict()
        self.sb = co.stampi_bombelli.to_dict()

    def q_co2_dry(self, p, T):
        return toth(p, T, **self.co2)

    def q_h2o(self, p_w, T, p_sat):
        return gab(smooth_cap(p_w / p_sat, self.rh_cap), T, **self.h2o)

    def q_co2(self, p, T, qw):
        pp = np.maximum(p, 0.0)
        return self._raw(pp, T, qw) * (pp / (pp + self.p_reg))

    def _raw(self, p, T, qw):
        qw = np.maximum(qw, 1e-12)
        m = self.model
        if m == "none":
            return self.q_co2_dry(p, T)
        if m == "wadst":
            wd = self.wadst
            w = np.exp(-wd["A"] / qw)
            q_wet = toth(p, T, wd["q_inf0_wet"], wd["chi_wet"], wd["b0_wet"], wd["dH_wet"],
                         wd["tau0_wet"], wd["alpha_wet"], self.co2["T0"])
            return (1.0 - w) * self.q_co2_dry(p, T) + w * q_wet
        if m == "mechanistic":
            mm = self.mech
            w = np.exp(-mm["A"] / qw)
            f_bl = mm["f_blocked_max"] * (1.0 - np.exp(-(mm["k"] * qw) ** mm["n"]))
            phi = mm["phi_dry"] + ((1.0 - f_bl) - mm["phi_dry"]) * w
            pars = dict(self.co2)
            pars["dH0"] = (1.0 - w) * self.co2["dH0"] + w * mm["dH_wet"]
            return (phi / mm["phi_dry"]) * toth(p, T, **pars)
        if m == "stampi_bombelli":
            pars = dict(self.co2)
            pars["q_inf0"] = self.co2["q_inf0"] / (1.0 - self.sb["gamma"] * qw)
            pars["b0"] = self.co2["b0"] * (1.0 + self.sb["beta"] * qw)
            return toth(p, T, **pars)
        raise ValueError(m)

    def equilibrium(self, p_co2, p_h2o, T, q_h2o_state, p_sat):
        if not self.water:
            return self.q_co2(p_co2, T, None)[None]
        qw_eq = self.q_h2o(p_h2o, T, p_sat)
        qc = self.q_co2(p_co2, T, q_h2o_state if self.basis == "state" else qw_eq)
        return np.stack([qc, qw_eq])


def zehner_schlunder(lam_g, lam_s, eps, Cf=1.25):
    """Stagnant bed conductivity (physics/heat_transfer.py:23-30)."""
    kap = lam_s / lam_g
    B = Cf * ((1.0 - eps) / eps) ** (10.0 / 9.0)
    s = np.sqrt(1.0 - eps)
    one_m = 1.0 - B / kap
    kc = (2.0 / one_m) * ((1.0 - 1.0 / kap) * B / one_m ** 2 * np.log(kap / B)
                          - 0.5 * (B + 1.0) - (B - 1.0) / one_m)
    return lam_g * (1.0 - s + s * kc)


def ergun_velocity(g, a, b):
    """Superficial velocity from dP/dn (physics/pressure.py:39-42)."""
    return -2.0 * g / (a + np.sqrt(a * a + 4.0 * b * np.abs(g)))


def ergun_dp_dn(u, a, b):
    return -(a * u + b * np.abs(u) * u)


def ramp(t, P_start, P_target, tau):
    return P_target + (P_start - P_target) * np.exp(-t / tau)


def smooth_positive(u, delta=1e-7):
    return 0.5 * (u + np.sqrt(u * u + delta * delta)) - 0.5 * delta


def softplus(x, width):
    """width * log(1 + exp(x / width)) (boundaries/conditions.py:46-48)."""
    return width * np.logaddexp(0.0, x / width)


U_SMOOTH = 1e-4                          # numerics/stencils.py:42
VALVE_DP_SMOOTH = 0.2                    # boundaries/conditions.py:43


def upwind_weight(u):
    return 0.5 * (1.0 + np.tanh(u / U_SMOOTH))


def limited_slope(dm, dp, limiter):
    """Limited slope from backward/forward differences (numerics/stencils.py:59-76)."""
    if limiter == "upwind":
        return np.zeros_like(dm)
    prod = dm * dp
    if limiter == "van_albada":
        return prod * (dm + dp) / (dm * dm + dp * dp + 1e-300)
    if limiter == "van_albada_tvd":
        return np.where(prod > 0.0, prod * (dm + dp) / (dm * dm + dp * dp + 1e-300), 0.0)
    if limiter == "van_leer":
        return np.where(prod > 0.0, 2.0 * prod / (dm + dp + 1e-300), 0.0)
    if limiter == "minmod":
        return np.where(prod > 0.0, np.sign(dm) * np.minimum(np.abs(dm), np.abs(dp)), 0.0)
    raise ValueError(limiter)


def face_values_z(phi, u, limiter):
    """MUSCL face values on interior z faces (numerics/stencils.py:79-95)."""
    w = upwind_weight(u)
    if phi.shape[-1] < 3 or limiter == "upwind":
        return w * phi[..., :-1] + (1.0 - w) * phi[..., 1:]
    d = np.diff(phi, axis=-1)
    s = np.zeros_like(phi)
    s[..., 1:-1] = limited_slope(d[..., :-1], d[..., 1:], limiter)
    return w * (phi[..., :-1] + 0.5 * s[..., :-1]) + (1.0 - w) * (phi[..., 1:] - 0.5 * s[..., 1:])


def upwind(phi_minus, phi_plus, u):
    w = upwind_weight(u)
    return w * phi_minus + (1.0 - w) * phi_plus


def tnext(a):
    n = a.shape[-2]
    return np.take(a, np.r_[1:n, 0], axis=-2)


def tprev(a):
    n = a.shape[-2]
    return np.take(a, np.r_[n - 1, 0:n - 1], axis=-2)


def divergence(Fz_full, Fr, Ft, nth):
    """Net outflow per cell (physics/mass.py:61-74)."""
    out = Fz_full[..., 1:] - Fz_full[..., :-1]
    if Fr is not None and Fr.shape[-3] > 0:
        out[..., :-1, :, :] += Fr
        out[..., 1:, :, :] -= Fr
    if Ft is not None and nth > 1:
        out += Ft - tprev(Ft)
    return out


def _wsum(w, a):
    out = w[0] * a[0]
    for k in range(1, len(w)):
        out = out + w[k] * a[k]
    return out


# ------------------------------------------------------------------ valve conditions
@dataclass
class EndCond:
    """Boundary condition at one end (valves/valve_logic.py:47-60)."""
    end: str
    valve: Optional[str]
    state: str
    kind: str
    P_target: float = 101325.0
    y_ext: Optional[np.ndarray] = None
    T_ext: float = 298.15
    check_valve: bool = False
    is_product: bool = False
    is_vacuum: bool = False
    is_feed: bool = False
    speed: float = 0.0


def humid_air(species, P, T, RH, y_co2_dry, p_sat):
    """valves/valve_logic.py:63-68."""
    y_w = RH * p_sat / P if "H2O" in species else 0.0
    y = {"H2O": y_w, "CO2": y_co2_dry * (1.0 - y_w)}
    y["N2"] = 1.0 - y["CO2"] - y_w
    return np.array([y[s] for s in species])


def end_conditions(cfg, step, species, gas):
    """valves/valve_logic.py:71-109."""
    fd, amb = cfg.feed, cfg.ambient
    water = "H2O" in species
    y_feed = humid_air(species, fd.P, fd.T, fd.RH if water else 0.0, fd.y_co2_dry, float(gas.p_sat(fd.T)))
    y_amb = humid_air(species, amb.P, amb.T, amb.RH if water else 0.0, amb.y_co2_dry, float(gas.p_sat(amb.T)))
    out = {}
    for end in ("top", "bottom"):
        valve = None
        for v in ("V1", "V2", "V3"):
            if cfg.ports[v] == end and step["valves"][v] != "CLOSED":
                valve = v
        if valve is None:
            out[end] = EndCond(end, None, "CLOSED", "closed")
            continue
        st = step["valves"][valve]
        is_prod = step.get("product_port", "V3") == valve
        op = cfg.operation
        if st == "OPEN":
            out[end] = EndCond(end, valve, st, "feed", fd.P, y_feed, fd.T, is_feed=True)
        elif st == "OPEN_ATMOSPHERIC":
            kind = "valve" if op.vent_model == "valve" else "pressure"
            if valve == "V1":
                out[end] = EndCond(end, valve, st, kind, fd.P, y_feed, fd.T, is_product=is_prod,
                                   is_feed=True, speed=op.vent_conductance)
            else:
                out[end] = EndCond(end, valve, st, kind, amb.P, y_amb, amb.T, is_product=is_prod,
                                   speed=op.vent_conductance)
        else:
            kind = "pump" if op.vacuum_model == "pump" else "pressure"
            out[end] = EndCond(end, valve, st, kind, op.P_vacuum, y_amb, amb.T,
                               check_valve=bool(op.vacuum_check_valve), is_product=is_prod,
                               is_vacuum=True, speed=op.pump_speed_per_area)
    return out


def feed_molar_flux(cfg):
    return cfg.feed.velocity * cfg.feed.P / (R_GAS * cfg.feed.T)


def feed_profile(cfg, g):
    """Inlet distribution with area-weighted mean 1 (valves/valve_logic.py:117-148)."""
    fd = cfg.feed
    A = fd.profile_amplitude
    rr, th = np.meshgrid(g.r_c, g.th_c, indexing="ij")
    if fd.profile == "uniform":
        f = np.ones_like(rr)
    elif fd.profile == "cosine_theta":
        f = 1.0 + A * np.cos(th - np.radians(fd.profile_theta0_deg))
    elif fd.profile == "linear_r":
        f = 1.0 + A * (2.0 * rr / g.R - 1.0)
    elif fd.profile == "central_disc":
        r_in = fd.profile_radius_fraction * g.R
        lo, hi = g.r_f[:-1], g.r_f[1:]
        frac = np.clip((np.minimum(hi, r_in) ** 2 - lo ** 2) / (hi ** 2 - lo ** 2), 0.0, 1.0)
        f = np.repeat(frac[:, None], g.Nth, axis=1)
    elif fd.profile == "sector":
        half = 0.5 * np.radians(fd.profile_sector_deg)
        t0 = np.radians(fd.profile_theta0_deg)
        lo, hi = g.th_f[:-1] - t0, g.th_f[1:] - t0
        ov = np.zeros(g.Nth)
        for sh in (-2 * np.pi, 0.0, 2 * np.pi):
            ov += np.clip(np.minimum(hi + sh, half) - np.maximum(lo + sh, -half), 0.0, None)
        f = np.repeat((ov / g.dth)[None, :], g.Nr, axis=0)
    else:
        raise ValueError(fd.profile)
    w = g.Az * np.ones_like(f)
    return f * w.sum() / (f * w).sum()


def end_face_fluxes(cond, t, P_start, lay, Az, half_dz, ka, kb, cpg, T_ref, feed_flux, tau):
    """Fluxes through one end of the column (boundaries/conditions.py:75-160)."""
    c, T, P = lay["c"], lay["T"], lay["P"]
    ng = c.shape[0]
    if cond.kind == "closed":
        z = np.zeros_like(T)
        return {"F_in": np.zeros_like(c), "H_in": z, "u_out": z.copy(), "P_face": P.copy()}
    a, b = ka * lay["mu"], kb * lay["rho"]
    if cond.kind == "feed":
        F_in = feed_flux[None] * cond.y_ext.reshape(ng, 1, 1)
        H_in = np.tensordot(cpg, F_in, axes=(0, 0)) * (cond.T_ext - T_ref)
        C_cell = np.maximum(c.sum(axis=0), 1e-12)
        u_in = feed_flux / (Az * C_cell)
        P_face = P + (a * u_in + b * u_in * u_in) * half_dz
        return {"F_in": F_in, "H_in": H_in, "u_out": -u_in, "P_face": P_face}
    if cond.kind == "pump":
        dp = P - cond.P_target
        C_cell = np.maximum(c.sum(axis=0), 1e-12)
        y_cell = c / C_cell
        if cond.check_valve:
            dp = softplus(dp, 1.0)
            w_out = np.ones_like(dp)
        else:
            w_out = 0.5 * (1.0 + np.tanh(dp / VALVE_DP_SMOOTH))
        N_out = cond.speed * dp / (R_GAS * T)
        y_use = w_out[None] * y_cell + (1.0 - w_out[None]) * cond.y_ext.reshape(ng, 1, 1)
        F_out = (N_out * Az)[None] * y_use
        T_up = w_out * T + (1.0 - w_out) * cond.T_ext
        H_out = np.tensordot(cpg, F_out, axes=(0, 0)) * (T_up - T_ref)
        u_out = N_out / C_cell
        P_face = P - (a * u_out + b * np.abs(u_out) * u_out) * half_dz
        return {"F_in": -F_out, "H_in": -H_out, "u_out": u_out, "P_face": P_face}
    if cond.kind == "valve":
        P_res = ramp(t, P_start, cond.P_target, tau) if tau > 0 else cond.P_target
        dp = P_res - P
        C_cell = np.maximum(c.sum(axis=0), 1e-12)
        p_in = softplus(dp, VALVE_DP_SMOOTH)
        p_out = softplus(-dp, VALVE_DP_SMOOTH)
        N_in = cond.speed * p_in / (R_GAS * cond.T_ext)
        N_out = cond.speed * p_out / (R_GAS * T)
        F_inp = (N_in * Az)[None] * cond.y_ext.reshape(ng, 1, 1)
        F_outp = (N_out * Az)[None] * (c / C_cell)
        H_inp = np.tensordot(cpg, F_inp, axes=(0, 0)) * (cond.T_ext - T_ref)
        H_outp = np.tensordot(cpg, F_outp, axes=(0, 0)) * (T - T_ref)
        u_out = -(N_in - N_out) / C_cell
        P_face = P - (a * u_out + b * np.abs(u_out) * u_out) * half_dz
        return {"F_in": F_inp - F_outp, "H_in": H_inp - H_outp, "u_out": u_out, "P_face": P_face,
                "F_inpart": F_inp, "F_outpart": F_outp, "H_inpart": H_inp, "H_outpart": H_outp}
    P_face = ramp(t, P_start, cond.P_target, tau) * np.ones_like(P)
    u_out = ergun_velocity((P_face - P) / half_dz, a, b)
    if cond.check_valve:
        u_out = smooth_positive(u_out)
    c_ext = cond.y_ext.reshape(ng, 1, 1) * P_face[None] / (R_GAS * cond.T_ext)
    w = upwind_weight(u_out)
    F_out = u_out[None] * Az[None] * (w[None] * c + (1.0 - w[None]) * c_ext)
    T_up = w * T + (1.0 - w) * cond.T_ext
    H_out = np.tensordot(cpg, F_out, axes=(0, 0)) * (T_up - T_ref)
    return {"F_in": -F_out, "H_in": -H_out, "u_out": u_out, "P_face": P_face}


# ------------------------------------------------------------------ the model
class OracleModel:
    """Restatement of ``TVSAModel`` (core/residual.py:40-482) for one cycle step."""

    def __init__(self, cfg, grid, gas, step: Dict, state0, T_jacket_prev: Optional[float] = None):
        self.cfg, self.grid, self.gas, self.step = cfg, grid, gas, step
        g = grid
        self.species = list(gas.species)
        self.ng = len(self.species)
        self.water = bool(cfg.components.water)
        self.adsorbed = ["CO2", "H2O"] if self.water else ["CO2"]
        self.na = len(self.adsorbed)
        self.i_co2 = self.species.index("CO2")
        self.i_h2o = self.species.index("H2O") if self.water else None
        self.i_inert = self.species.index("N2")
        self.ads_to_gas = [self.species.index(s) for s in self.adsorbed]
        sb, tr, th = cfg.sorbent, cfg.transport, cfg.thermo
        self.eps_b, self.eps_t = sb.bed_voidage, sb.total_voidage
        self.rho_b = cfg.derived.bulk_density
        self.cp_s = sb.cp_solid
        self.cpg = np.asarray(gas.cp)
        self.cpa = np.array([th.cp_ads_co2, th.cp_ads_h2o] if self.water else [th.cp_ads_co2])
        self.dHa = np.array([th.dH_ads_co2, th.dH_ads_h2o] if self.water else [th.dH_ads_co2])
        self.T_ref = th.T_ref
        self.limiter = cfg.solver.limiter
        self.lumped = tr.heat_transfer_mode == "lumped"
        self.eq = Equilibrium(cfg)
        kn = cfg.kinetics
        self.k_ref = np.array([kn.k_ldf_co2, kn.k_ldf_h2o] if self.water else [kn.k_ldf_co2])
        self.Ea = np.array([kn.Ea_co2, kn.Ea_h2o] if self.water else [kn.Ea_co2])
        self.k_Tref = kn.T_ref
        dp, eps = sb.particle_diameter, sb.bed_voidage
        self.dp = dp
        self.ka = tr.ergun_A * (1.0 - eps) ** 2 / (eps ** 3 * dp ** 2)        # physics/pressure.py:36-42
        self.kb = tr.ergun_B * (1.0 - eps) / (eps ** 3 * dp)
        self.disp_model = tr.dispersion_model
        self.g1, self.g2, self.beta, self.Pe_T = tr.er_gamma1, tr.er_gamma2, tr.er_beta, tr.Pe_mass_radial_inf
        self.bed_model, self.lam_const, self.Cf = tr.bed_conductivity_model, tr.bed_conductivity_constant, tr.zs_shape_factor
        self.lam_s = sb.particle_conductivity
        self.Pe_r, self.Pe_z = tr.Pe_heat_radial_inf, tr.Pe_heat_axial_inf
        self.wall_model, self.alpha_const = tr.wall_htc_model, tr.wall_htc_constant
        self.N_ratio = 2.0 * cfg.geometry.bed_radius / dp
        w = cfg.wall
        self.wall_on = bool(w.enabled)
        self.w_lam = w.conductivity
        self.wall_C = w.density * w.cp * g.V_wall
        # boundary conditions of this step (core/residual.py:70-85)
        self.conds = end_conditions(cfg, step, self.species, gas)
        self.tau_valve = cfg.operation.valve_time_constant
        P0 = state0.c.sum(axis=0) * R_GAS * state0.T
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
        self.hA_v = cfg.internal_hx.hA_per_volume
        self.U_lumped = tr.lumped_U
        en = cfg.energy
        self.gam, self.eta_vac, self.vac_model = en.gamma_gas, en.eta_vacuum, en.vacuum_efficiency_model
        self.eta_bl, self.pump_water = en.eta_blower, bool(en.pump_handles_water)
        n = g.ncell
        self.n_c = self.ng * n
        self.off_q = self.n_c
        self.off_e = self.n_c + self.na * n
        self.off_w = self.off_e + n
        self.n_w = g.Nth * g.Nz if self.wall_on else 0
        self.n_core = self.off_w + self.n_w
        self.acc_names = []
        for v in ("V1", "V2", "V3"):
            self.acc_names += [f"n_in_{v}_{s}" for s in self.species]
            self.acc_names += [f"n_out_{v}_{s}" for s in self.species]
            self.acc_names += [f"H_in_{v}", f"H_out_{v}", f"W_pump_{v}", f"W_fan_{v}"]
        self.acc_names += ["Q_jac_pos", "Q_jac_neg", "Q_amb", "Q_wb", "Q_hx_pos", "Q_hx_neg"]
        self.n_acc = len(self.acc_names)
        self.acc_index = {nm: i for i, nm in enumerate(self.acc_names)}
        s = cfg.solver
        tol_c = {"CO2": s.atol_c_co2, "H2O": s.atol_c_h2o, "N2": s.atol_c_inert}
        parts = [np.full(n, tol_c[sp]) for sp in self.species]
        parts += [np.full(n, s.atol_q) for _ in range(self.na)]
        parts.append(np.full(n, self.rho_b * self.cp_s * s.atol_T))
        if self.n_w:
            parts.append(np.full(self.n_w, s.atol_T))
        self.atol = np.concatenate(parts)
        self.nfev = 0

    # ---------------------------------------------------------------- thermo (physics/energy.py:46-60)
    def _terms(self, c, q):
        Cg = c[0] + c[1] if len(c) == 2 else c[0] + c[1] + c[2]
        A = self.eps_t * _wsum(self.cpg, c) + self.rho_b * (self.cp_s + _wsum(self.cpa, q))
        B = self.eps_t * R_GAS * Cg
        Hq = self.rho_b * _wsum(self.dHa, q)
        return A, B, Hq

    def temperature(self, e, c, q):
        A, B, Hq = self._terms(c, q)
        return (e - Hq + A * self.T_ref) / (A - B)

    def energy(self, T, c, q):
        A, B, Hq = self._terms(c, q)
        return A * (T - self.T_ref) - B * T + Hq

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
            vals.append(np.ravel(self.wall_C) * np.ones(self.n_w))
        return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                             shape=(self.ng + 1, self.n_core))

    def pack(self, st):
        parts = [st.c.ravel(), st.q.ravel(), self.energy(st.T, st.c, st.q).ravel()]
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

    def T_jacket(self, t):
        if self.jac_tau <= 0:
            return self.T_jac_set
        return self.T_jac_set + (self.T_jac_prev - self.T_jac_set) * np.exp(-t / self.jac_tau)

    def eta_vacuum(self, P):
        if self.vac_model == "krishnamurthy":
            pb = P / 1e5
            return np.maximum(0.8 * 19.55 * pb / (1.0 + 19.55 * pb), 1e-3)
        return self.eta_vac

    # ---------------------------------------------------------------- closures bound to the step
    def _stagnant(self, kg):
        if self.bed_model == "constant":
            return self.lam_const * np.ones_like(kg)
        if self.bed_model == "parallel":
            return self.eps_b * kg + (1.0 - self.eps_b) * self.lam_s
        return zehner_schlunder(kg, self.lam_s, self.eps_b, self.Cf)

    def _radial(self, lam0, C, cp, u):
        return lam0 + C * cp * np.abs(u) * self.dp / self.Pe_r

    def _axial(self, lam0, C, cp, u):
        return lam0 + C * cp * np.abs(u) * self.dp / self.Pe_z

    def _wall_coef(self, lam0, lam_g, rho, mu, cp_mass, u):
        """Martin-Nilles bed-to-wall coefficient (physics/heat_transfer.py:63-76)."""
        if self.wall_model == "constant":
            return self.alpha_const * np.ones_like(lam0)
        Re = rho * np.abs(u) * self.dp / mu
        Pr = cp_mass * mu / lam_g
        nu0 = (1.3 + 5.0 / self.N_ratio) * lam0 / lam_g
        if np.all(Re <= 0):
            dyn = 0.0
        else:
            a = 0.3 * Pr ** (1.0 / 3.0) * np.maximum(Re, 1e-12) ** 0.75
            b = 0.054 * Pr * np.maximum(Re, 1e-12)
            dyn = 1.0 / (1.0 / a + 1.0 / b)
        return (nu0 + dyn) * lam_g / self.dp

    def _disp_axial(self, Dm, u, rho, mu):
        """physics/dispersion.py:32-43."""
        v = np.abs(u) / self.eps_b
        dp = self.dp
        if self.disp_model == "edwards_richardson":
            return self.g1 * Dm + self.g2 * dp * v / (1.0 + self.beta * self.g1 * Dm / (dp * v + 1e-30))
        if self.disp_model == "young2021":
            return 0.7 * Dm + 0.5 * dp * v
        if self.disp_model == "wakao_kagei":
            Re = rho * np.abs(u) * dp / mu
            Sc = mu / (rho * Dm)
            return Dm / self.eps_b * (20.0 + 0.5 * Re * Sc)
        raise ValueError(self.disp_model)

    def _disp_trans(self, Dm, u):
        return self.g1 * Dm + self.dp * (np.abs(u) / self.eps_b) / self.Pe_T

    def _k_ldf(self, T):
        shape = (len(self.k_ref),) + (1,) * np.ndim(T)
        kr = self.k_ref.reshape(shape)
        if np.all(self.Ea == 0.0):
            return kr * np.ones_like(T)
        return kr * np.exp(-self.Ea.reshape(shape) / R_GAS * (1.0 / T - 1.0 / self.k_Tref))

    def _wall_conduction(self, Tw):
        """physics/energy.py:98-113."""
        g = self.grid
        out = np.zeros_like(Tw)
        if g.Nz > 1:
            fz = self.w_lam * g.A_wall_axial / g.dz * np.diff(Tw, axis=1)
            out[:, :-1] += fz
            out[:, 1:] -= fz
        if g.Nth > 1:
            kt = self.w_lam * g.wall_thickness * g.dz / (g.R_wall_mid * g.dth)
            n = Tw.shape[0]
            ft = kt * (np.take(Tw, np.r_[1:n, 0], axis=0) - Tw)
            out += ft
            out -= np.take(ft, np.r_[n - 1, 0:n - 1], axis=0)
        return out

    # ---------------------------------------------------------------- rhs (core/residual.py:211-417)
    def rhs(self, t, y, want_acc=True, diag=False):
        self.nfev += 1
        g = self.grid
        c, q, e, Tw = self.unpack(y)
        gas = self.gas
        Ctot = c.sum(axis=0)
        T = self.temperature(e, c, q)
        P = Ctot * R_GAS * T
        Tp = np.clip(T, 150.0, 1500.0)
        Cpos = np.maximum(Ctot, 1e-9)
        yk = c / Cpos
        ycl = np.clip(yk, 0.0, 1.0)
        mu, kg = gas.mu_k_mix(Tp, ycl)
        rho = Cpos * gas.molar_mass_mix(ycl)
        cpm = gas.cp_mix(ycl)
        Dm = gas.D_co2_mix(Tp, np.maximum(P, 1.0), ycl)
        lam0 = self._stagnant(kg)
        pk = c * R_GAS * Tp
        psat = gas.p_sat(Tp)
        if self.water:
            qstar = self.eq.equilibrium(pk[self.i_co2], pk[self.i_h2o], Tp, q[1], psat)
        else:
            qstar = self.eq.equilibrium(pk[self.i_co2], None, Tp, None, psat)
        rate = self._k_ldf(Tp) * (qstar - q)
        ka, kb = self.ka, self.kb
        mu_z = 0.5 * (mu[..., 1:] + mu[..., :-1])
        rho_z = 0.5 * (rho[..., 1:] + rho[..., :-1])
        uz = ergun_velocity(np.diff(P, axis=-1) / g.dz, ka * mu_z, kb * rho_z)
        if g.Nr > 1:
            ur = ergun_velocity(np.diff(P, axis=0) / g.drc, ka * (0.5 * (mu[1:] + mu[:-1])),
                                kb * (0.5 * (rho[1:] + rho[:-1])))
        else:
            ur = None
        if g.Nth > 1:
            uth = ergun_velocity((tnext(P) - P) / g.arc, ka * (0.5 * (mu + tnext(mu))),
                                 kb * (0.5 * (rho + tnext(rho))))
        else:
            uth = None
        bnd = {}
        for end, k in (("top", 0), ("bottom", -1)):
            lay = {"c": c[..., k], "T": T[..., k], "P": P[..., k], "mu": mu[..., k], "rho": rho[..., k]}
            bnd[end] = end_face_fluxes(self.conds[end], t, self.P_start[end], lay, g.Az, 0.5 * g.dz,
                                       ka, kb, self.cpg, self.T_ref, self.feed_flux, self.tau_valve)
        Cz = 0.5 * (Cpos[..., 1:] + Cpos[..., :-1])
        u_disp = self.N_nom / Cz
        Dz = self._disp_axial(0.5 * (Dm[..., 1:] + Dm[..., :-1]), u_disp, rho_z, mu_z)
        Fz_conv = uz * g.Az[:, :, None] * face_values_z(c, uz, self.limiter)
        Jz = -self.eps_b * Dz * Cz * np.diff(yk, axis=-1) / g.dz * g.Az[:, :, None]
        Fz_full = np.concatenate([bnd["top"]["F_in"][..., None], Fz_conv + Jz,
                                  -bnd["bottom"]["F_in"][..., None]], axis=-1)
        if ur is not None:
            Cr = 0.5 * (Cpos[1:] + Cpos[:-1])
            Dr = self._disp_trans(0.5 * (Dm[1:] + Dm[:-1]), self.u_nom)
            Fr_conv = ur * g.Ar * upwind(c[:, :-1], c[:, 1:], ur)
            Jr = -self.eps_b * Dr * Cr * np.diff(yk, axis=1) / g.drc * g.Ar
            Fr = Fr_conv + Jr
        else:
            Fr = None
        if uth is not None:
            Ct = 0.5 * (Cpos + tnext(Cpos))
            Dt = self._disp_trans(0.5 * (Dm + tnext(Dm)), self.u_nom)
            Ft_conv = uth * g.Ath * upwind(c, tnext(c), uth)
            Jt = -self.eps_b * Dt * Ct * (tnext(yk) - yk) / g.arc * g.Ath
            Ft = Ft_conv + Jt
        else:
            Ft = None
        div_c = divergence(Fz_full, Fr, Ft, g.Nth)
        sink = np.zeros_like(c)
        for j, kgas in enumerate(self.ads_to_gas):
            sink[kgas] = rate[j]
        dcdt = -div_c / (self.eps_t * g.V) - (self.rho_b / self.eps_t) * sink
        cpg, Tr = self.cpg, self.T_ref
        Tz_f = face_values_z(T, uz, self.limiter)
        Tz_avg = 0.5 * (T[..., 1:] + T[..., :-1])
        lam_z = self._axial(0.5 * (lam0[..., 1:] + lam0[..., :-1]), Cz, 0.5 * (cpm[..., 1:] + cpm[..., :-1]), u_disp)
        Hz_int = (_wsum(cpg, Fz_conv) * (Tz_f - Tr) + _wsum(cpg, Jz) * (Tz_avg - Tr)
                  - lam_z * np.diff(T, axis=-1) / g.dz * g.Az[:, :, None])
        Hz_full = np.concatenate([bnd["top"]["H_in"][..., None], Hz_int, -bnd["bottom"]["H_in"][..., None]], axis=-1)
        if ur is not None:
            Tr_up = upwind(T[:-1], T[1:], ur)
            lam_r = self._radial(0.5 * (lam0[1:] + lam0[:-1]), Cr, 0.5 * (cpm[1:] + cpm[:-1]), self.u_nom)
            Hr = (_wsum(cpg, Fr_conv) * (Tr_up - Tr) + _wsum(cpg, Jr) * (0.5 * (T[1:] + T[:-1]) - Tr)
                  - lam_r * np.diff(T, axis=0) / g.drc * g.Ar)
        else:
            Hr = None
        if uth is not None:
            Tn = tnext(T)
            lam_t = self._radial(0.5 * (lam0 + tnext(lam0)), Ct, 0.5 * (cpm + tnext(cpm)), self.u_nom)
            Ht = (_wsum(cpg, Ft_conv) * (upwind(T, Tn, uth) - Tr) + _wsum(cpg, Jt) * (0.5 * (T + Tn) - Tr)
                  - lam_t * (Tn - T) / g.arc * g.Ath)
        else:
            Ht = None
        div_e = divergence(Hz_full, Hr, Ht, g.Nth)
        To = T[-1]
        lam_out = self._radial(lam0[-1], Cpos[-1], cpm[-1], self.u_nom)
        if self.lumped:
            R_bed = 1.0 / self.U_lumped
        else:
            alpha_w = self._wall_coef(lam0[-1], kg[-1], rho[-1], mu[-1],
                                      cpm[-1] / np.maximum(rho[-1] / Cpos[-1], 1e-6), self.u_nom)
            R_bed = 1.0 / alpha_w + g.half_dr_outer / lam_out
        Tj = self.T_jacket(t)
        if self.wall_on:
            Q_jac = self.cov * self.h_jac * g.A_wall_out * (Tj - Tw) if self.jacket_on else np.zeros_like(Tw)
            Q_amb = (1.0 - self.cov) * self.h_amb * g.A_wall_out * (self.T_amb - Tw)
            Q_wb = g.A_wall_in * (Tw - To) / R_bed
            dTw = (Q_jac + Q_amb - Q_wb + self._wall_conduction(Tw)) / self.wall_C
        else:
            if self.jacket_on and self.h_jac > 0:
                R_j = R_bed if self.lumped else R_bed + 1.0 / self.h_jac
                Q_jac = self.cov * g.A_wall_in * (Tj - To) / R_j
            else:
                Q_jac = np.zeros_like(To)
            Q_amb = ((1.0 - self.cov) * g.A_wall_in * (self.T_amb - To) / (R_bed + 1.0 / self.h_amb)
                     if self.h_amb > 0 else np.zeros_like(To))
            Q_wb = Q_jac + Q_amb
            dTw = None
        src = np.zeros_like(T)
        src[-1] += Q_wb
        if self.hx_on:
            Q_hx = self.hA_v * g.V * (Tj - T)
            src += Q_hx
        else:
            Q_hx = None
        dedt = (-div_e + src) / g.V
        parts = [dcdt.ravel(), rate.ravel(), dedt.ravel()]
        if dTw is not None:
            parts.append(dTw.ravel())
        f = np.concatenate(parts)
        if diag:
            return {"T": T, "P": P, "y": yk, "qstar": qstar, "uz": uz, "ur": ur, "uth": uth, "bnd": bnd,
                    "Q_jac": Q_jac, "Q_wb": Q_wb, "Q_amb": Q_amb, "Q_hx": Q_hx, "R_bed": R_bed}
        if not want_acc:
            return f, None
        return f, self._acc(bnd, T, Q_jac, Q_amb, Q_wb, Q_hx)

    def _acc(self, bnd, T, Q_jac, Q_amb, Q_wb, Q_hx):
        """Accumulator rates (core/residual.py:371-417)."""
        acc = np.zeros(self.n_acc)
        ai = self.acc_index
        for end in ("top", "bottom"):
            cond = self.conds[end]
            if cond.valve is None:
                continue
            v, b = cond.valve, bnd[end]
            inflow = b["u_out"] < 0.0
            Fin = b["F_in"]
            if "F_inpart" in b:
                for k, s in enumerate(self.species):
                    acc[ai[f"n_in_{v}_{s}"]] = np.sum(b["F_inpart"][k])
                    acc[ai[f"n_out_{v}_{s}"]] = np.sum(b["F_outpart"][k])
                acc[ai[f"H_in_{v}"]] = np.sum(b["H_inpart"])
                acc[ai[f"H_out_{v}"]] = np.sum(b["H_outpart"])
            else:
                for k, s in enumerate(self.species):
                    acc[ai[f"n_in_{v}_{s}"]] = np.sum(np.where(inflow, Fin[k], 0.0))
                    acc[ai[f"n_out_{v}_{s}"]] = np.sum(np.where(inflow, 0.0, -Fin[k]))
                acc[ai[f"H_in_{v}"]] = np.sum(np.where(inflow, b["H_in"], 0.0))
                acc[ai[f"H_out_{v}"]] = np.sum(np.where(inflow, 0.0, -b["H_in"]))
            ex = (self.gam - 1.0) / self.gam
            pre = self.gam / (self.gam - 1.0) * R_GAS
            if cond.is_vacuum:
                Fout = np.where(inflow, 0.0, -Fin)
                if not self.pump_water and self.i_h2o is not None:
                    Fout = np.delete(Fout, self.i_h2o, axis=0)
                k_lay = 0 if end == "top" else -1
                ratio = np.maximum(self.P_amb / np.maximum(b["P_face"], 1.0), 1.0)
                acc[ai[f"W_pump_{v}"]] = np.sum(pre * T[..., k_lay] * Fout.sum(axis=0) * (ratio ** ex - 1.0)
                                                / self.eta_vacuum(b["P_face"]))
            if cond.is_feed:
                Ftot = b["F_inpart"].sum(axis=0) if "F_inpart" in b else np.where(inflow, Fin, 0.0).sum(axis=0)
                ratio = np.maximum(b["P_face"] / self.P_amb, 1.0)
                acc[ai[f"W_fan_{v}"]] = np.sum(pre * cond.T_ext * Ftot * (ratio ** ex - 1.0) / self.eta_bl)
        acc[ai["Q_jac_pos"]] = np.sum(np.maximum(Q_jac, 0.0))
        acc[ai["Q_jac_neg"]] = np.sum(np.minimum(Q_jac, 0.0))
        acc[ai["Q_amb"]] = np.sum(Q_amb)
        acc[ai["Q_wb"]] = np.sum(Q_wb)
        if Q_hx is not None:
            acc[ai["Q_hx_pos"]] = np.sum(np.maximum(Q_hx, 0.0))
            acc[ai["Q_hx_neg"]] = np.sum(np.minimum(Q_hx, 0.0))
        return acc

    def stage_residual(self, t, y, gamma_h, rhs_known):
        f, _ = self.rhs(t, y, want_acc=False)
        return y - gamma_h * f - rhs_known

    def inventories(self, y):
        """core/residual.py:426-443."""
        g = self.grid
        c, q, e, Tw = self.unpack(y)
        out = {}
        for k, s in enumerate(self.species):
            gas = float(np.sum(self.eps_t * c[k] * g.V))
            ads = float(np.sum(self.rho_b * q[self.adsorbed.index(s)] * g.V)) if s in self.adsorbed else 0.0
            out[f"gas_{s}"], out[f"ads_{s}"], out[f"tot_{s}"] = gas, ads, gas + ads
        out["E_bed"] = float(np.sum(e * g.V))
        out["E_wall"] = float(np.sum(self.wall_C * (Tw - self.T_ref))) if Tw is not None else 0.0
        out["E_total"] = out["E_bed"] + out["E_wall"]
        return out

    def q_feed_eq(self):
        cfg = self.cfg
        T = cfg.feed.T
        psat = float(self.gas.p_sat(T))
        yv = humid_air(self.species, cfg.feed.P, T, cfg.feed.RH if self.water else 0.0, cfg.feed.y_co2_dry, psat)
        p = yv * cfg.feed.P
        if self.water:
            return float(self.eq.q_co2(p[self.i_co2], T, self.eq.q_h2o(p[self.i_h2o], T, psat)))
        return float(self.eq.q_co2_dry(p[self.i_co2], T))

    def stop_metric(self, t, y):
        """core/residual.py:445-469."""
        stop = self.step.get("stop_on")
        if not stop:
            return None
        crit = stop["criterion"]
        g = self.grid
        c, q, e, Tw = self.unpack(y)
        if crit in ("min_temperature", "max_temperature"):
            T = self.temperature(e, c, q)
            return float(np.sum(T * g.V) / np.sum(g.V * np.ones_like(T)))
        if crit == "end_saturation":
            return float(np.mean(q[0][..., -1]) / self.q_feed_eq())
        if crit == "outlet_ratio":
            d = self.rhs(t, y, diag=True)
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

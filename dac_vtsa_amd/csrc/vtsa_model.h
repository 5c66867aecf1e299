// Device-side model of one cycle step of the TVSA bed: parameters, per-cell closures, face
// fluxes and boundary fluxes of the finite-volume right-hand side f(t, y).
//
// State layout in HBM (variable-major, cell index (i*Nth + j)*Nz + k, z fastest):
//   y = [c_CO2 | c_H2O | c_N2 | q_CO2 | q_H2O | e | Tw (Nth*Nz, wall only)]
// which is the upstream pack order (core/residual.py:170-183), so host and device share it.
//
// Every formula below restates the upstream physics in the upstream evaluation order
// (operands of + and * are associated exactly as the upstream NumPy expressions), so that the
// FP64 results agree with the CPU oracle to round-off; the code is compiled with
// -ffp-contract=off for the same reason.  Upstream references are path:line.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/vtsa.h"

namespace vtsa {

constexpr int NG = 3;        // gas species CO2, H2O, N2
constexpr int NA = 2;        // adsorbed CO2, H2O
constexpr int NV = 6;        // unknowns per cell
constexpr int NPROP = 12;    // cached per-cell properties (see Props)
constexpr int NSLOT = 9;     // stencil slots of a flux row (self, z-1, z+1, z-2, z+2, r-1, r+1, t-1, t+1)
constexpr double U_SMOOTH = 1e-4;         // numerics/stencils.py:42
constexpr double VALVE_DP_SMOOTH = 0.2;   // boundaries/conditions.py:43

enum Slot { S_SELF = 0, S_ZM1, S_ZP1, S_ZM2, S_ZP2, S_RM1, S_RP1, S_TM1, S_TP1 };
enum Limiter { LIM_VAN_ALBADA = 0, LIM_VAN_ALBADA_TVD, LIM_VAN_LEER, LIM_MINMOD, LIM_UPWIND };
enum EndKind { END_CLOSED = 0, END_FEED, END_PUMP, END_VALVE, END_PRESSURE };

// Device copy of the model: the public descriptor plus device pointers to the per-ring,
// per-column and per-wall-cell arrays.
struct Model {
  vtsa_model_desc d;
  int n, nw, N;                 // cells, wall cells, unknowns
  const double* V;              // [Nr] cell volume per ring
  const double* Az;             // [Nr] axial face area per ring
  const double* Ath;            // [Nr] theta face area per ring
  const double* arc;            // [Nr] centre-to-centre arc length per ring
  const double* Ar;             // [Nr-1] radial face area
  const double* drc;            // [Nr-1] centre-to-centre radial distance
  const double* feed_flux;      // [Nr*Nth] feed molar flow per inlet face (mol/s)
  const double* N_nom;          // [Nr*Nth] nominal molar flux per column (mol/m2/s)
  const double* cov;            // [Nth*Nz] jacket coverage per wall cell
};

// Cached per-cell quantities (SoA buffer of NPROP * n doubles, property-major).
struct Props {
  double T, P, Cpos, y[NG], mu, kg, rho, cpm, Dm, lam0;
};

__host__ __device__ inline double dmax(double a, double b) { return a > b ? a : (b != b ? b : a); }
__host__ __device__ inline double dmin(double a, double b) { return a < b ? a : (b != b ? b : a); }
// np.maximum / np.minimum propagate NaN from either operand
__device__ inline double npmax(double a, double b) { return (a != a || b != b) ? (a + b) : (a > b ? a : b); }
__device__ inline double npmin(double a, double b) { return (a != a || b != b) ? (a + b) : (a < b ? a : b); }
__device__ inline double npclip(double x, double lo, double hi) { return npmin(npmax(x, lo), hi); }

// ---------------------------------------------------------------- elementary closures
// np.logaddexp(0, x) (numpy npy_logaddexp with one argument 0)
__device__ inline double logaddexp0(double x) {
  if (x == 0.0) return 0.6931471805599453094;   // x + log(2) with x == 0
  double tmp = 0.0 - x;
  if (tmp > 0.0) return 0.0 + log1p(exp(-tmp));
  if (tmp <= 0.0) return x + log1p(exp(tmp));
  return tmp;   // NaN
}
__device__ inline double softplus(double x, double width) { return width * logaddexp0(x / width); }

__device__ inline double upwind_weight(double u) { return 0.5 * (1.0 + tanh(u / U_SMOOTH)); }

// Ergun inversion u = -2 g / (a + sqrt(a^2 + 4 b |g|)) (physics/pressure.py:39-42)
__device__ inline double ergun_velocity(double g, double a, double b) {
  return (-2.0 * g) / (a + sqrt(a * a + (4.0 * b) * fabs(g)));
}

// limited slope (numerics/stencils.py:59-76)
__device__ inline double limited_slope(double dm, double dp, int lim) {
  double prod = dm * dp;
  switch (lim) {
    case LIM_VAN_ALBADA: return (prod * (dm + dp)) / ((dm * dm + dp * dp) + 1e-300);
    case LIM_VAN_ALBADA_TVD:
      return prod > 0.0 ? (prod * (dm + dp)) / ((dm * dm + dp * dp) + 1e-300) : 0.0;
    case LIM_VAN_LEER: return prod > 0.0 ? (2.0 * prod) / ((dm + dp) + 1e-300) : 0.0;
    case LIM_MINMOD: {
      if (!(prod > 0.0)) return 0.0;
      double sg = dm > 0.0 ? 1.0 : (dm < 0.0 ? -1.0 : 0.0);
      return sg * npmin(fabs(dm), fabs(dp));
    }
    default: return 0.0;
  }
}

// Arden-Buck saturation pressure (properties/ideal_gas_backend.py:48-51)
__device__ inline double arden_buck(double T) {
  double Tc = T - 273.15;
  return 611.21 * exp((18.678 - Tc / 234.5) * (Tc / (257.14 + Tc)));
}

// Toth isotherm (physics/isotherms.py:166-172)
__device__ inline double toth(double p, double T, double q_inf0, double chi, double b0, double dH0,
                              double tau0, double alpha, double T0, double R) {
  p = npmax(p, 0.0);
  double qinf = q_inf0 * exp(chi * (1.0 - T / T0));
  double b = b0 * exp((-dH0) / (R * T));
  double tau = tau0 + alpha * (1.0 - T0 / T);
  double bp = b * p;
  return (qinf * bp) / pow(1.0 + pow(bp, tau), 1.0 / tau);
}

// GAB with Anderson temperature dependence (physics/isotherms.py:175-183)
__device__ inline double gab(double x, double T, const vtsa_model_desc& d) {
  double E1 = d.h2o_C - exp(d.h2o_D * T);
  double E29 = d.h2o_F + d.h2o_G * T;
  double E10 = d.h2o_E10_slope * T + d.h2o_E10_intercept;
  double RT = d.R * T;
  double c = exp((E1 - E10) / RT);
  double k = exp((E29 - E10) / RT);
  double kx = k * x;
  return ((d.h2o_q_m * c) * kx) / ((1.0 - kx) * (1.0 + (c - 1.0) * kx));
}

// C1 smoothed min(x, cap) (physics/isotherms.py:186-196)
__device__ inline double smooth_cap(double x, double cap) {
  const double width = 0.02;
  x = npmax(x, -0.2);
  double lo = cap - width;
  double over = npmax(x - lo, 0.0);
  double blend = (lo + over) - (over * over) / (4.0 * width);
  return x <= lo ? x : (over >= 2.0 * width ? cap : blend);
}

__device__ inline double q_h2o(double pw, double T, double psat, const vtsa_model_desc& d) {
  return gab(smooth_cap(pw / psat, d.h2o_rh_cap), T, d);
}

// CO2 equilibrium loading with co-adsorption and trace-level regularisation
// (physics/isotherms.py:228-259)
__device__ inline double q_co2(double p, double T, double qw, const vtsa_model_desc& d) {
  double pp = npmax(p, 0.0);
  double qwc = npmax(qw, 1e-12);
  double raw;
  switch (d.coads_model) {
    case VTSA_COADS_NONE:
      raw = toth(pp, T, d.co2_q_inf0, d.co2_chi, d.co2_b0, d.co2_dH0, d.co2_tau0, d.co2_alpha, d.co2_T0, d.R);
      break;
    case VTSA_COADS_WADST: {
      double w = exp((-d.wadst_A) / qwc);
      double q_wet = toth(pp, T, d.wadst_q_inf0, d.wadst_chi, d.wadst_b0, d.wadst_dH, d.wadst_tau0,
                          d.wadst_alpha, d.co2_T0, d.R);
      double q_dry = toth(pp, T, d.co2_q_inf0, d.co2_chi, d.co2_b0, d.co2_dH0, d.co2_tau0, d.co2_alpha, d.co2_T0, d.R);
      raw = (1.0 - w) * q_dry + w * q_wet;
      break;
    }
    case VTSA_COADS_MECHANISTIC: {
      double w = exp((-d.mech_A) / qwc);
      double f_bl = d.mech_f_blocked_max * (1.0 - exp(-pow(d.mech_k * qwc, d.mech_n)));
      double phi = d.mech_phi_dry + ((1.0 - f_bl) - d.mech_phi_dry) * w;
      double dH_ave = (1.0 - w) * d.co2_dH0 + w * d.mech_dH_wet;
      raw = (phi / d.mech_phi_dry) *
            toth(pp, T, d.co2_q_inf0, d.co2_chi, d.co2_b0, dH_ave, d.co2_tau0, d.co2_alpha, d.co2_T0, d.R);
      break;
    }
    default: {   // Stampi-Bombelli
      double qi = d.co2_q_inf0 / (1.0 - d.sb_gamma * qwc);
      double b0 = d.co2_b0 * (1.0 + d.sb_beta * qwc);
      raw = toth(pp, T, qi, d.co2_chi, b0, d.co2_dH0, d.co2_tau0, d.co2_alpha, d.co2_T0, d.R);
    }
  }
  return raw * (pp / (pp + d.co2_p_reg));
}

// Fuller binary diffusivity (utils/diffusion.py:19-26) with precomputed sqrt(M_ab) and
// (Sv_a^(1/3) + Sv_b^(1/3))
__device__ inline double fuller(double T, double P, double sqrtM, double s) {
  P = npmax(P, 1.0);
  return (1.43e-7 * pow(T, 1.75)) / ((((P / 1e5) * sqrtM) * s) * s);
}

// ---------------------------------------------------------------- per-cell properties
// Thermodynamic state, gas properties, stagnant conductivity and the LDF rates of one cell
// (core/residual.py:218-238 with physics/energy.py:46-56, properties/ideal_gas_backend.py:94-137,
// physics/heat_transfer.py:23-30,72-77, physics/isotherms.py:262-269, physics/kinetics.py:29-39).
// Split in three so that the transport and equilibrium halves can run in different wavefronts:
// cell_state (temperature, pressure, mole fractions), cell_transport (mixture viscosity and
// conductivity, diffusivity, stagnant conductivity) and cell_equilibrium (q* and LDF rates).
struct CellState {
  double T, P, Tp, Cpos, y[NG], ycl[NG];
};

__device__ __forceinline__ void cell_state(const vtsa_model_desc& d, const double c[NG], const double q[NA], double e,
                                  CellState& cs) {
  double Ctot = (c[0] + c[1]) + c[2];
  double gas_cp = (d.cpg[0] * c[0] + d.cpg[1] * c[1]) + d.cpg[2] * c[2];
  double A = d.eps_t * gas_cp + d.rho_b * (d.cp_s + (d.cpa[0] * q[0] + d.cpa[1] * q[1]));
  double B = (d.eps_t * d.R) * Ctot;
  double Hq = d.rho_b * (d.dHa[0] * q[0] + d.dHa[1] * q[1]);
  double T = ((e - Hq) + A * d.T_ref) / (A - B);
  double P = (Ctot * d.R) * T;
  double Tp = npclip(T, 150.0, 1500.0);
  double Cpos = npmax(Ctot, 1e-9);
  double y[NG], ycl[NG];
  for (int s = 0; s < NG; ++s) {
    y[s] = c[s] / Cpos;
    ycl[s] = npclip(y[s], 0.0, 1.0);
  }
  cs.T = T; cs.P = P; cs.Tp = Tp; cs.Cpos = Cpos;
  for (int s = 0; s < NG; ++s) { cs.y[s] = y[s]; cs.ycl[s] = ycl[s]; }
}

__device__ __forceinline__ void cell_transport(const vtsa_model_desc& d, const CellState& cs, Props& p) {
  const double Tp = cs.Tp, Cpos = cs.Cpos, P = cs.P;
  const double* ycl = cs.ycl;
  // Wilke viscosity / Mason-Saxena conductivity
  double Tr = Tp / 300.0;
  double mu_i[NG], k_i[NG], sq[NG];
  for (int s = 0; s < NG; ++s) {
    mu_i[s] = d.mu_f1[s] * (d.mu_a1[s] * Tp + d.mu_b1[s]);
    k_i[s] = d.k_c1[s] * pow(Tr, d.k_n1[s]);
    if (d.mu_two[s]) {
      mu_i[s] = mu_i[s] + d.mu_f2[s] * (d.mu_a2[s] * Tp + d.mu_b2[s]);
      k_i[s] = k_i[s] + d.k_c2[s] * pow(Tr, d.k_n2[s]);
    }
    sq[s] = sqrt(mu_i[s]);
  }
  double mu = 0.0, kg = 0.0;
  for (int i = 0; i < NG; ++i) {
    double den = 0.0;
    for (int j = 0; j < NG; ++j) {
      if (i == j) {
        den = den + ycl[j];
      } else {
        double x = 1.0 + (sq[i] / sq[j]) * d.wilke_mq[i][j];
        den = den + ycl[j] * ((x * x) / d.wilke_den[i][j]);
      }
    }
    den = npmax(den, 1e-30);
    mu = mu + (ycl[i] * mu_i[i]) / den;
    kg = kg + (ycl[i] * k_i[i]) / den;
  }
  double Mmix = (d.M[0] * ycl[0] + d.M[1] * ycl[1]) + d.M[2] * ycl[2];
  double rho = Cpos * Mmix;
  double cpm = (d.cpg[0] * ycl[0] + d.cpg[1] * ycl[1]) + d.cpg[2] * ycl[2];
  double Pc = npmax(P, 1.0);
  double D_ca = fuller(Tp, Pc, d.fuller_sqrtM_air, d.fuller_s_air);
  double D_cw = fuller(Tp, Pc, d.fuller_sqrtM_h2o, d.fuller_s_h2o);
  double yi = ycl[2], yw = ycl[1];
  double Dm = ((yi + yw) + 1e-9) / (((yi / D_ca) + (yw / D_cw)) + (1e-9 / D_ca));
  double lam0;
  if (d.bed_model == VTSA_BED_CONSTANT) {
    lam0 = d.lam_const * 1.0;
  } else if (d.bed_model == VTSA_BED_PARALLEL) {
    lam0 = d.eps_b * kg + (1.0 - d.eps_b) * d.lam_s;
  } else {   // Zehner-Schluender
    double kap = d.lam_s / kg;
    double B_ = d.zs_B;
    double one_m = 1.0 - B_ / kap;
    double kc = (2.0 / one_m) * (((((1.0 - 1.0 / kap) * B_) / (one_m * one_m)) * log(kap / B_) -
                                  0.5 * (B_ + 1.0)) - (B_ - 1.0) / one_m);
    lam0 = kg * ((1.0 - d.zs_s) + d.zs_s * kc);
  }
  p.T = cs.T; p.P = P; p.Cpos = Cpos;
  p.y[0] = cs.y[0]; p.y[1] = cs.y[1]; p.y[2] = cs.y[2];
  p.mu = mu; p.kg = kg; p.rho = rho; p.cpm = cpm; p.Dm = Dm; p.lam0 = lam0;
}

__device__ __forceinline__ void cell_equilibrium(const vtsa_model_desc& d, const double c[NG], const double q[NA],
                                        const CellState& cs, double rate[NA], double qstar[NA]) {
  const double Tp = cs.Tp;
  double pk0 = (c[0] * d.R) * Tp, pk1 = (c[1] * d.R) * Tp;
  double psat = arden_buck(Tp);
  double qw_eq = q_h2o(pk1, Tp, psat, d);
  double qc = q_co2(pk0, Tp, d.water_basis_eq ? qw_eq : q[1], d);
  qstar[0] = qc;
  qstar[1] = qw_eq;
  for (int j = 0; j < NA; ++j) {
    double k = d.arrhenius ? d.k_ldf[j] * exp(((-d.Ea[j]) / d.R) * (1.0 / Tp - 1.0 / d.k_T_ref))
                           : d.k_ldf[j] * 1.0;
    rate[j] = k * (qstar[j] - q[j]);
  }
}

__device__ inline void cell_props(const vtsa_model_desc& d, const double c[NG], const double q[NA],
                                  double e, Props& p, double rate[NA], double qstar[NA]) {
  CellState cs;
  cell_state(d, c, q, e, cs);
  cell_transport(d, cs, p);
  cell_equilibrium(d, c, q, cs, rate, qstar);
}

__device__ inline void load_props(const double* __restrict__ pb, int n, int cell, Props& p) {
  p.T = pb[cell]; p.P = pb[n + cell]; p.Cpos = pb[2 * n + cell];
  p.y[0] = pb[3 * n + cell]; p.y[1] = pb[4 * n + cell]; p.y[2] = pb[5 * n + cell];
  p.mu = pb[6 * n + cell]; p.kg = pb[7 * n + cell]; p.rho = pb[8 * n + cell];
  p.cpm = pb[9 * n + cell]; p.Dm = pb[10 * n + cell]; p.lam0 = pb[11 * n + cell];
}

__device__ inline void store_props(double* __restrict__ pb, int n, int cell, const Props& p) {
  pb[cell] = p.T; pb[n + cell] = p.P; pb[2 * n + cell] = p.Cpos;
  pb[3 * n + cell] = p.y[0]; pb[4 * n + cell] = p.y[1]; pb[5 * n + cell] = p.y[2];
  pb[6 * n + cell] = p.mu; pb[7 * n + cell] = p.kg; pb[8 * n + cell] = p.rho;
  pb[9 * n + cell] = p.cpm; pb[10 * n + cell] = p.Dm; pb[11 * n + cell] = p.lam0;
}

// ---------------------------------------------------------------- transport closures
// axial dispersion (physics/dispersion.py:32-43)
__device__ inline double disp_axial(const vtsa_model_desc& d, double Dm, double u, double rho, double mu) {
  double v = fabs(u) / d.eps_b;
  double dp = d.dp;
  if (d.disp_model == VTSA_DISP_EDWARDS_RICHARDSON)
    return d.er_g1 * Dm + ((d.er_g2 * dp) * v) / (1.0 + ((d.er_beta * d.er_g1) * Dm) / (dp * v + 1e-30));
  if (d.disp_model == VTSA_DISP_YOUNG2021) return 0.7 * Dm + (0.5 * dp) * v;
  double Re = ((rho * fabs(u)) * dp) / mu;
  double Sc = mu / (rho * Dm);
  return (Dm / d.eps_b) * (20.0 + (0.5 * Re) * Sc);
}
// transverse dispersion with the nominal velocity (physics/dispersion.py:45-47)
__device__ inline double disp_trans(const vtsa_model_desc& d, double Dm) {
  return d.er_g1 * Dm + (d.dp * (fabs(d.u_nom) / d.eps_b)) / d.Pe_T;
}
// radial / axial effective conductivity (physics/heat_transfer.py:79-83)
__device__ inline double lam_radial(const vtsa_model_desc& d, double lam0, double C, double cp) {
  return lam0 + (((C * cp) * fabs(d.u_nom)) * d.dp) / d.Pe_r;
}
__device__ inline double lam_axial(const vtsa_model_desc& d, double lam0, double C, double cp, double u) {
  return lam0 + (((C * cp) * fabs(u)) * d.dp) / d.Pe_z;
}
// bed-side resistance of the outer ring (core/residual.py:329-335, physics/heat_transfer.py:85-97)
__device__ inline double r_bed(const vtsa_model_desc& d, double half_dr_outer, const Props& o) {
  if (d.lumped) return 1.0 / d.U_lumped;
  double lam_out = lam_radial(d, o.lam0, o.Cpos, o.cpm);
  double alpha;
  if (d.wall_model == VTSA_WALLHTC_CONSTANT) {
    alpha = d.alpha_const * 1.0;
  } else {
    double cp_mass = o.cpm / npmax(o.rho / o.Cpos, 1e-6);
    double Re = ((o.rho * fabs(d.u_nom)) * d.dp) / o.mu;
    double Pr = (cp_mass * o.mu) / o.kg;
    double nu0 = ((1.3 + 5.0 / d.N_ratio) * o.lam0) / o.kg;
    double dyn = 0.0;
    if (d.mn_dynamic) {
      double Rc = npmax(Re, 1e-12);
      double a = (0.3 * pow(Pr, 1.0 / 3.0)) * pow(Rc, 0.75);
      double b = (0.054 * Pr) * Rc;
      dyn = 1.0 / (1.0 / a + 1.0 / b);
    }
    alpha = ((nu0 + dyn) * o.kg) / d.dp;
  }
  return 1.0 / alpha + half_dr_outer / lam_out;
}

__device__ inline double T_jacket(const vtsa_model_desc& d, double t) {
  if (d.jac_tau <= 0.0) return d.T_jac_set;
  return d.T_jac_set + (d.T_jac_prev - d.T_jac_set) * exp((-t) / d.jac_tau);
}

// ---------------------------------------------------------------- end-face fluxes
// Result of one end face (boundaries/conditions.py:75-160).  F_in/H_in are INTO the bed.
struct EndFlux {
  double F_in[NG], H_in, u_out, P_face;
  double F_inp[NG], F_outp[NG], H_inp, H_outp;   // valve flux split (kind == valve only)
};

__device__ inline double cpg_dot(const vtsa_model_desc& d, const double F[NG]) {
  return (d.cpg[0] * F[0] + d.cpg[1] * F[1]) + d.cpg[2] * F[2];
}

__device__ inline void end_flux(const vtsa_model_desc& d, int end, double t, double Az, double ff,
                                const double c[NG], const Props& p, EndFlux& b) {
  const int kind = d.end_kind[end];
  for (int s = 0; s < NG; ++s) { b.F_in[s] = 0.0; b.F_inp[s] = 0.0; b.F_outp[s] = 0.0; }
  b.H_in = 0.0; b.H_inp = 0.0; b.H_outp = 0.0; b.u_out = 0.0; b.P_face = p.P;
  if (kind == END_CLOSED) return;
  const double hdz = 0.5 * d.dz;
  const double T = p.T, P = p.P;
  double a = d.ka * p.mu, bb = d.kb * p.rho;
  double Ctot = (c[0] + c[1]) + c[2];
  double C_cell = npmax(Ctot, 1e-12);
  const double* ye = d.end_y_ext[end];
  if (kind == END_FEED) {
    for (int s = 0; s < NG; ++s) b.F_in[s] = ff * ye[s];
    b.H_in = cpg_dot(d, b.F_in) * (d.end_T_ext[end] - d.T_ref);
    double u_in = ff / (Az * C_cell);
    b.P_face = P + ((a * u_in + (bb * u_in) * u_in) * hdz);
    b.u_out = -u_in;
    return;
  }
  if (kind == END_PUMP) {
    double dp = P - d.end_P_target[end];
    double w_out;
    if (d.end_check[end]) {
      dp = softplus(dp, 1.0);
      w_out = 1.0;
    } else {
      w_out = 0.5 * (1.0 + tanh(dp / VALVE_DP_SMOOTH));
    }
    double N_out = (d.end_speed[end] * dp) / (d.R * T);
    double F_out[NG];
    for (int s = 0; s < NG; ++s) {
      double yu = w_out * (c[s] / C_cell) + (1.0 - w_out) * ye[s];
      F_out[s] = (N_out * Az) * yu;
      b.F_in[s] = -F_out[s];
    }
    double T_up = w_out * T + (1.0 - w_out) * d.end_T_ext[end];
    b.H_in = -(cpg_dot(d, F_out) * (T_up - d.T_ref));
    b.u_out = N_out / C_cell;
    b.P_face = P - ((a * b.u_out + (bb * fabs(b.u_out)) * b.u_out) * hdz);
    return;
  }
  if (kind == END_VALVE) {
    double P_res = d.tau_valve > 0.0
        ? d.end_P_target[end] + (d.end_P_start[end] - d.end_P_target[end]) * exp((-t) / d.tau_valve)
        : d.end_P_target[end];
    double dp = P_res - P;
    double p_in = softplus(dp, VALVE_DP_SMOOTH);
    double p_out = softplus(-dp, VALVE_DP_SMOOTH);
    double N_in = (d.end_speed[end] * p_in) / (d.R * d.end_T_ext[end]);
    double N_out = (d.end_speed[end] * p_out) / (d.R * T);
    for (int s = 0; s < NG; ++s) {
      b.F_inp[s] = (N_in * Az) * ye[s];
      b.F_outp[s] = (N_out * Az) * (c[s] / C_cell);
      b.F_in[s] = b.F_inp[s] - b.F_outp[s];
    }
    b.H_inp = cpg_dot(d, b.F_inp) * (d.end_T_ext[end] - d.T_ref);
    b.H_outp = cpg_dot(d, b.F_outp) * (T - d.T_ref);
    b.H_in = b.H_inp - b.H_outp;
    b.u_out = (-(N_in - N_out)) / C_cell;
    b.P_face = P - ((a * b.u_out + (bb * fabs(b.u_out)) * b.u_out) * hdz);
    return;
  }
  // pressure-controlled port (Dirichlet P ramp)
  double P_face = (d.end_P_target[end] + (d.end_P_start[end] - d.end_P_target[end]) * exp((-t) / d.tau_valve)) * 1.0;
  double u = ergun_velocity((P_face - P) / hdz, a, bb);
  if (d.end_check[end]) {
    const double delta = 1e-7;
    u = 0.5 * (u + sqrt(u * u + delta * delta)) - 0.5 * delta;
  }
  double w = upwind_weight(u);
  double F_out[NG];
  for (int s = 0; s < NG; ++s) {
    double c_ext = (ye[s] * P_face) / (d.R * d.end_T_ext[end]);
    F_out[s] = (u * Az) * (w * c[s] + (1.0 - w) * c_ext);
    b.F_in[s] = -F_out[s];
  }
  double T_up = w * T + (1.0 - w) * d.end_T_ext[end];
  b.H_in = -(cpg_dot(d, F_out) * (T_up - d.T_ref));
  b.u_out = u;
  b.P_face = P_face;
}

// ---------------------------------------------------------------- stencil access
// Access to the state and cached properties of stencil cells.  The residual kernels read
// both from HBM; the finite-difference Jacobian substitutes one perturbed cell.
struct BaseAccess {
  const double* __restrict__ y;
  const double* __restrict__ pb;
  int n;
  __device__ inline void props(int cell, Props& p) const { load_props(pb, n, cell, p); }
  __device__ inline double c(int s, int cell) const { return y[s * n + cell]; }
  __device__ inline double T(int cell) const { return pb[cell]; }
};

// The perturbed cell's values are held by value (no pointer to a thread's local arrays, which
// would put them in scratch memory).
struct SwapAccess {
  const double* __restrict__ y;
  const double* __restrict__ pb;
  int n;
  int cell_swap;
  Props ps;                 // properties of cell_swap
  double cs0, cs1, cs2;     // perturbed c of cell_swap
  __device__ inline void props(int cell, Props& p) const {
    if (cell == cell_swap) p = ps; else load_props(pb, n, cell, p);
  }
  __device__ inline double c(int s, int cell) const {
    return cell == cell_swap ? (s == 0 ? cs0 : (s == 1 ? cs1 : cs2)) : y[s * n + cell];
  }
  __device__ inline double T(int cell) const { return cell == cell_swap ? ps.T : pb[cell]; }
};

// MUSCL face value on the z face k+1/2 from phi at k-1..k+2 (numerics/stencils.py:79-95)
__device__ inline double muscl_face(double pm1, double p0, double p1, double p2, bool have_m1, bool have_p2,
                                    double w, int lim, bool plain) {
  if (plain) return w * p0 + (1.0 - w) * p1;
  double sL = have_m1 ? limited_slope(p0 - pm1, p1 - p0, lim) : 0.0;
  double sR = have_p2 ? limited_slope(p1 - p0, p2 - p1, lim) : 0.0;
  return w * (p0 + 0.5 * sL) + (1.0 - w) * (p1 - 0.5 * sR);
}

// Species and enthalpy flux through the interior z face between k and k+1 (+z direction),
// core/residual.py:241-245,271-276,300-306 with physics/mass.py:34-50.
template <class Acc>
__device__ inline void z_face(const Model& m, const Acc& acc, int i, int j, int k, double F[NG], double& H,
                              double* u_out = nullptr) {
  const vtsa_model_desc& d = m.d;
  const int Nz = d.Nz;
  const int base = (i * d.Nth + j) * Nz;
  const int cL = base + k, cR = base + k + 1;
  Props L, R;
  acc.props(cL, L);
  acc.props(cR, R);
  double mu_z = 0.5 * (R.mu + L.mu), rho_z = 0.5 * (R.rho + L.rho);
  double az = d.ka * mu_z, bz = d.kb * rho_z;
  double uz = ergun_velocity((R.P - L.P) / d.dz, az, bz);
  if (u_out) *u_out = uz;
  double Cz = 0.5 * (R.Cpos + L.Cpos);
  double u_disp = m.N_nom[i * d.Nth + j] / Cz;
  double Dz = disp_axial(d, 0.5 * (R.Dm + L.Dm), u_disp, rho_z, mu_z);
  double w = upwind_weight(uz);
  const bool plain = (Nz < 3) || (d.limiter == LIM_UPWIND);
  const bool hm = k >= 1, hp = (k + 1) <= Nz - 2;
  const double Az = m.Az[i];
  double Fc[NG], Jd[NG];
  for (int s = 0; s < NG; ++s) {
    double pm1 = hm ? acc.c(s, cL - 1) : 0.0;
    double p2 = hp ? acc.c(s, cR + 1) : 0.0;
    double fv = muscl_face(pm1, acc.c(s, cL), acc.c(s, cR), p2, hm, hp, w, d.limiter, plain);
    Fc[s] = (uz * Az) * fv;
    Jd[s] = (((((-d.eps_b) * Dz) * Cz) * (R.y[s] - L.y[s])) / d.dz) * Az;
    F[s] = Fc[s] + Jd[s];
  }
  double Tm1 = hm ? acc.T(cL - 1) : 0.0;
  double Tp2 = hp ? acc.T(cR + 1) : 0.0;
  double Tf = muscl_face(Tm1, L.T, R.T, Tp2, hm, hp, w, d.limiter, plain);
  double Tavg = 0.5 * (R.T + L.T);
  double lz = lam_axial(d, 0.5 * (R.lam0 + L.lam0), Cz, 0.5 * (R.cpm + L.cpm), u_disp);
  H = (cpg_dot(d, Fc) * (Tf - d.T_ref) + cpg_dot(d, Jd) * (Tavg - d.T_ref)) - ((lz * (R.T - L.T)) / d.dz) * Az;
}

// Flux through the interior radial face between rings i and i+1 (outward),
// core/residual.py:247-251,278-283,309-314.
template <class Acc>
__device__ inline void r_face(const Model& m, const Acc& acc, int i, int j, int k, double F[NG], double& H,
                              double* u_out = nullptr) {
  const vtsa_model_desc& d = m.d;
  const int cL = (i * d.Nth + j) * d.Nz + k, cR = ((i + 1) * d.Nth + j) * d.Nz + k;
  Props L, R;
  acc.props(cL, L);
  acc.props(cR, R);
  double ar = d.ka * (0.5 * (R.mu + L.mu)), br = d.kb * (0.5 * (R.rho + L.rho));
  const double drc = m.drc[i], Ar = m.Ar[i];
  double ur = ergun_velocity((R.P - L.P) / drc, ar, br);
  if (u_out) *u_out = ur;
  double w = upwind_weight(ur);
  double Cr = 0.5 * (R.Cpos + L.Cpos);
  double Dr = disp_trans(d, 0.5 * (R.Dm + L.Dm));
  double Fc[NG], Jd[NG];
  for (int s = 0; s < NG; ++s) {
    Fc[s] = (ur * Ar) * (w * acc.c(s, cL) + (1.0 - w) * acc.c(s, cR));
    Jd[s] = (((((-d.eps_b) * Dr) * Cr) * (R.y[s] - L.y[s])) / drc) * Ar;
    F[s] = Fc[s] + Jd[s];
  }
  double Tup = w * L.T + (1.0 - w) * R.T;
  double Tavg = 0.5 * (R.T + L.T);
  double lr = lam_radial(d, 0.5 * (R.lam0 + L.lam0), Cr, 0.5 * (R.cpm + L.cpm));
  H = (cpg_dot(d, Fc) * (Tup - d.T_ref) + cpg_dot(d, Jd) * (Tavg - d.T_ref)) - ((lr * (R.T - L.T)) / drc) * Ar;
}

// Flux through the theta face between sectors j and j+1 (periodic), core/residual.py:255-259,
// 286-291,317-323.
template <class Acc>
__device__ inline void t_face(const Model& m, const Acc& acc, int i, int j, int k, double F[NG], double& H,
                              double* u_out = nullptr) {
  const vtsa_model_desc& d = m.d;
  const int jn = (j + 1 == d.Nth) ? 0 : j + 1;
  const int cL = (i * d.Nth + j) * d.Nz + k, cR = (i * d.Nth + jn) * d.Nz + k;
  Props L, R;
  acc.props(cL, L);
  acc.props(cR, R);
  double at = d.ka * (0.5 * (L.mu + R.mu)), bt = d.kb * (0.5 * (L.rho + R.rho));
  const double arc = m.arc[i], Ath = m.Ath[i];
  double ut = ergun_velocity((R.P - L.P) / arc, at, bt);
  if (u_out) *u_out = ut;
  double w = upwind_weight(ut);
  double Ct = 0.5 * (L.Cpos + R.Cpos);
  double Dt = disp_trans(d, 0.5 * (L.Dm + R.Dm));
  double Fc[NG], Jd[NG];
  for (int s = 0; s < NG; ++s) {
    Fc[s] = (ut * Ath) * (w * acc.c(s, cL) + (1.0 - w) * acc.c(s, cR));
    Jd[s] = (((((-d.eps_b) * Dt) * Ct) * (R.y[s] - L.y[s])) / arc) * Ath;
    F[s] = Fc[s] + Jd[s];
  }
  double Tup = w * L.T + (1.0 - w) * R.T;
  double Tavg = 0.5 * (L.T + R.T);
  double lt = lam_radial(d, 0.5 * (L.lam0 + R.lam0), Ct, 0.5 * (L.cpm + R.cpm));
  H = (cpg_dot(d, Fc) * (Tup - d.T_ref) + cpg_dot(d, Jd) * (Tavg - d.T_ref)) - ((lt * (R.T - L.T)) / arc) * Ath;
}

// Heat flow from the wall (or jacket) into outer-ring cell (j, k); also the jacket and
// ambient heat flows (core/residual.py:328-351).
struct WallFlows { double Q_jac, Q_amb, Q_wb; };

__device__ inline WallFlows wall_flows(const Model& m, const Props& o, double Tw, double cov, double t) {
  const vtsa_model_desc& d = m.d;
  WallFlows w;
  double Rb = r_bed(d, d.half_dr_outer, o);
  double Tj = T_jacket(d, t);
  double To = o.T;
  if (d.wall_on) {
    w.Q_jac = d.jacket_on ? ((cov * d.h_jac) * d.A_wall_out) * (Tj - Tw) : 0.0;
    w.Q_amb = (((1.0 - cov) * d.h_amb) * d.A_wall_out) * (d.T_amb - Tw);
    w.Q_wb = (d.A_wall_in * (Tw - To)) / Rb;
  } else {
    if (d.jacket_on && d.h_jac > 0.0) {
      double Rj = d.lumped ? Rb : Rb + 1.0 / d.h_jac;
      w.Q_jac = ((cov * d.A_wall_in) * (Tj - To)) / Rj;
    } else {
      w.Q_jac = 0.0;
    }
    w.Q_amb = d.h_amb > 0.0 ? (((1.0 - cov) * d.A_wall_in) * (d.T_amb - To)) / (Rb + 1.0 / d.h_amb) : 0.0;
    w.Q_wb = w.Q_jac + w.Q_amb;
  }
  return w;
}

// Conduction into wall cell (j, k) from its wall neighbours (physics/energy.py:98-113).  Entry
// `swap` of the wall field (if >= 0) reads as `swap_val` (finite-difference perturbation).
__device__ inline double wall_conduction(const vtsa_model_desc& d, const double* __restrict__ Tw, int j, int k,
                                         int swap = -1, double swap_val = 0.0) {
  const int Nz = d.Nz, Nth = d.Nth;
  auto rd = [&](int idx) { return idx == swap ? swap_val : Tw[idx]; };
  double out = 0.0;
  double t0 = rd(j * Nz + k);
  if (Nz > 1) {
    double kz = (d.lam_w * d.A_wall_axial) / d.dz;
    if (k < Nz - 1) out = out + kz * (rd(j * Nz + k + 1) - t0);
    if (k > 0) out = out - kz * (t0 - rd(j * Nz + k - 1));
  }
  if (Nth > 1) {
    double kt = ((d.lam_w * d.wall_thickness) * d.dz) / (d.R_wall_mid * d.dth);
    int jn = (j + 1 == Nth) ? 0 : j + 1;
    int jp = (j == 0) ? Nth - 1 : j - 1;
    out = out + kt * (rd(jn * Nz + k) - t0);
    out = out - kt * (t0 - rd(jp * Nz + k));
  }
  return out;
}

// ---------------------------------------------------------------- one cell's flux rows
// d(c_s)/dt (s = 0..2) and de/dt of cell (i, j, k) from the divergence of its face fluxes:
// sink, wall and exchanger sources (core/residual.py:294-298,326,352-359).  rate[] are the cell's
// LDF rates; Tw_self is the wall temperature next to the cell (used for the outer ring when the
// wall is resolved).
template <class Acc>
__device__ inline void cell_rows(const Model& m, const Acc& acc, const double div[NG + 1], double Tw_self, int i,
                                 int j, int k, double t, const double rate[NA], double out[4]) {
  const vtsa_model_desc& d = m.d;
  const int cell = (i * d.Nth + j) * d.Nz + k;
  const double V = m.V[i];
  const double eV = d.eps_t * V;
  const double rfac = d.rho_b / d.eps_t;
  out[0] = (-div[0]) / eV - rfac * rate[0];
  out[1] = (-div[1]) / eV - rfac * rate[1];
  out[2] = (-div[2]) / eV - rfac * 0.0;
  double src = 0.0;
  if (i == d.Nr - 1) {
    Props o;
    acc.props(cell, o);
    WallFlows wf = wall_flows(m, o, Tw_self, m.cov[j * d.Nz + k], t);
    src = src + wf.Q_wb;
  }
  if (d.hx_on) {
    double Tj = T_jacket(d, t);
    src = src + (d.hA_v * V) * (Tj - acc.T(cell));
  }
  out[3] = ((-div[NG]) + src) / V;
}

// The boundary flux of an end face of column (i, j) with the sign the telescoping divergence
// uses: inflow at the inlet end (k = 0, face below the cell), minus the inflow at the outlet end
// (face above cell Nz - 1).
template <class Acc>
__device__ inline void end_face(const Model& m, const Acc& acc, int end, int i, int j, double t, double F[NG],
                                double& H) {
  const vtsa_model_desc& d = m.d;
  const int cell = (i * d.Nth + j) * d.Nz + (end == 0 ? 0 : d.Nz - 1);
  Props p;
  acc.props(cell, p);
  double c[NG] = {acc.c(0, cell), acc.c(1, cell), acc.c(2, cell)};
  EndFlux b;
  end_flux(d, end, t, m.Az[i], m.feed_flux[i * d.Nth + j], c, p, b);
  for (int s = 0; s < NG; ++s) F[s] = end == 0 ? b.F_in[s] : -b.F_in[s];
  H = end == 0 ? b.H_in : -b.H_in;
}

// Telescoping divergence of the face fluxes in the upstream order (physics/mass.py:61-74), the
// faces given by face(kind, i, j, k, F, H): kind 0 the z face below cell k, 3 the z face above
// cell k, 1 the radial face between rings i and i+1, 2 the theta face between j and j+1.
template <class FaceFn>
__device__ inline void cell_divergence(const vtsa_model_desc& d, int i, int j, int k, FaceFn face,
                                       double div[NG + 1]) {
  double Fl[NG], Fu[NG], Hl, Hu;
  face(0, i, j, k, Fl, Hl);
  face(3, i, j, k, Fu, Hu);
  for (int s = 0; s < NG; ++s) div[s] = Fu[s] - Fl[s];
  div[NG] = Hu - Hl;
  if (d.Nr > 1) {
    double F[NG], H;
    if (i < d.Nr - 1) {
      face(1, i, j, k, F, H);
      for (int s = 0; s < NG; ++s) div[s] = div[s] + F[s];
      div[NG] = div[NG] + H;
    }
    if (i > 0) {
      face(1, i - 1, j, k, F, H);
      for (int s = 0; s < NG; ++s) div[s] = div[s] - F[s];
      div[NG] = div[NG] - H;
    }
  }
  if (d.Nth > 1) {
    double Fp[NG], Hp, Fm[NG], Hm;
    face(2, i, j, k, Fp, Hp);
    face(2, i, (j == 0 ? d.Nth - 1 : j - 1), k, Fm, Hm);
    for (int s = 0; s < NG; ++s) div[s] = div[s] + (Fp[s] - Fm[s]);
    div[NG] = div[NG] + (Hp - Hm);
  }
}

// One face of the divergence computed in place from the stencil (acc).
template <class Acc>
__device__ inline void any_face(const Model& m, const Acc& acc, int kind, int i, int j, int k, double t,
                                double F[NG], double& H) {
  const int Nz = m.d.Nz;
  if (kind == 0) {
    if (k == 0) end_face(m, acc, 0, i, j, t, F, H);
    else z_face(m, acc, i, j, k - 1, F, H);
  } else if (kind == 3) {
    if (k == Nz - 1) end_face(m, acc, 1, i, j, t, F, H);
    else z_face(m, acc, i, j, k, F, H);
  } else if (kind == 1) {
    r_face(m, acc, i, j, k, F, H);
  } else {
    t_face(m, acc, i, j, k, F, H);
  }
}

// The flux rows of one cell, every face computed by this thread (the finite-difference
// Jacobian's perturbed evaluations).
template <class Acc>
__device__ inline void cell_flux(const Model& m, const Acc& acc, double Tw_self, int i, int j, int k, double t,
                                 const double rate[NA], double out[4]) {
  double div[NG + 1];
  cell_divergence(m.d, i, j, k,
                  [&](int kind, int fi, int fj, int fk, double* F, double& H) {
                    any_face(m, acc, kind, fi, fj, fk, t, F, H);
                  },
                  div);
  cell_rows(m, acc, div, Tw_self, i, j, k, t, rate, out);
}

// ---------------------------------------------------------------- face-parallel flux
// The residual's faces, each computed once by one thread into a face table (SoA, 4 rows of nf:
// F_0, F_1, F_2, H), then summed per cell.  Face index space (col = i Nth + j):
//   [0, ncol (Nz - 1))                interior z faces: col (Nz - 1) + k - 1 = the face below cell k
//   [.., + ncol)                      bottom end faces, by col
//   [.., + ncol)                      top end faces, by col
//   [.., + (Nr - 1) Nth Nz)           radial faces (i < Nr - 1, j, k)
//   [.., + Nth > 1 ? n : 0)           theta faces (i, j, k)
// The end faces run the boundary-condition code; grouped apart, they share wavefronts only
// with each other instead of putting that code into every z-face wavefront.
struct FaceTable {
  int ncol, nzi, nzf, nrf, ntf;
  __device__ __host__ inline int count() const { return nzf + nrf + ntf; }
  __device__ __host__ static FaceTable of(const vtsa_model_desc& d) {
    FaceTable f;
    f.ncol = d.Nr * d.Nth;
    f.nzi = f.ncol * (d.Nz - 1);
    f.nzf = f.ncol * (d.Nz + 1);
    f.nrf = (d.Nr - 1) * d.Nth * d.Nz;
    f.ntf = d.Nth > 1 ? d.Nr * d.Nth * d.Nz : 0;
    return f;
  }
  // kind 0: the face below cell k, kind 3: the face above it, 1: outer r face, 2: theta+ face
  __device__ inline int index(const vtsa_model_desc& d, int kind, int i, int j, int k) const {
    const int col = i * d.Nth + j;
    if (kind == 0 || kind == 3) {
      const int zf = k + (kind == 3 ? 1 : 0);   // 0 .. Nz along the column
      if (zf == 0) return nzi + col;
      if (zf == d.Nz) return nzi + ncol + col;
      return col * (d.Nz - 1) + zf - 1;
    }
    if (kind == 1) return nzf + col * d.Nz + k;
    return nzf + nrf + col * d.Nz + k;
  }
};

template <class Acc>
__device__ inline void face_by_index(const Model& m, const Acc& acc, const FaceTable& ft, int f, double t,
                                     double F[NG], double& H) {
  const vtsa_model_desc& d = m.d;
  int kind, col, k;
  if (f < ft.nzi) {
    col = f / (d.Nz - 1);
    k = f - col * (d.Nz - 1) + 1;
    kind = 0;
  } else if (f < ft.nzi + ft.ncol) {
    col = f - ft.nzi;
    k = 0;
    kind = 0;
  } else if (f < ft.nzf) {
    col = f - ft.nzi - ft.ncol;
    k = d.Nz - 1;
    kind = 3;   // the face above the last cell
  } else {
    const int g = f < ft.nzf + ft.nrf ? f - ft.nzf : f - ft.nzf - ft.nrf;
    kind = f < ft.nzf + ft.nrf ? 1 : 2;
    col = g / d.Nz;
    k = g - col * d.Nz;
  }
  const int i = col / d.Nth, j = col - i * d.Nth;
  any_face(m, acc, kind, i, j, k, t, F, H);
}

// The flux rows of one cell from the face table.
template <class Acc>
__device__ inline void cell_flux_table(const Model& m, const Acc& acc, const double* __restrict__ FF,
                                       const FaceTable& ft, double Tw_self, int i, int j, int k, double t,
                                       const double rate[NA], double out[4]) {
  const int nf = ft.count();
  double div[NG + 1];
  cell_divergence(m.d, i, j, k,
                  [&](int kind, int fi, int fj, int fk, double* F, double& H) {
                    const int x = ft.index(m.d, kind, fi, fj, fk);
                    for (int s = 0; s < NG; ++s) F[s] = FF[(size_t)s * nf + x];
                    H = FF[(size_t)NG * nf + x];
                  },
                  div);
  cell_rows(m, acc, div, Tw_self, i, j, k, t, rate, out);
}

}  // namespace vtsa

// Residual, diagnostics, accumulator and finite-difference Jacobian kernels.
//
// f(t, y) is evaluated in two cell-parallel passes (one thread per cell, z fastest so that
// a wavefront reads 64 consecutive cells of a z column with unit stride):
//   k_props : thermodynamic state, gas properties, equilibrium and LDF rates of each cell
//             (transport and equilibrium in two wavefronts of one block), cached as 12 SoA
//             property arrays (96 B/cell) and the loading rows of f;
//   k_faces : every face flux once (one thread per face) from the cached properties of the
//             face's stencil, into a face table;
//   k_flux  : per cell the telescoping divergence of its 6 faces from the table, sinks and
//             wall/exchanger sources (species + energy rows), plus the wall rows (threads
//             n .. n+nw-1).
// The finite-difference Jacobian perturbs one column per thread and recomputes only the
// affected row cell (k_fd_cells / k_fd_wall): the values equal the upstream coloured forward
// differences (core/jacobian.py:71-86) because a row reads only the columns of its stencil.  The
// perturbed cell's properties are formed once per column (k_fd_props) and a row re-evaluates
// only the faces that read the perturbed cell.
#include "vtsa_internal.h"

namespace vtsa {

__device__ inline void cell_ijk(const vtsa_model_desc& d, int cell, int& i, int& j, int& k) {
  k = cell % d.Nz;
  int col = cell / d.Nz;
  j = col % d.Nth;
  i = col / d.Nth;
}

__device__ inline void load_state(const double* __restrict__ y, int n, int cell, double c[NG], double q[NA], double& e) {
  c[0] = y[cell]; c[1] = y[n + cell]; c[2] = y[2 * n + cell];
  q[0] = y[3 * n + cell]; q[1] = y[4 * n + cell];
  e = y[5 * n + cell];
}

// ------------------------------------------------------------------ residual passes
// Property passes: 64 cells per block of two wavefronts.  Both waves form the cell's state;
// wave 0 then computes the transport properties, wave 1 the equilibrium loadings and LDF rates,
// so the two transcendental chains of a cell run on different SIMDs.
constexpr int PCELLS = 64;

__global__ void __launch_bounds__(2 * PCELLS) k_props(Model m, const double* __restrict__ y, double* __restrict__ pb,
                                                      double* __restrict__ f, double* __restrict__ qstar) {
  const int cell = blockIdx.x * PCELLS + (threadIdx.x & (PCELLS - 1));
  const bool transport = threadIdx.x < PCELLS;   // uniform per wavefront
  const int n = m.n;
  if (cell >= n) return;
  double c[NG], q[NA], e;
  load_state(y, n, cell, c, q, e);
  CellState cs;
  cell_state(m.d, c, q, e, cs);
  if (transport) {
    Props p;
    cell_transport(m.d, cs, p);
    store_props(pb, n, cell, p);
  } else {
    double rate[NA], qs[NA];
    cell_equilibrium(m.d, c, q, cs, rate, qs);
    if (f) { f[3 * n + cell] = rate[0]; f[4 * n + cell] = rate[1]; }
    if (qstar) { qstar[cell] = qs[0]; qstar[n + cell] = qs[1]; }
  }
}

__device__ inline double wall_row(const Model& m, const double* __restrict__ Tw_all, const Props& o, int j, int k,
                                  double t, int swap = -1, double swap_val = 0.0) {
  const vtsa_model_desc& d = m.d;
  const int w = j * d.Nz + k;
  double Tw = (w == swap) ? swap_val : Tw_all[w];
  WallFlows wf = wall_flows(m, o, Tw, m.cov[w], t);
  double cond = wall_conduction(d, Tw_all, j, k, swap, swap_val);
  return (((wf.Q_jac + wf.Q_amb) - wf.Q_wb) + cond) / d.C_wall;
}

// Face fluxes (4 rows of nf: F_0..F_2, H).  respect_done: skip when the GMRES cycle has ended.
__global__ void __launch_bounds__(BLK) k_faces(Model m, double t, const double* __restrict__ y,
                                               const double* __restrict__ pb, double* __restrict__ FF,
                                               const GmresScalars* __restrict__ gs, int respect_done) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (respect_done && gs->done) return;
  const FaceTable ft = FaceTable::of(m.d);
  const int nf = ft.count();
  if (f >= nf) return;
  BaseAccess acc{y, pb, m.n};
  double F[NG], H;
  face_by_index(m, acc, ft, f, t, F, H);
  for (int s = 0; s < NG; ++s) FF[(size_t)s * nf + f] = F[s];
  FF[(size_t)NG * nf + f] = H;
}

__global__ void __launch_bounds__(BLK) k_flux(Model m, double t, const double* __restrict__ y,
                                              const double* __restrict__ pb, const double* __restrict__ FF,
                                              double* __restrict__ f) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = m.n;
  const vtsa_model_desc& d = m.d;
  if (tid < n) {
    int i, j, k;
    cell_ijk(d, tid, i, j, k);
    BaseAccess acc{y, pb, n};
    double rate[NA] = {f[3 * n + tid], f[4 * n + tid]};
    double Tw = (d.wall_on && i == d.Nr - 1) ? y[6 * n + j * d.Nz + k] : 0.0;
    double out[4];
    cell_flux_table(m, acc, FF, FaceTable::of(d), Tw, i, j, k, t, rate, out);
    f[tid] = out[0]; f[n + tid] = out[1]; f[2 * n + tid] = out[2]; f[5 * n + tid] = out[3];
  } else if (tid < n + m.nw) {
    const int w = tid - n;
    const int j = w / d.Nz, k = w % d.Nz;
    Props o;
    load_props(pb, n, ((d.Nr - 1) * d.Nth + j) * d.Nz + k, o);
    f[6 * n + w] = wall_row(m, y + 6 * n, o, j, k, t);
  }
}

// ------------------------------------------------------------------ matvec residual
// Finite-difference Jacobian-vector product of the Newton operator (core/solver.py:196-203):
//   v = u*s, fp = f(t, z + sigma v), w = (v - dh (fp - f)/sigma)/s.
// Pass 1 forms the perturbed state, its properties and the loading rows of w.  Blocks
// 0 .. ceil(n/64)-1 take 64 cells each in two wavefronts (transport | equilibrium, as k_props);
// the blocks after them form the perturbed wall temperatures, 128 per block.
__global__ void __launch_bounds__(2 * PCELLS) k_mv_props(Model m, const double* __restrict__ z,
                                                         const double* __restrict__ u, const double* __restrict__ s,
                                                         const double* __restrict__ f,
                                                         const GmresScalars* __restrict__ gs, int respect_done,
                                                         double dh, double* __restrict__ yp, double* __restrict__ pb,
                                                         double* __restrict__ fp, double* __restrict__ w) {
  const int n = m.n;
  const int cblocks = (n + PCELLS - 1) / PCELLS;
  if ((int)blockIdx.x < cblocks) {
    const int cell = blockIdx.x * PCELLS + (threadIdx.x & (PCELLS - 1));
    const bool transport = threadIdx.x < PCELLS;   // uniform per wavefront
    if (cell >= n) return;
    // the cell's 18 inputs in one batch of loads (not one round trip per variable)
    double zv[NV], uv[NV], sv[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const int idx = v * n + cell;
      zv[v] = z[idx];
      uv[v] = u[idx];
      sv[v] = s[idx];
    }
    if (respect_done && gs->done) return;
    const double sig = gs->sigma;
    const bool zero = gs->zero_u != 0;
    double yl[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) yl[v] = zv[v] + sig * (uv[v] * sv[v]);
    if (transport) {
#pragma unroll
      for (int v = 0; v < NV; ++v) yp[v * n + cell] = yl[v];
    }
    CellState cs;
    cell_state(m.d, yl, yl + 3, yl[5], cs);
    if (transport) {
      Props p;
      cell_transport(m.d, cs, p);
      store_props(pb, n, cell, p);
    } else {
      double rate[NA], qs[NA];
      cell_equilibrium(m.d, yl, yl + 3, cs, rate, qs);
      for (int a = 0; a < NA; ++a) {
        int idx = (3 + a) * n + cell;
        fp[idx] = rate[a];
        double vv = uv[3 + a] * sv[3 + a];
        w[idx] = zero ? 0.0 : (vv - (dh * (rate[a] - f[idx])) / sig) / sv[3 + a];
      }
    }
  } else {
    if (respect_done && gs->done) return;
    const double sig = gs->sigma;
    const int wi = (blockIdx.x - cblocks) * (2 * PCELLS) + threadIdx.x;
    if (wi < m.nw) {
      int idx = 6 * n + wi;
      yp[idx] = z[idx] + sig * (u[idx] * s[idx]);
    }
  }
}

__global__ void __launch_bounds__(BLK) k_mv_flux(Model m, double t, const double* __restrict__ yp,
                                                 const double* __restrict__ pb, const double* __restrict__ FF,
                                                 const double* __restrict__ fp,
                                                 const double* __restrict__ u, const double* __restrict__ s,
                                                 const double* __restrict__ f, const GmresScalars* __restrict__ gs,
                                                 int respect_done, double dh, double* __restrict__ w) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = m.n;
  const vtsa_model_desc& d = m.d;
  if (respect_done && gs->done) return;
  const double sig = gs->sigma;
  const bool zero = gs->zero_u != 0;
  if (tid < n) {
    int i, j, k;
    cell_ijk(d, tid, i, j, k);
    BaseAccess acc{yp, pb, n};
    double rate[NA] = {fp[3 * n + tid], fp[4 * n + tid]};
    double Tw = (d.wall_on && i == d.Nr - 1) ? yp[6 * n + j * d.Nz + k] : 0.0;
    double out[4];
    cell_flux_table(m, acc, FF, FaceTable::of(d), Tw, i, j, k, t, rate, out);
    const int rows[4] = {0, 1, 2, 5};
    for (int r = 0; r < 4; ++r) {
      int idx = rows[r] * n + tid;
      double vv = u[idx] * s[idx];
      w[idx] = zero ? 0.0 : (vv - (dh * (out[r] - f[idx])) / sig) / s[idx];
    }
  } else if (tid < n + m.nw) {
    const int wi = tid - n;
    const int j = wi / d.Nz, k = wi % d.Nz;
    Props o;
    load_props(pb, n, ((d.Nr - 1) * d.Nth + j) * d.Nz + k, o);
    double fw = wall_row(m, yp + 6 * n, o, j, k, t);
    int idx = 6 * n + wi;
    double vv = u[idx] * s[idx];
    w[idx] = zero ? 0.0 : (vv - (dh * (fw - f[idx])) / sig) / s[idx];
  }
}

// ------------------------------------------------------------------ accumulator rates
// 36 accumulator rates (core/residual.py:371-417): boundary molar and enthalpy flows split by
// direction per valve, pump and fan work, jacket / ambient / wall / exchanger heat flows.
// One block; sums in a fixed order (deterministic).
__device__ inline double block_sum(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (t < off) sh[t] += sh[t + off];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(BLK) k_acc(Model m, double t, const double* __restrict__ y,
                                             const double* __restrict__ pb, double* __restrict__ acc) {
  __shared__ double sh[BLK];
  const vtsa_model_desc& d = m.d;
  const int n = m.n, Nth = d.Nth, Nz = d.Nz, Nr = d.Nr;
  const int nf = Nr * Nth;
  double out[NACC];
  for (int a = 0; a < NACC; ++a) out[a] = 0.0;
  const double ex = (d.gam - 1.0) / d.gam;
  const double pre = (d.gam / (d.gam - 1.0)) * d.R;
  for (int end = 0; end < 2; ++end) {
    const int v = d.end_valve[end];
    if (v < 0) continue;
    double loc[10];
    for (int a = 0; a < 10; ++a) loc[a] = 0.0;
    const int kl = end == 0 ? 0 : Nz - 1;
    for (int fidx = threadIdx.x; fidx < nf; fidx += blockDim.x) {
      const int i = fidx / Nth, j = fidx % Nth;
      const int cell = (i * Nth + j) * Nz + kl;
      Props p;
      load_props(pb, n, cell, p);
      double c[NG] = {y[cell], y[n + cell], y[2 * n + cell]};
      EndFlux b;
      end_flux(d, end, t, m.Az[i], m.feed_flux[fidx], c, p, b);
      const bool inflow = b.u_out < 0.0;
      const bool split = d.end_kind[end] == END_VALVE;
      for (int s = 0; s < NG; ++s) {
        loc[s] += split ? b.F_inp[s] : (inflow ? b.F_in[s] : 0.0);
        loc[3 + s] += split ? b.F_outp[s] : (inflow ? 0.0 : -b.F_in[s]);
      }
      loc[6] += split ? b.H_inp : (inflow ? b.H_in : 0.0);
      loc[7] += split ? b.H_outp : (inflow ? 0.0 : -b.H_in);
      if (d.end_is_vacuum[end]) {
        double Fo[NG];
        for (int s = 0; s < NG; ++s) Fo[s] = inflow ? 0.0 : -b.F_in[s];
        double Ftot = d.pump_water ? (Fo[0] + Fo[1]) + Fo[2] : Fo[0] + Fo[2];
        double ratio = npmax(d.P_amb / npmax(b.P_face, 1.0), 1.0);
        double eta = d.eta_vac;
        if (d.vac_krish) {
          double pbar = b.P_face / 1e5;
          eta = npmax(((0.8 * 19.55) * pbar) / (1.0 + 19.55 * pbar), 1e-3);
        }
        loc[8] += (((pre * p.T) * Ftot) * (pow(ratio, ex) - 1.0)) / eta;
      }
      if (d.end_is_feed[end]) {
        double Ftot = split ? (b.F_inp[0] + b.F_inp[1]) + b.F_inp[2]
                            : (inflow ? (b.F_in[0] + b.F_in[1]) + b.F_in[2] : 0.0);
        double ratio = npmax(b.P_face / d.P_amb, 1.0);
        loc[9] += (((pre * d.end_T_ext[end]) * Ftot) * (pow(ratio, ex) - 1.0)) / d.eta_bl;
      }
    }
    for (int a = 0; a < 10; ++a) out[v * 10 + a] = block_sum(loc[a], sh);
  }
  // heat flows through the wall / jacket (per wall cell, or per outer bed cell without a wall)
  double qj_pos = 0.0, qj_neg = 0.0, q_amb = 0.0, q_wb = 0.0, qh_pos = 0.0, qh_neg = 0.0;
  for (int w = threadIdx.x; w < Nth * Nz; w += blockDim.x) {
    const int j = w / Nz, k = w % Nz;
    Props o;
    load_props(pb, n, ((Nr - 1) * Nth + j) * Nz + k, o);
    double Tw = d.wall_on ? y[6 * n + w] : 0.0;
    WallFlows wf = wall_flows(m, o, Tw, m.cov[w], t);
    qj_pos += npmax(wf.Q_jac, 0.0);
    qj_neg += npmin(wf.Q_jac, 0.0);
    q_amb += wf.Q_amb;
    q_wb += wf.Q_wb;
  }
  if (d.hx_on) {
    const double Tj = T_jacket(d, t);
    for (int cell = threadIdx.x; cell < n; cell += blockDim.x) {
      const int i = cell / (Nth * Nz);
      double q = (d.hA_v * m.V[i]) * (Tj - pb[cell]);
      qh_pos += npmax(q, 0.0);
      qh_neg += npmin(q, 0.0);
    }
  }
  out[30] = block_sum(qj_pos, sh);
  out[31] = block_sum(qj_neg, sh);
  out[32] = block_sum(q_amb, sh);
  out[33] = block_sum(q_wb, sh);
  out[34] = block_sum(qh_pos, sh);
  out[35] = block_sum(qh_neg, sh);
  if (threadIdx.x == 0)
    for (int a = 0; a < NACC; ++a) acc[a] = out[a];
}

// ------------------------------------------------------------------ diagnostics
__global__ void __launch_bounds__(BLK) k_diag(Model m, double t, const double* __restrict__ y,
                                              const double* __restrict__ pb, const double* __restrict__ qstar,
                                              double* __restrict__ cellout, double* __restrict__ uz,
                                              double* __restrict__ ur, double* __restrict__ uth,
                                              double* __restrict__ bnd, double* __restrict__ wall) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = m.n;
  const vtsa_model_desc& d = m.d;
  const int Nr = d.Nr, Nth = d.Nth, Nz = d.Nz;
  if (tid < n) {
    int i, j, k;
    cell_ijk(d, tid, i, j, k);
    Props p;
    load_props(pb, n, tid, p);
    cellout[tid] = p.T;
    cellout[n + tid] = p.P;
    cellout[2 * n + tid] = p.y[0];
    cellout[3 * n + tid] = p.y[1];
    cellout[4 * n + tid] = p.y[2];
    cellout[5 * n + tid] = qstar[tid];
    cellout[6 * n + tid] = qstar[n + tid];
    cellout[7 * n + tid] = d.hx_on ? (d.hA_v * m.V[i]) * (T_jacket(d, t) - p.T) : 0.0;
    BaseAccess acc{y, pb, n};
    double F[NG], H, u;
    if (k < Nz - 1) { z_face(m, acc, i, j, k, F, H, &u); uz[((i * Nth + j) * (Nz - 1)) + k] = u; }
    if (i < Nr - 1) { r_face(m, acc, i, j, k, F, H, &u); ur[tid] = u; }
    if (Nth > 1) { t_face(m, acc, i, j, k, F, H, &u); uth[tid] = u; }
    if (k == 0 || k == Nz - 1) {
      double c[NG] = {y[tid], y[n + tid], y[2 * n + tid]};
      const int nf = Nr * Nth, fidx = i * Nth + j;
      for (int end = 0; end < 2; ++end) {
        if ((end == 0 && k != 0) || (end == 1 && k != Nz - 1)) continue;
        EndFlux b;
        end_flux(d, end, t, m.Az[i], m.feed_flux[fidx], c, p, b);
        double* o = bnd + (size_t)end * 14 * nf;
        for (int s = 0; s < NG; ++s) {
          o[s * nf + fidx] = b.F_in[s];
          o[(6 + s) * nf + fidx] = b.F_inp[s];
          o[(9 + s) * nf + fidx] = b.F_outp[s];
        }
        o[3 * nf + fidx] = b.H_in;
        o[4 * nf + fidx] = b.u_out;
        o[5 * nf + fidx] = b.P_face;
        o[12 * nf + fidx] = b.H_inp;
        o[13 * nf + fidx] = b.H_outp;
      }
    }
  } else if (tid < n + Nth * Nz) {
    const int w = tid - n;
    const int j = w / Nz, k = w % Nz;
    Props o;
    load_props(pb, n, ((Nr - 1) * Nth + j) * Nz + k, o);
    double Tw = d.wall_on ? y[6 * n + w] : 0.0;
    WallFlows wf = wall_flows(m, o, Tw, m.cov[w], t);
    const int nw2 = Nth * Nz;
    wall[w] = wf.Q_jac;
    wall[nw2 + w] = wf.Q_wb;
    wall[2 * nw2 + w] = wf.Q_amb;
    wall[3 * nw2 + w] = r_bed(d, d.half_dr_outer, o);
  }
}

// ------------------------------------------------------------------ equilibrium service
__global__ void k_equilibrium(Model m, int count, const double* __restrict__ pc, const double* __restrict__ pw,
                              const double* __restrict__ T, const double* __restrict__ qw, double* __restrict__ qc,
                              double* __restrict__ qwe) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const vtsa_model_desc& d = m.d;
  double psat = arden_buck(T[i]);
  double we = q_h2o(pw[i], T[i], psat, d);
  qwe[i] = we;
  qc[i] = q_co2(pc[i], T[i], d.water_basis_eq ? we : qw[i], d);
}

// ------------------------------------------------------------------ step hand-off
// e <- energy(temperature(e)) per cell: what a hand-off through the state object does, which stores
// T = (e - Hq + A T_ref) / (A - B) and has the next step form e = A (T - T_ref) - B T + Hq again
// (physics/energy.py:46-62 in the association order of the host Thermo and of cell_state).  The
// round trip is not the identity in floating point; applying it on the device keeps a hand-off in
// device memory bit-identical to one through the host arrays.
__global__ void k_energy_roundtrip(Model m, double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.n) return;
  const vtsa_model_desc& d = m.d;
  const int n = m.n;
  const double c0 = y[i], c1 = y[n + i], c2 = y[2 * n + i], q0 = y[3 * n + i], q1 = y[4 * n + i];
  const double e = y[5 * n + i];
  const double Ctot = (c0 + c1) + c2;
  const double gas_cp = (d.cpg[0] * c0 + d.cpg[1] * c1) + d.cpg[2] * c2;
  const double A = d.eps_t * gas_cp + d.rho_b * (d.cp_s + (d.cpa[0] * q0 + d.cpa[1] * q1));
  const double B = (d.eps_t * d.R) * Ctot;
  const double Hq = d.rho_b * (d.dHa[0] * q0 + d.dHa[1] * q1);
  const double T = ((e - Hq) + A * d.T_ref) / (A - B);
  y[5 * n + i] = (A * (T - d.T_ref) - B * T) + Hq;
}

void launch_energy_roundtrip(Handle& h, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_energy_roundtrip, dim3((h.n + BLK - 1) / BLK), dim3(BLK), 0, s, h.m, y);
}

// ------------------------------------------------------------------ finite-difference Jacobian
// FD step of one column (core/jacobian.py:75-76): h = (y + 1.5e-8 max(|y|, typ)) - y
__device__ inline double fd_step(double yv, double typ) {
  double h = 1.5e-8 * npmax(fabs(yv), typ);
  return (yv + h) - yv;
}

// Stencil slot -> neighbour cell of (i, j, k), or -1 if it does not exist.
__device__ inline int slot_cell(const vtsa_model_desc& d, int i, int j, int k, int slot) {
  const int Nr = d.Nr, Nth = d.Nth, Nz = d.Nz;
  int ii = i, jj = j, kk = k;
  switch (slot) {
    case S_SELF: break;
    case S_ZM1: kk = k - 1; break;
    case S_ZP1: kk = k + 1; break;
    case S_ZM2: kk = k - 2; break;
    case S_ZP2: kk = k + 2; break;
    case S_RM1: ii = i - 1; break;
    case S_RP1: ii = i + 1; break;
    case S_TM1: if (Nth == 1) return -1; jj = (j == 0) ? Nth - 1 : j - 1; break;
    case S_TP1: if (Nth == 1) return -1; jj = (j + 1 == Nth) ? 0 : j + 1; break;
  }
  if (kk < 0 || kk >= Nz || ii < 0 || ii >= Nr) return -1;
  return (ii * Nth + jj) * Nz + kk;
}

// Values of one finite-difference column's record: the perturbed cell's Props in member order,
// then its two LDF rates.
constexpr int FDREC = 14;

// Thread (A, v): cell A with variable v moved by its finite-difference step, its properties and
// LDF rates into record Pfd[A NV + v] (each is read by the up to 9 rows whose stencil holds A).
__global__ void __launch_bounds__(BLK) k_fd_props(Model m, const double* __restrict__ y,
                                                  double* __restrict__ Pfd) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = m.n;
  if (tid >= n * NV) return;
  const vtsa_model_desc& d = m.d;
  const int v = tid % NV, A = tid / NV;
  double ya[NV];
  for (int vv = 0; vv < NV; ++vv) ya[vv] = y[vv * n + A];
  const double h = fd_step(ya[v], d.typ[v]);
  ya[v] = ya[v] + h;
  Props p;
  double rate[NA], qs[NA];
  cell_props(d, ya, ya + 3, ya[5], p, rate, qs);
  double* r = Pfd + (size_t)tid * FDREC;
  r[0] = p.T; r[1] = p.P; r[2] = p.Cpos; r[3] = p.y[0]; r[4] = p.y[1]; r[5] = p.y[2];
  r[6] = p.mu; r[7] = p.kg; r[8] = p.rho; r[9] = p.cpm; r[10] = p.Dm; r[11] = p.lam0;
  r[12] = rate[0]; r[13] = rate[1];
}

// Whether face (kind, fi, fj, fk) of a cell's divergence reads cell (ai, aj, ak): the z face
// between cells k and k+1 of a column reads its cells k-1 .. k+2, an end face its end cell, a
// radial or theta face the two cells it separates.
__device__ inline bool face_reads(const vtsa_model_desc& d, int kind, int fi, int fj, int fk, int ai, int aj,
                                  int ak) {
  if (kind == 0 || kind == 3) {
    if (ai != fi || aj != fj) return false;
    if (kind == 0 && fk == 0) return ak == 0;
    if (kind == 3 && fk == d.Nz - 1) return ak == d.Nz - 1;
    const int lo = kind == 0 ? fk - 1 : fk;   // lower cell of the face
    return ak >= lo - 1 && ak <= lo + 2;
  }
  if (ak != fk) return false;
  if (kind == 1) return aj == fj && (ai == fi || ai == fi + 1);
  const int jn = (fj + 1 == d.Nth) ? 0 : fj + 1;
  return ai == fi && (aj == fj || aj == jn);
}

// Thread (group, slot, X, v): column v of stencil cell A = slot(X); writes Jc[X][slot][0..5][v].
// Cells go in groups of FDG; a group's threads are its 9 slots one after another, FDG NV threads
// (whole wavefronts) each, so a wavefront holds one slot and the faces of X that read A, the only
// ones evaluated here, are the same for all its lanes, while the 9 slots of a group run side by
// side and share its cells' data in cache.  The other faces come from the face table of the base
// point (the values the base residual f0 was summed from), and the divergence is summed in the
// order of the residual.  A's properties and rates come from its k_fd_props record.
constexpr int FDG = 32;
static_assert((FDG * NV) % 64 == 0, "a slot's share of a cell group is whole wavefronts");
__global__ void __launch_bounds__(BLK) k_fd_cells(Model m, double t, const double* __restrict__ y,
                                                  const double* __restrict__ pb, const double* __restrict__ FF,
                                                  const double* __restrict__ Pfd, const double* __restrict__ f0,
                                                  double* __restrict__ Jc) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = m.n;
  const int grp = tid / (NSLOT * FDG * NV), r = tid - grp * (NSLOT * FDG * NV);
  const int slot = r / (FDG * NV), q = r - slot * (FDG * NV);
  const int X = grp * FDG + q / NV, v = q % NV;
  if (X >= n) return;
  const vtsa_model_desc& d = m.d;
  int i, j, k;
  cell_ijk(d, X, i, j, k);
  double* out = Jc + ((size_t)X * NSLOT + slot) * 36;
  const int A = slot_cell(d, i, j, k, slot);
  if (A < 0) {
    for (int r = 0; r < NV; ++r) out[r * NV + v] = 0.0;
    return;
  }
  const double yv = y[v * n + A];
  const double h = fd_step(yv, d.typ[v]);
  const double* rec = Pfd + ((size_t)A * NV + v) * FDREC;
  SwapAccess acc;
  acc.y = y; acc.pb = pb; acc.n = n; acc.cell_swap = A;
  Props& pa = acc.ps;
  pa.T = rec[0]; pa.P = rec[1]; pa.Cpos = rec[2]; pa.y[0] = rec[3]; pa.y[1] = rec[4]; pa.y[2] = rec[5];
  pa.mu = rec[6]; pa.kg = rec[7]; pa.rho = rec[8]; pa.cpm = rec[9]; pa.Dm = rec[10]; pa.lam0 = rec[11];
  acc.cs0 = v == 0 ? yv + h : y[A];
  acc.cs1 = v == 1 ? yv + h : y[n + A];
  acc.cs2 = v == 2 ? yv + h : y[2 * n + A];
  double rate[NA];
  if (A == X) { rate[0] = rec[12]; rate[1] = rec[13]; }
  else { rate[0] = f0[3 * n + X]; rate[1] = f0[4 * n + X]; }
  double Tw = (d.wall_on && i == d.Nr - 1) ? y[6 * n + j * d.Nz + k] : 0.0;
  int ai, aj, ak;
  cell_ijk(d, A, ai, aj, ak);
  const FaceTable ft = FaceTable::of(d);
  const int nf = ft.count();
  double div[NG + 1];
  cell_divergence(d, i, j, k,
                  [&](int kind, int fi, int fj, int fk, double* F, double& H) {
                    if (face_reads(d, kind, fi, fj, fk, ai, aj, ak)) {
                      any_face(m, acc, kind, fi, fj, fk, t, F, H);
                    } else {
                      const int x = ft.index(d, kind, fi, fj, fk);
                      for (int s = 0; s < NG; ++s) F[s] = FF[(size_t)s * nf + x];
                      H = FF[(size_t)NG * nf + x];
                    }
                  },
                  div);
  double fl[4];
  cell_rows(m, acc, div, Tw, i, j, k, t, rate, fl);
  out[0 * NV + v] = (fl[0] - f0[X]) / h;
  out[1 * NV + v] = (fl[1] - f0[n + X]) / h;
  out[2 * NV + v] = (fl[2] - f0[2 * n + X]) / h;
  out[5 * NV + v] = (fl[3] - f0[5 * n + X]) / h;
  if (A == X) {
    out[3 * NV + v] = (rec[12] - f0[3 * n + X]) / h;
    out[4 * NV + v] = (rec[13] - f0[4 * n + X]) / h;
  } else {
    out[3 * NV + v] = 0.0;
    out[4 * NV + v] = 0.0;
  }
}

// Thread (W, col): col 0..4 = wall columns (self, z-1, z+1, t-1, t+1), col 5..10 = outer-cell
// variables.  Writes Jw[W][col], Jwc[W][col-5] and, for col 0, Jcw[W][0..3].
// The outer cell's rows wrt its wall temperature change no face: their divergence is the face
// table's of the base point.  The wall row wrt an outer-cell variable takes the perturbed cell's
// properties from its k_fd_props record.
__global__ void __launch_bounds__(BLK) k_fd_wall(Model m, double t, const double* __restrict__ y,
                                                 const double* __restrict__ pb, const double* __restrict__ FF,
                                                 const double* __restrict__ Pfd, const double* __restrict__ f0,
                                                 double* __restrict__ Jcw, double* __restrict__ Jw,
                                                 double* __restrict__ Jwc) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = m.n, nw = m.nw;
  if (tid >= nw * 11) return;
  const vtsa_model_desc& d = m.d;
  const int col = tid % 11, W = tid / 11;
  const int Nz = d.Nz, Nth = d.Nth;
  const int j = W / Nz, k = W % Nz;
  const int X = ((d.Nr - 1) * Nth + j) * Nz + k;
  const double* Tw_all = y + 6 * n;
  const double f0w = f0[6 * n + W];
  if (col < 5) {
    int jj = j, kk = k;
    bool ok = true;
    if (col == 1) { kk = k - 1; ok = kk >= 0; }
    if (col == 2) { kk = k + 1; ok = kk < Nz; }
    if (col == 3) { ok = Nth > 1; jj = (j == 0) ? Nth - 1 : j - 1; }
    if (col == 4) { ok = Nth > 1; jj = (j + 1 == Nth) ? 0 : j + 1; }
    if (!ok) { Jw[W * 5 + col] = 0.0; return; }
    const int Wp = jj * Nz + kk;
    const double h = fd_step(Tw_all[Wp], d.typ[6]);
    const double tw_new = Tw_all[Wp] + h;
    Props o;
    load_props(pb, n, X, o);
    double fp = wall_row(m, Tw_all, o, j, k, t, Wp, tw_new);
    Jw[W * 5 + col] = (fp - f0w) / h;
    if (col == 0) {
      // outer cell rows wrt its wall temperature
      BaseAccess acc{y, pb, n};
      double rate[NA] = {f0[3 * n + X], f0[4 * n + X]};
      double fl[4];
      cell_flux_table(m, acc, FF, FaceTable::of(d), tw_new, d.Nr - 1, j, k, t, rate, fl);
      Jcw[W * 4 + 0] = (fl[0] - f0[X]) / h;
      Jcw[W * 4 + 1] = (fl[1] - f0[n + X]) / h;
      Jcw[W * 4 + 2] = (fl[2] - f0[2 * n + X]) / h;
      Jcw[W * 4 + 3] = (fl[3] - f0[5 * n + X]) / h;
    }
  } else {
    const int v = col - 5;
    const double h = fd_step(y[v * n + X], d.typ[v]);
    const double* rec = Pfd + ((size_t)X * NV + v) * FDREC;
    Props po;
    po.T = rec[0]; po.P = rec[1]; po.Cpos = rec[2]; po.y[0] = rec[3]; po.y[1] = rec[4]; po.y[2] = rec[5];
    po.mu = rec[6]; po.kg = rec[7]; po.rho = rec[8]; po.cpm = rec[9]; po.Dm = rec[10]; po.lam0 = rec[11];
    double fp = wall_row(m, Tw_all, po, j, k, t);
    Jwc[W * 6 + v] = (fp - f0w) / h;
  }
}

// ------------------------------------------------------------------ launchers
static inline int nblk(long long work, int b = BLK) { return (int)((work + b - 1) / b); }
// The residual passes have one thread per cell or face and too few cells to fill the device
// (8,000 at 10x16x50): one-wavefront blocks spread them over four times as many CUs.
constexpr int RB = 64;

void launch_props(Handle& h, const double* y, double* pb, double* f, double* qstar, hipStream_t s) {
  hipLaunchKernelGGL(k_props, dim3(nblk(h.n, PCELLS)), dim3(2 * PCELLS), 0, s, h.m, y, pb, f, qstar);
}

void launch_flux(Handle& h, double t, const double* y, const double* pb, double* f, hipStream_t s) {
  const int nf = FaceTable::of(h.desc).count();
  hipLaunchKernelGGL(k_faces, dim3(nblk(nf, RB)), dim3(RB), 0, s, h.m, t, y, pb, h.FF, h.gs, 0);
  hipLaunchKernelGGL(k_flux, dim3(nblk(h.n + h.nw, RB)), dim3(RB), 0, s, h.m, t, y, pb, h.FF, f);
}

void launch_acc(Handle& h, double t, const double* y, const double* pb, double* acc, hipStream_t s) {
  hipLaunchKernelGGL(k_acc, dim3(1), dim3(BLK), 0, s, h.m, t, y, pb, acc);
}

void launch_diag(Handle& h, double t, const double* y, const double* pb, double* cell, double* uz, double* ur,
                 double* uth, double* bnd, double* wall, hipStream_t s) {
  // qstar of the current point: recompute the properties with the q* side output
  double* qs = h.vtmp2;
  hipLaunchKernelGGL(k_props, dim3(nblk(h.n, PCELLS)), dim3(2 * PCELLS), 0, s, h.m, y, h.pb, nullptr, qs);
  hipLaunchKernelGGL(k_diag, dim3(nblk(h.n + h.desc.Nth * h.desc.Nz)), dim3(BLK), 0, s, h.m, t, y, h.pb, qs,
                     cell, uz, ur, uth, bnd, wall);
}

void launch_fdjac(Handle& h, double t, const double* y, const double* f0, hipStream_t s) {
  // properties at y (the base point of every column)
  launch_props(h, y, h.pb, nullptr, nullptr, s);
  // its face table (the faces a column does not reach are read from it)
  const int nf = FaceTable::of(h.desc).count();
  hipLaunchKernelGGL(k_faces, dim3(nblk(nf, RB)), dim3(RB), 0, s, h.m, t, y, h.pb, h.FF, h.gs, 0);
  hipLaunchKernelGGL(k_fd_props, dim3(nblk((long long)h.n * NV)), dim3(BLK), 0, s, h.m, y, h.Pfd);
  const long long ngrp = (h.n + FDG - 1) / FDG;
  hipLaunchKernelGGL(k_fd_cells, dim3(nblk(ngrp * NSLOT * FDG * NV)), dim3(BLK), 0, s, h.m, t, y, h.pb, h.FF,
                     h.Pfd, f0, h.Jc);
  if (h.nw > 0)
    hipLaunchKernelGGL(k_fd_wall, dim3(nblk((long long)h.nw * 11)), dim3(BLK), 0, s, h.m, t, y, h.pb, h.FF, h.Pfd, f0,
                       h.Jcw, h.Jw, h.Jwc);
}

void launch_equilibrium(Handle& h, int count, const double* pc, const double* pw, const double* T,
                        const double* qw, double* qc, double* qwe, hipStream_t s) {
  hipLaunchKernelGGL(k_equilibrium, dim3(nblk(count)), dim3(BLK), 0, s, h.m, count, pc, pw, T, qw, qc, qwe);
}

void launch_matvec_residual(Handle& h, double t, const double* z, const double* u, const double* s,
                            const double* f, double dh, double* w, bool respect_done, hipStream_t st) {
  const int rd = respect_done ? 1 : 0;
  hipLaunchKernelGGL(k_mv_props, dim3(nblk(h.n, PCELLS) + nblk(h.nw, 2 * PCELLS)), dim3(2 * PCELLS), 0, st, h.m, z, u,
                     s, f, h.gs, rd, dh, h.ytmp, h.pb2, h.ftmp, w);
  const int nf = FaceTable::of(h.desc).count();
  hipLaunchKernelGGL(k_faces, dim3(nblk(nf, RB)), dim3(RB), 0, st, h.m, t, h.ytmp, h.pb2, h.FF, h.gs, rd);
  hipLaunchKernelGGL(k_mv_flux, dim3(nblk(h.n + h.nw, RB)), dim3(RB), 0, st, h.m, t, h.ytmp, h.pb2, h.FF, h.ftmp, u,
                     s, f, h.gs, rd, dh, w);
}

// ------------------------------------------------------------------ outlet rows
// The sums one outlet row of the step runner is formed from (steps/_runner.py:104-121 forms the
// row from the full diagnostic fields on the host; here one block reduces them on the device and
// the host divides).  Per end e (0 top, 1 bottom; 11 values at 11 e): the molar flow leaving
// through all faces (Fout = -F_in), sum P_face A, then over the faces with net outflow (Fout =
// F_outpart for a valve end, bit e of valve_mask): their flow, their flow per gas (3), sum T tot;
// then over all faces sum y_k A (3) and sum T A, at the layer next to that end.  At 22, 23: sum
// V T and sum V P over the bed.  One block of 1024 threads, fixed order: deterministic.
__global__ void __launch_bounds__(1024) k_outlet_sums(Model m, const double* __restrict__ cell,
                                                      const double* __restrict__ bnd, int valve_mask,
                                                      double* __restrict__ row) {
  constexpr int NS = 24;
  __shared__ double sh[16][NS];
  const int Nr = m.d.Nr, Nth = m.d.Nth, Nz = m.d.Nz, n = m.n, nf = Nr * Nth;
  double acc[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) acc[q] = 0.0;
  for (int f = threadIdx.x; f < nf; f += blockDim.x) {
    const int i = f / Nth, j = f - i * Nth;
    const double A = m.Az[i];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const double* B = bnd + (size_t)e * 14 * nf;
      const int idx = (i * Nth + j) * Nz + (e == 0 ? 0 : Nz - 1);
      const double T = cell[idx];
      double F[NG];
#pragma unroll
      for (int k = 0; k < NG; ++k) F[k] = -B[(size_t)k * nf + f];
      double tot = (F[0] + F[1]) + F[2];
      acc[11 * e] += tot;
      if ((valve_mask >> e) & 1) {
#pragma unroll
        for (int k = 0; k < NG; ++k) F[k] = B[(size_t)(9 + k) * nf + f];
        tot = (F[0] + F[1]) + F[2];
      }
      acc[11 * e + 1] += B[(size_t)5 * nf + f] * A;
      if (tot > 0.0) {
        acc[11 * e + 2] += tot;
#pragma unroll
        for (int k = 0; k < NG; ++k) acc[11 * e + 3 + k] += F[k];
        acc[11 * e + 6] += T * tot;
      }
#pragma unroll
      for (int k = 0; k < NG; ++k) acc[11 * e + 7 + k] += cell[(size_t)(2 + k) * n + idx] * A;
      acc[11 * e + 10] += T * A;
    }
  }
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    const double V = m.V[c / (Nth * Nz)];
    acc[22] += V * cell[c];
    acc[23] += V * cell[n + c];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    double o[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) o[q] = __shfl_xor(acc[q], off, 64);
#pragma unroll
    for (int q = 0; q < NS; ++q) acc[q] += o[q];
  }
  const int wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int q = 0; q < NS; ++q) sh[wv][q] = acc[q];
  __syncthreads();
  if ((int)threadIdx.x < NS) {
    double v = 0.0;
    for (int w = 0; w < nwv; ++w) v += sh[w][threadIdx.x];
    row[threadIdx.x] = v;
  }
}

void launch_outlet_sums(Handle& h, const double* cell, const double* bnd, int valve_mask, double* row, hipStream_t s) {
  hipLaunchKernelGGL(k_outlet_sums, dim3(1), dim3(1024), 0, s, h.m, cell, bnd, valve_mask, row);
}

}  // namespace vtsa

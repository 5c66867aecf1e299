/* C ABI of the MI355X implicit step engine for the dac_vtsa TVSA bed model.
 *
 * One opaque handle per cycle step (valve states, jacket and grid fixed), as the upstream
 * TVSAModel is built once per step (core/residual.py:41-123).  All vectors are device pointers
 * (FP64) in the upstream state layout (core/residual.py:170-183):
 *     y = [c_CO2 | c_H2O | c_N2 | q_CO2 | q_H2O | e | Tw]   cell = (i*Ntheta + j)*Nz + k
 * Every call takes a HIP stream (NULL = default stream) and returns 0 on success, > 0 for a
 * numerical failure (non-finite values) and < 0 for an API or HIP error; no exception crosses
 * the ABI.  A handle is not thread-safe: calls on one handle must not overlap.
 *
 * The upstream simulator has no FFI of its own; its Python call sites are the boundary and the
 * comment of each entry point names the upstream function it replaces (path:line).
 */
#ifndef VTSA_H
#define VTSA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vtsa_handle vtsa_handle;
typedef void* vtsa_stream;   /* hipStream_t */

enum { VTSA_COADS_NONE = 0, VTSA_COADS_WADST = 1, VTSA_COADS_MECHANISTIC = 2, VTSA_COADS_SB = 3 };
enum { VTSA_DISP_EDWARDS_RICHARDSON = 0, VTSA_DISP_YOUNG2021 = 1, VTSA_DISP_WAKAO_KAGEI = 2 };
enum { VTSA_BED_ZS = 0, VTSA_BED_PARALLEL = 1, VTSA_BED_CONSTANT = 2 };
enum { VTSA_WALLHTC_MN = 0, VTSA_WALLHTC_CONSTANT = 1 };
enum { VTSA_PREC_BLOCK_JACOBI = 0, VTSA_PREC_ZLINE = 1, VTSA_PREC_ZLINE_GS = 2 };
enum { VTSA_FP64 = 0, VTSA_FP32 = 1 };

/* Everything the residual of one cycle step depends on, as scalars.  Filled by the host from
 * the validated configuration (upstream config/schema.py keys) and the step's boundary
 * conditions (valves/valve_logic.py:71-148).  Ends: 0 = top (z = 0), 1 = bottom (z = L). */
typedef struct vtsa_model_desc {
  int32_t Nr, Nth, Nz;
  int32_t wall_on, lumped, hx_on, jacket_on;
  int32_t limiter, disp_model, coads_model, water_basis_eq, bed_model, wall_model;
  int32_t arrhenius, mn_dynamic, vac_krish, pump_water;
  int32_t end_kind[2], end_check[2], end_valve[2], end_is_vacuum[2], end_is_feed[2];
  int32_t pad_;
  double end_P_target[2], end_T_ext[2], end_speed[2], end_P_start[2], end_y_ext[2][3];
  double dz, dth, A_wall_in, A_wall_out, A_wall_axial, R_wall_mid, wall_thickness, half_dr_outer, C_wall;
  double R, eps_b, eps_t, rho_b, cp_s, T_ref;
  double cpg[3], cpa[2], dHa[2], M[3];
  double mu_f1[3], mu_a1[3], mu_b1[3], mu_f2[3], mu_a2[3], mu_b2[3];
  double k_c1[3], k_n1[3], k_c2[3], k_n2[3];
  int32_t mu_two[3], pad2_;
  double wilke_mq[3][3], wilke_den[3][3];
  double fuller_sqrtM_air, fuller_s_air, fuller_sqrtM_h2o, fuller_s_h2o;
  double lam_const, lam_s, zs_B, zs_s;
  double co2_q_inf0, co2_chi, co2_b0, co2_dH0, co2_tau0, co2_alpha, co2_T0, co2_p_reg;
  double h2o_q_m, h2o_C, h2o_D, h2o_F, h2o_G, h2o_E10_slope, h2o_E10_intercept, h2o_rh_cap;
  double wadst_b0, wadst_q_inf0, wadst_tau0, wadst_chi, wadst_alpha, wadst_dH, wadst_A;
  double mech_f_blocked_max, mech_k, mech_phi_dry, mech_A, mech_dH_wet, mech_n;
  double sb_gamma, sb_beta;
  double k_ldf[2], Ea[2], k_T_ref;
  double ka, kb, dp, er_g1, er_g2, er_beta, Pe_T, Pe_r, Pe_z, N_ratio, alpha_const, U_lumped;
  double u_nom, tau_valve;
  double h_jac, h_amb, T_amb, P_amb, T_jac_set, T_jac_prev, jac_tau, hA_v, lam_w;
  double gam, eta_vac, eta_bl;
  double typ[7];          /* FD perturbation floors per variable (core/jacobian.py:40-48) */
  double atol[7];         /* absolute tolerances per variable (core/residual.py:135-144) */
} vtsa_model_desc;

/* GMRES outcome (scipy.sparse.linalg.gmres semantics as called at core/solver.py:223-224). */
typedef struct vtsa_gmres_info {
  int32_t info;          /* 0 converged, 1 not converged within one restart cycle */
  int32_t iters;         /* Arnoldi iterations (the upstream callback count) */
  int32_t breakdown;
  int32_t pad_;
  double bnorm, presid, rnorm, atol_eff;
} vtsa_gmres_info;

/* Build / free the device model of one step.  Replaces TVSAModel.__init__
 * (core/residual.py:41-123).  Ring arrays have Nr entries (Ar, drc: Nr-1); feed_flux and N_nom
 * have Nr*Nth entries; cov has Nth*Nz.  All are host pointers and are copied. */
int vtsa_create(const vtsa_model_desc* desc, const double* V, const double* Az, const double* Ath,
                const double* arc, const double* Ar, const double* drc, const double* feed_flux,
                const double* N_nom, const double* cov, vtsa_handle** out);
int vtsa_destroy(vtsa_handle* h);
/* Re-initialise a handle for the next step of a cycle on the same grid (same Nr, Nth, Nz and
 * wall; -2 otherwise): the descriptor and the nine arrays of vtsa_create are replaced, the
 * preconditioner factorisation and low-rank correction are dropped, the engine options and the
 * sampled timing return to their defaults, and every device buffer is kept.  The copies are queued on stream s behind the handle's earlier work and waited for.
 * Replaces the per-step TVSAModel construction of vtsa_cycle.py:80-86 / core/residual.py:41-123
 * when the steps of a cycle hand the bed on in device memory. */
int vtsa_reset(vtsa_handle* h, const vtsa_model_desc* desc, const double* V, const double* Az,
               const double* Ath, const double* arc, const double* Ar, const double* drc,
               const double* feed_flux, const double* N_nom, const double* cov, vtsa_stream s);
int vtsa_sizes(const vtsa_handle* h, int32_t* n_cells, int32_t* n_wall, int32_t* n_unknowns);

/* f = dy/dt at (t, y); acc (36 accumulator rates, device) when non-NULL.
 * Replaces TVSAModel.rhs (core/residual.py:211-417). */
int vtsa_rhs(vtsa_handle* h, double t, const double* y, double* f, double* acc, vtsa_stream s);

/* R = y - gamma_h f(t, y) - rhs_known.  Replaces TVSAModel.stage_residual (core/residual.py:420-423). */
int vtsa_stage_residual(vtsa_handle* h, double t, const double* y, double gamma_h, const double* rhs_known,
                        double* R, vtsa_stream s);

/* Diagnostic fields of rhs(diag=True) (core/residual.py:366-368), device outputs:
 *  cell[8*n]  = T, P, y_CO2, y_H2O, y_N2, q*_CO2, q*_H2O, Q_hx
 *  uz[Nr*Nth*(Nz-1)], ur[(Nr-1)*Nth*Nz], uth[Nr*Nth*Nz]
 *  bnd[2][14][Nr*Nth] = per end: F_in(3), H_in, u_out, P_face, F_inpart(3), F_outpart(3), H_inpart, H_outpart
 *  wall[4][Nth*Nz]    = Q_jac, Q_wb, Q_amb, R_bed */
int vtsa_diag(vtsa_handle* h, double t, const double* y, double* cell, double* uz, double* ur, double* uth,
              double* bnd, double* wall, vtsa_stream s);

/* Equilibrium loadings for arbitrary points (isotherm service; physics/isotherms.py:262-269).
 * q_h2o_state is used as the CO2 water basis when the configuration asks for it. */
int vtsa_equilibrium(vtsa_handle* h, int32_t count, const double* p_co2, const double* p_h2o, const double* T,
                     const double* q_h2o_state, double* q_co2_eq, double* q_h2o_eq, vtsa_stream s);

/* Finite-difference Jacobian of f at (t, y) given f0 = f(t, y), into handle-owned block storage:
 *  Jc [n][9][6][6]  cell rows x (self, z-1, z+1, z-2, z+2, r-1, r+1, t-1, t+1) cell columns
 *  Jcw[Nth*Nz][4]   outer-ring flux rows (c_CO2, c_H2O, c_N2, e) x own wall column
 *  Jw [Nth*Nz][5]   wall rows x wall columns (self, z-1, z+1, t-1, t+1)
 *  Jwc[Nth*Nz][6]   wall rows x outer-cell columns
 * Entry values equal the upstream coloured forward differences (core/jacobian.py:71-86) with the
 * same steps h = (y + 1.5e-8 max(|y|, typ)) - y; all columns are evaluated in one launch. */
int vtsa_fd_jacobian(vtsa_handle* h, double t, const double* y, const double* f0, vtsa_stream s);
/* Copy the block Jacobian to host buffers (any may be NULL). */
int vtsa_jacobian_get(vtsa_handle* h, double* Jc, double* Jcw, double* Jw, double* Jwc);

/* Factor the preconditioner M ~ I - dh J in the variables scaled by `scale` (device, length N).
 * Replaces TRBDF2._factor / splu (core/solver.py:334-337).  kind: VTSA_PREC_BLOCK_JACOBI (6x6
 * cell blocks), VTSA_PREC_ZLINE (block LU of every z-line) or VTSA_PREC_ZLINE_GS (line
 * Gauss-Seidel over the z-lines of the cells and the wall lines, coloured so that lines of one
 * colour do not couple; `sweeps` symmetric sweeps, 1..16, ignored by the other kinds);
 * precision: VTSA_FP64 or VTSA_FP32 (factors stored and applied in FP32, vectors stay FP64). */
int vtsa_precond_setup(vtsa_handle* h, double dh, const double* scale, int32_t kind, int32_t precision,
                       int32_t sweeps, vtsa_stream s);
/* out = M^-1 r (unscaled); replaces lu.solve (core/solver.py:182,209,460).  out may alias r. */
int vtsa_precond_apply(vtsa_handle* h, const double* r, double* out, vtsa_stream s);

/* Right-preconditioned GMRES(restart), one cycle, x0 = 0, on the scaled Newton operator
 *   A u = (v - dh (f(t, z + sigma v) - f)/sigma)/s,  v = u*s,  sigma = sqrt(eps) zs_norm / rms(u)
 *   P^-1 w = M^-1(w*s)/s
 * solving (A P^-1) w = b with tolerance max(atol, rtol |b|) (core/solver.py:196-224).
 * The step uses max(1, zs_norm); zs_norm < 0: max(1, rms(z/s)) is formed on the device instead
 * (no wait before the cycle).
 * On return w holds the GMRES solution (in the preconditioned variable, as upstream) and, when pw
 * is not NULL, pw = M^-1 (w*s): the Newton step dz that upstream forms from it with one more solve
 * (u = prec(w), dz = u*s: core/solver.py:240-241), taken from the apply of the cycle's closing
 * residual, which solves the same system (pw may not alias w). */
int vtsa_gmres(vtsa_handle* h, double t, const double* z, const double* f, const double* s, double dh,
               double zs_norm, const double* b, double* w, double* pw, double rtol, double atol,
               int32_t restart, vtsa_gmres_info* info, vtsa_stream s_);

/* In place on a state vector y (device, N values): every cell's energy e replaced by
 * energy(temperature(e)), the round trip a hand-off through CycleState makes (it stores T, the
 * next step re-forms e: physics/energy.py:46-62 as called by core/residual.py:175-196 to_state /
 * pack).  Lets the next step start from the device vector with the same bits as from the host
 * state.  Concentrations, loadings and wall temperatures are untouched. */
int vtsa_state_roundtrip(vtsa_handle* h, double* y, vtsa_stream s);

/* ---- Newton vector operations (core/solver.py:84-108,169-259); scalars returned on the host. */
/* s = atol + rtol |z| */
int vtsa_scale(vtsa_handle* h, const double* z, double rtol, double* s, vtsa_stream st);
/* out[0] = sqrt(mean((a/s)^2)) (s NULL: plain rms) */
int vtsa_rms(vtsa_handle* h, const double* a, const double* s, int32_t len, double* out, vtsa_stream st);
/* Low-rank correction of the GMRES right preconditioner (no upstream counterpart): GMRES then
 * solves with P^-1 Q, Q = I + U D U^T in its scaled variables, U = V_m G (m x k, row-major, k <= 16)
 * built on the device from the basis of the last GMRES cycle, D (k x k row-major).  k = 0 removes
 * it; every vtsa_precond_setup removes it too.  vtsa_gmres_hessenberg copies the last cycle's
 * Arnoldi Hessenberg as computed ((MAX 65) x 64 doubles, row-major, before Givens rotations) and
 * its column count m (host memory, no device access).  vtsa_defl_apply: out = Q in. */
int vtsa_gmres_hessenberg(vtsa_handle* h, double* H, int32_t* m);
int vtsa_defl_set(vtsa_handle* h, int32_t k, int32_t m, const double* G, const double* D, vtsa_stream st);
int vtsa_defl_apply(vtsa_handle* h, const double* in, double* out, vtsa_stream st);
/* No counterpart in the upstream code; for callers that cannot allocate device memory during
 * the call (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * *out (a DEVICE double) = sqrt(mean((a/s)^2)) over len entries (s NULL: plain rms), the same
 * bits as vtsa_rms.  No handle, no hipMalloc/hipFree: the reduction's per-block partials and
 * ticket counter live in a fixed file-scope device array of 4.3 KB per GPU, held until the process
 * ends.  len must be in 1 .. VTSA_NOALLOC_MAX_LEN (else -2).  Asynchronous on st; the first call
 * on a device looks the array up and clears it (do that call before any capture).  Calls that
 * overlap on one GPU (two streams, or two threads) race on the shared array. */
#define VTSA_NOALLOC_MAX_LEN (1 << 30)
int vtsa_rms_noalloc(const double* a, const double* s, int32_t len, double* out, vtsa_stream st);
/* f = f(t,z); R = z - dh f - rhs_known; rn = rms(R/s); pn = rms(M^-1 R / s) */
int vtsa_newton_eval(vtsa_handle* h, double t, const double* z, double dh, const double* rhs_known,
                     const double* s, double* f, double* R, double* rn, double* pn, vtsa_stream st);
/* The residual half of vtsa_newton_eval: f = f(t,z); R = z - dh f - rhs_known; with rn non-NULL
 * rn = rms(R/s) on the host (synchronises the stream), with rn NULL nothing is waited for.  No
 * preconditioner apply.  Used where pn is not needed (the conservation projection's residuals). */
int vtsa_newton_residual(vtsa_handle* h, double t, const double* z, double dh, const double* rhs_known,
                         const double* s, double* f, double* R, double* rn, vtsa_stream st);
/* z_try = truncate(z + lam dz, z); f_try, R_try at z_try; rn, pn as above; step = rms((z_try-z)/s),
 * or NaN when dz holds a non-finite entry (the test of core/solver.py:241-242, with no wait of its
 * own; finite z and s make step NaN in no other case). */
int vtsa_newton_try(vtsa_handle* h, double t, const double* z, const double* dz, double lam, double dh,
                    const double* rhs_known, const double* s, double* z_try, double* f_try, double* R_try,
                    double* rn, double* pn, double* step, vtsa_stream st);
/* conservation test |W R| <= tol W s for every conserved quantity (core/solver.py:88-96);
 * wr[4] = W R and ws[4] = W s returned on the host. */
int vtsa_cons(vtsa_handle* h, const double* R, const double* s, double* wr, double* ws, vtsa_stream st);
/* Conservation projection (no upstream counterpart; optional in the Newton stage solve):
 * out[k] = sum over the entries of conserved row k of (W_ki s_i)^2, k = CO2, H2O, N2, energy. */
int vtsa_cons_norm2(vtsa_handle* h, const double* s, double* out, vtsa_stream st);
/* out = z - S^2 W^T lambda (lambda: 4 host values), the minimum-norm shift in the s-scaled
 * variables that changes the conserved totals W z by -W S^2 W^T lambda.  out may alias z.
 * mode bits: VTSA_SHIFT_TRUNCATE: a non-negative concentration or loading that would fall below
 * -s (one tolerance unit under zero) keeps 0.1 of its value instead, as a Newton update does
 * (core/solver.py:99-108); with VTSA_SHIFT_AT_ZERO as well, any that would fall below zero.
 * Returns 1 when such an entry exists (truncated or not) and synchronises the stream, except with
 * VTSA_SHIFT_NO_REPORT: returns 0 without waiting. */
enum { VTSA_SHIFT_TRUNCATE = 1, VTSA_SHIFT_AT_ZERO = 2, VTSA_SHIFT_NO_REPORT = 4 };
int vtsa_cons_shift(vtsa_handle* h, const double* z, const double* s, const double* lambda, int32_t mode,
                    double* out, vtsa_stream st);
/* As vtsa_cons, with W R and W s kept on the device in slot 0..3 (nothing is waited for);
 * vtsa_cons_fetch returns slots 0..n-1 on the host, wr[4 n] and ws[4 n] (synchronises). */
int vtsa_cons_stash(vtsa_handle* h, const double* R, const double* s, int32_t slot, vtsa_stream st);
int vtsa_cons_fetch(vtsa_handle* h, int32_t n, double* wr, double* ws, vtsa_stream st);
/* One outlet row of the step record, reduced on the device into row[32] (device memory, no
 * host wait), replacing the host reductions of steps/_runner.py:104-121 over the full
 * diagnostic fields: per end (top at 0, bottom at 11) the flow leaving through all faces, sum
 * P_face A, over the faces with net outflow (the valve's outflow part for an end whose bit is set
 * in valve_mask) their flow, their flow per gas (3) and sum T flow, over all faces sum y_k A (3)
 * and sum T A; at 22, 23 sum V T and sum V P; at 24 .. 30 the inventories of vtsa_inventories. */
int vtsa_outlet_row(vtsa_handle* h, double t, const double* y, int32_t valve_mask, double* row, vtsa_stream st);
/* inventories: out[0..2] gas moles, out[3..4] adsorbed moles, out[5] E_bed, out[6] E_wall
 * (core/residual.py:426-443) */
int vtsa_inventories(vtsa_handle* h, const double* y, double* out, vtsa_stream st);

/* Integrator vector formulas (core/solver.py:419-476), upstream operation order; npos entries
 * (c and q) are truncated for ops 1 and 2 (core/solver.py:99-108):
 *  0: out = x + a y            (y + dh f_n)
 *  1: out = trunc(x + a y, x)  (stage predictor y + (c h) slope)
 *  2: out = trunc(x + (y - x)/a, x)   (stage-2 predictor)
 *  3: out = x + a ((y - x)/b)  (SDIRK2 stage-2 right-hand side)
 *  4: out = a y - b x          (TR-BDF2 stage-2 right-hand side)
 *  5: out = x + a ((b y + c z) + e w)  (third-order quadrature y3)
 *  6: out = (y - x)/a          (slope)
 *  7: out = y - x
 *  8: out = (-x)/y             (GMRES right-hand side -R/s)
 *  9: out = x y */
int vtsa_vec_op(vtsa_handle* h, int32_t op, double a, double b, double c, double e, const double* x, const double* y,
                const double* z, const double* w, double* out, vtsa_stream st);
/* en = rms(M^-1 (z2 - y3) / (atol + rtol max(|y|, |z2|))), the filtered local error
 * (core/solver.py:460-463). */
int vtsa_error_norm(vtsa_handle* h, const double* y, const double* z2, const double* y3, double rtol, double* en,
                    vtsa_stream st);

/* ---- Device memory, copies, streams and timing (plumbing for callers without a HIP runtime
 * of their own; no upstream counterpart).  Copies to the host are synchronous. */
int vtsa_device_count(int32_t* count);
int vtsa_device_alloc(size_t bytes, void** out);
int vtsa_device_free(void* ptr);
int vtsa_copy_h2d(void* dst, const void* src, size_t bytes, vtsa_stream st);
int vtsa_copy_d2h(void* dst, const void* src, size_t bytes, vtsa_stream st);
int vtsa_copy_d2d(void* dst, const void* src, size_t bytes, vtsa_stream st);
int vtsa_fill(double* dst, double value, size_t count, vtsa_stream st);
int vtsa_sync(vtsa_stream st);
/* A stream of its own for one case (default priority, not synchronised with the null stream),
 * so that independent cases run concurrently on one GPU.  Handles may be used from any stream;
 * calls on one handle must stay on one stream. */
int vtsa_stream_create(vtsa_stream* out);
int vtsa_stream_destroy(vtsa_stream st);
/* Elapsed milliseconds of `repeat` back-to-back residual evaluations f = f(t, y) on stream st,
 * measured with HIP events recorded on that stream (kernel timing for the roofline). */
int vtsa_time_rhs(vtsa_handle* h, double t, const double* y, double* f, int32_t repeat, double* ms, vtsa_stream st);
/* The same for `repeat` finite-difference Jacobian evaluations at (t, y) with f0 = f(t, y) (all
 * colours' perturbed residuals in one launch, then the wall columns). */
int vtsa_time_fdjac(vtsa_handle* h, double t, const double* y, const double* f0, int32_t repeat, double* ms,
                    vtsa_stream st);

/* Engine options (no upstream counterpart).  VTSA_OPT_GS_WHOLE_LINES, how a line Gauss-Seidel
 * half-sweep stages its line's factor in LDS: 1 (default) automatic -- the first of the forms
 * below whose blocks the CUs hold all at once, else the first the device grants at all; 3 the
 * whole pair factor, 2 the two recurrence matrices of each pair (the rest read from global
 * memory), 4 the forward matrix, then the backward one in its place, 5 with FP32 factors the
 * same with both matrices widened to FP64 in LDS and the off-line products summed per neighbour
 * first, with FP64 factors as 4 (each when granted); 0 cell factors in segments during the sweep.
 * Automatic tries 5 first for FP32 factors.  Levels 1-4 give bit-identical results, 5 with FP32
 * factors the same recurrences on a right-hand side summed in another order; level 0 solves the
 * same line systems cell by cell (other roundings).  VTSA_OPT_GS_FORWARD: 0 (default) symmetric
 * sweeps (colours 0 1 0 1 0 for two sweeps of two colours), 1 forward sweeps only (0 1 0 1).
 * VTSA_OPT_GS_TWIST: 1 (default) lines of an even number of cells (at least 6) are solved by a
 * twisted factorisation, top half from the first cell down and bottom half from the last cell up,
 * meeting in a middle pair (half as many dependent steps; other roundings), 0 the classic top-down
 * factorisation; read by the next preconditioner set-up, which keeps the classic form when the
 * staging level is 0 or no pair form is granted.  A twisted set-up keeps the cell factors only
 * down to the twist pair: staging levels 1-5 run the twisted form, and level 0 set after it runs
 * the automatic pair form. */
enum { VTSA_OPT_GS_WHOLE_LINES = 1, VTSA_OPT_GS_FORWARD = 2, VTSA_OPT_GS_TWIST = 3 };
int vtsa_set_option(vtsa_handle* h, int32_t option, int32_t value);

/* Sampled timing of the line Gauss-Seidel half-sweep kernel, the dominant kernel of a step (no
 * upstream counterpart: instrumentation for the roofline).  Every `every`-th launch is bracketed
 * by HIP events on its stream (0 disables); the counters are reset. */
int vtsa_profile(vtsa_handle* h, int32_t every);
/* Sampled launches so far, their total milliseconds (events) and algorithmic bytes. */
int vtsa_profile_read(vtsa_handle* h, int64_t* launches, double* total_ms, double* total_bytes);

/* ---- Learned stage guess (no upstream counterpart: the upstream stage predictor is the linear
 * extrapolation of core/solver.py:420,432-435, which Newton-Krylov then corrects).
 *
 * A per-cell network, 32 inputs -> 128 -> 128 -> 128 -> 16 outputs (6 used), ReLU, BF16
 * weights and activations on the matrix cores with FP32 accumulation, FP32 master weights and
 * Adam moments (beta1 0.9, beta2 0.999, eps 1e-8).  It predicts the converged stage value
 * relative to the predictor, (z - z_pred) / s with s = atol + rtol |y|, for the six cell
 * variables; the Newton-Krylov solve still runs to the usual tolerance from the guess.
 * Parameters (FP32, flat, W[out][in]): W1[128][32] b1[128] W2[128][128] b2[128] W3[128][128]
 * b3[128] W4[16][128] b4[16].  Normalisation (FP32[96]): input mean[32], input 1/std[32],
 * target mean[16], target std[16].  Feature rows and targets are device arrays (row-major,
 * 32 and 8 floats per row); `rows` is at most the capacity given at creation. */
typedef struct vtsa_mlp vtsa_mlp;
int vtsa_mlp_sizes(int32_t* n_param, int32_t* n_in, int32_t* n_hidden, int32_t* n_out, int32_t* n_norm);
int vtsa_mlp_create(int32_t max_rows, vtsa_mlp** out);
int vtsa_mlp_destroy(vtsa_mlp* m);
/* Host parameters in; resets the Adam moments and step counter. */
int vtsa_mlp_set_params(vtsa_mlp* m, const float* params, vtsa_stream st);
int vtsa_mlp_get_params(vtsa_mlp* m, float* params, vtsa_stream st);
int vtsa_mlp_set_norm(vtsa_mlp* m, const float* norm, vtsa_stream st);
/* out[rows][16] = network(feat[rows][32]) (normalised target units). */
int vtsa_mlp_forward(vtsa_mlp* m, const float* feat, int32_t rows, float* out, vtsa_stream st);
/* One training step on one minibatch: forward, mean-squared error over the six used outputs
 * against the normalised targets, backward, and an Adam step of size lr (lr = 0: gradient only).
 * loss (host, may be NULL): the minibatch loss before the step; grad (host, may be NULL): the
 * gradient in the parameter layout.  Either output synchronises the stream. */
int vtsa_mlp_train_step(vtsa_mlp* m, const float* feat, const float* target, int32_t rows, float lr, float* loss,
                        float* grad, vtsa_stream st);
/* Feature rows of one stage for the n cells of a handle: the cell's six variables and those of
 * its z-neighbours at the step start y, (z_pred - y)/s, r/s, log(h) and the stage index; the
 * integrator passes r = M^-1 R(z_pred), the preconditioned stage residual at the predictor. */
int vtsa_mlp_features(vtsa_handle* h, const double* y, const double* z_pred, const double* s, const double* r,
                      double log_h, double stage, float* feat, vtsa_stream st);
/* Training targets (z_converged - z_base)/s of the six cell variables (8 floats per cell), with
 * z_base = z_pred - r, or z_base = z_pred when r is NULL. */
int vtsa_mlp_targets(vtsa_handle* h, const double* z_conv, const double* z_pred, const double* r, const double* s,
                     float* target, vtsa_stream st);
/* Stage guess z = z_base + s (out std + mean) on the cell variables (c and q truncated to at least
 * 0.1 z_pred as core/solver.py:99-108 does); wall entries = z_base. */
int vtsa_mlp_apply(vtsa_mlp* m, vtsa_handle* h, const float* out, const double* z_pred, const double* r,
                   const double* s, double* z, vtsa_stream st);

/* Library version string and last error message (host memory owned by the library). */
const char* vtsa_version(void);
const char* vtsa_last_error(void);

#ifdef __cplusplus
}
#endif
#endif

// Internal declarations shared by the translation units of libvtsa (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "vtsa_model.h"

namespace vtsa {

constexpr int NACC = 36;
constexpr int BLK = 256;          // default threads per block for cell-parallel kernels
constexpr int MAX_RESTART = 64;
constexpr int DEFL_MAX = 16;      // columns of the low-rank correction of the GMRES preconditioner

// Device-resident scalars of one GMRES cycle and of the reductions around it.
struct GmresScalars {
  double H[MAX_RESTART + 1][MAX_RESTART];   // H[row][col] (Hessenberg after Givens)
  double Hr[MAX_RESTART + 1][MAX_RESTART];  // the Arnoldi Hessenberg as computed (before Givens)
  double cs[MAX_RESTART], sn[MAX_RESTART], S[MAX_RESTART + 1], yv[MAX_RESTART];
  double h_tmp[MAX_RESTART + 1];            // Gram-Schmidt coefficients of the current column
  double h0sq;                              // |w|^2 before orthogonalisation
  double zs_rms;                            // rms(z/s) of the cycle (the FD step uses max(1, .))
  double sigma;                             // FD step of the current matvec
  double unorm;                             // rms(u) of the current matvec
  double inv_h1;                            // 1 / |w| after orthogonalisation
  double presid, ptol, bnorm, atol_eff, rnorm;
  int done, iters, breakdown, zero_u;
};

// Reduction scratch: per-block partial sums and a completion counter (last-block pattern).
struct Reduce {
  double* partial = nullptr;   // [max_blocks * max_k]
  unsigned* counter = nullptr; // [8]
  double* result = nullptr;    // [64] device results
  int max_blocks = 0, max_k = 0;
};

struct Handle {
  Model m{};
  vtsa_model_desc desc{};
  int n = 0, nw = 0, N = 0, ncol = 0;
  // device model arrays
  double *V = nullptr, *Az = nullptr, *Ath = nullptr, *arc = nullptr, *Ar = nullptr, *drc = nullptr;
  double *feed_flux = nullptr, *N_nom = nullptr, *cov = nullptr;
  // work buffers
  double* pb = nullptr;      // properties at the last residual point [12n]
  double* pb2 = nullptr;     // properties of a second point (matvec) [12n]
  double* FF = nullptr;      // face fluxes of the last residual pass [4][FaceTable::count]
  double* ytmp = nullptr;    // perturbed state [N]
  double* ftmp = nullptr;    // residual at perturbed state [N]
  double* vtmp = nullptr;    // scratch [N]
  double* vtmp2 = nullptr;   // scratch [N]
  double* accbuf = nullptr;  // [NACC]
  // Jacobian blocks
  double *Jc = nullptr, *Jcw = nullptr, *Jw = nullptr, *Jwc = nullptr;
  double* Pfd = nullptr;      // [n][NV][14] per finite-difference column: the perturbed cell's Props, LDF rates
  double* Jo = nullptr;       // [144][n] off-line blocks (r-1, r+1, theta-1, theta+1), cell fastest
  // preconditioner factors (FP64 or FP32), scale at factorisation
  void* F = nullptr;          // [n][5][36] blocks (L2, L1, Dinv, U1, U2) in the chosen precision
  void* P = nullptr;          // line Gauss-Seidel pair form of F, gs_pairs_count values
  void* Fr = nullptr;         // as F for each z-line factored from its last cell upwards (twisted solve)
  double* Fw = nullptr;       // [nw] wall diagonal of M
  double* Tw = nullptr;       // [nw][3] wall-line tridiagonal LU: lo, 1 / den, c (line Gauss-Seidel)
  double* s0 = nullptr;       // [N] scale of the factorisation
  int prec_kind = -1, prec_prec = 0, prec_ready = 0;
  double prec_dh = 0.0;
  // line Gauss-Seidel: lines grouped by colour (cell z-lines 0..ncol-1, wall lines after)
  int* gs_order = nullptr;           // device [ncol + wall lines]
  std::vector<int> gs_cstart;        // colour c owns gs_order[gs_cstart[c] .. gs_cstart[c + 1])
  int gs_sweeps = 1;
  int gs_full = 1;                   // LDS staging of a half-sweep: 0 segments, 1 whole line
  int gs_forward = 0;                // line Gauss-Seidel order: 0 symmetric sweeps, 1 forward sweeps only
  int gs_twist = 1;                  // twisted pair solve when the line allows it (option, read at set-up)
  int gs_tw_m = -1;                  // twist pair of the last set-up's pair form, -1: classic pair form
  int lds_max = 65536;               // LDS bytes one block may use on this device
  int lds_per_cu = 65536;            // LDS bytes of one CU
  int n_cu = 256;                    // compute units of the device
  double* gsb = nullptr;             // [N] scaled right-hand side
  double* gsu = nullptr;             // [N] unscaled iterate
  std::vector<int> gs_ncell, gs_nouter, gs_nwall;   // per colour: cell lines, outer-ring lines, wall lines
  // sampled timing of the half-sweep launches: every prof_every-th launch between HIP events
  int prof_every = 0;
  long long prof_seen = 0, prof_launches = 0;
  double prof_ms = 0.0, prof_bytes = 0.0;
  std::vector<hipEvent_t> prof_pending, prof_pool;  // (start, stop) pairs awaiting read; spare events
  std::vector<double> prof_pending_bytes;
  // GMRES
  double* Vb = nullptr;       // basis [(MAX_RESTART+1) * N]
  int vb_cols = 0;
  GmresScalars* gs = nullptr; // device
  GmresScalars* gs_host = nullptr;
  Reduce red;
  double* host_scalars = nullptr;   // pinned, mapped [64]: reduction results the host reads; the
                                    // kernels write them here directly (red.result is its device
                                    // address), so a wait for the stream is all a read needs
  hipStream_t gs_stream = nullptr;  // stream of the last GMRES cycle (its Hessenberg is copied on demand)
  double* dg = nullptr;              // diagnostics scratch of the outlet rows (allocated on first use)
  int* progress = nullptr;          // coherent pinned host word: GMRES columns completed (< 0: done)
  int* progress_dev = nullptr;      // its device address
  // low-rank correction Q = I + Ud Dd Ud^T of the GMRES right preconditioner (scaled variables):
  // Ud [defl_k][N] orthonormal columns, Dd [defl_k][defl_k]; defl_k = 0: none
  int defl_k = 0;
  double* Ud = nullptr;             // [DEFL_MAX][N] (allocated on first use)
  double* defl_small = nullptr;     // device: Dd [DEFL_MAX^2] | d [DEFL_MAX] | G [MAX_RESTART * DEFL_MAX] |
                                    //   uv [DEFL_MAX][MAX_RESTART + 1] (Ud_j . v_k of the current cycle)
  double* defl_partial = nullptr;   // [DEFL_MAX][RMAXB] reduction partials
  unsigned* defl_counter = nullptr; // [2]
  double* defl_tmp = nullptr;       // [N] the corrected Arnoldi vector
};

void set_error(const std::string& msg);
int check(hipError_t e, const char* what);

// residual kernels (vtsa_kernels.hip)
void launch_props(Handle& h, const double* y, double* pb, double* f, double* qstar, hipStream_t s);
void launch_flux(Handle& h, double t, const double* y, const double* pb, double* f, hipStream_t s);
void launch_acc(Handle& h, double t, const double* y, const double* pb, double* acc, hipStream_t s);
void launch_energy_roundtrip(Handle& h, double* y, hipStream_t s);
constexpr int OUTLET_ROW = 32;      // doubles per device outlet row
void launch_outlet_sums(Handle& h, const double* cell, const double* bnd, int valve_mask, double* row, hipStream_t s);
int reduce_weighted_dev(Handle& h, const double* a, int mode, double* result, hipStream_t st);
void launch_diag(Handle& h, double t, const double* y, const double* pb, double* cell, double* uz, double* ur,
                 double* uth, double* bnd, double* wall, hipStream_t s);
void launch_fdjac(Handle& h, double t, const double* y, const double* f0, hipStream_t s);
void launch_equilibrium(Handle& h, int count, const double* pc, const double* pw, const double* T,
                        const double* qw, double* qc, double* qwe, hipStream_t s);
// Perturbed matvec residual: props and fluxes at z + sigma*(u*s), epilogue
//   w = ((u*s) - dh (fp - f)/sigma)/s   (sigma, zero flag read from device scalars)
void launch_matvec_residual(Handle& h, double t, const double* z, const double* u, const double* s,
                            const double* f, double dh, double* w, bool respect_done, hipStream_t st);

// linear algebra (vtsa_linalg.hip)
void launch_prec_setup(Handle& h, double dh, const double* scale, int kind, int prec, int sweeps, hipStream_t s);
void gs_colour_lines(int Nr, int Nth, int nwall_lines, std::vector<int>& order, std::vector<int>& cstart);
int prof_resolve(Handle& h);
// returns 1 when a non-negative c or q entry would fall below -s; truncate = 1: such entries keep
// 0.1 of their value
void cons_stash_impl(Handle& h, const double* R, const double* s, int slot, hipStream_t st);
int cons_fetch_impl(Handle& h, int n, double* out, hipStream_t st);
int cons_shift_impl(Handle& h, const double* z, const double* s, const double lam[4], int truncate, double* out,
                    hipStream_t st);
// out = post^-1 * M^-1 (pre * in * (in_scale ? *in_scale : 1)); optionally stores the scaled
// input to `store_in`; optionally accumulates sum(out^2) and sets sigma in gs (matvec use).
void launch_prec_apply(Handle& h, const double* in, const double* in_scale, double* store_in, const double* pre,
                       const double* post, double* out, bool sigma_for_matvec, hipStream_t s);
int defl_set_impl(Handle& h, int k, int m, const double* G, const double* D, hipStream_t st);
void defl_apply_launch(Handle& h, const double* in, const double* in_scale, double* store, double* out,
                       int respect_done, hipStream_t st);
int gmres_cycle(Handle& h, double t, const double* z, const double* f, const double* s, double dh, double zs_norm,
                const double* b, double* w, double* pw, double rtol, double atol, int restart,
                vtsa_gmres_info* info, hipStream_t st);
// reductions returning host values (synchronising)
int reduce_rms(Handle& h, const double* a, const double* s, int len, double* out, hipStream_t st);
int reduce_weighted(Handle& h, const double* a, const double* w, int len, int mode, double* out, hipStream_t st);
void launch_residual_combine(Handle& h, const double* z, const double* f, double dh, const double* rhs_known,
                             double* R, hipStream_t st);
int newton_residual_impl(Handle& h, double t, const double* z, double dh, const double* rk, const double* s,
                         double* f, double* R, double* rn, hipStream_t st);
int newton_eval_impl(Handle& h, double t, const double* z, double dh, const double* rk, const double* s, double* f,
                     double* R, double* rn, double* pn, hipStream_t st);
int newton_try_impl(Handle& h, double t, const double* z, const double* dz, double lam, double dh, const double* rk,
                    const double* s, double* zt, double* ft, double* Rt, double* rn, double* pn, double* step,
                    hipStream_t st);
void scale_impl(Handle& h, const double* z, double rtol, double* s, hipStream_t st);
int error_norm_impl(Handle& h, const double* y, const double* z2, const double* y3, double rtol, double* en,
                    hipStream_t st);
void vec_op_impl(Handle& h, int op, int npos, double a, double b, double c, double e, const double* x,
                 const double* y, const double* z, const double* w, double* out, hipStream_t st);
size_t partial_size(int ncol, int nw);
bool zline_fits(int Nz);
size_t gs_pairs_count(int ncol, int Nz);   // values of Handle::P

// Preconditioner apply: out = post^-1 * S0 M~^-1 S0^-1 (pre * scale * in); optional side outputs:
//   store_in  <- scale * in (normalises a GMRES basis vector in place; may alias in)
//   red_out   <- rms(out / red_s) (or rms(out))
//   gs        <- unorm, sigma = sqrt(eps) max(1, *zs_rms) / unorm, zero_u (matvec finite-difference step)
// respect_done: return at once when the GMRES cycle in gs has finished.
struct ApplyArgs {
  const double* in; const double* in_scale; double* store_in; const double* pre; const double* post; double* out;
  const double* red_s; double* red_out; GmresScalars* gs; const double* zs_rms; int respect_done;
  double* out_raw;   // optional: the output before the division by post, i.e. M^-1 (pre * scale * in)
};
void apply_launch(Handle& h, const ApplyArgs& a, hipStream_t s);

}  // namespace vtsa

// The opaque handle of the C ABI.
struct vtsa_handle {
  vtsa::Handle h;
};

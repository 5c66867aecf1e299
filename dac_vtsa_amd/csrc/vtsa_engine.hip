// C ABI of libvtsa: handle lifetime, residual / Jacobian / preconditioner / GMRES / Newton entry
// points.  Host code here only sizes buffers, launches kernels and moves scalars.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "vtsa_internal.h"

namespace vtsa {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int check(hipError_t e, const char* what) {
  if (e == hipSuccess) return 0;
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return -1;
}

template <typename T>
static int dalloc(T** p, size_t count) {
  if (count == 0) count = 1;
  return check(hipMalloc((void**)p, sizeof(T) * count), "hipMalloc");
}

static int upload(double** dst, const double* src, size_t count) {
  if (dalloc(dst, count)) return -1;
  if (src && count) return check(hipMemcpy(*dst, src, sizeof(double) * count, hipMemcpyHostToDevice), "hipMemcpy");
  return check(hipMemset(*dst, 0, sizeof(double) * (count ? count : 1)), "hipMemset");
}

static void free_handle(Handle* h) {
  if (!h) return;
  double* bufs[] = {h->V, h->Az, h->Ath, h->arc, h->Ar, h->drc, h->feed_flux, h->N_nom, h->cov, h->pb, h->pb2, h->FF,
                    h->ytmp, h->ftmp, h->vtmp, h->vtmp2, h->accbuf, h->Jc, h->Pfd, h->Jo, h->Jcw, h->Jw, h->Jwc, h->Fw, h->Tw, h->s0,
                    h->Vb, h->red.partial, h->Ud, h->defl_small, h->defl_partial, h->defl_tmp, h->dg};
  for (double* b : bufs)
    if (b) (void)hipFree(b);
  if (h->F) (void)hipFree(h->F);
  if (h->P) (void)hipFree(h->P);
  if (h->Fr) (void)hipFree(h->Fr);
  if (h->gs) (void)hipFree(h->gs);
  if (h->red.counter) (void)hipFree(h->red.counter);
  if (h->defl_counter) (void)hipFree(h->defl_counter);
  if (h->gs_host) (void)hipHostFree(h->gs_host);
  if (h->host_scalars) (void)hipHostFree(h->host_scalars);
  if (h->progress) (void)hipHostFree(h->progress);
  if (h->gs_order) (void)hipFree(h->gs_order);
  if (h->gsb) (void)hipFree(h->gsb);
  if (h->gsu) (void)hipFree(h->gsu);
  for (hipEvent_t e : h->prof_pending) (void)hipEventDestroy(e);
  for (hipEvent_t e : h->prof_pool) (void)hipEventDestroy(e);
  delete h;
}

static int grow_basis(Handle& h, int restart) {
  if (restart + 1 <= h.vb_cols) return 0;
  if (h.Vb) (void)hipFree(h.Vb);
  h.Vb = nullptr;
  h.vb_cols = 0;
  if (dalloc(&h.Vb, (size_t)(restart + 1) * h.N)) return -1;
  h.vb_cols = restart + 1;
  return 0;
}

}  // namespace vtsa

using namespace vtsa;

static inline hipStream_t S(vtsa_stream s) { return (hipStream_t)s; }

static int finish(hipStream_t s) {
  (void)s;
  hipError_t e = hipGetLastError();
  return check(e, "kernel launch");
}

extern "C" {

const char* vtsa_version(void) { return "vtsa 0.1.0 (gfx950, FP64 residual)"; }
const char* vtsa_last_error(void) { return g_err.c_str(); }

int vtsa_create(const vtsa_model_desc* desc, const double* V, const double* Az, const double* Ath, const double* arc,
                const double* Ar, const double* drc, const double* feed_flux, const double* N_nom, const double* cov,
                vtsa_handle** out) {
  if (!desc || !out) { set_error("vtsa_create: null argument"); return -1; }
  const int Nr = desc->Nr, Nth = desc->Nth, Nz = desc->Nz;
  if (Nr < 1 || Nth < 1 || Nz < 2 || Nth == 2) { set_error("vtsa_create: invalid grid"); return -2; }
  if (!zline_fits(Nz)) { set_error("vtsa_create: Nz too large for the z-line preconditioner (LDS)"); return -2; }
  vtsa_handle* H = new vtsa_handle();
  Handle& h = H->h;
  h.desc = *desc;
  h.n = Nr * Nth * Nz;
  h.nw = desc->wall_on ? Nth * Nz : 0;
  h.N = 6 * h.n + h.nw;
  h.ncol = Nr * Nth;
  const size_t nwall = (size_t)Nth * Nz;
  int rc = 0;
  {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess) {
      if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
        h.lds_max = v;
      if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) == hipSuccess && v > 0)
        h.lds_per_cu = v;
      if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
        h.n_cu = v;
    }
    (void)hipGetLastError();
  }
  rc |= upload(&h.V, V, Nr);
  rc |= upload(&h.Az, Az, Nr);
  rc |= upload(&h.Ath, Ath, Nr);
  rc |= upload(&h.arc, arc, Nr);
  rc |= upload(&h.Ar, Nr > 1 ? Ar : nullptr, Nr > 1 ? Nr - 1 : 1);
  rc |= upload(&h.drc, Nr > 1 ? drc : nullptr, Nr > 1 ? Nr - 1 : 1);
  rc |= upload(&h.feed_flux, feed_flux, (size_t)Nr * Nth);
  rc |= upload(&h.N_nom, N_nom, (size_t)Nr * Nth);
  rc |= upload(&h.cov, cov, nwall);
  rc |= dalloc(&h.pb, (size_t)NPROP * h.n);
  rc |= dalloc(&h.pb2, (size_t)NPROP * h.n);
  rc |= dalloc(&h.FF, (size_t)4 * FaceTable::of(h.desc).count());
  rc |= dalloc(&h.ytmp, h.N);
  rc |= dalloc(&h.ftmp, h.N);
  rc |= dalloc(&h.vtmp, h.N);
  rc |= dalloc(&h.vtmp2, h.N);
  rc |= dalloc(&h.accbuf, NACC);
  rc |= dalloc(&h.Jc, (size_t)h.n * NSLOT * 36);
  rc |= dalloc(&h.Pfd, (size_t)h.n * NV * 14);
  rc |= dalloc(&h.Jcw, nwall * 4);
  rc |= dalloc(&h.Jw, nwall * 5);
  rc |= dalloc(&h.Jwc, nwall * 6);
  rc |= dalloc(&h.Tw, nwall * 3);
  rc |= check(hipMalloc(&h.F, sizeof(double) * (size_t)h.n * 180), "hipMalloc");
  rc |= check(hipMalloc(&h.P, sizeof(double) * gs_pairs_count(h.ncol, h.desc.Nz)), "hipMalloc");
  rc |= check(hipMalloc(&h.Fr, sizeof(double) * (size_t)h.n * 180), "hipMalloc");
  rc |= dalloc(&h.Fw, nwall);
  rc |= dalloc(&h.s0, h.N);
  rc |= dalloc(&h.gsb, h.N);
  rc |= dalloc(&h.gsu, h.N);
  rc |= dalloc(&h.Jo, (size_t)h.n * 144);
  {
    std::vector<int> order;
    gs_colour_lines(Nr, Nth, h.nw > 0 ? Nth : 0, order, h.gs_cstart);
    const int ncol_ = h.ncol;
    const int C = (int)h.gs_cstart.size() - 1;
    h.gs_ncell.assign(C, 0);
    h.gs_nouter.assign(C, 0);
    h.gs_nwall.assign(C, 0);
    for (int c = 0; c < C; ++c)
      for (int q = h.gs_cstart[c]; q < h.gs_cstart[c + 1]; ++q) {
        const int L = order[q];
        if (L >= ncol_) ++h.gs_nwall[c];
        else { ++h.gs_ncell[c]; if (h.nw > 0 && L / Nth == Nr - 1) ++h.gs_nouter[c]; }
      }
    rc |= check(hipMalloc((void**)&h.gs_order, sizeof(int) * order.size()), "hipMalloc");
    if (rc == 0)
      rc |= check(hipMemcpy(h.gs_order, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice), "hipMemcpy");
  }
  rc |= dalloc(&h.gs, 1);
  rc |= check(hipHostMalloc((void**)&h.gs_host, sizeof(GmresScalars)), "hipHostMalloc");
  rc |= check(hipHostMalloc((void**)&h.host_scalars, sizeof(double) * 64, hipHostMallocCoherent | hipHostMallocMapped),
              "hipHostMalloc");
  if (rc == 0) rc |= check(hipHostGetDevicePointer((void**)&h.red.result, h.host_scalars, 0), "hipHostGetDevicePointer");
  rc |= check(hipHostMalloc((void**)&h.progress, 64, hipHostMallocCoherent | hipHostMallocMapped), "hipHostMalloc");
  if (rc == 0) rc |= check(hipHostGetDevicePointer((void**)&h.progress_dev, h.progress, 0), "hipHostGetDevicePointer");
  h.red.max_blocks = 512;
  rc |= dalloc(&h.red.partial, partial_size(h.ncol, h.nw));
  rc |= dalloc(&h.red.counter, 16);
  if (rc == 0) {
    rc |= check(hipMemset(h.red.counter, 0, sizeof(unsigned) * 16), "hipMemset");
    rc |= check(hipMemset(h.gs, 0, sizeof(GmresScalars)), "hipMemset");
    rc |= check(hipMemset(h.Jc, 0, sizeof(double) * (size_t)h.n * NSLOT * 36), "hipMemset");
  }
  if (rc == 0) rc = grow_basis(h, 30);
  // the set-up copies and fills above run on the null stream: finish them before the handle is
  // used from a stream that does not synchronise with it
  if (rc == 0) rc |= check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
  if (rc != 0) {
    free_handle(new Handle(h));
    delete H;
    return -3;
  }
  h.m.d = h.desc;
  h.m.n = h.n;
  h.m.nw = h.nw;
  h.m.N = h.N;
  h.m.V = h.V; h.m.Az = h.Az; h.m.Ath = h.Ath; h.m.arc = h.arc; h.m.Ar = h.Ar; h.m.drc = h.drc;
  h.m.feed_flux = h.feed_flux; h.m.N_nom = h.N_nom; h.m.cov = h.cov;
  *out = H;
  return 0;
}

int vtsa_reset(vtsa_handle* H, const vtsa_model_desc* desc, const double* V, const double* Az, const double* Ath,
               const double* arc, const double* Ar, const double* drc, const double* feed_flux, const double* N_nom,
               const double* cov, vtsa_stream s) {
  if (!H || !desc || !V || !Az || !Ath || !arc || !feed_flux || !N_nom || !cov) {
    set_error("vtsa_reset: null argument");
    return -1;
  }
  Handle& h = H->h;
  if (desc->Nr != h.desc.Nr || desc->Nth != h.desc.Nth || desc->Nz != h.desc.Nz ||
      (desc->wall_on != 0) != (h.desc.wall_on != 0)) {
    set_error("vtsa_reset: the handle was created for another grid");
    return -2;
  }
  if (desc->Nr > 1 && (!Ar || !drc)) { set_error("vtsa_reset: null argument"); return -1; }
  const hipStream_t st = S(s);
  // the previous model's work on this stream reads these arrays: the copies queue behind it
  const int Nr = desc->Nr;
  const size_t nrt = (size_t)Nr * desc->Nth, nwall = (size_t)desc->Nth * desc->Nz;
  struct { double* dst; const double* src; size_t count; } cp[] = {
      {h.V, V, (size_t)Nr}, {h.Az, Az, (size_t)Nr}, {h.Ath, Ath, (size_t)Nr}, {h.arc, arc, (size_t)Nr},
      {h.Ar, Nr > 1 ? Ar : nullptr, (size_t)(Nr - 1)}, {h.drc, Nr > 1 ? drc : nullptr, (size_t)(Nr - 1)},
      {h.feed_flux, feed_flux, nrt}, {h.N_nom, N_nom, nrt}, {h.cov, cov, nwall}};
  for (const auto& c : cp)
    if (c.src && c.count &&
        check(hipMemcpyAsync(c.dst, c.src, sizeof(double) * c.count, hipMemcpyHostToDevice, st), "hipMemcpyAsync"))
      return -1;
  // the host arrays may go out of scope when this returns
  if (check(hipStreamSynchronize(st), "hipStreamSynchronize")) return -1;
  if (prof_resolve(h)) return -1;
  h.desc = *desc;
  h.m.d = h.desc;
  h.prec_kind = -1;
  h.prec_ready = 0;
  h.prec_dh = 0.0;
  h.defl_k = 0;
  // options and sampling back to their defaults, as for a handle just created
  const Handle fresh{};
  h.gs_full = fresh.gs_full;
  h.gs_forward = fresh.gs_forward;
  h.gs_twist = fresh.gs_twist;
  h.gs_tw_m = -1;
  h.gs_sweeps = fresh.gs_sweeps;
  h.prof_every = 0;
  h.prof_seen = h.prof_launches = 0;
  h.prof_ms = h.prof_bytes = 0.0;
  return 0;
}

int vtsa_destroy(vtsa_handle* H) {
  if (!H) return 0;
  (void)hipDeviceSynchronize();
  Handle* copy = new Handle(H->h);
  free_handle(copy);
  delete H;
  return 0;
}

int vtsa_sizes(const vtsa_handle* H, int32_t* n_cells, int32_t* n_wall, int32_t* n_unknowns) {
  if (!H) return -1;
  if (n_cells) *n_cells = H->h.n;
  if (n_wall) *n_wall = H->h.nw;
  if (n_unknowns) *n_unknowns = H->h.N;
  return 0;
}

int vtsa_rhs(vtsa_handle* H, double t, const double* y, double* f, double* acc, vtsa_stream s) {
  if (!H || !y || !f) { set_error("vtsa_rhs: null argument"); return -1; }
  Handle& h = H->h;
  launch_props(h, y, h.pb, f, nullptr, S(s));
  launch_flux(h, t, y, h.pb, f, S(s));
  if (acc) launch_acc(h, t, y, h.pb, acc, S(s));
  return finish(S(s));
}

int vtsa_stage_residual(vtsa_handle* H, double t, const double* y, double gamma_h, const double* rhs_known, double* R,
                        vtsa_stream s) {
  if (!H || !y || !rhs_known || !R) { set_error("vtsa_stage_residual: null argument"); return -1; }
  Handle& h = H->h;
  launch_props(h, y, h.pb, h.ftmp, nullptr, S(s));
  launch_flux(h, t, y, h.pb, h.ftmp, S(s));
  launch_residual_combine(h, y, h.ftmp, gamma_h, rhs_known, R, S(s));
  return finish(S(s));
}

int vtsa_diag(vtsa_handle* H, double t, const double* y, double* cell, double* uz, double* ur, double* uth,
              double* bnd, double* wall, vtsa_stream s) {
  if (!H || !y || !cell || !uz || !ur || !uth || !bnd || !wall) { set_error("vtsa_diag: null argument"); return -1; }
  Handle& h = H->h;
  launch_diag(h, t, y, h.pb, cell, uz, ur, uth, bnd, wall, S(s));
  return finish(S(s));
}

int vtsa_equilibrium(vtsa_handle* H, int32_t count, const double* pc, const double* pw, const double* T,
                     const double* qw, double* qc, double* qwe, vtsa_stream s) {
  if (!H || count < 0) { set_error("vtsa_equilibrium: bad argument"); return -1; }
  if (count == 0) return 0;
  if (!pc || !pw || !T || !qw || !qc || !qwe) { set_error("vtsa_equilibrium: null argument"); return -1; }
  launch_equilibrium(H->h, count, pc, pw, T, qw, qc, qwe, S(s));
  return finish(S(s));
}

int vtsa_fd_jacobian(vtsa_handle* H, double t, const double* y, const double* f0, vtsa_stream s) {
  if (!H || !y || !f0) { set_error("vtsa_fd_jacobian: null argument"); return -1; }
  launch_fdjac(H->h, t, y, f0, S(s));
  return finish(S(s));
}

int vtsa_jacobian_get(vtsa_handle* H, double* Jc, double* Jcw, double* Jw, double* Jwc) {
  if (!H) return -1;
  Handle& h = H->h;
  const size_t nwall = (size_t)h.desc.Nth * h.desc.Nz;
  if (check(hipDeviceSynchronize(), "sync")) return -1;
  if (Jc && check(hipMemcpy(Jc, h.Jc, sizeof(double) * (size_t)h.n * NSLOT * 36, hipMemcpyDeviceToHost), "memcpy"))
    return -1;
  if (h.nw > 0) {
    if (Jcw && check(hipMemcpy(Jcw, h.Jcw, sizeof(double) * nwall * 4, hipMemcpyDeviceToHost), "memcpy")) return -1;
    if (Jw && check(hipMemcpy(Jw, h.Jw, sizeof(double) * nwall * 5, hipMemcpyDeviceToHost), "memcpy")) return -1;
    if (Jwc && check(hipMemcpy(Jwc, h.Jwc, sizeof(double) * nwall * 6, hipMemcpyDeviceToHost), "memcpy")) return -1;
  }
  return 0;
}

int vtsa_precond_setup(vtsa_handle* H, double dh, const double* scale, int32_t kind, int32_t precision,
                       int32_t sweeps, vtsa_stream s) {
  if (!H || !scale) { set_error("vtsa_precond_setup: null argument"); return -1; }
  if (kind != VTSA_PREC_BLOCK_JACOBI && kind != VTSA_PREC_ZLINE && kind != VTSA_PREC_ZLINE_GS) {
    set_error("unknown preconditioner");
    return -2;
  }
  if (precision != VTSA_FP64 && precision != VTSA_FP32) { set_error("unknown precision"); return -2; }
  if (sweeps < 1 || sweeps > 16) { set_error("vtsa_precond_setup: sweeps must be in 1..16"); return -2; }
  launch_prec_setup(H->h, dh, scale, kind, precision, sweeps, S(s));
  H->h.defl_k = 0;   // a new factorisation invalidates the low-rank correction built for the old one
  return finish(S(s));
}

int vtsa_gmres_hessenberg(vtsa_handle* H, double* Hr, int32_t* m) {
  if (!H || !Hr || !m) { set_error("vtsa_gmres_hessenberg: null argument"); return -1; }
  Handle& h = H->h;
  if (check(hipMemcpyAsync(h.gs_host, h.gs, sizeof(GmresScalars), hipMemcpyDeviceToHost, h.gs_stream), "memcpy"))
    return -1;
  if (check(hipStreamSynchronize(h.gs_stream), "sync")) return -1;
  const GmresScalars& G = *h.gs_host;
  *m = G.iters;
  for (int r = 0; r <= MAX_RESTART; ++r)
    for (int c = 0; c < MAX_RESTART; ++c) Hr[r * MAX_RESTART + c] = G.Hr[r][c];
  return 0;
}

int vtsa_defl_set(vtsa_handle* H, int32_t k, int32_t m, const double* G, const double* D, vtsa_stream s) {
  if (!H) { set_error("vtsa_defl_set: null handle"); return -1; }
  if (k < 0 || k > DEFL_MAX || (k > 0 && (!G || !D || m < 1 || m > H->h.vb_cols - 1))) {
    set_error("vtsa_defl_set: k must be in 0..16 with G (m x k) and D (k x k), m within the last basis");
    return -2;
  }
  const int rc = defl_set_impl(H->h, k, m, G, D, S(s));
  if (rc) return rc;
  return finish(S(s));
}

int vtsa_defl_apply(vtsa_handle* H, const double* in, double* out, vtsa_stream s) {
  if (!H || !in || !out) { set_error("vtsa_defl_apply: null argument"); return -1; }
  if (H->h.defl_k == 0) {
    if (in != out && check(hipMemcpyAsync(out, in, sizeof(double) * H->h.N, hipMemcpyDeviceToDevice, S(s)), "copy"))
      return -1;
    return finish(S(s));
  }
  defl_apply_launch(H->h, in, nullptr, nullptr, out, 0, S(s));
  return finish(S(s));
}

int vtsa_precond_apply(vtsa_handle* H, const double* r, double* out, vtsa_stream s) {
  if (!H || !r || !out) { set_error("vtsa_precond_apply: null argument"); return -1; }
  if (!H->h.prec_ready) { set_error("vtsa_precond_apply: no factorisation"); return -2; }
  launch_prec_apply(H->h, r, nullptr, nullptr, nullptr, nullptr, out, false, S(s));
  return finish(S(s));
}

int vtsa_gmres(vtsa_handle* H, double t, const double* z, const double* f, const double* sc, double dh,
               double zs_norm, const double* b, double* w, double* pw, double rtol, double atol, int32_t restart,
               vtsa_gmres_info* info, vtsa_stream s) {
  if (!H || !z || !f || !sc || !b || !w || !info) { set_error("vtsa_gmres: null argument"); return -1; }
  Handle& h = H->h;
  if (!h.prec_ready) { set_error("vtsa_gmres: no factorisation"); return -2; }
  if (restart < 1 || restart > MAX_RESTART) { set_error("vtsa_gmres: restart out of range"); return -2; }
  if (grow_basis(h, restart)) return -3;
  int rc = gmres_cycle(h, t, z, f, sc, dh, zs_norm, b, w, pw, rtol, atol, restart, info, S(s));
  if (rc) return rc;
  return finish(S(s));
}

int vtsa_scale(vtsa_handle* H, const double* z, double rtol, double* s, vtsa_stream st) {
  if (!H || !z || !s) return -1;
  scale_impl(H->h, z, rtol, s, S(st));
  return finish(S(st));
}

int vtsa_rms(vtsa_handle* H, const double* a, const double* s, int32_t len, double* out, vtsa_stream st) {
  if (!H || !a || !out) return -1;
  return reduce_rms(H->h, a, s, len, out, S(st));
}

int vtsa_newton_eval(vtsa_handle* H, double t, const double* z, double dh, const double* rk, const double* s,
                     double* f, double* R, double* rn, double* pn, vtsa_stream st) {
  if (!H || !z || !rk || !s || !f || !R || !rn || !pn) return -1;
  if (!H->h.prec_ready) { set_error("vtsa_newton_eval: no factorisation"); return -2; }
  return newton_eval_impl(H->h, t, z, dh, rk, s, f, R, rn, pn, S(st));
}

int vtsa_newton_residual(vtsa_handle* H, double t, const double* z, double dh, const double* rk, const double* s,
                         double* f, double* R, double* rn, vtsa_stream st) {
  if (!H || !z || !rk || !s || !f || !R) { set_error("vtsa_newton_residual: null argument"); return -1; }
  const int rc = newton_residual_impl(H->h, t, z, dh, rk, s, f, R, rn, S(st));
  if (rc) return rc;
  return finish(S(st));
}

int vtsa_newton_try(vtsa_handle* H, double t, const double* z, const double* dz, double lam, double dh,
                    const double* rk, const double* s, double* zt, double* ft, double* Rt, double* rn, double* pn,
                    double* step, vtsa_stream st) {
  if (!H || !z || !dz || !rk || !s || !zt || !ft || !Rt) return -1;
  if (!H->h.prec_ready) { set_error("vtsa_newton_try: no factorisation"); return -2; }
  return newton_try_impl(H->h, t, z, dz, lam, dh, rk, s, zt, ft, Rt, rn, pn, step, S(st));
}

int vtsa_cons(vtsa_handle* H, const double* R, const double* s, double* wr, double* ws, vtsa_stream st) {
  if (!H || !R || !s || !wr || !ws) return -1;
  double out[8];
  int rc = reduce_weighted(H->h, R, s, 0, 0, out, S(st));
  for (int q = 0; q < 4; ++q) { wr[q] = out[q]; ws[q] = out[4 + q]; }
  return rc;
}

int vtsa_cons_norm2(vtsa_handle* H, const double* s, double* out, vtsa_stream st) {
  if (!H || !s || !out) { set_error("vtsa_cons_norm2: null argument"); return -1; }
  return reduce_weighted(H->h, s, nullptr, 0, 2, out, S(st));
}

int vtsa_cons_shift(vtsa_handle* H, const double* z, const double* s, const double* lambda, int32_t mode,
                    double* out, vtsa_stream st) {
  if (!H || !z || !s || !lambda || !out) { set_error("vtsa_cons_shift: null argument"); return -1; }
  if (mode & ~(VTSA_SHIFT_TRUNCATE | VTSA_SHIFT_AT_ZERO | VTSA_SHIFT_NO_REPORT)) {
    set_error("vtsa_cons_shift: unknown mode bits");
    return -2;
  }
  const int rc = cons_shift_impl(H->h, z, s, lambda, mode, out, S(st));
  if (rc < 0) return rc;
  const int e = finish(S(st));
  return e != 0 ? e : rc;
}

int vtsa_cons_stash(vtsa_handle* H, const double* R, const double* s, int32_t slot, vtsa_stream st) {
  if (!H || !R || !s) { set_error("vtsa_cons_stash: null argument"); return -1; }
  if (slot < 0 || slot >= 4) { set_error("vtsa_cons_stash: slot out of range"); return -2; }
  cons_stash_impl(H->h, R, s, slot, S(st));
  return finish(S(st));
}

int vtsa_cons_fetch(vtsa_handle* H, int32_t n, double* wr, double* ws, vtsa_stream st) {
  if (!H || !wr || !ws) { set_error("vtsa_cons_fetch: null argument"); return -1; }
  if (n < 0 || n > 4) { set_error("vtsa_cons_fetch: slot count out of range"); return -2; }
  double out[32];
  if (cons_fetch_impl(H->h, n, out, S(st))) return -1;
  for (int q = 0; q < n; ++q)
    for (int k = 0; k < 4; ++k) { wr[4 * q + k] = out[8 * q + k]; ws[4 * q + k] = out[8 * q + 4 + k]; }
  return 0;
}

int vtsa_outlet_row(vtsa_handle* H, double t, const double* y, int32_t valve_mask, double* row, vtsa_stream st) {
  if (!H || !y || !row) { set_error("vtsa_outlet_row: null argument"); return -1; }
  Handle& h = H->h;
  const int Nr = h.desc.Nr, Nth = h.desc.Nth, Nz = h.desc.Nz;
  const size_t n = (size_t)h.n, nf = (size_t)Nr * Nth, nw2 = (size_t)Nth * Nz;
  const size_t o_uz = 8 * n, o_ur = o_uz + std::max<size_t>(1, nf * (Nz - 1)),
               o_uth = o_ur + std::max<size_t>(1, (size_t)(Nr - 1) * Nth * Nz), o_bnd = o_uth + n,
               o_wall = o_bnd + 2 * 14 * nf, total = o_wall + 4 * nw2;
  if (!h.dg && dalloc(&h.dg, total)) return -1;
  launch_diag(h, t, y, h.pb, h.dg, h.dg + o_uz, h.dg + o_ur, h.dg + o_uth, h.dg + o_bnd, h.dg + o_wall, S(st));
  launch_outlet_sums(h, h.dg, h.dg + o_bnd, valve_mask, row, S(st));
  if (reduce_weighted_dev(h, y, 1, row + 24, S(st))) return -1;
  return finish(S(st));
}

int vtsa_state_roundtrip(vtsa_handle* H, double* y, vtsa_stream st) {
  if (!H || !y) { set_error("vtsa_state_roundtrip: null argument"); return -1; }
  launch_energy_roundtrip(H->h, y, S(st));
  return finish(S(st));
}

int vtsa_inventories(vtsa_handle* H, const double* y, double* out, vtsa_stream st) {
  if (!H || !y || !out) return -1;
  return reduce_weighted(H->h, y, nullptr, 0, 1, out, S(st));
}

int vtsa_vec_op(vtsa_handle* H, int32_t op, double a, double b, double c, double e, const double* x, const double* y,
                const double* z, const double* w, double* out, vtsa_stream st) {
  if (!H || !out || op < 0 || op > 9) { set_error("vtsa_vec_op: bad argument"); return -1; }
  // every op reads x and y; op 5 (the third-order quadrature) reads z and w as well
  if (!x || !y || (op == 5 && (!z || !w))) {
    set_error("vtsa_vec_op: null input for this op");
    return -1;
  }
  vec_op_impl(H->h, op, 5 * H->h.n, a, b, c, e, x, y, z, w, out, S(st));
  return finish(S(st));
}

int vtsa_error_norm(vtsa_handle* H, const double* y, const double* z2, const double* y3, double rtol, double* en,
                    vtsa_stream st) {
  if (!H || !y || !z2 || !y3 || !en) return -1;
  if (!H->h.prec_ready) { set_error("vtsa_error_norm: no factorisation"); return -2; }
  return error_norm_impl(H->h, y, z2, y3, rtol, en, S(st));
}

// ---------------------------------------------------------------- plumbing
__global__ void k_fill(double* p, double v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

int vtsa_device_count(int32_t* count) {
  int c = 0;
  if (check(hipGetDeviceCount(&c), "hipGetDeviceCount")) { *count = 0; return -1; }
  *count = c;
  return 0;
}

int vtsa_device_alloc(size_t bytes, void** out) {
  if (!out) return -1;
  *out = nullptr;
  return check(hipMalloc(out, bytes ? bytes : 8), "hipMalloc");
}

int vtsa_device_free(void* p) { return p ? check(hipFree(p), "hipFree") : 0; }

int vtsa_copy_h2d(void* dst, const void* src, size_t bytes, vtsa_stream st) {
  if (check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(st)), "hipMemcpyAsync")) return -1;
  return check(hipStreamSynchronize(S(st)), "hipStreamSynchronize");
}

int vtsa_copy_d2h(void* dst, const void* src, size_t bytes, vtsa_stream st) {
  if (check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(st)), "hipMemcpyAsync")) return -1;
  return check(hipStreamSynchronize(S(st)), "hipStreamSynchronize");
}

int vtsa_copy_d2d(void* dst, const void* src, size_t bytes, vtsa_stream st) {
  return check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(st)), "hipMemcpyAsync");
}

int vtsa_fill(double* dst, double value, size_t count, vtsa_stream st) {
  if (count == 0) return 0;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, S(st), dst, value, count);
  return finish(S(st));
}

int vtsa_sync(vtsa_stream st) { return check(hipStreamSynchronize(S(st)), "hipStreamSynchronize"); }

int vtsa_stream_create(vtsa_stream* out) {
  if (!out) { set_error("vtsa_stream_create: null argument"); return -1; }
  hipStream_t s = nullptr;
  // default priority; does not synchronise with the null stream
  if (check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags")) return -1;
  *out = (vtsa_stream)s;
  return 0;
}

int vtsa_stream_destroy(vtsa_stream st) {
  if (!st) return 0;
  if (check(hipStreamSynchronize(S(st)), "hipStreamSynchronize")) return -1;
  return check(hipStreamDestroy(S(st)), "hipStreamDestroy");
}

extern "C++" {
namespace {
template <typename F>
int time_launches(F&& launch, int32_t repeat, double* ms, hipStream_t st) {
  hipEvent_t e0, e1;
  if (check(hipEventCreate(&e0), "hipEventCreate")) return -1;
  if (check(hipEventCreate(&e1), "hipEventCreate")) {
    (void)hipEventDestroy(e0);
    return -1;
  }
  (void)hipEventRecord(e0, st);
  for (int r = 0; r < repeat; ++r) launch();
  (void)hipEventRecord(e1, st);
  int rc = check(hipEventSynchronize(e1), "hipEventSynchronize");
  float m = 0.f;
  if (!rc) rc = check(hipEventElapsedTime(&m, e0, e1), "hipEventElapsedTime");
  *ms = m;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc ? rc : finish(st);
}
}  // namespace
}

int vtsa_time_rhs(vtsa_handle* H, double t, const double* y, double* f, int32_t repeat, double* ms, vtsa_stream st) {
  if (!H || !y || !f || !ms || repeat < 1) { set_error("vtsa_time_rhs: null argument or repeat < 1"); return -1; }
  Handle& h = H->h;
  return time_launches([&] {
    launch_props(h, y, h.pb, f, nullptr, S(st));
    launch_flux(h, t, y, h.pb, f, S(st));
  }, repeat, ms, S(st));
}

int vtsa_time_fdjac(vtsa_handle* H, double t, const double* y, const double* f0, int32_t repeat, double* ms,
                    vtsa_stream st) {
  if (!H || !y || !f0 || !ms || repeat < 1) { set_error("vtsa_time_fdjac: null argument or repeat < 1"); return -1; }
  Handle& h = H->h;
  return time_launches([&] { launch_fdjac(h, t, y, f0, S(st)); }, repeat, ms, S(st));
}

int vtsa_set_option(vtsa_handle* H, int32_t option, int32_t value) {
  if (!H) { set_error("vtsa_set_option: null handle"); return -1; }
  switch (option) {
    case VTSA_OPT_GS_WHOLE_LINES:
      if (value < 0 || value > 5) { set_error("vtsa_set_option: staging level must be 0 .. 5"); return -2; }
      H->h.gs_full = value;
      return 0;
    case VTSA_OPT_GS_FORWARD:
      if (value < 0 || value > 1) { set_error("vtsa_set_option: sweep order must be 0 or 1"); return -2; }
      H->h.gs_forward = value;
      return 0;
    case VTSA_OPT_GS_TWIST:
      if (value < 0 || value > 1) { set_error("vtsa_set_option: twisted solve must be 0 or 1"); return -2; }
      H->h.gs_twist = value;
      return 0;
    default: set_error("vtsa_set_option: unknown option"); return -2;
  }
}

int vtsa_profile(vtsa_handle* H, int32_t every) {
  if (!H || every < 0) { set_error("vtsa_profile: bad argument"); return -1; }
  Handle& h = H->h;
  if (prof_resolve(h)) return -1;
  h.prof_every = every;
  h.prof_seen = h.prof_launches = 0;
  h.prof_ms = h.prof_bytes = 0.0;
  return 0;
}

int vtsa_profile_read(vtsa_handle* H, int64_t* launches, double* total_ms, double* total_bytes) {
  if (!H || !launches || !total_ms || !total_bytes) { set_error("vtsa_profile_read: null argument"); return -1; }
  Handle& h = H->h;
  if (prof_resolve(h)) return -1;
  *launches = h.prof_launches;
  *total_ms = h.prof_ms;
  *total_bytes = h.prof_bytes;
  return 0;
}

}  // extern "C"

// Reductions, Newton vector operations and the device-resident GMRES cycle.
//
// Reductions use one launch: per-block partial sums, then the last block to finish (ticket
// counter) adds the partials in a fixed order, so results are deterministic.
#include <chrono>
#include <mutex>
#include <thread>

#include "vtsa_internal.h"
#include "vtsa_reduce.h"

namespace vtsa {

// ------------------------------------------------------------------ elementwise + reduction kernels
// R = z - dh f - rhs_known and sum((R/s)^2)  (core/solver.py:172-173)
__global__ void __launch_bounds__(RBLK) k_residual_combine(int N, const double* __restrict__ z,
                                                           const double* __restrict__ f, double dh,
                                                           const double* __restrict__ rk, const double* __restrict__ s,
                                                           double* __restrict__ R, double* __restrict__ partial,
                                                           unsigned* counter, double* result) {
  __shared__ double sh[RBLK];
  __shared__ bool slast;
  double loc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    double r = (z[i] - dh * f[i]) - rk[i];
    R[i] = r;
    if (s) { double q = r / s[i]; loc += q * q; }
  }
  double bs = wave_block_sum(loc, sh);
  if (threadIdx.x == 0) store_partial(partial + blockIdx.x, bs);
  double tot;
  if (last_block_total(counter, &slast, partial, &tot)) {
    *result = sqrt(tot / (double)N);
  }
}

__global__ void k_set_scalar(double* p, double v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

// generic reduction: mode 0: rms(a/s); mode 1: rms(a)

__global__ void __launch_bounds__(RBLK) k_rms(int N, const double* __restrict__ a, const double* __restrict__ s,
                                              double* __restrict__ partial, unsigned* counter, double* result) {
  __shared__ double sh[RBLK];
  __shared__ bool slast;
  double loc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    double q = s ? a[i] / s[i] : a[i];
    loc += q * q;
  }
  double bs = wave_block_sum(loc, sh);
  if (threadIdx.x == 0) store_partial(partial + blockIdx.x, bs);
  double tot;
  if (last_block_total(counter, &slast, partial, &tot)) {
    *result = sqrt(tot / (double)N);
  }
}

// z_try = truncate(z + lam dz, z) on the first npos entries (c and q), step = rms((z_try - z)/s)
// (core/solver.py:99-108,245,255)
__global__ void __launch_bounds__(RBLK) k_newton_update(int N, int npos, const double* __restrict__ z,
                                                        const double* __restrict__ dz, double lam,
                                                        const double* __restrict__ s, double* __restrict__ zt,
                                                        double* __restrict__ partial, unsigned* counter,
                                                        double* result) {
  __shared__ double sh[RBLK];
  __shared__ bool slast;
  double loc = 0.0;
  bool finite = true;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    double zo = z[i];
    const double d = dz[i];
    finite = finite && __builtin_isfinite(d);
    double v = zo + lam * d;
    if (i < npos) {
      double lo = 0.1 * (zo > 0.0 ? zo : (zo != zo ? zo : 0.0));
      v = (v != v || lo != lo) ? v + lo : (v > lo ? v : lo);
    }
    zt[i] = v;
    double q = (v - zo) / s[i];
    loc += q * q;
  }
  // a non-finite step direction turns the step norm into NaN (an overflowing sum gives +inf)
  if (!finite) loc = __builtin_nan("");
  double bs = wave_block_sum(loc, sh);
  if (threadIdx.x == 0) store_partial(partial + blockIdx.x, bs);
  double tot;
  if (last_block_total(counter, &slast, partial, &tot)) {
    *result = sqrt(tot / (double)N);
  }
}

// s = atol + rtol |z|   (TRBDF2.scale, core/solver.py:331-332)
__global__ void k_scale(Model m, const double* __restrict__ z, double rtol, double* __restrict__ s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.N) return;
  int v = i / m.n;
  if (v > 6) v = 6;
  s[i] = m.d.atol[v] + rtol * fabs(z[i]);
}

// sc = atol + rtol max(|y|, |z2|)   (core/solver.py:461)
__global__ void k_scale2(Model m, const double* __restrict__ y, const double* __restrict__ z, double rtol,
                         double* __restrict__ s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.N) return;
  int v = i / m.n;
  if (v > 6) v = 6;
  double a = fabs(y[i]), b = fabs(z[i]);
  double mx = (a != a || b != b) ? a + b : (a > b ? a : b);
  s[i] = m.d.atol[v] + rtol * mx;
}

// Weighted sums: conservation (W R, W s) and inventories.  mode 0: cons (8 outputs),
// mode 1: inventories (7 outputs).  V per ring, C_wall, eps_t, rho_b from the model.
__global__ void __launch_bounds__(RBLK) k_weighted(Model m, int mode, const double* __restrict__ a,
                                                   const double* __restrict__ b, double* __restrict__ partial,
                                                   unsigned* counter, double* result) {
  __shared__ double sh[RBLK];
  __shared__ bool slast;
  const vtsa_model_desc& d = m.d;
  const int n = m.n, nw = m.nw;
  const int per_ring = d.Nth * d.Nz;
  double loc[8];
  for (int q = 0; q < 8; ++q) loc[q] = 0.0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n + nw; c += gridDim.x * blockDim.x) {
    if (c < n) {
      const double V = m.V[c / per_ring];
      if (mode == 0) {
        for (int src = 0; src < 2; ++src) {
          const double* x = src == 0 ? a : b;
          double g0 = (d.eps_t * V) * x[c], g1 = (d.eps_t * V) * x[n + c], g2 = (d.eps_t * V) * x[2 * n + c];
          double a0 = (d.rho_b * V) * x[3 * n + c], a1 = (d.rho_b * V) * x[4 * n + c];
          loc[src * 4 + 0] += g0 + a0;
          loc[src * 4 + 1] += g1 + a1;
          loc[src * 4 + 2] += g2;
          loc[src * 4 + 3] += V * x[5 * n + c];
        }
      } else if (mode == 1) {
        for (int s = 0; s < 3; ++s) loc[s] += (d.eps_t * a[s * n + c]) * V;
        loc[3] += (d.rho_b * a[3 * n + c]) * V;
        loc[4] += (d.rho_b * a[4 * n + c]) * V;
        loc[5] += a[5 * n + c] * V;
      } else {
        // sum over each conserved row of (W_i a_i)^2
        const double g0 = (d.eps_t * V) * a[c], g1 = (d.eps_t * V) * a[n + c], g2 = (d.eps_t * V) * a[2 * n + c];
        const double a0 = (d.rho_b * V) * a[3 * n + c], a1 = (d.rho_b * V) * a[4 * n + c];
        const double e = V * a[5 * n + c];
        loc[0] += g0 * g0 + a0 * a0;
        loc[1] += g1 * g1 + a1 * a1;
        loc[2] += g2 * g2;
        loc[3] += e * e;
      }
    } else {
      const int w = c - n;
      if (mode == 0) {
        loc[3] += d.C_wall * a[6 * n + w];
        loc[7] += d.C_wall * b[6 * n + w];
      } else if (mode == 1) {
        loc[6] += d.C_wall * (a[6 * n + w] - d.T_ref);
      } else {
        const double e = d.C_wall * a[6 * n + w];
        loc[3] += e * e;
      }
    }
  }
  const int nout = mode == 0 ? 8 : (mode == 1 ? 7 : 4);
  __shared__ double tot[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {   // constant trip count keeps loc[] in registers
    if (q < nout) {
      double bs = wave_block_sum(loc[q], sh);
      if (threadIdx.x == 0) store_partial(partial + q * RMAXB + blockIdx.x, bs);
    }
  }
  if (last_block_totals(counter, &slast, partial, nout, tot) && threadIdx.x < nout) result[threadIdx.x] = tot[threadIdx.x];
}

__global__ void k_fill_u32(unsigned* p, unsigned v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// out = z - S^2 W^T lambda: each entry moved along its conserved row k (gas species, adsorbed
// species into the same row, energy of bed and wall) by s_i^2 W_ki lambda_k.  *flag is set
// when a non-negative c or q entry would fall below -s (below 0 with truncate = 3); with truncate
// (1 or 3) such entries keep 0.1 of their value.
__global__ void k_cons_shift(Model m, const double* z, const double* __restrict__ s, double l0, double l1,
                             double l2, double l3, double* out,   // out may alias z (elementwise)
                             int truncate, unsigned* flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.N) return;
  const vtsa_model_desc& d = m.d;
  const int n = m.n;
  double w, lam;
  if (i < 6 * n) {
    const int v = i / n, c = i - v * n;
    const double V = m.V[c / (d.Nth * d.Nz)];
    const int k = v < 3 ? v : (v < 5 ? v - 3 : 3);
    w = v < 3 ? d.eps_t * V : (v < 5 ? d.rho_b * V : V);
    lam = k == 0 ? l0 : (k == 1 ? l1 : (k == 2 ? l2 : l3));
  } else {
    w = d.C_wall;
    lam = l3;
  }
  const double zo = z[i];
  const double si = s[i];
  double v = zo - (si * si) * w * lam;
  // c and q: an excursion below zero by more than the tolerance scale s (any, with truncate = 3)
  // is flagged; truncated, such entries keep at least 0.1 of their value, as a Newton update
  // (core/solver.py:99-108)
  if (i < 5 * n && zo >= 0.0 && v < (truncate == 3 ? 0.0 : -si)) {
    if (truncate) v = 0.1 * zo;
    atomicOr(flag, 1u);
  }
  out[i] = v;
}

// Integrator vector formulas (core/solver.py:419-476), exact upstream operation order.
__global__ void k_vec_op(int op, int N, int npos, double a, double b, double c, double e,
                         const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
                         const double* __restrict__ w, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double r;
  switch (op) {
    case 0: r = x[i] + a * y[i]; break;                                // y + dh f
    case 1: r = x[i] + a * y[i]; break;                                // y + (c h) slope, truncated
    case 2: r = x[i] + (y[i] - x[i]) / a; break;                       // y + (z1 - y)/gamma, truncated
    case 3: r = x[i] + a * ((y[i] - x[i]) / b); break;                 // y + ((1-g)h) (z1 - y)/dh
    case 4: r = a * y[i] - b * x[i]; break;                            // C1 z1 - C2 y
    case 5: r = x[i] + a * ((b * y[i] + c * z[i]) + e * w[i]); break;  // y3
    case 6: r = (y[i] - x[i]) / a; break;                              // (z2 - y)/h
    case 7: r = y[i] - x[i]; break;                                    // z2 - y3
    case 8: r = (-x[i]) / y[i]; break;                                 // b = -R/s
    case 9: r = x[i] * y[i]; break;                                    // w*s
    default: r = 0.0;
  }
  if ((op == 1 || op == 2) && i < npos) {
    double zo = x[i];
    double lo = 0.1 * (zo > 0.0 ? zo : (zo != zo ? zo : 0.0));
    r = (r != r || lo != lo) ? r + lo : (r > lo ? r : lo);
  }
  out[i] = r;
}

// ------------------------------------------------------------------ GMRES kernels
// The low-rank correction's dots folded into a pass that already walks the vector x that the next
// Arnoldi column is formed from (k_gm_init: x = b; k_gm_orth: x = w after orthogonalisation):
// per thread cl[j] = sum Ud_j[i] x_i for j < dk, reduced to per-block partial rows row0 .. row0 +
// dk - 1 (stride RMAXB; the rows before are the pass's own sums), then, in the last block, c_j =
// row sums scaled by 1 / |x| and d = Dd c, which the next column's k_defl_axpy uses (Q v = v + Ud d
// with v = x / |x|).  c_j = Ud_j . v is also kept per basis vector (uv[j][k], UVS per row), so
// that the pass that ends an Arnoldi column can form the next vector's dots without a pass of its
// own.
constexpr int UVS = MAX_RESTART + 1;   // row stride of uv (Ud_j . v_k per basis vector k)

__device__ inline void defl_block_partials(double (&cl)[DEFL_MAX], int dk, double* shd, double* partial,
                                           int row0) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    double o[DEFL_MAX];
#pragma unroll
    for (int j = 0; j < DEFL_MAX; ++j) o[j] = __shfl_xor(cl[j], off, 64);
#pragma unroll
    for (int j = 0; j < DEFL_MAX; ++j) cl[j] += o[j];
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int j = 0; j < DEFL_MAX; ++j) shd[wv * DEFL_MAX + j] = cl[j];
  __syncthreads();
  if ((int)threadIdx.x < dk) {
    double bs = 0.0;
    for (int q = 0; q < RBLK / 64; ++q) bs += shd[q * DEFL_MAX + threadIdx.x];
    store_partial(partial + (size_t)(row0 + threadIdx.x) * RMAXB + blockIdx.x, bs);
  }
}

// Last block: csd[j] = total of partial row 1 + j (j < dk), one wavefront per row.
__device__ inline void defl_last_rows(int dk, const double* __restrict__ partial, double* csd) {
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int r = wv; r < dk; r += RBLK / 64) {
    const double v = wave_sum_strided(partial + (size_t)(1 + r) * RMAXB, gridDim.x);
    if (lane == 0) csd[r] = v;
  }
}

// d = Dd (csd * inv) by threads 0 .. dk-1 (after a barrier that orders csd and inv).
__device__ inline void defl_form_d(int dk, const double* csd, double inv, const double* __restrict__ Dd,
                                   double* __restrict__ dvec) {
  const int r = threadIdx.x;
  if (r >= dk) return;
  double acc = 0.0;
  for (int j = 0; j < dk; ++j) acc = fma(Dd[r * DEFL_MAX + j], csd[j] * inv, acc);
  dvec[r] = acc;
}

// Initialisation: |b|, tolerances (scipy gmres with x0 = 0, M = I), V[0] unnormalised = b.  DEFL:
// the correction's dots of b as above (uv[j][0] = Ud_j . v_0).
template <bool DEFL>
__global__ void __launch_bounds__(RBLK) k_gm_init(int N, const double* __restrict__ b, double rtol, double atol,
                                                  GmresScalars* __restrict__ gs, double* __restrict__ partial,
                                                  unsigned* counter, int* __restrict__ progress,
                                                  const double* __restrict__ Ud, int dk,
                                                  const double* __restrict__ Dd, double* __restrict__ dvec,
                                                  double* __restrict__ uv) {
  __shared__ double sh[RBLK];
  __shared__ double shd[DEFL ? (RBLK / 64) * DEFL_MAX : 1];
  __shared__ double csd[DEFL_MAX];
  __shared__ double s_inv;
  __shared__ bool slast;
  double loc = 0.0;
  double cl[DEFL_MAX];
#pragma unroll
  for (int j = 0; j < DEFL_MAX; ++j) cl[j] = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    const double bi = b[i];
    loc += bi * bi;
    if constexpr (DEFL) {
#pragma unroll
      for (int j = 0; j < DEFL_MAX; ++j)
        if (j < dk) cl[j] += Ud[(size_t)j * N + i] * bi;
    }
  }
  double bs = wave_block_sum(loc, sh);
  if (threadIdx.x == 0) store_partial(partial + blockIdx.x, bs);
  if constexpr (DEFL) defl_block_partials(cl, dk, shd, partial, 1);
  if (!ticket_last(counter, &slast)) return;
  if constexpr (DEFL) defl_last_rows(dk, partial, csd);
  if (threadIdx.x < 64) {
    const double tot = wave_sum_strided(partial, gridDim.x);
    if (threadIdx.x == 0) {
      ticket_reset(counter);
      double bn = sqrt(tot);
      gs->bnorm = bn;
      double ae = atol > rtol * bn ? atol : rtol * bn;
      gs->atol_eff = ae;
      gs->iters = 0;
      gs->breakdown = 0;
      gs->presid = 0.0;
      gs->rnorm = bn;
      gs->done = (bn == 0.0 || bn < ae) ? 2 : 0;   // 2: exit before any iteration
      gs->ptol = bn > 0.0 ? bn * (1.0 < ae / bn ? 1.0 : ae / bn) : 0.0;
      gs->S[0] = bn;
      gs->inv_h1 = bn > 0.0 ? 1.0 / bn : 0.0;
      s_inv = gs->inv_h1;
      // pacing only: the host reads no data on the strength of this word, so relaxed suffices
      __hip_atomic_store(progress, gs->done ? -1 : 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
  if constexpr (DEFL) {
    __syncthreads();
    defl_form_d(dk, csd, s_inv, Dd, dvec);
    if ((int)threadIdx.x < dk) uv[threadIdx.x * UVS] = csd[threadIdx.x] * s_inv;
  }
}

// Classical Gram-Schmidt with one reorthogonalisation (CGS2) in three passes per Arnoldi step, two
// of them grid reductions: k_gm_dots (h = V^T w, |w|^2), k_gm_orth (w' = w - V h, h2 = V^T w',
// |w'|^2, and in its last block the end of the column), k_gm_project (w'' = w' - V h2, no
// reduction; with the low-rank correction the next column's k_defl_axpy forms w'' as it reads it).  The norm of w'' comes from |w''|^2 = |w'|^2 - |h2|^2 (V orthonormal), so no pass
// reduces over w''.  Every pass reads the K = col + 1 basis vectors once per element; per-vector
// block sums go through wavefront butterflies (no per-thread arrays indexed by K), and the last
// block sums the per-block partials with one wavefront per row.
constexpr int GMW = MAX_RESTART + 2;   // partial rows per reduction (K dots + |w|^2)
constexpr int GMC = 8;                 // basis vectors whose loads are in flight together

// Wavefront sums of V_k . w over this thread's elements for k < nk (w is stored as the column
// after the last basis vector, so k = col + 1 gives |w|^2); lane 0 writes them to out[k].
__device__ inline void block_dots(const double* __restrict__ Vb, size_t ld, const double* __restrict__ w, int nk,
                                  int N, int i0, int stride, double* out) {
  for (int k0 = 0; k0 < nk; k0 += GMC) {
    double p[GMC];
#pragma unroll
    for (int j = 0; j < GMC; ++j) p[j] = 0.0;
    for (int i = i0; i < N; i += stride) {
      const double wi = w[i];
      double v[GMC];
#pragma unroll
      for (int j = 0; j < GMC; ++j) v[j] = (k0 + j < nk) ? Vb[(size_t)(k0 + j) * ld + i] : 0.0;
#pragma unroll
      for (int j = 0; j < GMC; ++j) p[j] += v[j] * wi;
    }
    // GMC butterflies side by side: each step issues all its lane exchanges before waiting
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      double o[GMC];
#pragma unroll
      for (int j = 0; j < GMC; ++j) o[j] = __shfl_xor(p[j], off, 64);
#pragma unroll
      for (int j = 0; j < GMC; ++j) p[j] += o[j];
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
      for (int j = 0; j < GMC; ++j)
        if (k0 + j < nk) out[k0 + j] = p[j];
    }
  }
}

// w_i - sum_k h_k V_k[i], subtracting in k order, with GMC loads in flight.
__device__ inline double project_out(const double* __restrict__ Vb, size_t ld, const double* hk, int K, int i,
                                     double wi) {
  for (int k0 = 0; k0 < K; k0 += GMC) {
    double v[GMC];
#pragma unroll
    for (int j = 0; j < GMC; ++j) v[j] = (k0 + j < K) ? Vb[(size_t)(k0 + j) * ld + i] : 0.0;
#pragma unroll
    for (int j = 0; j < GMC; ++j)
      if (k0 + j < K) wi -= hk[k0 + j] * v[j];
  }
  return wi;
}

// Pass 1: h_k = V_k . w (k <= col), |w|^2.
__global__ void __launch_bounds__(RBLK) k_gm_dots(int N, int col, const double* __restrict__ Vb, size_t ld,
                                                  const double* __restrict__ w, GmresScalars* __restrict__ gs,
                                                  double* __restrict__ partial, unsigned* counter) {
  __shared__ double sh[(RBLK / 64) * GMW];
  __shared__ double tot[GMW];
  __shared__ bool slast;
  if (gs->done) return;
  const int K = col + 1;
  const int wv = threadIdx.x >> 6;
  const int i0 = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
  block_dots(Vb, ld, w, K + 1, N, i0, stride, sh + wv * GMW);
  __syncthreads();
  if ((int)threadIdx.x <= K) {
    double s = 0.0;
    for (int q = 0; q < RBLK / 64; ++q) s += sh[q * GMW + threadIdx.x];
    store_partial(partial + threadIdx.x * RMAXB + blockIdx.x, s);
  }
  if (last_block_totals(counter, &slast, partial, K + 1, tot) && (int)threadIdx.x <= K) {
    const int k = threadIdx.x;
    if (k < K) {
      gs->h_tmp[k] = tot[k];
      gs->H[k][col] = tot[k];
    } else {
      gs->h0sq = tot[K];
    }
  }
}

// LAPACK-style Givens rotation: [c s; -s c] [f; g] = [r; 0]
__device__ inline void lartg(double f, double g, double& c, double& s, double& r) {
  if (g == 0.0) { c = 1.0; s = 0.0; r = f; return; }
  if (f == 0.0) { c = 0.0; s = g > 0.0 ? 1.0 : -1.0; r = fabs(g); return; }
  double dd = sqrt(f * f + g * g);
  c = fabs(f) / dd;
  r = f >= 0.0 ? dd : -dd;
  s = g / r;
}

// Pass 2: w' = w - V h_tmp (stored over w); h2 = V . w' and |w'|^2; DEFL: Ud . w' folded in.
// Its last block then ends the column (scipy gmres inner loop): H[:, col] = h + h2 (kept raw in
// Hr), |w''| = sqrt(|w'|^2 - sum_k h2_k^2) with w'' = w' - V h2 (formed by k_gm_project or by the
// next column's k_defl_axpy), the Givens sweep of the column, the convergence test and the
// host-visible progress; DEFL: the next
// vector's dots Ud . w'' = Ud . w' - sum_k (Ud . v_k) h2_k, scaled by 1 / |w''|, into uv[., col + 1]
// and d = Dd c.  Everything the tail reads from gs is loaded at kernel start with the
// coefficients, so that the serial tail waits for no load.
template <bool DEFL>
__global__ void __launch_bounds__(RBLK) k_gm_orth(int N, int col, const double* __restrict__ Vb, size_t ld,
                                                  double* __restrict__ w, GmresScalars* __restrict__ gs,
                                                  double* __restrict__ partial, unsigned* counter,
                                                  int* __restrict__ progress, const double* __restrict__ Ud,
                                                  int dk, const double* __restrict__ Dd, double* __restrict__ dvec,
                                                  double* __restrict__ uv) {
  __shared__ double sh[(RBLK / 64) * GMW];
  __shared__ double shd[DEFL ? (RBLK / 64) * DEFL_MAX : 1];
  __shared__ double hk[GMW];
  __shared__ double tot[GMW + DEFL_MAX];
  __shared__ double hc[GMW], csl[GMW], snl[GMW];
  __shared__ double csd[DEFL_MAX];
  __shared__ double s_h0sq, s_S, s_ptol, s_inv;
  __shared__ bool slast;
  if (gs->done) return;
  const int K = col + 1;
  const int t = threadIdx.x, wv = t >> 6;
  if (t < K) {
    hk[t] = gs->h_tmp[t];
    hc[t] = gs->H[t][col];
  }
  if (t < col) {
    csl[t] = gs->cs[t];
    snl[t] = gs->sn[t];
  }
  if (t == 0) {
    s_h0sq = gs->h0sq;
    s_S = gs->S[col];
    s_ptol = gs->ptol;
  }
  __syncthreads();
  const int i0 = blockIdx.x * blockDim.x + t, stride = gridDim.x * blockDim.x;
  double cl[DEFL_MAX];
#pragma unroll
  for (int j = 0; j < DEFL_MAX; ++j) cl[j] = 0.0;
  for (int i = i0; i < N; i += stride) {
    // DEFL: the element's correction vectors are requested before its projection, whose loads
    // they then share a round trip with
    double ud[DEFL ? DEFL_MAX : 1];
    if constexpr (DEFL) {
#pragma unroll
      for (int j = 0; j < DEFL_MAX; ++j) ud[j] = j < dk ? Ud[(size_t)j * N + i] : 0.0;
    }
    const double wi = project_out(Vb, ld, hk, K, i, w[i]);
    w[i] = wi;
    if constexpr (DEFL) {
#pragma unroll
      for (int j = 0; j < DEFL_MAX; ++j)
        if (j < dk) cl[j] += ud[j] * wi;
    }
  }
  // rows 0 .. K-1: V_k . w'; row K: |w'|^2 (w' is stored as column K); DEFL rows K+1 ..: Ud . w'
  block_dots(Vb, ld, w, K + 1, N, i0, stride, sh + wv * GMW);
  __syncthreads();
  if (t <= K) {
    double sm = 0.0;
    for (int q = 0; q < RBLK / 64; ++q) sm += sh[q * GMW + t];
    store_partial(partial + t * RMAXB + blockIdx.x, sm);
  }
  if constexpr (DEFL) defl_block_partials(cl, dk, shd, partial, K + 1);
  if (!last_block_totals(counter, &slast, partial, K + 1 + (DEFL ? dk : 0), tot)) return;
  // last block: H[:, col] = h + h2 (raw copy in Hr), h2 for k_gm_project
  for (int k = t; k < K; k += blockDim.x) {
    const double hn = hc[k] + tot[k];
    gs->h_tmp[k] = tot[k];
    gs->Hr[k][col] = hn;
    hc[k] = hn;
  }
  __syncthreads();
  if (t == 0) {
    double s2 = 0.0;
    for (int k = 0; k < K; ++k) s2 += tot[k] * tot[k];
    const double h1sq = tot[K] - s2;
    const double h1 = h1sq > 0.0 ? sqrt(h1sq) : 0.0;
    const double h0 = sqrt(s_h0sq);
    const bool brk = h1 <= EPS64 * h0;
    hc[col + 1] = brk ? 0.0 : h1;
    gs->Hr[col + 1][col] = hc[col + 1];
    gs->inv_h1 = brk ? 0.0 : 1.0 / h1;
    s_inv = gs->inv_h1;
    for (int k = 0; k < col; ++k) {
      double c = csl[k], s = snl[k];
      double n0 = hc[k], n1 = hc[k + 1];
      hc[k] = c * n0 + s * n1;
      hc[k + 1] = -s * n0 + c * n1;
    }
    double c, s, mag;
    lartg(hc[col], hc[col + 1], c, s, mag);
    gs->cs[col] = c;
    gs->sn[col] = s;
    hc[col] = mag;
    hc[col + 1] = 0.0;
    double tmp = -s * s_S;
    gs->S[col] = c * s_S;
    gs->S[col + 1] = tmp;
    const double presid = fabs(tmp);
    gs->presid = presid;
    gs->iters = col + 1;
    if (brk) gs->breakdown = 1;
    const bool fin = presid <= s_ptol || brk;
    if (fin) gs->done = 1;
    // host-visible progress: completed columns, negative once the cycle is done
    __hip_atomic_store(progress, fin ? -(col + 1) : (col + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  __syncthreads();
  for (int k = t; k <= col + 1; k += blockDim.x) gs->H[k][col] = hc[k];
  if constexpr (DEFL) {
    if (t < dk) {
      double a = tot[K + 1 + t];
      for (int k = 0; k < K; ++k) a -= uv[t * UVS + k] * tot[k];
      csd[t] = a;
    }
    __syncthreads();
    defl_form_d(dk, csd, s_inv, Dd, dvec);
    if (t < dk && col + 1 < UVS) uv[t * UVS + col + 1] = csd[t] * s_inv;
  }
}

// Pass 3: w'' = w' - V h2 (the reorthogonalised vector the next column starts from), elementwise;
// nothing to do once the cycle is done.
__global__ void __launch_bounds__(256) k_gm_project(int N, int col, const double* __restrict__ Vb, size_t ld,
                                                    double* __restrict__ w, const GmresScalars* __restrict__ gs) {
  __shared__ double hk[GMW];
  if (gs->done) return;
  const int K = col + 1;
  if ((int)threadIdx.x < K) hk[threadIdx.x] = gs->h_tmp[threadIdx.x];
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) w[i] = project_out(Vb, ld, hk, K, i, w[i]);
}

// Back substitution of the least-squares problem (scipy's pseudo-solve of the triangular system).
// Back substitution H y = S (scipy's column-oriented loop, same operation order), one wavefront:
// the triangle is staged in LDS, lane i owns y_i.  Launch with one block of 64 threads.
__global__ void __launch_bounds__(64) k_gm_solve(GmresScalars* gs) {
  __shared__ double Hs[MAX_RESTART][MAX_RESTART + 1];
  __shared__ double ys[MAX_RESTART];
  if (gs->done == 2) return;
  const int col = gs->iters - 1;
  if (col < 0) return;
  const int lane = threadIdx.x;
  for (int e = lane; e < (col + 1) * (col + 1); e += 64) {
    const int i = e / (col + 1), k = e - i * (col + 1);
    Hs[i][k] = gs->H[i][k];
  }
  if (lane <= col) ys[lane] = (lane == col && gs->H[col][col] == 0.0) ? 0.0 : gs->S[lane];
  __syncthreads();
  for (int k = col; k > 0; --k) {
    const double yk = ys[k];
    if (yk != 0.0) {
      const double q = yk / Hs[k][k];
      if (lane < k) ys[lane] -= q * Hs[lane][k];
      __syncthreads();
      if (lane == 0) ys[k] = q;
    }
    __syncthreads();
  }
  if (lane == 0 && ys[0] != 0.0) ys[0] /= Hs[0][0];
  __syncthreads();
  if (lane <= col) gs->yv[lane] = ys[lane];
}

// x = sum_k y_k V_k
__global__ void k_gm_update(int N, const double* __restrict__ Vb, size_t ld, const GmresScalars* __restrict__ gs,
                            double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double acc = 0.0;
  if (gs->done != 2) {
    const int K = gs->iters;
    for (int k = 0; k < K; ++k) acc += gs->yv[k] * Vb[k * ld + i];
  }
  x[i] = acc;
}

// r = b - A P^-1 x and |r|
__global__ void __launch_bounds__(RBLK) k_gm_resid(int N, const double* __restrict__ b, const double* __restrict__ ax,
                                                   GmresScalars* __restrict__ gs, double* __restrict__ partial,
                                                   unsigned* counter, double* __restrict__ hout) {
  __shared__ double sh[RBLK];
  __shared__ bool slast;
  double loc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    double r = b[i] - ax[i];
    loc += r * r;
  }
  double bs = wave_block_sum(loc, sh);
  if (threadIdx.x == 0) store_partial(partial + blockIdx.x, bs);
  double tot;
  if (last_block_total(counter, &slast, partial, &tot)) {
    const double rn = sqrt(tot);
    gs->rnorm = rn;
    // the cycle's outcome for the host (mapped host memory): rnorm, bnorm, presid, atol_eff, done,
    // iterations, breakdown
    hout[0] = rn;
    hout[1] = gs->bnorm;
    hout[2] = gs->presid;
    hout[3] = gs->atol_eff;
    hout[4] = (double)gs->done;
    hout[5] = (double)gs->iters;
    hout[6] = (double)gs->breakdown;
  }
}

// ------------------------------------------------------------------ host-side drivers
static inline int eblk(long long N) { return (int)((N + 255) / 256); }

int reduce_rms(Handle& h, const double* a, const double* s, int len, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_rms, dim3(rgrid(len)), dim3(RBLK), 0, st, len, a, s, h.red.partial, h.red.counter + 2,
                     h.red.result);
  if (check(hipStreamSynchronize(st), "sync")) return -1;
  *out = h.host_scalars[0];
  return 0;
}

void launch_residual_combine(Handle& h, const double* z, const double* f, double dh, const double* rhs_known,
                             double* R, hipStream_t st) {
  hipLaunchKernelGGL(k_residual_combine, dim3(rgrid(h.N)), dim3(RBLK), 0, st, h.N, z, f, dh, rhs_known,
                     (const double*)nullptr, R, h.red.partial, h.red.counter + 2, h.red.result);
}

// The same reduction with its results left in device memory (inventories: mode 1, 7 values).
int reduce_weighted_dev(Handle& h, const double* a, int mode, double* result, hipStream_t st) {
  hipLaunchKernelGGL(k_weighted, dim3(rgrid(h.n + h.nw)), dim3(RBLK), 0, st, h.m, mode, a, nullptr, h.red.partial,
                     h.red.counter + 3, result);
  return check(hipGetLastError(), "k_weighted");
}

int reduce_weighted(Handle& h, const double* a, const double* w, int len, int mode, double* out, hipStream_t st) {
  (void)len;
  hipLaunchKernelGGL(k_weighted, dim3(rgrid(h.n + h.nw)), dim3(RBLK), 0, st, h.m, mode, a, w, h.red.partial,
                     h.red.counter + 3, h.red.result);
  const int nout = mode == 0 ? 8 : (mode == 1 ? 7 : 4);
  if (check(hipStreamSynchronize(st), "sync")) return -1;
  for (int q = 0; q < nout; ++q) out[q] = h.host_scalars[q];
  return 0;
}

// W R, W s of mode 0 into device slot q (result[16 + 8 q ..]); fetched together by cons_fetch_impl.
void cons_stash_impl(Handle& h, const double* R, const double* s, int slot, hipStream_t st) {
  hipLaunchKernelGGL(k_weighted, dim3(rgrid(h.n + h.nw)), dim3(RBLK), 0, st, h.m, 0, R, s, h.red.partial,
                     h.red.counter + 3, h.red.result + 16 + 8 * slot);
}

int cons_fetch_impl(Handle& h, int n, double* out, hipStream_t st) {
  if (n <= 0) return 0;
  if (check(hipStreamSynchronize(st), "sync")) return -1;
  for (int q = 0; q < 8 * n; ++q) out[q] = h.host_scalars[16 + q];
  return 0;
}

int cons_shift_impl(Handle& h, const double* z, const double* s, const double lam[4], int mode, double* out,
                    hipStream_t st) {
  // kernel truncate argument: 0 none, 1 below -s, 3 below 0
  const int truncate = (mode & VTSA_SHIFT_TRUNCATE) ? ((mode & VTSA_SHIFT_AT_ZERO) ? 3 : 1) : 0;
  if (mode & VTSA_SHIFT_NO_REPORT) {   // the flag is written but not read
    hipLaunchKernelGGL(k_cons_shift, dim3(eblk(h.N)), dim3(256), 0, st, h.m, z, s, lam[0], lam[1], lam[2], lam[3],
                       out, truncate, h.red.counter + 12);
    return 0;
  }
  unsigned* flag = h.red.counter + 12;
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, flag, 0u, 1);
  hipLaunchKernelGGL(k_cons_shift, dim3(eblk(h.N)), dim3(256), 0, st, h.m, z, s, lam[0], lam[1], lam[2], lam[3], out,
                     truncate, flag);
  unsigned* host = reinterpret_cast<unsigned*>(h.host_scalars + 63);
  if (check(hipMemcpyAsync(host, flag, sizeof(unsigned), hipMemcpyDeviceToHost, st), "memcpy")) return -1;
  if (check(hipStreamSynchronize(st), "sync")) return -1;
  return *host != 0u ? 1 : 0;
}

// Newton building blocks (host drivers; each returns host scalars)
int newton_residual_impl(Handle& h, double t, const double* z, double dh, const double* rk, const double* s,
                         double* f, double* R, double* rn, hipStream_t st) {
  launch_props(h, z, h.pb, f, nullptr, st);
  launch_flux(h, t, z, h.pb, f, st);
  hipLaunchKernelGGL(k_residual_combine, dim3(rgrid(h.N)), dim3(RBLK), 0, st, h.N, z, f, dh, rk, s, R,
                     h.red.partial, h.red.counter + 2, h.red.result + 0);
  if (!rn) return 0;
  if (check(hipStreamSynchronize(st), "sync")) return -1;
  *rn = h.host_scalars[0];
  return 0;
}

int newton_eval_impl(Handle& h, double t, const double* z, double dh, const double* rk, const double* s, double* f,
                     double* R, double* rn, double* pn, hipStream_t st) {
  launch_props(h, z, h.pb, f, nullptr, st);
  launch_flux(h, t, z, h.pb, f, st);
  hipLaunchKernelGGL(k_residual_combine, dim3(rgrid(h.N)), dim3(RBLK), 0, st, h.N, z, f, dh, rk, s, R,
                     h.red.partial, h.red.counter + 2, h.red.result + 0);
  ApplyArgs a{R, nullptr, nullptr, nullptr, nullptr, h.vtmp, s, h.red.result + 1, nullptr, nullptr, 0};
  apply_launch(h, a, st);
  if (check(hipStreamSynchronize(st), "sync")) return -1;
  *rn = h.host_scalars[0];
  *pn = h.host_scalars[1];
  return 0;
}

int newton_try_impl(Handle& h, double t, const double* z, const double* dz, double lam, double dh, const double* rk,
                    const double* s, double* zt, double* ft, double* Rt, double* rn, double* pn, double* step,
                    hipStream_t st) {
  hipLaunchKernelGGL(k_newton_update, dim3(rgrid(h.N)), dim3(RBLK), 0, st, h.N, 5 * h.n, z, dz, lam, s, zt,
                     h.red.partial + RMAXB * 40, h.red.counter + 4, h.red.result + 2);
  launch_props(h, zt, h.pb, ft, nullptr, st);
  launch_flux(h, t, zt, h.pb, ft, st);
  hipLaunchKernelGGL(k_residual_combine, dim3(rgrid(h.N)), dim3(RBLK), 0, st, h.N, zt, ft, dh, rk, s, Rt,
                     h.red.partial, h.red.counter + 2, h.red.result + 0);
  ApplyArgs a{Rt, nullptr, nullptr, nullptr, nullptr, h.vtmp, s, h.red.result + 1, nullptr, nullptr, 0};
  apply_launch(h, a, st);
  if (check(hipStreamSynchronize(st), "sync")) return -1;
  *rn = h.host_scalars[0];
  *pn = h.host_scalars[1];
  *step = h.host_scalars[2];
  return 0;
}

void scale_impl(Handle& h, const double* z, double rtol, double* s, hipStream_t st) {
  hipLaunchKernelGGL(k_scale, dim3(eblk(h.N)), dim3(256), 0, st, h.m, z, rtol, s);
}

int error_norm_impl(Handle& h, const double* y, const double* z2, const double* y3, double rtol, double* en,
                    hipStream_t st) {
  hipLaunchKernelGGL(k_vec_op, dim3(eblk(h.N)), dim3(256), 0, st, 7, h.N, 0, 0.0, 0.0, 0.0, 0.0, y3, z2,
                     (const double*)nullptr, (const double*)nullptr, h.vtmp);
  hipLaunchKernelGGL(k_scale2, dim3(eblk(h.N)), dim3(256), 0, st, h.m, y, z2, rtol, h.vtmp2);
  ApplyArgs a{h.vtmp, nullptr, nullptr, nullptr, nullptr, h.vtmp, h.vtmp2, h.red.result + 3, nullptr, nullptr, 0};
  apply_launch(h, a, st);
  if (check(hipStreamSynchronize(st), "sync")) return -1;
  *en = h.host_scalars[3];
  return 0;
}

void vec_op_impl(Handle& h, int op, int npos, double a, double b, double c, double e, const double* x,
                 const double* y, const double* z, const double* w, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_vec_op, dim3(eblk(h.N)), dim3(256), 0, st, op, h.N, npos, a, b, c, e, x, y, z, w, out);
}

// ------------------------------------------------------------------ low-rank correction
// Q v = v + Ud (Dd (Ud^T v)) in the scaled variables of GMRES.  Ud holds orthonormal approximations
// of the invariant subspace of the right-preconditioned operator B = S^-1 A M^-1 S that belongs to
// its eigenvalues nearest zero (harmonic Ritz vectors of an earlier cycle); with
// Dd = T^-1 - I, T = Ud^T B Ud, the operator B Q maps that subspace onto itself with eigenvalues
// near 1, so GMRES no longer spends iterations on those modes.
//
// k_defl_dots: v = in * scale (stored to `store` when given), c_j = Ud_j . v for j < K; the last
// block forms d = Dd c.  The K block sums share one exchange through LDS.  With respect_done, a
// queued call after the cycle converged returns at once.
template <int K>
__global__ void __launch_bounds__(RBLK) k_defl_dots(int N, const double* in, const double* __restrict__ in_scale,
                                                    double* store, const double* __restrict__ Ud,
                                                    const double* __restrict__ Dd, double* __restrict__ d,
                                                    const GmresScalars* __restrict__ gs, int respect_done,
                                                    double* __restrict__ partial, unsigned* counter) {
  __shared__ double sh[(RBLK / 64) * K];
  __shared__ double cs[K];
  __shared__ bool slast;
  if (respect_done && gs->done) return;
  const double sc = in_scale ? *in_scale : 1.0;
  double loc[K];
#pragma unroll
  for (int j = 0; j < K; ++j) loc[j] = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    const double v = in[i] * sc;
    if (store) store[i] = v;
    double u[K];
#pragma unroll
    for (int j = 0; j < K; ++j) u[j] = Ud[(size_t)j * N + i];
#pragma unroll
    for (int j = 0; j < K; ++j) loc[j] = fma(u[j], v, loc[j]);
  }
  // K butterflies side by side, then one LDS exchange of the per-wave sums
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    double o[K];
#pragma unroll
    for (int j = 0; j < K; ++j) o[j] = __shfl_xor(loc[j], off, 64);
#pragma unroll
    for (int j = 0; j < K; ++j) loc[j] += o[j];
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int j = 0; j < K; ++j) sh[wv * K + j] = loc[j];
  __syncthreads();
  if ((int)threadIdx.x < K) {
    double bs = 0.0;
    for (int w = 0; w < RBLK / 64; ++w) bs += sh[w * K + threadIdx.x];
    store_partial(partial + (size_t)threadIdx.x * RMAXB + blockIdx.x, bs);
  }
  if (!last_block_totals(counter, &slast, partial, K, cs)) return;
  if ((int)threadIdx.x < K) {
    const int r = threadIdx.x;
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < K; ++j) acc = fma(Dd[r * DEFL_MAX + j], cs[j], acc);
    d[r] = acc;
  }
}

// out = v + Ud d with v = in * scale, v also stored to `store` when given (out and store may alias
// in).
// K > 0: `in` is the previous Arnoldi column's w' and is first reorthogonalised, w'' = w' -
// sum_k h2_k V_k (k < K, h2 = gs->h_tmp), with the arithmetic of k_gm_project, which the GMRES
// cycle then skips for that column.
__global__ void __launch_bounds__(256) k_defl_axpy(int N, int k, const double* in, const double* __restrict__ in_scale,
                                                   double* store, const double* __restrict__ Ud,
                                                   const double* __restrict__ d, double* out,
                                                   const GmresScalars* __restrict__ gs, int respect_done,
                                                   const double* __restrict__ Vb, size_t ld, int K) {
  __shared__ double hk[GMW];
  if (respect_done && gs->done) return;
  if (K > 0) {
    if ((int)threadIdx.x < K) hk[threadIdx.x] = gs->h_tmp[threadIdx.x];
    __syncthreads();
  }
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double x = in[i];
  if (K > 0) x = project_out(Vb, ld, hk, K, i, x);
  double v = x * (in_scale ? *in_scale : 1.0);
  if (store) store[i] = v;
  for (int j = 0; j < k; ++j) v = fma(Ud[(size_t)j * N + i], d[j], v);
  out[i] = v;
}

// Ud_j = sum_i V_i G[i][j] over the first m basis vectors (G row-major m x k, ld DEFL_MAX).
__global__ void k_defl_build(int N, int m, int k, const double* __restrict__ Vb, size_t ld,
                             const double* __restrict__ G, double* __restrict__ Ud) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double acc[DEFL_MAX];
#pragma unroll
  for (int j = 0; j < DEFL_MAX; ++j) acc[j] = 0.0;
  for (int r = 0; r < m; ++r) {
    const double v = Vb[(size_t)r * ld + i];
#pragma unroll
    for (int j = 0; j < DEFL_MAX; ++j)
      if (j < k) acc[j] = fma(v, G[r * DEFL_MAX + j], acc[j]);
  }
  for (int j = 0; j < k; ++j) Ud[(size_t)j * N + i] = acc[j];
}

int defl_set_impl(Handle& h, int k, int m, const double* G, const double* D, hipStream_t st) {
  if (k == 0) {
    h.defl_k = 0;
    return 0;
  }
  if (!h.Ud) {
    if (check(hipMalloc((void**)&h.Ud, sizeof(double) * (size_t)DEFL_MAX * h.N), "hipMalloc")) return -1;
    if (check(hipMalloc((void**)&h.defl_small, sizeof(double) * (DEFL_MAX * DEFL_MAX + DEFL_MAX +
                                                                  MAX_RESTART * DEFL_MAX + DEFL_MAX * UVS)),
              "hipMalloc"))
      return -1;
    if (check(hipMalloc((void**)&h.defl_partial, sizeof(double) * DEFL_MAX * RMAXB), "hipMalloc")) return -1;
    if (check(hipMalloc((void**)&h.defl_counter, sizeof(unsigned) * 2), "hipMalloc")) return -1;
    if (check(hipMemset(h.defl_counter, 0, sizeof(unsigned) * 2), "hipMemset")) return -1;
    if (check(hipMalloc((void**)&h.defl_tmp, sizeof(double) * h.N), "hipMalloc")) return -1;
  }
  std::vector<double> small(DEFL_MAX * DEFL_MAX + DEFL_MAX + MAX_RESTART * DEFL_MAX, 0.0);
  for (int r = 0; r < k; ++r)
    for (int c = 0; c < k; ++c) small[r * DEFL_MAX + c] = D[r * k + c];
  double* Gp = small.data() + DEFL_MAX * DEFL_MAX + DEFL_MAX;
  for (int r = 0; r < m; ++r)
    for (int c = 0; c < k; ++c) Gp[r * DEFL_MAX + c] = G[r * k + c];
  if (check(hipMemcpyAsync(h.defl_small, small.data(), sizeof(double) * small.size(), hipMemcpyHostToDevice, st),
            "memcpy"))
    return -1;
  // the copy reads host memory that goes out of scope: wait for it before returning
  if (check(hipStreamSynchronize(st), "sync")) return -1;
  hipLaunchKernelGGL(k_defl_build, dim3(eblk(h.N)), dim3(256), 0, st, h.N, m, k, h.Vb, (size_t)h.N,
                     h.defl_small + DEFL_MAX * DEFL_MAX + DEFL_MAX, h.Ud);
  h.defl_k = k;
  return check(hipGetLastError(), "k_defl_build");
}

void defl_apply_launch(Handle& h, const double* in, const double* in_scale, double* store, double* out,
                       int respect_done, hipStream_t st) {
  double* d = h.defl_small + DEFL_MAX * DEFL_MAX;
  const dim3 grid(rgrid(h.N)), block(RBLK);
  switch (h.defl_k) {
#define VTSA_DEFL_DOTS(K)                                                                                   \
  case K:                                                                                                   \
    hipLaunchKernelGGL(k_defl_dots<K>, grid, block, 0, st, h.N, in, in_scale, store, h.Ud, h.defl_small, d, \
                       h.gs, respect_done, h.defl_partial, h.defl_counter);                                 \
    break;
    VTSA_DEFL_DOTS(1) VTSA_DEFL_DOTS(2) VTSA_DEFL_DOTS(3) VTSA_DEFL_DOTS(4) VTSA_DEFL_DOTS(5) VTSA_DEFL_DOTS(6)
    VTSA_DEFL_DOTS(7) VTSA_DEFL_DOTS(8) VTSA_DEFL_DOTS(9) VTSA_DEFL_DOTS(10) VTSA_DEFL_DOTS(11)
    VTSA_DEFL_DOTS(12) VTSA_DEFL_DOTS(13) VTSA_DEFL_DOTS(14) VTSA_DEFL_DOTS(15) VTSA_DEFL_DOTS(16)
#undef VTSA_DEFL_DOTS
    default: return;   // defl_set_impl admits 1 .. DEFL_MAX only
  }
  const double* v = store ? store : in;
  const double* vs = store ? nullptr : in_scale;
  hipLaunchKernelGGL(k_defl_axpy, dim3(eblk(h.N)), dim3(256), 0, st, h.N, h.defl_k, v, vs, (double*)nullptr, h.Ud, d,
                     out, h.gs, respect_done, (const double*)nullptr, (size_t)0, 0);
}

int gmres_cycle(Handle& h, double t, const double* z, const double* f, const double* s, double dh, double zs_norm,
                const double* b, double* w, double* pw, double rtol, double atol, int restart,
                vtsa_gmres_info* info, hipStream_t st) {
  if (restart > MAX_RESTART) restart = MAX_RESTART;
  if (restart > h.N) restart = h.N;
  const int N = h.N;
  const size_t ld = (size_t)N;
  double* Vb = h.Vb;
  double* U = h.vtmp2;
  const int g = rgrid(N);
  // The host queues at most GM_LOOKAHEAD Arnoldi columns beyond the last one the device reports
  // complete (a coherent host word written by k_gm_init / k_gm_orth), and stops queueing once
  // the device reports convergence; queued columns past convergence return at once.
  constexpr int GM_LOOKAHEAD = 2;
  constexpr int NOT_STARTED = -0x7fffffff;
  __atomic_store_n(h.progress, NOT_STARTED, __ATOMIC_RELEASE);
  // rms(z/s) for the matvec's finite-difference step: given by the caller, or (zs_norm < 0) here
  if (zs_norm < 0.0)
    hipLaunchKernelGGL(k_rms, dim3(rgrid(N)), dim3(RBLK), 0, st, N, z, s, h.red.partial, h.red.counter + 2,
                       &h.gs->zs_rms);
  else
    hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(64), 0, st, &h.gs->zs_rms, zs_norm);
  // with the low-rank correction, the dots of each next Arnoldi vector ride on the pass that forms
  // it (k_gm_init for b, k_gm_orth for w), so a column's correction is one axpy
  const int dk = h.defl_k;
  double* dvec = h.defl_small + DEFL_MAX * DEFL_MAX;
  double* uv = dk > 0 ? h.defl_small + DEFL_MAX * DEFL_MAX + DEFL_MAX + MAX_RESTART * DEFL_MAX : nullptr;
  if (dk > 0)
    hipLaunchKernelGGL(k_gm_init<true>, dim3(g), dim3(RBLK), 0, st, N, b, rtol, atol, h.gs, h.red.partial,
                       h.red.counter + 5, h.progress_dev, (const double*)h.Ud, dk, (const double*)h.defl_small, dvec,
                       uv);
  else
    hipLaunchKernelGGL(k_gm_init<false>, dim3(g), dim3(RBLK), 0, st, N, b, rtol, atol, h.gs, h.red.partial,
                       h.red.counter + 5, h.progress_dev, (const double*)nullptr, 0, (const double*)nullptr,
                       (double*)nullptr, (double*)nullptr);
  if (check(hipGetLastError(), "k_gm_init")) return -1;
  for (int col = 0; col < restart; ++col) {
    if (col >= GM_LOOKAHEAD) {
      const auto t0 = std::chrono::steady_clock::now();
      int v;
      for (;;) {
        v = __atomic_load_n(h.progress, __ATOMIC_ACQUIRE);
        if (v != NOT_STARTED && (v < 0 || v >= col - GM_LOOKAHEAD + 1)) break;
        if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(120)) {
          set_error("vtsa_gmres: no device progress for 120 s");
          return -1;
        }
        std::this_thread::yield();
      }
      if (v < 0) break;
    }
    // u = P^-1 v_col with v_col = w_prev / |w_prev| normalised in place (u = P^-1 Q v_col with the
    // low-rank correction, whose axpy also forms w_prev = w' - V h2 of the previous column)
    const double* vin = col == 0 ? b : Vb + (size_t)col * ld;
    if (dk > 0) {
      hipLaunchKernelGGL(k_defl_axpy, dim3(eblk(N)), dim3(256), 0, st, N, dk, vin, &h.gs->inv_h1,
                         Vb + (size_t)col * ld, (const double*)h.Ud, (const double*)dvec, h.defl_tmp, h.gs, 1,
                         (const double*)Vb, ld, col);
      ApplyArgs a{h.defl_tmp, nullptr, nullptr, s, s, U, nullptr, nullptr, h.gs, &h.gs->zs_rms, 1};
      apply_launch(h, a, st);
    } else {
      ApplyArgs a{vin, &h.gs->inv_h1, Vb + (size_t)col * ld, s, s, U, nullptr, nullptr, h.gs, &h.gs->zs_rms, 1};
      apply_launch(h, a, st);
    }
    double* wcol = Vb + (size_t)(col + 1) * ld;
    launch_matvec_residual(h, t, z, U, s, f, dh, wcol, true, st);
    hipLaunchKernelGGL(k_gm_dots, dim3(g), dim3(RBLK), 0, st, N, col, Vb, ld, wcol, h.gs, h.red.partial,
                       h.red.counter + 5);
    if (dk > 0)
      hipLaunchKernelGGL(k_gm_orth<true>, dim3(g), dim3(RBLK), 0, st, N, col, Vb, ld, wcol, h.gs, h.red.partial,
                         h.red.counter + 5, h.progress_dev, (const double*)h.Ud, dk, (const double*)h.defl_small,
                         dvec, uv);
    else
      hipLaunchKernelGGL(k_gm_orth<false>, dim3(g), dim3(RBLK), 0, st, N, col, Vb, ld, wcol, h.gs, h.red.partial,
                         h.red.counter + 5, h.progress_dev, (const double*)nullptr, 0, (const double*)nullptr,
                         (double*)nullptr, (double*)nullptr);
    // w'' = w' - V h2; with the correction the next column's axpy forms it instead
    if (dk == 0 || col + 1 >= restart)
      hipLaunchKernelGGL(k_gm_project, dim3(eblk(N)), dim3(256), 0, st, N, col, Vb, ld, wcol, h.gs);
  }
  hipLaunchKernelGGL(k_gm_solve, dim3(1), dim3(64), 0, st, h.gs);
  hipLaunchKernelGGL(k_gm_update, dim3(eblk(N)), dim3(256), 0, st, N, Vb, ld, h.gs, w);
  // with the low-rank correction the solution is x = P^-1 Q w: w is returned as Q w
  if (h.defl_k > 0) defl_apply_launch(h, w, nullptr, nullptr, w, 0, st);
  // true residual r = b - A P^-1 w (one more matvec, as scipy does after the cycle); the apply
  // also leaves M^-1 (w*s) in pw when asked, the step the caller would form with one more apply
  {
    ApplyArgs a{w, nullptr, nullptr, s, s, U, nullptr, nullptr, h.gs, &h.gs->zs_rms, 0, pw};
    apply_launch(h, a, st);
  }
  launch_matvec_residual(h, t, z, U, s, f, dh, h.vtmp, false, st);
  const double* out = h.host_scalars + 48;
  hipLaunchKernelGGL(k_gm_resid, dim3(g), dim3(RBLK), 0, st, N, b, h.vtmp, h.gs, h.red.partial, h.red.counter + 5,
                     h.red.result + 48);
  h.gs_stream = st;
  if (check(hipStreamSynchronize(st), "sync")) return -1;
  const int done_flag = (int)out[4];
  info->iters = (int)out[5];
  info->breakdown = (int)out[6];
  info->bnorm = out[1];
  info->presid = out[2];
  info->atol_eff = out[3];
  if (done_flag == 2) {   // exited before the first iteration: x = 0 (or b == 0)
    info->rnorm = out[1];
    info->info = 0;
    info->iters = 0;
  } else {
    info->rnorm = out[0];
    info->info = (out[0] <= out[3]) ? 0 : 1;
  }
  return 0;
}

}  // namespace vtsa

// ------------------------------------------------------------------ no-allocation rms
// vtsa_rms_noalloc: the reduction of vtsa_rms (same kernel, same grid, so the same bits) with its
// per-block partials and ticket counter in a file-scope device array instead of a handle's
// buffers, and the result written to a caller-owned device double.  Each device holds its own
// copy of the array; its address is looked up once per device (hipGetSymbolAddress) with that
// device current, and only that address is handed to kernels.  The counter must start at zero:
// the first call on a device clears the array with a fill kernel on the caller's stream; every
// later reduction leaves the counter at zero again (its last block resets it).
namespace vtsa {
namespace {
constexpr int NOALLOC_PARTIALS = RMAXB;                     // one double per block
constexpr int NOALLOC_DOUBLES = NOALLOC_PARTIALS + 32;      // + the ticket counter, padded
constexpr int NOALLOC_MAX_DEVICES = 64;
__device__ __align__(256) double scratch[NOALLOC_DOUBLES];
std::mutex noalloc_mu;
double* noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

__global__ void k_noalloc_clear(double* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// Device address of this device's scratch array; cleared on first use.  nullptr on error.
double* noalloc_scratch(hipStream_t st) {
  int dev = -1;
  if (check(hipGetDevice(&dev), "hipGetDevice")) return nullptr;
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
    set_error("vtsa_rms_noalloc: device index out of range");
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(noalloc_mu);
  if (!noalloc_ptr[dev]) {
    double* p = nullptr;
    if (check(hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress")) return nullptr;
    hipLaunchKernelGGL(k_noalloc_clear, dim3((NOALLOC_DOUBLES + 255) / 256), dim3(256), 0, st, p, NOALLOC_DOUBLES);
    if (check(hipGetLastError(), "clear")) return nullptr;
    noalloc_ptr[dev] = p;
  }
  return noalloc_ptr[dev];
}
}  // namespace

}  // namespace vtsa

extern "C" int vtsa_rms_noalloc(const double* a, const double* s, int32_t len, double* out, vtsa_stream st) {
  using namespace vtsa;
  if (!a || !out) { set_error("vtsa_rms_noalloc: null argument"); return -1; }
  if (len < 1 || len > VTSA_NOALLOC_MAX_LEN) {
    set_error("vtsa_rms_noalloc: len must be in 1 .. VTSA_NOALLOC_MAX_LEN (2^30)");
    return -2;
  }
  hipStream_t hs = (hipStream_t)st;
  double* p = noalloc_scratch(hs);
  if (!p) return -1;
  static_assert(sizeof(unsigned) <= sizeof(double), "counter slot");
  hipLaunchKernelGGL(k_rms, dim3(rgrid(len)), dim3(RBLK), 0, hs, (int)len, a, s, p,
                     reinterpret_cast<unsigned*>(p + NOALLOC_PARTIALS), out);
  return check(hipGetLastError(), "vtsa_rms_noalloc launch") ? -1 : 0;
}

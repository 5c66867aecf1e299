// Preconditioner, reductions, Newton vector operations and the device-resident GMRES cycle.
//
// Preconditioner M ~ I - dh J in variables scaled by s0 (M~ = S0^-1 M S0):
//  * z-line: for every (r, theta) column the block-pentadiagonal matrix of the 6x6 cell blocks
//    coupling k with k-2..k+2 (MUSCL stencil) is factored by block LU without inter-block
//    pivoting, M~ = L U, storing per cell L_{k,k-2}, L_{k,k-1}, D_k^-1 (Gauss-Jordan with partial
//    pivoting), U_{k,k+1}, U_{k,k+2}.  The r and theta couplings (weak except next to the axis)
//    are left to GMRES.
//  * block Jacobi: D_k^-1 of the diagonal blocks only.
//  * wall rows: point Jacobi on the wall diagonal.
// The apply kernel gives each column six lanes (one per variable) of a 64-lane wavefront
// (10 columns per wavefront); the sweep along z exchanges vector components with wavefront
// shuffles, so a column's forward and backward substitution needs no LDS or barriers.
//
// Reductions use one launch: per-block partial sums, then the last block to finish (ticket
// counter) adds the partials in block order, so results are deterministic.
#include "vtsa_internal.h"

namespace vtsa {

constexpr int RBLK = 256;        // reduction block size
constexpr int RMAXB = 512;       // max blocks of a reduction grid
constexpr int PW = 64;           // preconditioner apply: threads per block (one wavefront)
constexpr int PCOLS = 10;        // columns per wavefront (6 lanes each)
constexpr double EPS64 = 2.220446049250313e-16;

// ------------------------------------------------------------------ reduction helpers
__device__ inline double wave_block_sum(double v, double* sh) {
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

// Returns true in the last block to finish; that block may then read every block's partials.
__device__ inline bool ticket_last(unsigned* counter, bool* sflag) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    unsigned tk = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    *sflag = (tk == gridDim.x - 1);
  }
  __syncthreads();
  return *sflag;
}

__device__ inline void ticket_reset(unsigned* counter) {
  __hip_atomic_store(counter, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

static inline int rgrid(long long work) {
  long long b = (work + RBLK - 1) / RBLK;
  if (b < 1) b = 1;
  if (b > RMAXB) b = RMAXB;
  return (int)b;
}

// ------------------------------------------------------------------ 6x6 helpers
template <typename T>
__device__ inline void mat_inv6(T* A, T* Ainv) {
  // Gauss-Jordan with partial pivoting, A is destroyed.
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) Ainv[r * 6 + c] = (r == c) ? T(1) : T(0);
  for (int p = 0; p < 6; ++p) {
    int piv = p;
    T best = fabs(A[p * 6 + p]);
    for (int r = p + 1; r < 6; ++r) {
      T v = fabs(A[r * 6 + p]);
      if (v > best) { best = v; piv = r; }
    }
    if (piv != p) {
      for (int c = 0; c < 6; ++c) {
        T a = A[p * 6 + c]; A[p * 6 + c] = A[piv * 6 + c]; A[piv * 6 + c] = a;
        T b = Ainv[p * 6 + c]; Ainv[p * 6 + c] = Ainv[piv * 6 + c]; Ainv[piv * 6 + c] = b;
      }
    }
    T inv = T(1) / A[p * 6 + p];
    for (int c = 0; c < 6; ++c) { A[p * 6 + c] *= inv; Ainv[p * 6 + c] *= inv; }
    for (int r = 0; r < 6; ++r) {
      if (r == p) continue;
      T fct = A[r * 6 + p];
      if (fct == T(0)) continue;
      for (int c = 0; c < 6; ++c) { A[r * 6 + c] -= fct * A[p * 6 + c]; Ainv[r * 6 + c] -= fct * Ainv[p * 6 + c]; }
    }
  }
}

// C = A * B
template <typename T>
__device__ inline void mat_mul6(const T* A, const T* B, T* C) {
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) {
      T acc = T(0);
      for (int m = 0; m < 6; ++m) acc += A[r * 6 + m] * B[m * 6 + c];
      C[r * 6 + c] = acc;
    }
}
// C -= A * B
template <typename T>
__device__ inline void mat_submul6(const T* A, const T* B, T* C) {
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) {
      T acc = T(0);
      for (int m = 0; m < 6; ++m) acc += A[r * 6 + m] * B[m * 6 + c];
      C[r * 6 + c] -= acc;
    }
}

// Scaled block of M~ = S0^-1 (I - dh J) S0 between row cell X and column cell A (slot).
template <typename T>
__device__ inline void load_mblock(const double* __restrict__ Jc, const double* __restrict__ s0, int n, int X, int A,
                                   int slot, double dh, T* B) {
  const double* J = Jc + ((size_t)X * NSLOT + slot) * 36;
  for (int r = 0; r < 6; ++r) {
    double sr = s0[r * n + X];
    for (int c = 0; c < 6; ++c) {
      double v = (slot == S_SELF && r == c ? 1.0 : 0.0) - dh * J[r * 6 + c];
      B[r * 6 + c] = (T)(v * (s0[c * n + A] / sr));
    }
  }
}

// One thread per (r, theta) column: block LU along z.  F blocks per cell: 0 L2, 1 L1, 2 Dinv, 3 U1, 4 U2.
template <typename T>
__global__ void k_prec_factor(int n, int Nz, int ncol, int kind, const double* __restrict__ Jc,
                              const double* __restrict__ s0, double dh, T* __restrict__ F) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  T A[36], B[36], D[36];
  for (int k = 0; k < Nz; ++k) {
    const int X = col * Nz + k;
    T* FX = F + (size_t)X * 5 * 36;
    if (kind == VTSA_PREC_BLOCK_JACOBI) {
      load_mblock<T>(Jc, s0, n, X, X, S_SELF, dh, D);
      mat_inv6<T>(D, FX + 2 * 36);
      continue;
    }
    // L2 = A_{k,k-2} D_{k-2}^-1
    if (k >= 2) {
      load_mblock<T>(Jc, s0, n, X, X - 2, S_ZM2, dh, A);
      mat_mul6<T>(A, F + (size_t)(X - 2) * 180 + 2 * 36, FX + 0);
    }
    // L1 = (A_{k,k-1} - L2 U1_{k-2}) D_{k-1}^-1
    if (k >= 1) {
      load_mblock<T>(Jc, s0, n, X, X - 1, S_ZM1, dh, A);
      if (k >= 2) mat_submul6<T>(FX + 0, F + (size_t)(X - 2) * 180 + 3 * 36, A);
      mat_mul6<T>(A, F + (size_t)(X - 1) * 180 + 2 * 36, FX + 36);
    }
    // D = A_kk - L1 U1_{k-1} - L2 U2_{k-2}
    load_mblock<T>(Jc, s0, n, X, X, S_SELF, dh, D);
    if (k >= 1) mat_submul6<T>(FX + 36, F + (size_t)(X - 1) * 180 + 3 * 36, D);
    if (k >= 2) mat_submul6<T>(FX + 0, F + (size_t)(X - 2) * 180 + 4 * 36, D);
    for (int e = 0; e < 36; ++e) B[e] = D[e];
    mat_inv6<T>(B, FX + 2 * 36);
    // U1 = A_{k,k+1} - L1 U2_{k-1}
    if (k + 1 < Nz) {
      load_mblock<T>(Jc, s0, n, X, X + 1, S_ZP1, dh, A);
      if (k >= 1) mat_submul6<T>(FX + 36, F + (size_t)(X - 1) * 180 + 4 * 36, A);
      for (int e = 0; e < 36; ++e) FX[3 * 36 + e] = A[e];
    }
    // U2 = A_{k,k+2}
    if (k + 2 < Nz) {
      load_mblock<T>(Jc, s0, n, X, X + 2, S_ZP2, dh, A);
      for (int e = 0; e < 36; ++e) FX[4 * 36 + e] = A[e];
    }
  }
}

__global__ void k_prec_wall(int nw, const double* __restrict__ Jw, double dh, double* __restrict__ Fw) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < nw) Fw[w] = 1.0 - dh * Jw[w * 5 + 0];
}

// Apply: out = post^-1 * S0 M~^-1 S0^-1 (pre * scale * in); optional side outputs:
//   store_in  <- scale * in (normalises a GMRES basis vector in place)
//   red_out   <- rms(out / red_s) (or rms(out))
//   gs        <- unorm, sigma = sqrt(eps) zs_norm / unorm, zero_u (matvec finite-difference step)
struct ApplyArgs {
  const double* in; const double* in_scale; double* store_in; const double* pre; const double* post; double* out;
  const double* red_s; double* red_out; GmresScalars* gs; double zs_norm; int respect_done;
};

template <typename T>
__global__ void __launch_bounds__(PW) k_prec_apply(int n, int Nz, int ncol, int nw, int kind, const T* __restrict__ F,
                                                   const double* __restrict__ Fw, const double* __restrict__ s0,
                                                   ApplyArgs a, double* __restrict__ partial, unsigned* counter) {
  __shared__ double sh[PW];
  __shared__ bool slast;
  if (a.respect_done && a.gs->done) return;
  const int ncb = (ncol + PCOLS - 1) / PCOLS;
  const double sc = a.in_scale ? *a.in_scale : 1.0;
  double local = 0.0;
  const int lane = threadIdx.x;
  if ((int)blockIdx.x < ncb) {
    const int grp = lane / 6, r = lane % 6;
    const int col = blockIdx.x * PCOLS + grp;
    const bool active = (grp < PCOLS) && (col < ncol);
    const int base = (grp < PCOLS ? grp : 0) * 6;
    // forward: z_k = b_k - L1 z_{k-1} - L2 z_{k-2}
    double zm1 = 0.0, zm2 = 0.0;
    for (int k = 0; k < Nz; ++k) {
      const int X = (active ? col : 0) * Nz + k;
      const int idx = r * n + X;
      double b = 0.0;
      if (active) {
        double x = a.in[idx] * sc;
        if (a.store_in) a.store_in[idx] = x;
        if (a.pre) x = x * a.pre[idx];
        b = x / s0[idx];
      }
      double zk = b;
      if (kind == VTSA_PREC_ZLINE) {
        const T* L2 = F + (size_t)X * 180 + 0 * 36 + r * 6;
        const T* L1 = F + (size_t)X * 180 + 1 * 36 + r * 6;
        for (int c = 0; c < 6; ++c) {
          double v1 = __shfl(zm1, base + c);
          double v2 = __shfl(zm2, base + c);
          if (active && k >= 1) zk -= (double)L1[c] * v1;
          if (active && k >= 2) zk -= (double)L2[c] * v2;
        }
      }
      if (active) a.out[idx] = zk;
      zm2 = zm1;
      zm1 = zk;
    }
    // backward: x_k = Dinv (z_k - U1 x_{k+1} - U2 x_{k+2})
    double xp1 = 0.0, xp2 = 0.0;
    for (int k = Nz - 1; k >= 0; --k) {
      const int X = (active ? col : 0) * Nz + k;
      const int idx = r * n + X;
      double tk = active ? a.out[idx] : 0.0;
      if (kind == VTSA_PREC_ZLINE) {
        const T* U1 = F + (size_t)X * 180 + 3 * 36 + r * 6;
        const T* U2 = F + (size_t)X * 180 + 4 * 36 + r * 6;
        for (int c = 0; c < 6; ++c) {
          double v1 = __shfl(xp1, base + c);
          double v2 = __shfl(xp2, base + c);
          if (active && k + 1 < Nz) tk -= (double)U1[c] * v1;
          if (active && k + 2 < Nz) tk -= (double)U2[c] * v2;
        }
      }
      const T* Di = F + (size_t)X * 180 + 2 * 36 + r * 6;
      double xk = 0.0;
      for (int c = 0; c < 6; ++c) {
        double tc = __shfl(tk, base + c);
        if (active) xk += (double)Di[c] * tc;
      }
      if (active) {
        double o = xk * s0[idx];
        if (a.post) o = o / a.post[idx];
        a.out[idx] = o;
        double q = a.red_s ? o / a.red_s[idx] : o;
        local += q * q;
      }
      xp2 = xp1;
      xp1 = xk;
    }
  } else {
    // wall rows: point Jacobi
    const int wb = blockIdx.x - ncb;
    const int nwb = gridDim.x - ncb;
    for (int w = wb * PW + lane; w < nw; w += nwb * PW) {
      const int idx = 6 * n + w;
      double x = a.in[idx] * sc;
      if (a.store_in) a.store_in[idx] = x;
      if (a.pre) x = x * a.pre[idx];
      double o = x / Fw[w];
      if (a.post) o = o / a.post[idx];
      a.out[idx] = o;
      double q = a.red_s ? o / a.red_s[idx] : o;
      local += q * q;
    }
  }
  double bs = wave_block_sum(local, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = bs;
  if (ticket_last(counter, &slast) && threadIdx.x == 0) {
    double tot = 0.0;
    for (int b = 0; b < (int)gridDim.x; ++b) tot += partial[b];
    ticket_reset(counter);
    const int N = 6 * n + nw;
    double rms = sqrt(tot / (double)N);
    if (a.red_out) *a.red_out = rms;
    if (a.gs) {
      a.gs->unorm = rms;
      a.gs->zero_u = (rms == 0.0) ? 1 : 0;
      a.gs->sigma = rms == 0.0 ? 0.0 : (1.4901161193847656e-08 * a.zs_norm) / rms;
    }
  }
}

// ------------------------------------------------------------------ host launchers (preconditioner)
void launch_prec_setup(Handle& h, double dh, const double* scale, int kind, int prec, hipStream_t s) {
  hipMemcpyAsync(h.s0, scale, sizeof(double) * h.N, hipMemcpyDeviceToDevice, s);
  const int nb = (h.ncol + 63) / 64;
  if (prec == VTSA_FP32)
    hipLaunchKernelGGL(k_prec_factor<float>, dim3(nb), dim3(64), 0, s, h.n, h.desc.Nz, h.ncol, kind, h.Jc, h.s0, dh,
                       (float*)h.F);
  else
    hipLaunchKernelGGL(k_prec_factor<double>, dim3(nb), dim3(64), 0, s, h.n, h.desc.Nz, h.ncol, kind, h.Jc, h.s0, dh,
                       (double*)h.F);
  if (h.nw > 0) hipLaunchKernelGGL(k_prec_wall, dim3((h.nw + 255) / 256), dim3(256), 0, s, h.nw, h.Jw, dh, h.Fw);
  h.prec_kind = kind;
  h.prec_prec = prec;
  h.prec_ready = 1;
}

size_t partial_size(int ncol, int nw) {
  size_t a = (size_t)RMAXB * (MAX_RESTART + 2);
  size_t pg = (size_t)(ncol + PCOLS - 1) / PCOLS + (nw + PW * 4 - 1) / (PW * 4) + 1;
  size_t b = (size_t)RMAXB * 8 + pg;
  return a > b ? a : b;
}

static void apply_launch(Handle& h, const ApplyArgs& a, hipStream_t s) {
  const int ncb = (h.ncol + PCOLS - 1) / PCOLS;
  int nwb = h.nw > 0 ? (h.nw + PW * 4 - 1) / (PW * 4) : 0;
  const int grid = ncb + nwb;
  if (h.prec_prec == VTSA_FP32)
    hipLaunchKernelGGL(k_prec_apply<float>, dim3(grid), dim3(PW), 0, s, h.n, h.desc.Nz, h.ncol, h.nw, h.prec_kind,
                       (const float*)h.F, h.Fw, h.s0, a, h.red.partial + RMAXB * 8, h.red.counter + 1);
  else
    hipLaunchKernelGGL(k_prec_apply<double>, dim3(grid), dim3(PW), 0, s, h.n, h.desc.Nz, h.ncol, h.nw, h.prec_kind,
                       (const double*)h.F, h.Fw, h.s0, a, h.red.partial + RMAXB * 8, h.red.counter + 1);
}

void launch_prec_apply(Handle& h, const double* in, const double* in_scale, double* store_in, const double* pre,
                       const double* post, double* out, bool sigma_for_matvec, double zs_norm, hipStream_t s) {
  ApplyArgs a{in, in_scale, store_in, pre, post, out, nullptr, nullptr, sigma_for_matvec ? h.gs : nullptr, zs_norm,
              sigma_for_matvec ? 1 : 0};
  apply_launch(h, a, s);
}

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
  if (threadIdx.x == 0) partial[blockIdx.x] = bs;
  if (ticket_last(counter, &slast) && threadIdx.x == 0) {
    double tot = 0.0;
    for (int b = 0; b < (int)gridDim.x; ++b) tot += partial[b];
    ticket_reset(counter);
    *result = sqrt(tot / (double)N);
  }
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
  if (threadIdx.x == 0) partial[blockIdx.x] = bs;
  if (ticket_last(counter, &slast) && threadIdx.x == 0) {
    double tot = 0.0;
    for (int b = 0; b < (int)gridDim.x; ++b) tot += partial[b];
    ticket_reset(counter);
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
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    double zo = z[i];
    double v = zo + lam * dz[i];
    if (i < npos) {
      double lo = 0.1 * (zo > 0.0 ? zo : (zo != zo ? zo : 0.0));
      v = (v != v || lo != lo) ? v + lo : (v > lo ? v : lo);
    }
    zt[i] = v;
    double q = (v - zo) / s[i];
    loc += q * q;
  }
  double bs = wave_block_sum(loc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = bs;
  if (ticket_last(counter, &slast) && threadIdx.x == 0) {
    double tot = 0.0;
    for (int b = 0; b < (int)gridDim.x; ++b) tot += partial[b];
    ticket_reset(counter);
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
      } else {
        for (int s = 0; s < 3; ++s) loc[s] += (d.eps_t * a[s * n + c]) * V;
        loc[3] += (d.rho_b * a[3 * n + c]) * V;
        loc[4] += (d.rho_b * a[4 * n + c]) * V;
        loc[5] += a[5 * n + c] * V;
      }
    } else {
      const int w = c - n;
      if (mode == 0) {
        loc[3] += d.C_wall * a[6 * n + w];
        loc[7] += d.C_wall * b[6 * n + w];
      } else {
        loc[6] += d.C_wall * (a[6 * n + w] - d.T_ref);
      }
    }
  }
  const int nout = mode == 0 ? 8 : 7;
  for (int q = 0; q < nout; ++q) {
    double bs = wave_block_sum(loc[q], sh);
    if (threadIdx.x == 0) partial[q * RMAXB + blockIdx.x] = bs;
  }
  if (ticket_last(counter, &slast) && threadIdx.x == 0) {
    for (int q = 0; q < nout; ++q) {
      double tot = 0.0;
      for (int bb = 0; bb < (int)gridDim.x; ++bb) tot += partial[q * RMAXB + bb];
      result[q] = tot;
    }
    ticket_reset(counter);
  }
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
// Initialisation: |b|, tolerances (scipy gmres with x0 = 0, M = I), V[0] unnormalised = b.
__global__ void __launch_bounds__(RBLK) k_gm_init(int N, const double* __restrict__ b, double rtol, double atol,
                                                  GmresScalars* __restrict__ gs, double* __restrict__ partial,
                                                  unsigned* counter) {
  __shared__ double sh[RBLK];
  __shared__ bool slast;
  double loc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) loc += b[i] * b[i];
  double bs = wave_block_sum(loc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = bs;
  if (ticket_last(counter, &slast) && threadIdx.x == 0) {
    double tot = 0.0;
    for (int bb = 0; bb < (int)gridDim.x; ++bb) tot += partial[bb];
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
  }
}

// Pass 1 of classical Gram-Schmidt: h_k = V_k . w (k <= col), |w|^2.
__global__ void __launch_bounds__(RBLK) k_gm_dots(int N, int col, const double* __restrict__ Vb, size_t ld,
                                                  const double* __restrict__ w, GmresScalars* __restrict__ gs,
                                                  double* __restrict__ partial, unsigned* counter) {
  __shared__ double sh[RBLK];
  __shared__ bool slast;
  if (gs->done) return;
  double loc[MAX_RESTART + 1];
  const int K = col + 1;
  for (int k = 0; k <= K; ++k) loc[k] = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    double wi = w[i];
    for (int k = 0; k < K; ++k) loc[k] += Vb[k * ld + i] * wi;
    loc[K] += wi * wi;
  }
  for (int k = 0; k <= K; ++k) {
    double bs = wave_block_sum(loc[k], sh);
    if (threadIdx.x == 0) partial[k * RMAXB + blockIdx.x] = bs;
  }
  if (ticket_last(counter, &slast) && threadIdx.x == 0) {
    for (int k = 0; k <= K; ++k) {
      double tot = 0.0;
      for (int bb = 0; bb < (int)gridDim.x; ++bb) tot += partial[k * RMAXB + bb];
      if (k < K) gs->h_tmp[k] = tot; else gs->h0sq = tot;
    }
    for (int k = 0; k < K; ++k) gs->H[k][col] = gs->h_tmp[k];
    ticket_reset(counter);
  }
}

// Pass 2: w -= V h_tmp; h2 = V . w (reorthogonalisation); H[:, col] += h2.
__global__ void __launch_bounds__(RBLK) k_gm_orth(int N, int col, const double* __restrict__ Vb, size_t ld,
                                                  double* __restrict__ w, GmresScalars* __restrict__ gs,
                                                  double* __restrict__ partial, unsigned* counter) {
  __shared__ double sh[RBLK];
  __shared__ bool slast;
  if (gs->done) return;
  const int K = col + 1;
  double hk[MAX_RESTART];
  for (int k = 0; k < K; ++k) hk[k] = gs->h_tmp[k];
  double loc[MAX_RESTART];
  for (int k = 0; k < K; ++k) loc[k] = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    double wi = w[i];
    for (int k = 0; k < K; ++k) wi -= hk[k] * Vb[k * ld + i];
    w[i] = wi;
    for (int k = 0; k < K; ++k) loc[k] += Vb[k * ld + i] * wi;
  }
  for (int k = 0; k < K; ++k) {
    double bs = wave_block_sum(loc[k], sh);
    if (threadIdx.x == 0) partial[k * RMAXB + blockIdx.x] = bs;
  }
  if (ticket_last(counter, &slast) && threadIdx.x == 0) {
    for (int k = 0; k < K; ++k) {
      double tot = 0.0;
      for (int bb = 0; bb < (int)gridDim.x; ++bb) tot += partial[k * RMAXB + bb];
      gs->h_tmp[k] = tot;
      gs->H[k][col] += tot;
    }
    ticket_reset(counter);
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

// Pass 3: w -= V h2; |w|; Hessenberg column update with Givens rotations; convergence test
// (scipy gmres inner loop).
__global__ void __launch_bounds__(RBLK) k_gm_finish(int N, int col, const double* __restrict__ Vb, size_t ld,
                                                    double* __restrict__ w, GmresScalars* __restrict__ gs,
                                                    double* __restrict__ partial, unsigned* counter) {
  __shared__ double sh[RBLK];
  __shared__ bool slast;
  if (gs->done) return;
  const int K = col + 1;
  double hk[MAX_RESTART];
  for (int k = 0; k < K; ++k) hk[k] = gs->h_tmp[k];
  double loc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    double wi = w[i];
    for (int k = 0; k < K; ++k) wi -= hk[k] * Vb[k * ld + i];
    w[i] = wi;
    loc += wi * wi;
  }
  double bs = wave_block_sum(loc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = bs;
  if (ticket_last(counter, &slast) && threadIdx.x == 0) {
    double tot = 0.0;
    for (int bb = 0; bb < (int)gridDim.x; ++bb) tot += partial[bb];
    ticket_reset(counter);
    double h1 = sqrt(tot);
    double h0 = sqrt(gs->h0sq);
    bool brk = h1 <= EPS64 * h0;
    gs->H[col + 1][col] = brk ? 0.0 : h1;
    gs->inv_h1 = brk ? 0.0 : 1.0 / h1;
    for (int k = 0; k < col; ++k) {
      double c = gs->cs[k], s = gs->sn[k];
      double n0 = gs->H[k][col], n1 = gs->H[k + 1][col];
      gs->H[k][col] = c * n0 + s * n1;
      gs->H[k + 1][col] = -s * n0 + c * n1;
    }
    double c, s, mag;
    lartg(gs->H[col][col], gs->H[col + 1][col], c, s, mag);
    gs->cs[col] = c;
    gs->sn[col] = s;
    gs->H[col][col] = mag;
    gs->H[col + 1][col] = 0.0;
    double tmp = -s * gs->S[col];
    gs->S[col] = c * gs->S[col];
    gs->S[col + 1] = tmp;
    gs->presid = fabs(tmp);
    gs->iters = col + 1;
    if (brk) gs->breakdown = 1;
    if (gs->presid <= gs->ptol || brk) gs->done = 1;
  }
}

// Back substitution of the least-squares problem (scipy's pseudo-solve of the triangular system).
__global__ void k_gm_solve(GmresScalars* gs) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (gs->done == 2) return;
  const int col = gs->iters - 1;
  if (col < 0) return;
  if (gs->H[col][col] == 0.0) gs->S[col] = 0.0;
  for (int k = 0; k <= col; ++k) gs->yv[k] = gs->S[k];
  for (int k = col; k > 0; --k) {
    if (gs->yv[k] != 0.0) {
      gs->yv[k] /= gs->H[k][k];
      double tmp = gs->yv[k];
      for (int i = 0; i < k; ++i) gs->yv[i] -= tmp * gs->H[i][k];
    }
  }
  if (gs->yv[0] != 0.0) gs->yv[0] /= gs->H[0][0];
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
                                                   unsigned* counter) {
  __shared__ double sh[RBLK];
  __shared__ bool slast;
  double loc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    double r = b[i] - ax[i];
    loc += r * r;
  }
  double bs = wave_block_sum(loc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = bs;
  if (ticket_last(counter, &slast) && threadIdx.x == 0) {
    double tot = 0.0;
    for (int bb = 0; bb < (int)gridDim.x; ++bb) tot += partial[bb];
    ticket_reset(counter);
    gs->rnorm = sqrt(tot);
  }
}

// ------------------------------------------------------------------ host-side drivers
static inline int eblk(long long N) { return (int)((N + 255) / 256); }

int reduce_rms(Handle& h, const double* a, const double* s, int len, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_rms, dim3(rgrid(len)), dim3(RBLK), 0, st, len, a, s, h.red.partial, h.red.counter + 2,
                     h.red.result);
  if (check(hipMemcpyAsync(h.host_scalars, h.red.result, sizeof(double), hipMemcpyDeviceToHost, st), "memcpy")) return -1;
  if (check(hipStreamSynchronize(st), "sync")) return -1;
  *out = h.host_scalars[0];
  return 0;
}

void launch_residual_combine(Handle& h, const double* z, const double* f, double dh, const double* rhs_known,
                             double* R, hipStream_t st) {
  hipLaunchKernelGGL(k_residual_combine, dim3(rgrid(h.N)), dim3(RBLK), 0, st, h.N, z, f, dh, rhs_known,
                     (const double*)nullptr, R, h.red.partial, h.red.counter + 2, h.red.result);
}

int reduce_weighted(Handle& h, const double* a, const double* w, int len, int mode, double* out, hipStream_t st) {
  (void)len;
  hipLaunchKernelGGL(k_weighted, dim3(rgrid(h.n + h.nw)), dim3(RBLK), 0, st, h.m, mode, a, w, h.red.partial,
                     h.red.counter + 3, h.red.result);
  const int nout = mode == 0 ? 8 : 7;
  if (check(hipMemcpyAsync(h.host_scalars, h.red.result, sizeof(double) * nout, hipMemcpyDeviceToHost, st), "memcpy"))
    return -1;
  if (check(hipStreamSynchronize(st), "sync")) return -1;
  for (int q = 0; q < nout; ++q) out[q] = h.host_scalars[q];
  return 0;
}

// Newton building blocks (host drivers; each returns host scalars)
int newton_eval_impl(Handle& h, double t, const double* z, double dh, const double* rk, const double* s, double* f,
                     double* R, double* rn, double* pn, hipStream_t st) {
  launch_props(h, z, h.pb, f, nullptr, st);
  launch_flux(h, t, z, h.pb, f, st);
  hipLaunchKernelGGL(k_residual_combine, dim3(rgrid(h.N)), dim3(RBLK), 0, st, h.N, z, f, dh, rk, s, R,
                     h.red.partial, h.red.counter + 2, h.red.result + 0);
  ApplyArgs a{R, nullptr, nullptr, nullptr, nullptr, h.vtmp, s, h.red.result + 1, nullptr, 0.0, 0};
  apply_launch(h, a, st);
  if (check(hipMemcpyAsync(h.host_scalars, h.red.result, 2 * sizeof(double), hipMemcpyDeviceToHost, st), "memcpy"))
    return -1;
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
  ApplyArgs a{Rt, nullptr, nullptr, nullptr, nullptr, h.vtmp, s, h.red.result + 1, nullptr, 0.0, 0};
  apply_launch(h, a, st);
  if (check(hipMemcpyAsync(h.host_scalars, h.red.result, 3 * sizeof(double), hipMemcpyDeviceToHost, st), "memcpy"))
    return -1;
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
  ApplyArgs a{h.vtmp, nullptr, nullptr, nullptr, nullptr, h.vtmp, h.vtmp2, h.red.result + 3, nullptr, 0.0, 0};
  apply_launch(h, a, st);
  if (check(hipMemcpyAsync(h.host_scalars, h.red.result + 3, sizeof(double), hipMemcpyDeviceToHost, st), "memcpy"))
    return -1;
  if (check(hipStreamSynchronize(st), "sync")) return -1;
  *en = h.host_scalars[0];
  return 0;
}

void vec_op_impl(Handle& h, int op, int npos, double a, double b, double c, double e, const double* x,
                 const double* y, const double* z, const double* w, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_vec_op, dim3(eblk(h.N)), dim3(256), 0, st, op, h.N, npos, a, b, c, e, x, y, z, w, out);
}

int gmres_cycle(Handle& h, double t, const double* z, const double* f, const double* s, double dh, double zs_norm,
                const double* b, double* w, double rtol, double atol, int restart, vtsa_gmres_info* info,
                hipStream_t st) {
  if (restart > MAX_RESTART) restart = MAX_RESTART;
  if (restart > h.N) restart = h.N;
  const int N = h.N;
  const size_t ld = (size_t)N;
  double* Vb = h.Vb;
  double* U = h.vtmp2;
  const int g = rgrid(N);
  hipLaunchKernelGGL(k_gm_init, dim3(g), dim3(RBLK), 0, st, N, b, rtol, atol, h.gs, h.red.partial, h.red.counter + 5);
  for (int col = 0; col < restart; ++col) {
    // u = P^-1 v_col with v_col = w_prev / |w_prev| normalised in place
    const double* vin = col == 0 ? b : Vb + (size_t)col * ld;
    ApplyArgs a{vin, &h.gs->inv_h1, Vb + (size_t)col * ld, s, s, U, nullptr, nullptr, h.gs, zs_norm, 1};
    apply_launch(h, a, st);
    double* wcol = Vb + (size_t)(col + 1) * ld;
    launch_matvec_residual(h, t, z, U, s, f, dh, wcol, true, st);
    hipLaunchKernelGGL(k_gm_dots, dim3(g), dim3(RBLK), 0, st, N, col, Vb, ld, wcol, h.gs, h.red.partial,
                       h.red.counter + 5);
    hipLaunchKernelGGL(k_gm_orth, dim3(g), dim3(RBLK), 0, st, N, col, Vb, ld, wcol, h.gs, h.red.partial,
                       h.red.counter + 5);
    hipLaunchKernelGGL(k_gm_finish, dim3(g), dim3(RBLK), 0, st, N, col, Vb, ld, wcol, h.gs, h.red.partial,
                       h.red.counter + 5);
  }
  hipLaunchKernelGGL(k_gm_solve, dim3(1), dim3(1), 0, st, h.gs);
  hipLaunchKernelGGL(k_gm_update, dim3(eblk(N)), dim3(256), 0, st, N, Vb, ld, h.gs, w);
  // true residual r = b - A P^-1 w (one more matvec, as scipy does after the cycle)
  {
    ApplyArgs a{w, nullptr, nullptr, s, s, U, nullptr, nullptr, h.gs, zs_norm, 0};
    apply_launch(h, a, st);
  }
  launch_matvec_residual(h, t, z, U, s, f, dh, h.vtmp, false, st);
  hipLaunchKernelGGL(k_gm_resid, dim3(g), dim3(RBLK), 0, st, N, b, h.vtmp, h.gs, h.red.partial, h.red.counter + 5);
  if (check(hipMemcpyAsync(h.gs_host, h.gs, sizeof(GmresScalars), hipMemcpyDeviceToHost, st), "memcpy")) return -1;
  if (check(hipStreamSynchronize(st), "sync")) return -1;
  const GmresScalars& G = *h.gs_host;
  const int done_flag = G.done;
  info->iters = G.iters;
  info->breakdown = G.breakdown;
  info->bnorm = G.bnorm;
  info->presid = G.presid;
  info->atol_eff = G.atol_eff;
  if (done_flag == 2) {   // exited before the first iteration: x = 0 (or b == 0)
    info->rnorm = G.bnorm;
    info->info = 0;
    info->iters = 0;
  } else {
    info->rnorm = G.rnorm;
    info->info = (G.rnorm <= G.atol_eff) ? 0 : 1;
  }
  return 0;
}

}  // namespace vtsa

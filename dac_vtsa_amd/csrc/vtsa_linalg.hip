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
// The apply kernel stages the right-hand side of up to ten columns in LDS (coalesced loads and
// stores); six lanes per column (one per variable) sweep along z through LDS, with the factor
// rows of the next cell loaded while the current one is processed.
//
// Reductions use one launch: per-block partial sums, then the last block to finish (ticket
// counter) adds the partials in block order, so results are deterministic.
#include <chrono>
#include <thread>

#include "vtsa_internal.h"

namespace vtsa {

constexpr int RBLK = 256;        // reduction block size
constexpr int RMAXB = 512;       // max blocks of a reduction grid
constexpr int PW = 64;           // preconditioner apply: threads per block (one wavefront)
constexpr int PCOLS = 10;        // columns per wavefront (6 lanes each)
constexpr int AC = 4;            // apply gather/scatter: elements per lane whose loads go together
constexpr double EPS64 = 2.220446049250313e-16;

// ------------------------------------------------------------------ reduction helpers
// Butterfly sum over the 64 lanes of a wavefront.  a + b == b + a exactly, so every lane ends
// with the same bits, and the order is fixed, so the result is deterministic.
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block sum: wavefront sums, then the per-wave values in wave order.  sh needs one slot per wave.
__device__ inline double wave_block_sum(double v, double* sh) {
  const int nwv = (blockDim.x + 63) >> 6;
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  for (int w = 0; w < nwv; ++w) r += sh[w];
  __syncthreads();
  return r;
}

// Cross-block reductions without cache-wide fences.  Every block writes its partial sums with
// agent-scope stores (coherent across the XCDs' L2 caches), waits for their completion, and
// only then takes a ticket; the last block to take one reads the partials with agent-scope
// loads.  A device-scope release/acquire fence would write back and invalidate the whole L2 of
// the block's XCD instead, once per block.
__device__ inline void store_partial(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double load_partial(const double* p) {
  return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum of partial[0 .. cnt) by one wavefront (lane-strided, then butterfly): deterministic.
__device__ inline double wave_sum_strided(const double* __restrict__ partial, int cnt) {
  double acc = 0.0;
  for (int b = threadIdx.x & 63; b < cnt; b += 64) acc += load_partial(partial + b);
  return wave_sum(acc);
}

// Returns true in the last block to finish; that block may then read every block's partials.
__device__ inline bool ticket_last(unsigned* counter, bool* sflag) {
  __builtin_amdgcn_s_waitcnt(0);   // this wavefront's partial stores have completed
  __syncthreads();                 // ... and every other wavefront's
  if (threadIdx.x == 0) {
    unsigned tk = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *sflag = (tk == gridDim.x - 1);
  }
  __syncthreads();
  return *sflag;
}

__device__ inline void ticket_reset(unsigned* counter) {
  __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Last block of a one-value reduction: its first wavefront sums the per-block partials; returns
// true, with the total in *tot, in thread 0 of the last block only (counter already reset).
__device__ inline bool last_block_total(unsigned* counter, bool* sflag, const double* partial, double* tot) {
  if (!ticket_last(counter, sflag)) return false;
  if (threadIdx.x >= 64) return false;
  const double t = wave_sum_strided(partial, gridDim.x);
  if (threadIdx.x != 0) return false;
  ticket_reset(counter);
  *tot = t;
  return true;
}

// Last block of an nval-value reduction (partial rows of stride RMAXB): its wavefronts sum one
// row each into out[]; returns true in every thread of the last block, after a barrier.
__device__ inline bool last_block_totals(unsigned* counter, bool* sflag, const double* partial, int nval,
                                         double* out) {
  if (!ticket_last(counter, sflag)) return false;
  const int wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  for (int q = wv; q < nval; q += nwv) {
    const double t = wave_sum_strided(partial + (size_t)q * RMAXB, gridDim.x);
    if ((threadIdx.x & 63) == 0) out[q] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) ticket_reset(counter);
  return true;
}

// Barrier for one-wavefront blocks that exchange data through LDS only: orders LDS accesses
// (lgkmcnt) without waiting for outstanding global loads, so prefetches stay in flight.
__device__ inline void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

static inline int rgrid(long long work) {
  long long b = (work + RBLK - 1) / RBLK;
  if (b < 1) b = 1;
  if (b > RMAXB) b = RMAXB;
  return (int)b;
}

// ------------------------------------------------------------------ preconditioner factor
// Block LU of the z-line matrices (block Jacobi: the diagonal inverses only).  Six lanes per
// column, lane r computing row r of every block product, PCOLS columns per wavefront; the
// factors of the previous two cells stay in LDS.  The pivot blocks are inverted by Gauss-Jordan
// with partial pivoting (the first row of largest |entry| wins), rows in parallel through LDS.
// Block products accumulate over m in order in T, as C = A B and C -= A B of 6x6 matrices.
// F blocks per cell, rows of 6: 0 L2, 1 L1, 2 Dinv, 3 U1, 4 U2.
constexpr int FSTR = 258;   // LDS doubles per column: Dinv/U1/U2 ring of two cells + pivot scratch

template <typename T>
__global__ void __launch_bounds__(PW) k_prec_factor(int n, int Nz, int ncol, int kind,
                                                    const double* __restrict__ Jc, const double* __restrict__ s0,
                                                    double dh, T* __restrict__ F) {
  __shared__ double lds[PCOLS * FSTR];
  const int lane = threadIdx.x;
  const int grp = lane / 6, r = lane - grp * 6;
  const int col0 = blockIdx.x * PCOLS;
  const int nc = min(PCOLS, ncol - col0);
  const bool active = grp < nc;
  const int col = col0 + (active ? grp : 0);
  double* Lc = lds + (active ? grp : 0) * FSTR;
  double* pDi = Lc;          // [2][36] Dinv of cells k-2 / k-1 (ring slot k & 1)
  double* pU1 = Lc + 72;     // [2][36]
  double* pU2 = Lc + 144;    // [2][36]
  double* colp = Lc + 216;   // [6]  |pivot column|
  double* sw = Lc + 222;     // [24] row swap
  double* prow = Lc + 246;   // [12] normalised pivot row of D and of its inverse
  const bool zl = kind == VTSA_PREC_ZLINE;
  for (int k = 0; k < Nz; ++k) {
    const int X = col * Nz + k;
    const double sr = s0[r * n + X];
    // row r of the block of M~ = S0^-1 (I - dh J) S0 between cell X and cell X + dk
    auto mrow = [&](int slot, int dk, T* out) {
      const double* J = Jc + ((size_t)X * NSLOT + slot) * 36 + r * 6;
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        const double v = (slot == S_SELF && r == c ? 1.0 : 0.0) - dh * J[c];
        out[c] = (T)(v * (s0[c * n + X + dk] / sr));
      }
    };
    const int q = k & 1, q1 = (k + 1) & 1;   // ring slots of cells k-2 (and k) and k-1
    T l2[6], l1[6], d[6], u1[6], u2[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) { l2[c] = T(0); l1[c] = T(0); u1[c] = T(0); u2[c] = T(0); }
    if (zl && k >= 2) {   // L2 = A_{k,k-2} Dinv_{k-2}
      T a[6];
      mrow(S_ZM2, -2, a);
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        T acc = T(0);
#pragma unroll
        for (int m = 0; m < 6; ++m) acc += a[m] * (T)pDi[q * 36 + m * 6 + c];
        l2[c] = acc;
      }
    }
    if (zl && k >= 1) {   // L1 = (A_{k,k-1} - L2 U1_{k-2}) Dinv_{k-1}
      T a[6];
      mrow(S_ZM1, -1, a);
      if (k >= 2) {
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          T acc = T(0);
#pragma unroll
          for (int m = 0; m < 6; ++m) acc += l2[m] * (T)pU1[q * 36 + m * 6 + c];
          a[c] -= acc;
        }
      }
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        T acc = T(0);
#pragma unroll
        for (int m = 0; m < 6; ++m) acc += a[m] * (T)pDi[q1 * 36 + m * 6 + c];
        l1[c] = acc;
      }
    }
    mrow(S_SELF, 0, d);   // D = A_kk - L1 U1_{k-1} - L2 U2_{k-2}
    if (zl && k >= 1) {
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        T acc = T(0);
#pragma unroll
        for (int m = 0; m < 6; ++m) acc += l1[m] * (T)pU1[q1 * 36 + m * 6 + c];
        d[c] -= acc;
      }
    }
    if (zl && k >= 2) {
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        T acc = T(0);
#pragma unroll
        for (int m = 0; m < 6; ++m) acc += l2[m] * (T)pU2[q * 36 + m * 6 + c];
        d[c] -= acc;
      }
    }
    if (zl && k + 1 < Nz) {   // U1 = A_{k,k+1} - L1 U2_{k-1}
      mrow(S_ZP1, 1, u1);
      if (k >= 1) {
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          T acc = T(0);
#pragma unroll
          for (int m = 0; m < 6; ++m) acc += l1[m] * (T)pU2[q1 * 36 + m * 6 + c];
          u1[c] -= acc;
        }
      }
    }
    if (zl && k + 2 < Nz) mrow(S_ZP2, 2, u2);   // U2 = A_{k,k+2}
    lds_barrier();   // every lane has read ring slot q (cell k-2) before cell k overwrites it
    T* FX = F + (size_t)X * 180 + r * 6;
    if (active && zl) {
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        FX[c] = l2[c];
        FX[36 + c] = l1[c];
        FX[108 + c] = u1[c];
        FX[144 + c] = u2[c];
        pU1[q * 36 + r * 6 + c] = (double)u1[c];
        pU2[q * 36 + r * 6 + c] = (double)u2[c];
      }
    }
    // Gauss-Jordan inverse of D, lane r holding row r of D and of the inverse
    T iv[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) iv[c] = (r == c) ? T(1) : T(0);
#pragma unroll
    for (int p = 0; p < 6; ++p) {
      if (active) colp[r] = (double)fabs(d[p]);
      lds_barrier();
      int piv = p;
      double best = colp[p];
      for (int rr = p + 1; rr < 6; ++rr) {
        const double v = colp[rr];
        if (v > best) { best = v; piv = rr; }
      }
      if (active && (r == p || r == piv)) {
        double* dst = sw + (r == p ? 0 : 12);
#pragma unroll
        for (int c = 0; c < 6; ++c) { dst[c] = (double)d[c]; dst[6 + c] = (double)iv[c]; }
      }
      lds_barrier();
      if (piv != p && (r == p || r == piv)) {
        const double* src = sw + (r == p ? 12 : 0);
#pragma unroll
        for (int c = 0; c < 6; ++c) { d[c] = (T)src[c]; iv[c] = (T)src[6 + c]; }
      }
      if (r == p) {
        const T inv = T(1) / d[p];
#pragma unroll
        for (int c = 0; c < 6; ++c) { d[c] *= inv; iv[c] *= inv; }
        if (active) {
#pragma unroll
          for (int c = 0; c < 6; ++c) { prow[c] = (double)d[c]; prow[6 + c] = (double)iv[c]; }
        }
      }
      lds_barrier();
      if (r != p) {
        const T f = d[p];
        if (f != T(0)) {
#pragma unroll
          for (int c = 0; c < 6; ++c) { d[c] -= f * (T)prow[c]; iv[c] -= f * (T)prow[6 + c]; }
        }
      }
      lds_barrier();
    }
    if (active) {
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        FX[72 + c] = iv[c];
        pDi[q * 36 + r * 6 + c] = (double)iv[c];
      }
    }
    lds_barrier();
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
__global__ void __launch_bounds__(PW) k_prec_apply(int n, int Nz, int ncol, int nw, int kind, int cpb,
                                                   const T* __restrict__ F, const double* __restrict__ Fw,
                                                   const double* __restrict__ s0, ApplyArgs a,
                                                   double* __restrict__ partial, unsigned* counter) {
  // LDS: xs[cpb][Nz][6] (right-hand side, then z, then x, in place) and tb[cpb][6] (exchange)
  extern __shared__ double xs[];
  __shared__ double sh[1];
  __shared__ bool slast;
  if (a.respect_done && a.gs->done) return;
  const int ncb = (ncol + cpb - 1) / cpb;
  const double sc = a.in_scale ? *a.in_scale : 1.0;
  double local = 0.0;
  const int lane = threadIdx.x;
  if ((int)blockIdx.x < ncb) {
    const int col0 = blockIdx.x * cpb;
    const int nc = min(cpb, ncol - col0);
    const int X0 = col0 * Nz;          // the block's cells are X0 .. X0 + nc*Nz - 1
    const int cnt = nc * Nz;
    double* tb = xs + (size_t)cpb * Nz * 6;
    // 1. b = (pre * scale * in) / s0, coalesced per variable.  store_in may alias in, so every
    //    load of a chunk of AC elements is issued before the chunk's stores.
    for (int r = 0; r < 6; ++r)
      for (int e0 = lane; e0 < cnt; e0 += PW * AC) {
        double xin[AC], xpre[AC], xs0[AC];
#pragma unroll
        for (int j = 0; j < AC; ++j) {
          const int e = e0 + j * PW;
          const int idx = r * n + X0 + (e < cnt ? e : 0);
          xin[j] = a.in[idx];
          xpre[j] = a.pre ? a.pre[idx] : 1.0;
          xs0[j] = s0[idx];
        }
#pragma unroll
        for (int j = 0; j < AC; ++j) {
          const int e = e0 + j * PW;
          if (e < cnt) {
            double x = xin[j] * sc;
            if (a.store_in) a.store_in[r * n + X0 + e] = x;
            if (a.pre) x = x * xpre[j];
            xs[e * 6 + r] = x / xs0[j];
          }
        }
      }
    lds_barrier();
    if (kind == VTSA_PREC_ZLINE) {
      // 2. six lanes per column (lane r owns row r); factor rows of the next cell are loaded
      //    while the current cell is processed
      const int grp = lane / 6, r = lane - grp * 6;
      const bool active = grp < nc;
      const int g = active ? grp : 0;
      double* xc = xs + (size_t)g * Nz * 6;
      double* tg = tb + g * 6;
      const T* Fc = F + (size_t)(X0 + g * Nz) * 180 + r * 6;
      T c1[6], c2[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) { c1[q] = Fc[36 + q]; c2[q] = Fc[q]; }
      for (int k = 0; k < Nz; ++k) {          // forward: z_k = b_k - L1 z_{k-1} - L2 z_{k-2}
        T n1[6], n2[6];
        const T* Fn = Fc + (size_t)(k + 1 < Nz ? k + 1 : k) * 180;
#pragma unroll
        for (int q = 0; q < 6; ++q) { n1[q] = Fn[36 + q]; n2[q] = Fn[q]; }
        double zk = xc[k * 6 + r];
        if (k >= 1) {
          double s1 = 0.0;
#pragma unroll
          for (int c = 0; c < 6; ++c) s1 += (double)c1[c] * xc[(k - 1) * 6 + c];
          zk -= s1;
        }
        if (k >= 2) {
          double s2 = 0.0;
#pragma unroll
          for (int c = 0; c < 6; ++c) s2 += (double)c2[c] * xc[(k - 2) * 6 + c];
          zk -= s2;
        }
        if (active) xc[k * 6 + r] = zk;
        lds_barrier();
#pragma unroll
        for (int q = 0; q < 6; ++q) { c1[q] = n1[q]; c2[q] = n2[q]; }
      }
      T u1[6], u2[6], di[6];
      {
        const T* Fk = Fc + (size_t)(Nz - 1) * 180;
#pragma unroll
        for (int q = 0; q < 6; ++q) { u1[q] = Fk[108 + q]; u2[q] = Fk[144 + q]; di[q] = Fk[72 + q]; }
      }
      for (int k = Nz - 1; k >= 0; --k) {     // backward: x_k = Dinv (z_k - U1 x_{k+1} - U2 x_{k+2})
        T m1[6], m2[6], md[6];
        const T* Fn = Fc + (size_t)(k > 0 ? k - 1 : 0) * 180;
#pragma unroll
        for (int q = 0; q < 6; ++q) { m1[q] = Fn[108 + q]; m2[q] = Fn[144 + q]; md[q] = Fn[72 + q]; }
        double tk = xc[k * 6 + r];
        if (k + 1 < Nz) {
          double s1 = 0.0;
#pragma unroll
          for (int c = 0; c < 6; ++c) s1 += (double)u1[c] * xc[(k + 1) * 6 + c];
          tk -= s1;
        }
        if (k + 2 < Nz) {
          double s2 = 0.0;
#pragma unroll
          for (int c = 0; c < 6; ++c) s2 += (double)u2[c] * xc[(k + 2) * 6 + c];
          tk -= s2;
        }
        if (active) tg[r] = tk;
        lds_barrier();
        double xk = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) xk += (double)di[c] * tg[c];
        if (active) xc[k * 6 + r] = xk;
        lds_barrier();
#pragma unroll
        for (int q = 0; q < 6; ++q) { u1[q] = m1[q]; u2[q] = m2[q]; di[q] = md[q]; }
      }
    } else {
      // block Jacobi: every cell on its own, one lane per cell
      for (int e = lane; e < cnt; e += PW) {
        const T* Di = F + (size_t)(X0 + e) * 180 + 72;
        double bv[6], xv[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) bv[c] = xs[e * 6 + c];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
          double acc = 0.0;
#pragma unroll
          for (int c = 0; c < 6; ++c) acc += (double)Di[r * 6 + c] * bv[c];
          xv[r] = acc;
        }
#pragma unroll
        for (int r = 0; r < 6; ++r) xs[e * 6 + r] = xv[r];
      }
      lds_barrier();
    }
    // 3. out = x s0 / post, coalesced per variable, and the norm of out / red_s (loads of a
    //    chunk before its stores, as above)
    for (int r = 0; r < 6; ++r)
      for (int e0 = lane; e0 < cnt; e0 += PW * AC) {
        double xs0[AC], xpost[AC], xred[AC];
#pragma unroll
        for (int j = 0; j < AC; ++j) {
          const int e = e0 + j * PW;
          const int idx = r * n + X0 + (e < cnt ? e : 0);
          xs0[j] = s0[idx];
          xpost[j] = a.post ? a.post[idx] : 1.0;
          xred[j] = a.red_s ? a.red_s[idx] : 1.0;
        }
#pragma unroll
        for (int j = 0; j < AC; ++j) {
          const int e = e0 + j * PW;
          if (e < cnt) {
            double o = xs[e * 6 + r] * xs0[j];
            if (a.post) o = o / xpost[j];
            a.out[r * n + X0 + e] = o;
            double q = a.red_s ? o / xred[j] : o;
            local += q * q;
          }
        }
      }
  } else {
    // wall rows: point Jacobi
    const int wb = blockIdx.x - ncb;
    const int nwb = gridDim.x - ncb;
    for (int w0 = wb * PW + lane; w0 < nw; w0 += nwb * PW * AC) {
      double xin[AC], xpre[AC], xfw[AC], xpost[AC], xred[AC];
#pragma unroll
      for (int j = 0; j < AC; ++j) {
        const int w = w0 + j * nwb * PW;
        const int wi = w < nw ? w : 0;
        const int idx = 6 * n + wi;
        xin[j] = a.in[idx];
        xpre[j] = a.pre ? a.pre[idx] : 1.0;
        xfw[j] = Fw[wi];
        xpost[j] = a.post ? a.post[idx] : 1.0;
        xred[j] = a.red_s ? a.red_s[idx] : 1.0;
      }
#pragma unroll
      for (int j = 0; j < AC; ++j) {
        const int w = w0 + j * nwb * PW;
        if (w < nw) {
          const int idx = 6 * n + w;
          double x = xin[j] * sc;
          if (a.store_in) a.store_in[idx] = x;
          if (a.pre) x = x * xpre[j];
          double o = x / xfw[j];
          if (a.post) o = o / xpost[j];
          a.out[idx] = o;
          double q = a.red_s ? o / xred[j] : o;
          local += q * q;
        }
      }
    }
  }
  double bs = wave_block_sum(local, sh);
  if (threadIdx.x == 0) store_partial(partial + blockIdx.x, bs);
  double tot;
  if (last_block_total(counter, &slast, partial, &tot)) {
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
  (void)hipMemcpyAsync(h.s0, scale, sizeof(double) * h.N, hipMemcpyDeviceToDevice, s);
  const int nb = (h.ncol + PCOLS - 1) / PCOLS;
  if (prec == VTSA_FP32)
    hipLaunchKernelGGL(k_prec_factor<float>, dim3(nb), dim3(PW), 0, s, h.n, h.desc.Nz, h.ncol, kind, h.Jc, h.s0, dh,
                       (float*)h.F);
  else
    hipLaunchKernelGGL(k_prec_factor<double>, dim3(nb), dim3(PW), 0, s, h.n, h.desc.Nz, h.ncol, kind, h.Jc, h.s0, dh,
                       (double*)h.F);
  if (h.nw > 0) hipLaunchKernelGGL(k_prec_wall, dim3((h.nw + 255) / 256), dim3(256), 0, s, h.nw, h.Jw, dh, h.Fw);
  h.prec_kind = kind;
  h.prec_prec = prec;
  h.prec_ready = 1;
}

// Columns per apply block: up to PCOLS, as many as fit the 64 KiB of LDS (cpb (6 Nz + 6) doubles).
int apply_columns_per_block(int Nz) {
  const int fit = (int)(65536 / (8 * (6 * (long long)Nz + 6)));
  return fit < PCOLS ? fit : PCOLS;
}

size_t partial_size(int ncol, int nw) {
  size_t a = (size_t)RMAXB * (MAX_RESTART + 2);
  size_t pg = (size_t)ncol + (nw + PW * 4 - 1) / (PW * 4) + 1;   // apply grid with one column per block
  size_t b = (size_t)RMAXB * 8 + pg;
  return a > b ? a : b;
}

static void apply_launch(Handle& h, const ApplyArgs& a, hipStream_t s) {
  const int Nz = h.desc.Nz;
  const int cpb = apply_columns_per_block(Nz);
  const int ncb = (h.ncol + cpb - 1) / cpb;
  int nwb = h.nw > 0 ? (h.nw + PW * 4 - 1) / (PW * 4) : 0;
  const int grid = ncb + nwb;
  const size_t lds = sizeof(double) * ((size_t)cpb * Nz * 6 + (size_t)cpb * 6);
  if (h.prec_prec == VTSA_FP32)
    hipLaunchKernelGGL(k_prec_apply<float>, dim3(grid), dim3(PW), lds, s, h.n, Nz, h.ncol, h.nw, h.prec_kind, cpb,
                       (const float*)h.F, h.Fw, h.s0, a, h.red.partial + RMAXB * 8, h.red.counter + 1);
  else
    hipLaunchKernelGGL(k_prec_apply<double>, dim3(grid), dim3(PW), lds, s, h.n, Nz, h.ncol, h.nw, h.prec_kind, cpb,
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
  if (threadIdx.x == 0) store_partial(partial + blockIdx.x, bs);
  double tot;
  if (last_block_total(counter, &slast, partial, &tot)) {
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
                                                  unsigned* counter, int* __restrict__ progress) {
  __shared__ double sh[RBLK];
  __shared__ bool slast;
  double loc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) loc += b[i] * b[i];
  double bs = wave_block_sum(loc, sh);
  if (threadIdx.x == 0) store_partial(partial + blockIdx.x, bs);
  double tot;
  if (last_block_total(counter, &slast, partial, &tot)) {
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
    // pacing only: the host reads no data on the strength of this word, so relaxed suffices
    __hip_atomic_store(progress, gs->done ? -1 : 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Classical Gram-Schmidt with one reorthogonalisation (CGS2), three passes per Arnoldi step.
// Every pass reads the K = col + 1 basis vectors once per element; per-vector block sums go
// through wavefront butterflies (no per-thread arrays indexed by K), and the last block sums the
// per-block partials with one wavefront per basis vector.
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

// Pass 2: w -= V h_tmp; h2 = V . w (reorthogonalisation); H[:, col] += h2.
__global__ void __launch_bounds__(RBLK) k_gm_orth(int N, int col, const double* __restrict__ Vb, size_t ld,
                                                  double* __restrict__ w, GmresScalars* __restrict__ gs,
                                                  double* __restrict__ partial, unsigned* counter) {
  __shared__ double sh[(RBLK / 64) * GMW];
  __shared__ double hk[GMW];
  __shared__ double tot[GMW];
  __shared__ bool slast;
  if (gs->done) return;
  const int K = col + 1;
  const int wv = threadIdx.x >> 6;
  if ((int)threadIdx.x < K) hk[threadIdx.x] = gs->h_tmp[threadIdx.x];
  __syncthreads();
  const int i0 = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
  for (int i = i0; i < N; i += stride) w[i] = project_out(Vb, ld, hk, K, i, w[i]);
  block_dots(Vb, ld, w, K, N, i0, stride, sh + wv * GMW);
  __syncthreads();
  if ((int)threadIdx.x < K) {
    double s = 0.0;
    for (int q = 0; q < RBLK / 64; ++q) s += sh[q * GMW + threadIdx.x];
    store_partial(partial + threadIdx.x * RMAXB + blockIdx.x, s);
  }
  if (last_block_totals(counter, &slast, partial, K, tot) && (int)threadIdx.x < K) {
    const int k = threadIdx.x;
    gs->h_tmp[k] = tot[k];
    gs->H[k][col] += tot[k];
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
                                                    double* __restrict__ partial, unsigned* counter,
                                                    int* __restrict__ progress) {
  __shared__ double sh[RBLK / 64];
  __shared__ double hk[GMW];
  __shared__ double hc[GMW], csl[GMW], snl[GMW];
  __shared__ bool slast;
  if (gs->done) return;
  const int K = col + 1;
  if ((int)threadIdx.x < K) hk[threadIdx.x] = gs->h_tmp[threadIdx.x];
  __syncthreads();
  double loc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    const double wi = project_out(Vb, ld, hk, K, i, w[i]);
    w[i] = wi;
    loc += wi * wi;
  }
  double bs = wave_block_sum(loc, sh);
  if (threadIdx.x == 0) store_partial(partial + blockIdx.x, bs);
  if (!ticket_last(counter, &slast)) return;
  // last block: |w|, then the Givens sweep of column col on LDS copies
  const int t = threadIdx.x;
  double tot = 0.0;
  if (t < 64) tot = wave_sum_strided(partial, gridDim.x);
  for (int k = t; k <= col; k += blockDim.x) hc[k] = gs->H[k][col];
  for (int k = t; k < col; k += blockDim.x) { csl[k] = gs->cs[k]; snl[k] = gs->sn[k]; }
  __syncthreads();
  if (t == 0) {
    ticket_reset(counter);
    double h1 = sqrt(tot);
    double h0 = sqrt(gs->h0sq);
    bool brk = h1 <= EPS64 * h0;
    hc[col + 1] = brk ? 0.0 : h1;
    gs->inv_h1 = brk ? 0.0 : 1.0 / h1;
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
    double tmp = -s * gs->S[col];
    gs->S[col] = c * gs->S[col];
    gs->S[col + 1] = tmp;
    gs->presid = fabs(tmp);
    gs->iters = col + 1;
    if (brk) gs->breakdown = 1;
    const bool fin = gs->presid <= gs->ptol || brk;
    if (fin) gs->done = 1;
    // host-visible progress: completed columns, negative once the cycle is done
    __hip_atomic_store(progress, fin ? -(col + 1) : (col + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  __syncthreads();
  for (int k = t; k <= col + 1; k += blockDim.x) gs->H[k][col] = hc[k];
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
                                                   unsigned* counter) {
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
  // The host queues at most GM_LOOKAHEAD Arnoldi columns beyond the last one the device reports
  // complete (a coherent host word written by k_gm_init / k_gm_finish), and stops queueing once
  // the device reports convergence; queued columns past convergence return at once.
  constexpr int GM_LOOKAHEAD = 3;
  constexpr int NOT_STARTED = -0x7fffffff;
  __atomic_store_n(h.progress, NOT_STARTED, __ATOMIC_RELEASE);
  hipLaunchKernelGGL(k_gm_init, dim3(g), dim3(RBLK), 0, st, N, b, rtol, atol, h.gs, h.red.partial, h.red.counter + 5,
                     h.progress_dev);
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
                       h.red.counter + 5, h.progress_dev);
  }
  hipLaunchKernelGGL(k_gm_solve, dim3(1), dim3(64), 0, st, h.gs);
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

// Preconditioner of the Newton-Krylov solve: M ~ I - dh J in variables scaled by s0
// (M~ = S0^-1 M S0).
//  * z-line: for every (r, theta) column the block-pentadiagonal matrix of the 6x6 cell blocks
//    coupling k with k-2..k+2 (MUSCL stencil) is factored by block LU without inter-block
//    pivoting, M~ = L U, storing per cell L_{k,k-2}, L_{k,k-1}, D_k^-1 (Gauss-Jordan with partial
//    pivoting), U_{k,k+1}, U_{k,k+2}.  The r and theta couplings are left to GMRES.
//  * block Jacobi: D_k^-1 of the diagonal blocks only.
//  * wall rows: point Jacobi on the wall diagonal.
//  * line Gauss-Seidel: the z-line solves above, swept over coloured lines (cell z-lines and wall
//    lines) so that the r and theta couplings are taken into account as well.
#include <mutex>

#include <algorithm>

#include "vtsa_internal.h"
#include "vtsa_reduce.h"

namespace vtsa {

constexpr int PW = 64;           // threads per block (one wavefront)
constexpr int PCOLS = 10;        // factor: columns per wavefront (6 lanes each)
constexpr int AC = 4;            // apply gather/scatter: elements per lane whose loads go together

// ------------------------------------------------------------------ preconditioner factor
// Block LU of the z-line matrices (block Jacobi: the diagonal inverses only).  Six lanes per
// column, lane r computing row r of every block product, PCOLS columns per wavefront; the
// factors of the previous two cells stay in LDS.  The pivot blocks are inverted by Gauss-Jordan
// with partial pivoting (the first row of largest |entry| wins), rows in parallel through LDS.
// Block products accumulate over m in order in T, as C = A B and C -= A B of 6x6 matrices.
// F blocks per cell, rows of 6: 0 L2, 1 L1, 2 Dinv, 3 U1, 4 U2.
// Blocks before nbt factor cells 0 .. kend_top - 1 of each line (all of it, or down to the twist
// pair of the twisted form); blocks from nbt on factor the same lines from their last cell upwards
// (the line with its cells in reverse order: the blocks of slots z+1, z+2 become the lower ones),
// cells 0 .. kend_rev - 1 of the reversed line, into Fr at the reversed cell index.
constexpr int FSTR = 258;   // LDS doubles per column: Dinv/U1/U2 ring of two cells + pivot scratch

template <typename T>
__global__ void __launch_bounds__(PW) k_prec_factor(int n, int Nz, int ncol, int kind,
                                                    const double* __restrict__ Jc, const double* __restrict__ s0,
                                                    double dh, T* __restrict__ F, T* __restrict__ Fr, int nbt,
                                                    int kend_top, int kend_rev) {
  __shared__ double lds[PCOLS * FSTR];
  const int lane = threadIdx.x;
  const int grp = lane / 6, r = lane - grp * 6;
  const bool rev = (int)blockIdx.x >= nbt;
  const int col0 = ((int)blockIdx.x - (rev ? nbt : 0)) * PCOLS;
  const int kend = rev ? kend_rev : kend_top;
  T* const Fo = rev ? Fr : F;
  // slots and cell offsets of the blocks k-2, k-1, k+1, k+2 along the factorisation order
  const int sm2 = rev ? S_ZP2 : S_ZM2, sm1 = rev ? S_ZP1 : S_ZM1, sp1 = rev ? S_ZM1 : S_ZP1,
            sp2 = rev ? S_ZM2 : S_ZP2;
  const int dd = rev ? -1 : 1;
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
  for (int k = 0; k < kend; ++k) {
    const int X = col * Nz + (rev ? Nz - 1 - k : k);
    const double sr = s0[r * n + X];
    // row r of the block of M~ = S0^-1 (I - dh J) S0 between cell X and cell X + dk
    // loaded first (jr: row r of J's block, sc: the scales of the block's columns), formed after
    auto mload = [&](int slot, int dk, double* jr, double* sc) {
      const double* J = Jc + ((size_t)X * NSLOT + slot) * 36 + r * 6;
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        jr[c] = J[c];
        sc[c] = s0[c * n + X + dk];
      }
    };
    auto mform = [&](int slot, const double* jr, const double* sc, T* out) {
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        const double v = (slot == S_SELF && r == c ? 1.0 : 0.0) - dh * jr[c];
        out[c] = (T)(v * (sc[c] / sr));
      }
    };
    const int q = k & 1, q1 = (k + 1) & 1;   // ring slots of cells k-2 (and k) and k-1
    T l2[6], l1[6], d[6], u1[6], u2[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) { l2[c] = T(0); l1[c] = T(0); u1[c] = T(0); u2[c] = T(0); }
    // the cell's five block rows are requested together, one round trip per cell
    double jr[5][6], sc[5][6];
    if (zl && k >= 2) mload(sm2, -2 * dd, jr[0], sc[0]);
    if (zl && k >= 1) mload(sm1, -dd, jr[1], sc[1]);
    mload(S_SELF, 0, jr[2], sc[2]);
    if (zl && k + 1 < Nz) mload(sp1, dd, jr[3], sc[3]);
    if (zl && k + 2 < Nz) mload(sp2, 2 * dd, jr[4], sc[4]);
    T a2[6], a1[6];
    if (zl && k >= 2) mform(sm2, jr[0], sc[0], a2);
    if (zl && k >= 1) mform(sm1, jr[1], sc[1], a1);
    mform(S_SELF, jr[2], sc[2], d);
    if (zl && k + 1 < Nz) mform(sp1, jr[3], sc[3], u1);
    if (zl && k + 2 < Nz) mform(sp2, jr[4], sc[4], u2);
    if (zl && k >= 2) {   // L2 = A_{k,k-2} Dinv_{k-2}
      const T* a = a2;
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        T acc = T(0);
#pragma unroll
        for (int m = 0; m < 6; ++m) acc += a[m] * (T)pDi[q * 36 + m * 6 + c];
        l2[c] = acc;
      }
    }
    if (zl && k >= 1) {   // L1 = (A_{k,k-1} - L2 U1_{k-2}) Dinv_{k-1}
      T* a = a1;
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
    // D = A_kk - L1 U1_{k-1} - L2 U2_{k-2}
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
    // U2 = A_{k,k+2} as loaded
    lds_barrier();   // every lane has read ring slot q (cell k-2) before cell k overwrites it
    T* FX = Fo + (size_t)(col * Nz + k) * 180 + r * 6;
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

// Line Gauss-Seidel wall lines: LU (Thomas) of each wall line's tridiagonal block of M~ (axial
// conduction; entries scaled by s0), one thread per line, sequential in z.  Per cell
// Tw = (lo_k, 1 / den_k, c_k) with den_k = diag_k - lo_k c_{k-1}, c_k = up_k / den_k.
__global__ void k_gs_wall_factor(int Nth, int Nz, int n, const double* __restrict__ Jw,
                                 const double* __restrict__ s0, double dh, double* __restrict__ Tw) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= Nth) return;
  const double* s0w = s0 + 6 * (size_t)n + (size_t)j * Nz;
  const double* Jwl = Jw + (size_t)j * Nz * 5;
  double* T3 = Tw + (size_t)j * Nz * 3;
  double cprev = 0.0;
  for (int k = 0; k < Nz; ++k) {
    const double diag = 1.0 - dh * Jwl[k * 5 + 0];
    const double lo = k > 0 ? (-dh * Jwl[k * 5 + 1]) * (s0w[k - 1] / s0w[k]) : 0.0;
    const double up = k + 1 < Nz ? (-dh * Jwl[k * 5 + 2]) * (s0w[k + 1] / s0w[k]) : 0.0;
    const double den = diag - lo * cprev;
    cprev = up / den;
    T3[k * 3 + 0] = lo;
    T3[k * 3 + 1] = 1.0 / den;
    T3[k * 3 + 2] = cprev;
  }
}

// ------------------------------------------------------------------ preconditioner apply
// out = post^-1 * S0 M~^-1 S0^-1 (pre * scale * in), side outputs as described at ApplyArgs.
//
// One z-column per 64-lane block.  The column's right-hand side is gathered into LDS
// (xs[k][6]) with batched coalesced loads; the factor rows are staged into LDS by direct
// global-to-LDS copies in segments of ZSEG cells, double-buffered so the next segment is in
// flight while the current one is swept.  Lanes 0..5 sweep (lane r owns row r); each new
// 6-vector is broadcast to every lane with v_readlane, so the chain of cells along z carries
// no LDS round trip.  Wall rows: point Jacobi in blocks of their own.
constexpr int ZSEG = 16;   // cells per staged segment

template <typename T>
__host__ __device__ constexpr size_t zline_buf_bytes() {   // one staging buffer, whole 1 KiB copies
  return ((size_t)ZSEG * 108 * sizeof(T) + 1023) / 1024 * 1024;
}
template <typename T>
__host__ __device__ constexpr size_t zline_lds_bytes(int Nz) {
  return sizeof(double) * ((size_t)Nz * 6 + 8) + 2 * zline_buf_bytes<T>();
}

// Copy `per` factor elements (starting at element `off` of each cell's 180) of cells
// [kbeg, kbeg + kc) of the column starting at cell X0 into dst, packed per cell: 16-byte
// chunks, 64 per instruction, straight into LDS.  Chunks past the end re-read the first one.
template <typename T>
__device__ inline void stage_rows(const T* __restrict__ F, int X0, int kbeg, int kc, int off, int per, T* dst) {
  const int cpc = per * (int)sizeof(T) / 16;          // chunks per cell
  const int total = kc * cpc;
  const char* base = reinterpret_cast<const char*>(F + (size_t)(X0 + kbeg) * 180 + off);
  char* d = reinterpret_cast<char*>(dst);
  for (int j = 0; j * 64 < total; ++j) {
    const int c = j * 64 + (int)threadIdx.x;
    const int cc = c < total ? c : 0;
    const int cell = cc / cpc, q = cc - cell * cpc;
    const char* src = base + (size_t)cell * 180 * sizeof(T) + (size_t)q * 16;
    __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)(d + j * 1024), 16, 0, 0);
  }
}

// The apply's blocks are one wavefront, whose LDS instructions execute in issue order, so LDS
// traffic between its lanes needs only a compiler ordering point (no s_barrier, and no wait for
// the staged global-to-LDS copies still in flight).
static_assert(PW == 64, "the apply kernel assumes one-wavefront blocks");
__device__ inline void wave_lds_sync() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront"); }

// ---- block rows across lanes
// A cell's 6-vector lives one component per lane: lane l holds component line_row(l), i.e. lanes
// 0..5 of every 16-lane DPP row hold components 0..5 and lanes 6..15 repeat component 0, so all
// four rows of a wavefront carry the same values.  A dot product of a block row with a 6-vector
// then takes component c from lane c of the lane's own row with DPP `row_newbcast:c`, folded into
// the FP64 FMA (v_fmac_f64_dpp), with no broadcast instructions of its own.  Each block starts
// with `s_nop 1`: a DPP read of a VGPR needs two wait states after the VALU write of that VGPR.
__device__ inline int line_row(int lane) { return (lane & 15) < 6 ? (lane & 15) : 0; }

#define VTSA_FMAC_BC(acc, v, l, c) "v_fmac_f64_dpp " acc ", " v ", " l " row_newbcast:" #c " row_mask:0xf bank_mask:0xf\n\t"

// b - L1 v1 - L2 v2 for block rows l1, l2 (this lane's row) and vectors v1, v2 (this lane's
// component), as ((b - sum_{c<3} l2_c v2_c) + (-sum_{c>=3} l2_c v2_c))
//   + ((-sum_{c<3} l1_c v1_c) + (-sum_{c>=3} l1_c v1_c)),
// each partial sum accumulated with c in order.  The four chains are interleaved, so no FMA
// waits for the one before it.
__device__ inline double bsub12(double b, double v1, const double (&l1)[6], double v2, const double (&l2)[6]) {
  double a2 = 0.0, c1 = 0.0, c2 = 0.0;
  asm("s_nop 1\n\t"
      VTSA_FMAC_BC("%0", "-%5", "%12", 0) VTSA_FMAC_BC("%1", "-%5", "%15", 3)
      VTSA_FMAC_BC("%2", "-%4", "%6", 0)  VTSA_FMAC_BC("%3", "-%4", "%9", 3)
      VTSA_FMAC_BC("%0", "-%5", "%13", 1) VTSA_FMAC_BC("%1", "-%5", "%16", 4)
      VTSA_FMAC_BC("%2", "-%4", "%7", 1)  VTSA_FMAC_BC("%3", "-%4", "%10", 4)
      VTSA_FMAC_BC("%0", "-%5", "%14", 2) VTSA_FMAC_BC("%1", "-%5", "%17", 5)
      VTSA_FMAC_BC("%2", "-%4", "%8", 2)  VTSA_FMAC_BC("%3", "-%4", "%11", 5)
      : "+v"(b), "+v"(a2), "+v"(c1), "+v"(c2)
      : "v"(v1), "v"(v2), "v"(l1[0]), "v"(l1[1]), "v"(l1[2]), "v"(l1[3]), "v"(l1[4]), "v"(l1[5]),
        "v"(l2[0]), "v"(l2[1]), "v"(l2[2]), "v"(l2[3]), "v"(l2[4]), "v"(l2[5]));
  return (b + a2) + (c1 + c2);
}

// L v for block row l: ((l_0 v_0 + l_1 v_1) + (l_2 v_2 + l_3 v_3)) + (l_4 v_4 + l_5 v_5)
__device__ inline double bdot6(double v, const double (&l)[6]) {
  double p = 0.0, q = 0.0, w = 0.0;
  asm("s_nop 1\n\t"
      VTSA_FMAC_BC("%0", "%3", "%4", 0) VTSA_FMAC_BC("%1", "%3", "%6", 2) VTSA_FMAC_BC("%2", "%3", "%8", 4)
      VTSA_FMAC_BC("%0", "%3", "%5", 1) VTSA_FMAC_BC("%1", "%3", "%7", 3) VTSA_FMAC_BC("%2", "%3", "%9", 5)
      : "+v"(p), "+v"(q), "+v"(w)
      : "v"(v), "v"(l[0]), "v"(l[1]), "v"(l[2]), "v"(l[3]), "v"(l[4]), "v"(l[5]));
  return (p + q) + w;
}
#undef VTSA_FMAC_BC

// One cell of the forward block substitution: z_k = b_k - L1 z_{k-1} - L2 z_{k-2} (component
// line_row(lane) of each vector in this lane).
__device__ inline double fwd_cell(double b, const double (&l1)[6], const double (&l2)[6], double zm1, double zm2) {
  return bsub12(b, zm1, l1, zm2, l2);
}

// One cell of the backward substitution: x_k = Dinv (z_k - U1 x_{k+1} - U2 x_{k+2}).
__device__ inline double bwd_cell(double z, const double (&d)[6], const double (&u1)[6], const double (&u2)[6],
                                  double xp1, double xp2) {
  return bdot6(bsub12(z, xp1, u1, xp2, u2), d);
}

// Wait for this wavefront's staged copies, then barrier (one-wavefront blocks).
__device__ inline void stage_wait() {
  __builtin_amdgcn_s_waitcnt(0);
  lds_barrier();
}

// ------------------------------------------------------------------ z-line column solve
// Shared by the z-line apply and the line Gauss-Seidel half-sweeps.  xs[k][6] holds the column's
// right-hand side in LDS and is overwritten with M~_line^-1 xs.  All 64 lanes call; lanes 0..5
// own the rows.  zline_stage_first starts moving the first forward segment and may be issued
// before the caller fills xs.

template <typename T>
__device__ inline void zline_stage_first(const T* __restrict__ F, int X0, int Nz, T* fb0) {
  stage_rows<T>(F, X0, 0, min(ZSEG, Nz), 0, 72, fb0);
}

template <typename T>
__device__ inline void zline_sweeps(const T* __restrict__ F, int X0, int Nz, double* xs, T* fb0, T* fb1) {
  const int lane = threadIdx.x;
  const bool active = lane < 6;
  const int r = line_row(lane);
  const int nseg = (Nz + ZSEG - 1) / ZSEG;
  // component r of the last two cells' vectors
  double zm1 = 0.0, zm2 = 0.0;
  // forward: z_k = b_k - L1 z_{k-1} - L2 z_{k-2}
  for (int sg = 0; sg < nseg; ++sg) {
    const int k0 = sg * ZSEG, kc = min(ZSEG, Nz - k0);
    T* cur = (sg & 1) ? fb1 : fb0;
    stage_wait();
    if (sg + 1 < nseg) stage_rows<T>(F, X0, k0 + ZSEG, min(ZSEG, Nz - k0 - ZSEG), 0, 72, (sg & 1) ? fb0 : fb1);
    // row r of L2 at [0, 36), of L1 at [36, 72) of each cell; the next cell's rows and right-hand
    // side are read while the current cell is computed
    double l1[6], l2[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) { l1[c] = (double)cur[r * 6 + 36 + c]; l2[c] = (double)cur[r * 6 + c]; }
    double bk = xs[k0 * 6 + r];
    for (int kk = 0; kk < kc; ++kk) {
      const int k = k0 + kk;
      const int kn = kk + 1 < kc ? kk + 1 : kk;
      double n1[6], n2[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) { n1[c] = (double)cur[kn * 72 + r * 6 + 36 + c]; n2[c] = (double)cur[kn * 72 + r * 6 + c]; }
      const double bn = xs[(k0 + kn) * 6 + r];
      const double zk = fwd_cell(bk, l1, l2, zm1, zm2);   // the factor holds zero blocks where k < 2
      if (active) xs[k * 6 + r] = zk;
      zm2 = zm1;
      zm1 = zk;
#pragma unroll
      for (int c = 0; c < 6; ++c) { l1[c] = n1[c]; l2[c] = n2[c]; }
      bk = bn;
    }
  }
  // backward: x_k = Dinv (z_k - U1 x_{k+1} - U2 x_{k+2}); rows Dinv | U1 | U2 per cell
  {
    const int k0 = (nseg - 1) * ZSEG;
    stage_wait();   // the forward sweep is done with both buffers
    stage_rows<T>(F, X0, k0, Nz - k0, 72, 108, fb0);
  }
  double xp1 = 0.0, xp2 = 0.0;
  for (int sg = nseg - 1, b = 0; sg >= 0; --sg, b ^= 1) {
    const int k0 = sg * ZSEG, kc = min(ZSEG, Nz - k0);
    T* cur = b ? fb1 : fb0;
    stage_wait();
    if (sg > 0) stage_rows<T>(F, X0, k0 - ZSEG, ZSEG, 72, 108, b ? fb0 : fb1);
    // row r of Dinv at [0, 36), U1 at [36, 72), U2 at [72, 108) of each cell; the next (lower)
    // cell's rows and z are read while the current cell is computed
    double dr[6], u1[6], u2[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const T* R = cur + (kc - 1) * 108 + r * 6;
      dr[c] = (double)R[c]; u1[c] = (double)R[36 + c]; u2[c] = (double)R[72 + c];
    }
    double zk = xs[(k0 + kc - 1) * 6 + r];
    for (int kk = kc - 1; kk >= 0; --kk) {
      const int k = k0 + kk;
      const int kn = kk > 0 ? kk - 1 : kk;
      const T* Rn = cur + kn * 108 + r * 6;
      double nd[6], n1[6], n2[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) { nd[c] = (double)Rn[c]; n1[c] = (double)Rn[36 + c]; n2[c] = (double)Rn[72 + c]; }
      const double zn = xs[(k0 + kn) * 6 + r];
      const double xk = bwd_cell(zk, dr, u1, u2, xp1, xp2);   // zero blocks where k + 2 > Nz - 1
      if (active) xs[k * 6 + r] = xk;
      xp2 = xp1;
      xp1 = xk;
#pragma unroll
      for (int c = 0; c < 6; ++c) { dr[c] = nd[c]; u1[c] = n1[c]; u2[c] = n2[c]; }
      zk = zn;
    }
  }
  lds_barrier();
}

// ------------------------------------------------------------------ whole-line pair sweeps
// When the device allows a block that much LDS, a line Gauss-Seidel half-sweep copies its line's
// factor into LDS at kernel start (the copy runs under the right-hand-side gather, spread over
// the block's GSW threads) and solves the line two cells at a time.  Over pairs of cells
// (2K, 2K+1) the block-pentadiagonal L U of a line is block bidiagonal:
//   forward   z_K = b'_K - G_K (z_{2K-2}, z_{2K-1}),  b'_K = (b_{2K}, b_{2K+1} - L1_{2K+1} b_{2K})
//   backward  x_K = z'_K - H_K (x_{2K+2}, x_{2K+3}),  z'_K = (h_{2K} - ^U1_{2K} h_{2K+1}, h_{2K+1}),
//             h_k = Dinv_k z_k,  ^U = Dinv U
//   G_K = [ L2_{2K}               L1_{2K}                     ]
//         [ -L1_{2K+1} L2_{2K}    L2_{2K+1} - L1_{2K+1} L1_{2K} ]
//   H_K = [ ^U2_{2K} - ^U1_{2K} ^U1_{2K+1}    -^U1_{2K} ^U2_{2K+1} ]
//         [ ^U1_{2K+1}                        ^U2_{2K+1}           ]
// so each step is one 12-wide dot product per row (rows in lanes 0..11), half as many dependent
// steps as cell by cell, and b', h and z' are data-parallel passes of the whole block.  A line
// with odd Nz ends in a pair whose second cell is a zero phantom.  Per pair, PP values:
// G (12 x 12) | H (12 x 12) | L1_{2K+1} | Dinv_{2K} | Dinv_{2K+1} | ^U1_{2K}, rows of 12 or 6; the
// two recurrence matrices come first, so that a line whose whole factor does not fit the LDS
// budget can stage only them (PGH values per pair) and read the rest from global memory.
constexpr int GSW = 512;   // threads of a whole-line half-sweep block
constexpr int PP = 432;
constexpr int PGH = 288;
constexpr int PO_G = 0, PO_H = 144, PO_C = 288, PO_D = 324, PO_U = 396;

__host__ __device__ constexpr int npairs(int Nz) { return (Nz + 1) / 2; }
// xs and zs, [2 npairs + 3][6] each: the last cell a spare slot, then a junk pair for the twisted
// recurrences' padding steps; at least TW_AREA doubles, since zs holds the twist record until the
// forward recurrences end
constexpr int TW_SIG = 0, TW_SIGR = 144, TW_GINV = 288, TW_XCH = 432, TW_AREA = 456;
__host__ __device__ constexpr size_t pair_vec_doubles(int Nz) {
  return (size_t)(2 * npairs(Nz) + 3) * 6 > (size_t)TW_AREA ? (size_t)(2 * npairs(Nz) + 3) * 6 : (size_t)TW_AREA;
}
template <typename T>
__host__ __device__ constexpr size_t pair_buf_bytes(int Nz) {   // whole copies of GSW x 16 bytes
  return ((size_t)npairs(Nz) * PP * sizeof(T) + GSW * 16 - 1) / (GSW * 16) * (GSW * 16);
}
template <typename T>
__host__ __device__ constexpr size_t gs_full_lds_bytes(int Nz) {
  return sizeof(double) * 2 * pair_vec_doubles(Nz) + pair_buf_bytes<T>(Nz);
}
template <typename T>
__host__ __device__ constexpr size_t pair_gh_buf_bytes(int Nz) {   // whole copies of GSW x 16 bytes
  return ((size_t)npairs(Nz) * PGH * sizeof(T) + GSW * 16 - 1) / (GSW * 16) * (GSW * 16);
}
template <typename T>
__host__ __device__ constexpr size_t gs_gh_lds_bytes(int Nz) {
  return sizeof(double) * 2 * pair_vec_doubles(Nz) + pair_gh_buf_bytes<T>(Nz);
}
template <typename T>
__host__ __device__ constexpr size_t pair_g_buf_bytes(int Nz) {   // one 12 x 12 matrix per pair
  return ((size_t)npairs(Nz) * 144 * sizeof(T) + GSW * 16 - 1) / (GSW * 16) * (GSW * 16);
}
template <typename T>
__host__ __device__ constexpr size_t gs_g_lds_bytes(int Nz) {
  return sizeof(double) * 2 * pair_vec_doubles(Nz) + pair_g_buf_bytes<T>(Nz);
}

size_t gs_pairs_count(int ncol, int Nz) { return (size_t)ncol * npairs(Nz) * PP; }

// Pair form of the factors, one thread per (line, pair, row r of 6), PPB pairs per block: the
// products every row needs, ^U1_{2K+1} = Dinv U1 and ^U2_{2K+1} = Dinv U2 of the pair's second cell,
// are formed once per pair (row r by thread r) into LDS.  Products accumulate over m in order in
// FP64 and are stored in T.  Twisted form (tw_m >= 0): pairs K < tw_m from F, pairs K > tw_m from
// the reversed factor Fr (pair NP - 1 - K of the reversed line, its two cells in reverse order),
// pair tw_m written by k_gs_twist.
constexpr int PPB = 32;   // pairs per block of the pair-form builder (6 threads each)
template <typename T>
__global__ void __launch_bounds__(PPB * 6) k_gs_pairs(int Nz, int ncol, const T* __restrict__ F, T* __restrict__ P,
                                                      const T* __restrict__ Fr, int tw_m) {
  __shared__ double su1[PPB][36], su2[PPB][36];
  const int NP = npairs(Nz);
  const int lp = threadIdx.x / 6, r = threadIdx.x - lp * 6;
  const int pr = blockIdx.x * PPB + lp;
  const int L = pr / NP, K = pr - L * NP;
  const bool valid = pr < ncol * NP && !(tw_m >= 0 && K == tw_m);
  const bool bot = tw_m >= 0 && K > tw_m;
  const T* src = bot ? Fr : F;
  const int Ks = bot ? NP - 1 - K : K;
  const int k0 = 2 * Ks, k1 = 2 * Ks + 1;
  // entry (i, c) of block b (0 L2, 1 L1, 2 Dinv, 3 U1, 4 U2) of cell k; zero for the phantom
  auto fb = [&](int k, int b, int i, int c) -> double {
    return valid && k < Nz ? (double)src[(size_t)(L * Nz + k) * 180 + b * 36 + i * 6 + c] : 0.0;
  };
  {   // row r of ^U1_{2K+1}, ^U2_{2K+1}
    double d[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) d[m] = fb(k1, 2, r, m);
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double a1 = 0.0, a2 = 0.0;
#pragma unroll
      for (int m = 0; m < 6; ++m) { a1 += d[m] * fb(k1, 3, m, c); a2 += d[m] * fb(k1, 4, m, c); }
      su1[lp][r * 6 + c] = a1;
      su2[lp][r * 6 + c] = a2;
    }
  }
  __syncthreads();
  if (!valid) return;
  double w1[6], w2[6];   // row r of ^U1_{2K}, ^U2_{2K}
  {
    double d[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) d[m] = fb(k0, 2, r, m);
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double a1 = 0.0, a2 = 0.0;
#pragma unroll
      for (int m = 0; m < 6; ++m) { a1 += d[m] * fb(k0, 3, m, c); a2 += d[m] * fb(k0, 4, m, c); }
      w1[c] = a1;
      w2[c] = a2;
    }
  }
  double l1[6];   // row r of L1_{2K+1}
#pragma unroll
  for (int m = 0; m < 6; ++m) l1[m] = fb(k1, 1, r, m);
  T* out = P + (size_t)pr * PP;
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    double a = 0.0, b2 = 0.0, p = 0.0, q = 0.0;
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      a += l1[m] * fb(k0, 0, m, c);         // (L1_{2K+1} L2_{2K})[r][c]
      b2 += l1[m] * fb(k0, 1, m, c);        // (L1_{2K+1} L1_{2K})[r][c]
      p += w1[m] * su1[lp][m * 6 + c];      // (^U1_{2K} ^U1_{2K+1})[r][c]
      q += w1[m] * su2[lp][m * 6 + c];      // (^U1_{2K} ^U2_{2K+1})[r][c]
    }
    out[PO_G + r * 12 + c] = (T)fb(k0, 0, r, c);
    out[PO_G + r * 12 + 6 + c] = (T)fb(k0, 1, r, c);
    out[PO_G + (6 + r) * 12 + c] = (T)(-a);
    out[PO_G + (6 + r) * 12 + 6 + c] = (T)(fb(k1, 0, r, c) - b2);
    out[PO_C + r * 6 + c] = (T)l1[c];
    out[PO_H + r * 12 + c] = (T)(w2[c] - p);
    out[PO_H + r * 12 + 6 + c] = (T)(-q);
    out[PO_H + (6 + r) * 12 + c] = (T)su1[lp][r * 6 + c];
    out[PO_H + (6 + r) * 12 + 6 + c] = (T)su2[lp][r * 6 + c];
    out[PO_D + r * 6 + c] = (T)fb(k0, 2, r, c);
    out[PO_D + 36 + r * 6 + c] = (T)fb(k1, 2, r, c);
    out[PO_U + r * 6 + c] = (T)w1[c];
  }
}

// Copy the NP pairs of line L into dst: 16-byte chunks, GSW per pass (64 per wavefront
// instruction), straight into LDS.  Lanes past the end re-read the last chunk into the buffer's
// padding.  Called by all GSW threads of the block.
template <typename T>
__device__ inline void stage_pairs(const T* __restrict__ P, int L, int NP, T* dst) {
  const int total = NP * PP * (int)sizeof(T) / 16;
  const int tid = threadIdx.x, w = tid >> 6;
  const char* base = reinterpret_cast<const char*>(P + (size_t)L * NP * PP);
  char* d = reinterpret_cast<char*>(dst);
  for (int j = 0; j * GSW < total; ++j) {
    const int c = j * GSW + tid;
    const int cc = c < total ? c : total - 1;
    __builtin_amdgcn_global_load_lds(base + (size_t)cc * 16,
                                     (__attribute__((address_space(3))) void*)(d + (size_t)j * GSW * 16 + w * 1024),
                                     16, 0, 0);
  }
}

// Copy values [OFF, OFF + CNT) of each of the NP pairs of line L into dst, packed at CNT values
// per pair: 16-byte chunks, GSW per pass, straight into LDS; chunk c of the packed run comes from
// pair c / cpp, chunk c % cpp of that pair.  Lanes past the end re-read the last chunk.
template <typename T, int OFF, int CNT>
__device__ inline void stage_pairs_part(const T* __restrict__ P, int L, int NP, T* dst) {
  constexpr int cpp = CNT * (int)sizeof(T) / 16;
  static_assert(CNT * sizeof(T) % 16 == 0 && OFF * sizeof(T) % 16 == 0, "16-byte chunks");
  const int total = NP * cpp;
  const int tid = threadIdx.x, w = tid >> 6;
  const char* base = reinterpret_cast<const char*>(P + (size_t)L * NP * PP + OFF);
  char* d = reinterpret_cast<char*>(dst);
  for (int j = 0; j * GSW < total; ++j) {
    const int c = j * GSW + tid;
    const int cc = c < total ? c : total - 1;
    const int K = cc / cpp, q = cc - K * cpp;
    __builtin_amdgcn_global_load_lds(base + ((size_t)K * PP * sizeof(T) + (size_t)q * 16),
                                     (__attribute__((address_space(3))) void*)(d + (size_t)j * GSW * 16 + w * 1024),
                                     16, 0, 0);
  }
}

// b - sum_{c<12} g_c v_c, v_c taken from lane c of the lane's 16-lane row (DPP row_newbcast), as
// (b - g0 v0 - g1 v1 - .. - g5 v5) + (-g6 v6 - .. - g11 v11): two interleaved chains, so no FMA
// waits for the one before it and a single add joins them.
#define VTSA_FMAC_BC(acc, v, l, c) "v_fmac_f64_dpp " acc ", " v ", " l " row_newbcast:" #c " row_mask:0xf bank_mask:0xf\n\t"
__device__ inline double bsub_pair(double b, double v, const double (&g)[12]) {
  double a2 = 0.0;
  asm("s_nop 1\n\t"
      VTSA_FMAC_BC("%0", "-%2", "%3", 0) VTSA_FMAC_BC("%1", "-%2", "%9", 6)
      VTSA_FMAC_BC("%0", "-%2", "%4", 1) VTSA_FMAC_BC("%1", "-%2", "%10", 7)
      VTSA_FMAC_BC("%0", "-%2", "%5", 2) VTSA_FMAC_BC("%1", "-%2", "%11", 8)
      VTSA_FMAC_BC("%0", "-%2", "%6", 3) VTSA_FMAC_BC("%1", "-%2", "%12", 9)
      VTSA_FMAC_BC("%0", "-%2", "%7", 4) VTSA_FMAC_BC("%1", "-%2", "%13", 10)
      VTSA_FMAC_BC("%0", "-%2", "%8", 5) VTSA_FMAC_BC("%1", "-%2", "%14", 11)
      : "+v"(b), "+v"(a2)
      : "v"(v), "v"(g[0]), "v"(g[1]), "v"(g[2]), "v"(g[3]), "v"(g[4]), "v"(g[5]), "v"(g[6]), "v"(g[7]),
        "v"(g[8]), "v"(g[9]), "v"(g[10]), "v"(g[11]));
  return b + a2;
}
#undef VTSA_FMAC_BC

// One direction of the pair recurrence, s_K = rhs[K] - M_K s_{K -+ 1}, over the NP pairs of a line
// resident in LDS (M at offset `off` of each pair: G forward, H backward).  Run by lanes 0..11 of
// the block's first wavefront, lane rho holding row rho of a pair; the caller synchronises.  A
// pair's rows and right-hand side are read one pair ahead.  LDS order per pair: the next pair's
// rows, then this pair's store; the look-ahead index is clamped rather than tested, the
// scheduler may not move the reads (sched_barrier), and every path into a pair has one store
// between its rows and the next pair's, so a pair waits for its own rows only.
template <typename T, bool FORWARD, int STRIDE>
__device__ inline void pair_recurrence(const T* Pl, int NP, int off, const double* rhs, double* out) {
  const int rho = threadIdx.x;
  struct Row { double g[12]; double b; };
  const int first = FORWARD ? 0 : NP - 1, step = FORWARD ? 1 : -1;
  double s = 0.0;   // lane rho: component rho of the previous pair's solution
  if constexpr (sizeof(T) == sizeof(double)) {
    auto ld = [&](int K, Row& w) {
      const T* R = Pl + (size_t)K * STRIDE + off + rho * 12;
#pragma unroll
      for (int c = 0; c < 12; ++c) w.g[c] = (double)R[c];
      w.b = rhs[K * 12 + rho];
    };
    Row wa, wb;
    ld(first, wa);
    out[NP * 12 + (rho < 6 ? rho : rho - 6)] = 0.0;   // spare slot
    int K = first, left = NP;
    for (; left >= 2; K += 2 * step, left -= 2) {
      ld(K + step, wb);
      __builtin_amdgcn_sched_barrier(0);
      s = bsub_pair(wa.b, s, wa.g);
      out[K * 12 + rho] = s;
      __builtin_amdgcn_sched_barrier(0);
      ld(left > 2 ? K + 2 * step : K + step, wa);
      __builtin_amdgcn_sched_barrier(0);
      s = bsub_pair(wb.b, s, wb.g);
      out[(K + step) * 12 + rho] = s;
      __builtin_amdgcn_sched_barrier(0);
    }
    if (left == 1) out[K * 12 + rho] = bsub_pair(wa.b, s, wa.g);
  } else {
    // Narrow rows: loaded two pairs ahead and widened one pair ahead, so that the conversions
    // read registers whose loads were issued a whole pair step earlier instead of waiting for
    // the load just issued.  Look-ahead indices past the end are clamped to the last pair.
    struct Raw { T g[12]; double b; };
    auto issue = [&](int K, Raw& w) {
      const T* R = Pl + (size_t)K * STRIDE + off + rho * 12;
#pragma unroll
      for (int c = 0; c < 12; ++c) w.g[c] = R[c];
      w.b = rhs[K * 12 + rho];
    };
    auto widen = [&](const Raw& r, Row& w) {
#pragma unroll
      for (int c = 0; c < 12; ++c) w.g[c] = (double)r.g[c];
      w.b = r.b;
    };
    const int last = FORWARD ? NP - 1 : 0;
    auto ahead = [&](int K, int d) {   // K + d steps, clamped to the last pair
      const int k2 = K + d * step;
      return FORWARD ? (k2 > last ? last : k2) : (k2 < last ? last : k2);
    };
    Raw ra, rb;
    Row ca, cb;
    issue(first, ra);
    widen(ra, ca);
    issue(ahead(first, 1), rb);
    out[NP * 12 + (rho < 6 ? rho : rho - 6)] = 0.0;   // spare slot
    int K = first, left = NP;
    for (; left >= 2; K += 2 * step, left -= 2) {
      __builtin_amdgcn_sched_barrier(0);
      s = bsub_pair(ca.b, s, ca.g);
      out[K * 12 + rho] = s;
      __builtin_amdgcn_sched_barrier(0);
      widen(rb, cb);
      issue(ahead(K, 2), ra);
      __builtin_amdgcn_sched_barrier(0);
      s = bsub_pair(cb.b, s, cb.g);
      out[(K + step) * 12 + rho] = s;
      __builtin_amdgcn_sched_barrier(0);
      widen(ra, ca);
      issue(ahead(K, 3), rb);
    }
    if (left == 1) out[K * 12 + rho] = bsub_pair(ca.b, s, ca.g);
  }
}

// ------------------------------------------------------------------ twisted pair solve
// A line of an even number of cells is also factored from its last cell upwards (k_prec_factor,
// reversed blocks), and the two factorisations meet in the twist pair m = (NP - 1) / 2.  At pair
// level (pair K = cells 2K, 2K+1) the line matrix is block tridiagonal; with the top factor's
//   Sigma_K = [L2_{2K} L1_{2K}; 0 L2_{2K+1}]          (L, pair K on pair K-1)
//   Ups_K   = [U2_{2K} 0; U1_{2K+1} U2_{2K+1}]         (U, pair K on pair K+1)
// and ~Sigma, ~Ups the same of the reversed line, the forward pair recurrence runs from the top
// over pairs 0 .. m-1 and from the bottom over pairs NP-1 .. m+1 (in reversed cell order), then
//   x_m = Gamma^-1 (b_m - Sigma_m z_{m-1} - P ~Sigma_{m'} ~z_{m'-1}),
//   Gamma = A_m - Sigma_m Ups_{m-1} - P ~Sigma_{m'} ~Ups_{m'-1} P,
// with A_m pair m's diagonal block, m' = NP - 1 - m and P the swap of a pair's two cells, and the
// backward recurrences run from x_m outwards, top over pairs m-1 .. 0 and bottom over m+1 .. NP-1.
// Each half is the classic LU solve of its end of the line, so the pairs other than m keep their
// pair form (from Fr, in reversed cell order, below the twist); record m holds Sigma_m, ~Sigma_{m'}
// and Gamma^-1 (TW_SIG, TW_SIGR, TW_GINV).  Half as many dependent steps as the classic solve.

// element of a line vector holding component rho of a pair taken in reversed cell order
__device__ inline int tw_perm(int rho) { return rho < 6 ? rho + 6 : rho - 6; }

// Record m of every line: one 64-lane block per line, lanes 0..11 owning rows; Gamma and its
// Gauss-Jordan inverse (partial pivoting, the first row of largest |entry| wins) in FP64.
template <typename T>
__global__ void __launch_bounds__(64) k_gs_twist(int n, int Nz, const double* __restrict__ Jc,
                                                 const double* __restrict__ s0, double dh, const T* __restrict__ F,
                                                 const T* __restrict__ Fr, int m, T* __restrict__ P) {
  __shared__ double Sg[144], Up[144], Sr[144], Ur[144], Am[144];
  __shared__ double colp[12], sw[48], prow[24];
  const int L = blockIdx.x, lane = threadIdx.x;
  const int NP = npairs(Nz), mr = NP - 1 - m;
  auto fe = [&](const T* src, int k, int b, int i, int c) -> double {
    return (double)src[(size_t)(L * Nz + k) * 180 + b * 36 + i * 6 + c];
  };
  for (int e = lane; e < 144; e += 64) {
    const int row = e / 12, col = e - row * 12;
    const int ci = row / 6, i = row - ci * 6, cj = col / 6, c = col - cj * 6;
    auto sig = [&](const T* src, int K) -> double {   // Sigma_K (rows: cells 2K, 2K+1; columns: 2K-2, 2K-1)
      if (ci == 0) return fe(src, 2 * K, cj == 0 ? 0 : 1, i, c);
      return cj == 0 ? 0.0 : fe(src, 2 * K + 1, 0, i, c);
    };
    auto ups = [&](const T* src, int K) -> double {   // Ups_K (rows: cells 2K, 2K+1; columns: 2K+2, 2K+3)
      if (ci == 0) return cj == 0 ? fe(src, 2 * K, 4, i, c) : 0.0;
      return fe(src, 2 * K + 1, cj == 0 ? 3 : 4, i, c);
    };
    Sg[e] = sig(F, m);
    Up[e] = ups(F, m - 1);
    Sr[e] = sig(Fr, mr);
    Ur[e] = ups(Fr, mr - 1);
    // A_m = S0^-1 (I - dh J) S0 between cells 2m + ci and 2m + cj
    const int X = L * Nz + 2 * m + ci, dk = cj - ci;
    const int slot = dk == 0 ? S_SELF : dk > 0 ? S_ZP1 : S_ZM1;
    const double v = (dk == 0 && i == c ? 1.0 : 0.0) - dh * Jc[((size_t)X * NSLOT + slot) * 36 + i * 6 + c];
    Am[e] = v * (s0[c * n + X + dk] / s0[i * n + X]);
  }
  __syncthreads();
  const int rho = lane < 12 ? lane : 0, pr = tw_perm(rho);
  double g[12], iv[12];
#pragma unroll
  for (int c = 0; c < 12; ++c) {
    const int pc = tw_perm(c);
    double a = 0.0, b = 0.0;
    for (int q = 0; q < 12; ++q) {
      a += Sg[rho * 12 + q] * Up[q * 12 + c];
      b += Sr[pr * 12 + q] * Ur[q * 12 + pc];
    }
    g[c] = (Am[rho * 12 + c] - a) - b;
    iv[c] = rho == c ? 1.0 : 0.0;
  }
  const bool own = lane < 12;
#pragma unroll
  for (int p = 0; p < 12; ++p) {
    if (own) colp[lane] = fabs(g[p]);
    __syncthreads();
    int piv = p;
    double best = colp[p];
    for (int rr = p + 1; rr < 12; ++rr) {
      const double v = colp[rr];
      if (v > best) { best = v; piv = rr; }
    }
    if (own && (lane == p || lane == piv)) {
      double* dst = sw + (lane == p ? 0 : 24);
#pragma unroll
      for (int c = 0; c < 12; ++c) { dst[c] = g[c]; dst[12 + c] = iv[c]; }
    }
    __syncthreads();
    if (own && piv != p && (lane == p || lane == piv)) {
      const double* src = sw + (lane == p ? 24 : 0);
#pragma unroll
      for (int c = 0; c < 12; ++c) { g[c] = src[c]; iv[c] = src[12 + c]; }
    }
    if (lane == p) {
      const double inv = 1.0 / g[p];
#pragma unroll
      for (int c = 0; c < 12; ++c) { g[c] *= inv; iv[c] *= inv; prow[c] = g[c]; prow[12 + c] = iv[c]; }
    }
    __syncthreads();
    if (own && lane != p) {
      const double f = g[p];
      if (f != 0.0) {
#pragma unroll
        for (int c = 0; c < 12; ++c) { g[c] -= f * prow[c]; iv[c] -= f * prow[12 + c]; }
      }
    }
    __syncthreads();
  }
  T* R = P + ((size_t)L * NP + m) * PP;
  for (int e = lane; e < 144; e += 64) {
    R[TW_SIG + e] = (T)Sg[e];
    R[TW_SIGR + e] = (T)Sr[e];
  }
  if (own) {
#pragma unroll
    for (int c = 0; c < 12; ++c) R[TW_GINV + lane * 12 + c] = (T)iv[c];
  }
}

// The twisted recurrences, forward (top pairs 0 .. m-1, bottom NP-1 .. m+1, from zero) or backward
// (top m-1 .. 0, bottom m+1 .. NP-1, from s = x_m): lanes 0..11 the top rows, lanes 16..27 the
// bottom rows (reversed cell order), each group in its own 16-lane DPP row, so one instruction
// stream advances both.  The bottom group has NP-1-m steps, the top m, one fewer when NP is even:
// the top's extra step re-reads its last pair and stores into the junk pair after the spare slot.
// Loads run ahead as in pair_recurrence; the caller synchronises.
template <typename T, bool FORWARD, int STRIDE>
__device__ inline double pair_recurrence_tw(const T* Pl, int NP, int m, int off, const double* rhs, double* out,
                                            double s) {
  const int lane = threadIdx.x;
  const bool bot = lane >= 16;
  const int rho = lane & 15;
  const int S = NP - 1 - m, cnt = bot ? S : m;
  const int K0 = FORWARD ? (bot ? NP - 1 : 0) : (bot ? m + 1 : m - 1);
  const int dK = (FORWARD != bot) ? 1 : -1;
  const int el = bot ? tw_perm(rho) : rho;
  const int junk = NP * 12 + 6 + rho;
  auto pair_of = [&](int i) { return K0 + dK * (i < cnt ? i : cnt - 1); };
  auto oidx = [&](int i) { return i < cnt ? (K0 + dK * i) * 12 + el : junk; };
  struct Row { double g[12]; double b; };
  if constexpr (sizeof(T) == sizeof(double)) {
    auto ld = [&](int i, Row& w) {
      const int K = pair_of(i);
      const T* R = Pl + (size_t)K * STRIDE + off + rho * 12;
#pragma unroll
      for (int c = 0; c < 12; ++c) w.g[c] = (double)R[c];
      w.b = rhs[K * 12 + el];
    };
    Row wa, wb;
    ld(0, wa);
    int i = 0;
    for (; i + 2 <= S; i += 2) {
      ld(i + 1, wb);
      __builtin_amdgcn_sched_barrier(0);
      s = bsub_pair(wa.b, s, wa.g);
      out[oidx(i)] = s;
      __builtin_amdgcn_sched_barrier(0);
      ld(i + 2 < S ? i + 2 : i + 1, wa);
      __builtin_amdgcn_sched_barrier(0);
      s = bsub_pair(wb.b, s, wb.g);
      out[oidx(i + 1)] = s;
      __builtin_amdgcn_sched_barrier(0);
    }
    if (i < S) {
      s = bsub_pair(wa.b, s, wa.g);
      out[oidx(i)] = s;
    }
  } else {
    // narrow rows: loaded two steps ahead, widened one step ahead (as in pair_recurrence)
    struct Raw { T g[12]; double b; };
    auto issue = [&](int i, Raw& w) {
      const int K = pair_of(i < S ? i : S - 1);
      const T* R = Pl + (size_t)K * STRIDE + off + rho * 12;
#pragma unroll
      for (int c = 0; c < 12; ++c) w.g[c] = R[c];
      w.b = rhs[K * 12 + el];
    };
    auto widen = [&](const Raw& r, Row& w) {
#pragma unroll
      for (int c = 0; c < 12; ++c) w.g[c] = (double)r.g[c];
      w.b = r.b;
    };
    Raw ra, rb;
    Row ca, cb;
    issue(0, ra);
    widen(ra, ca);
    issue(1, rb);
    int i = 0;
    for (; i + 2 <= S; i += 2) {
      __builtin_amdgcn_sched_barrier(0);
      s = bsub_pair(ca.b, s, ca.g);
      out[oidx(i)] = s;
      __builtin_amdgcn_sched_barrier(0);
      widen(rb, cb);
      issue(i + 2, ra);
      __builtin_amdgcn_sched_barrier(0);
      s = bsub_pair(cb.b, s, cb.g);
      out[oidx(i + 1)] = s;
      __builtin_amdgcn_sched_barrier(0);
      widen(ra, ca);
      issue(i + 3, rb);
    }
    if (i < S) {
      s = bsub_pair(ca.b, s, ca.g);
      out[oidx(i)] = s;
    }
  }
  return s;
}

// The twist pair, right after the forward recurrences, by the same lanes (zs holds record m):
// top rows y = b_m - Sigma_m z_{m-1}, bottom rows -~Sigma ~z_{m'-1} exchanged through zs in
// pair m's cell order, then x_m = Gamma^-1 (y + exchanged) into xs by the top rows.  Every lane
// stores its y (the top rows after the exchange slots) and every lane runs the second product:
// lanes that exchange through LDS must not sit in branches of their own, whose order the
// compiler may choose.
__device__ inline void twist_pair(int m, double* xs, double* zs) {
  const int lane = threadIdx.x;
  const bool bot = lane >= 16;
  const int rho = lane & 15;
  wave_lds_sync();
  const double v = bot ? xs[(m + 1) * 12 + tw_perm(rho)] : xs[(m - 1) * 12 + rho];
  const double b = bot ? 0.0 : xs[m * 12 + rho];
  double g[12];
#pragma unroll
  for (int c = 0; c < 12; ++c) g[c] = zs[(bot ? TW_SIGR : TW_SIG) + rho * 12 + c];
  const double y = bsub_pair(b, v, g);
  zs[TW_XCH + (bot ? tw_perm(rho) : 12 + rho)] = y;
  wave_lds_sync();
  __builtin_amdgcn_wave_barrier();
  const double r = y + zs[TW_XCH + rho];
#pragma unroll
  for (int c = 0; c < 12; ++c) g[c] = zs[TW_GINV + rho * 12 + c];
  const double x = bsub_pair(0.0, -r, g);
  if (!bot) xs[m * 12 + rho] = x;
}

// FP32 recurrence matrices widened to FP64 in LDS, so that the pair recurrences run on FP64
// operands (no conversions on the dependent chain) while HBM moves FP32.  A matrix set (G or H,
// 144 values per pair) is copied as FP32 straight into the top of its FP64 buffer (no registers
// held during the copy), then widened in place: every thread reads its float4 chunks, a barrier,
// then every thread writes them as doubles; packed chunk c (pair c / 36) lands at buf + 4 c.
// One pass of WCH chunks per thread, so lines of at most WCH * GSW / 36 pairs.
constexpr int WCH = 4;
template <typename T>
__host__ __device__ constexpr bool widen_fits(int Nz) { return npairs(Nz) * 36 <= WCH * GSW; }
__host__ __device__ constexpr size_t narrow_off(int Nz) {
  return pair_g_buf_bytes<double>(Nz) - pair_g_buf_bytes<float>(Nz);
}
// the gather's partial sums, 4 x 6 x Nz doubles: below the FP32 copy of G when they fit there,
// else after the FP64 buffer
__host__ __device__ constexpr size_t widen_part_off(int Nz) {
  return (size_t)192 * Nz <= narrow_off(Nz) ? 0 : pair_g_buf_bytes<double>(Nz);
}
__host__ __device__ constexpr size_t gs_w_lds_bytes(int Nz) {
  return sizeof(double) * 2 * pair_vec_doubles(Nz) + pair_g_buf_bytes<double>(Nz) +
         ((size_t)192 * Nz <= narrow_off(Nz) ? 0 : (size_t)192 * Nz);
}
template <int OFF>
__device__ inline void stage_narrow(const float* __restrict__ P, int L, int NP, int Nz, double* buf) {
  stage_pairs_part<float, OFF, 144>(P, L, NP, reinterpret_cast<float*>(reinterpret_cast<char*>(buf) + narrow_off(Nz)));
}
// after the copies completed (stage_wait); the caller's barrier orders the writes before use
__device__ inline void widen_in_place(int NP, int Nz, double* buf) {
  const float4* src = reinterpret_cast<const float4*>(reinterpret_cast<const char*>(buf) + narrow_off(Nz));
  const int total = NP * 36;
  float4 v[WCH];
#pragma unroll
  for (int u = 0; u < WCH; ++u) {
    const int c = u * GSW + (int)threadIdx.x;
    if (c < total) v[u] = src[c];
  }
  lds_barrier();
#pragma unroll
  for (int u = 0; u < WCH; ++u) {
    const int c = u * GSW + (int)threadIdx.x;
    if (c < total) {
      double2* d = reinterpret_cast<double2*>(buf + 4 * (size_t)c);
      d[0] = make_double2((double)v[u].x, (double)v[u].y);
      d[1] = make_double2((double)v[u].z, (double)v[u].w);
    }
  }
}

// ------------------------------------------------------------------ z-line / block-Jacobi apply
// One z-column per 64-lane block; wall rows (point Jacobi) in blocks of their own.
template <typename T>
__global__ void __launch_bounds__(PW) k_prec_apply(int n, int Nz, int ncol, int nw, int kind,
                                                   const T* __restrict__ F, const double* __restrict__ Fw,
                                                   const double* __restrict__ s0, ApplyArgs a,
                                                   double* __restrict__ partial, unsigned* counter) {
  extern __shared__ __align__(16) double lds[];    // xs[Nz][6] | 8 spare | two staging buffers
  __shared__ double sh[1];
  __shared__ bool slast;
  if (a.respect_done && a.gs->done) return;
  const double sc = a.in_scale ? *a.in_scale : 1.0;
  double local = 0.0;
  const int lane = threadIdx.x;
  if ((int)blockIdx.x < ncol) {
    const int X0 = blockIdx.x * Nz;
    double* xs = lds;
    T* fb0 = reinterpret_cast<T*>(lds + (size_t)Nz * 6 + 8);
    T* fb1 = reinterpret_cast<T*>(reinterpret_cast<char*>(fb0) + zline_buf_bytes<T>());
    const int cnt = 6 * Nz;
    const bool zl = kind == VTSA_PREC_ZLINE;
    if (zl) zline_stage_first<T>(F, X0, Nz, fb0);
    // 1. b = (pre * scale * in) / s0.  store_in may alias in, so every load of a chunk of AC
    //    elements is issued before the chunk's stores.
    for (int e0 = lane; e0 < cnt; e0 += PW * AC) {
      double xin[AC], xpre[AC], xs0[AC];
#pragma unroll
      for (int j = 0; j < AC; ++j) {
        const int e = e0 + j * PW < cnt ? e0 + j * PW : 0;
        const int r = e / Nz, k = e - r * Nz;
        const int idx = r * n + X0 + k;
        xin[j] = a.in[idx];
        xpre[j] = a.pre ? a.pre[idx] : 1.0;
        xs0[j] = s0[idx];
      }
#pragma unroll
      for (int j = 0; j < AC; ++j) {
        const int e = e0 + j * PW;
        if (e < cnt) {
          const int r = e / Nz, k = e - r * Nz;
          double x = xin[j] * sc;
          if (a.store_in) a.store_in[r * n + X0 + k] = x;
          if (a.pre) x = x * xpre[j];
          xs[k * 6 + r] = x / xs0[j];
        }
      }
    }
    if (zl) {
      zline_sweeps<T>(F, X0, Nz, xs, fb0, fb1);
    } else {
      // block Jacobi: one lane per cell
      lds_barrier();
      for (int k = lane; k < Nz; k += PW) {
        const T* Di = F + (size_t)(X0 + k) * 180 + 72;
        double bv[6], xv[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) bv[c] = xs[k * 6 + c];
#pragma unroll
        for (int rr = 0; rr < 6; ++rr) {
          double acc = 0.0;
#pragma unroll
          for (int c = 0; c < 6; ++c) acc += (double)Di[rr * 6 + c] * bv[c];
          xv[rr] = acc;
        }
#pragma unroll
        for (int rr = 0; rr < 6; ++rr) xs[k * 6 + rr] = xv[rr];
      }
      lds_barrier();
    }
    // 2. out = x s0 / post and the norm of out / red_s (loads of a chunk before its stores)
    for (int e0 = lane; e0 < cnt; e0 += PW * AC) {
      double xs0[AC], xpost[AC], xred[AC];
#pragma unroll
      for (int j = 0; j < AC; ++j) {
        const int e = e0 + j * PW < cnt ? e0 + j * PW : 0;
        const int r = e / Nz, k = e - r * Nz;
        const int idx = r * n + X0 + k;
        xs0[j] = s0[idx];
        xpost[j] = a.post ? a.post[idx] : 1.0;
        xred[j] = a.red_s ? a.red_s[idx] : 1.0;
      }
#pragma unroll
      for (int j = 0; j < AC; ++j) {
        const int e = e0 + j * PW;
        if (e < cnt) {
          const int r = e / Nz, k = e - r * Nz;
          double o = xs[k * 6 + r] * xs0[j];
          if (a.out_raw) a.out_raw[r * n + X0 + k] = o;
          if (a.post) o = o / xpost[j];
          a.out[r * n + X0 + k] = o;
          double q = a.red_s ? o / xred[j] : o;
          local += q * q;
        }
      }
    }
  } else {
    // wall rows: point Jacobi
    const int wb = blockIdx.x - ncol;
    const int nwb = gridDim.x - ncol;
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
          if (a.out_raw) a.out_raw[idx] = o;
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
      a.gs->sigma = rms == 0.0 ? 0.0 : (1.4901161193847656e-08 * fmax(1.0, *a.zs_rms)) / rms;
    }
  }
}

// ------------------------------------------------------------------ line Gauss-Seidel apply
// Red/black (in general: greedy-coloured) Gauss-Seidel over lines: the z-lines of the cells
// (solved exactly by the z-line factor) and the wall lines (tridiagonal along z, solved by the
// Thomas algorithm).  Lines of one colour do not couple, so each colour is one launch.  The
// sequence of colours for S sweeps is 0, 1, .., C-1, C-2, .., 0, then 1, .., 0 again per extra
// sweep (a colour repeated back to back would reproduce its own result).  The iterate is kept
// unscaled, U = S0 x~, so the couplings to other lines need only J and U:
//     (M~ x~)_R = x~_R - (dh / s0_R) sum_C J_RC U_C      (R, C in different lines)
struct GsArgs {
  const void* Jo;   // off-line blocks in the factor type of the half-sweep (double or float)
  const double* Jcw; const double* Jw; const double* Jwc; const double* s0;
  double* bs; double* U; const int* order; double dh; int first; const double* Tw;
};

// Inclusive scan of the first-order linear recurrence y_l = a_l + b_l y_{l-1} over the 64 lanes of
// a wavefront (Hillis-Steele: log2(64) steps of (a, b) <- (a + b a', b b') with the pair `off`
// lanes below).  On return a holds y_l for y_{-1} = 0.  Called by all 64 lanes.
__device__ inline void scan_linear(double& a, double& b, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double ap = __shfl_up(a, off, 64), bp = __shfl_up(b, off, 64);
    if (lane >= off) {
      a = fma(b, ap, a);
      b = b * bp;
    }
  }
}

// b~ = (pre * scale * in) / s0 and U = 0 (store_in as in the z-line apply; may alias in).
__global__ void k_gs_gather(int N, ApplyArgs a, const double* __restrict__ s0, double* __restrict__ bs,
                            double* __restrict__ U) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (a.respect_done && a.gs->done) return;
  if (i >= N) return;
  const double sc = a.in_scale ? *a.in_scale : 1.0;
  double x = a.in[i] * sc;
  if (a.store_in) a.store_in[i] = x;
  if (a.pre) x = x * a.pre[i];
  bs[i] = x / s0[i];
  U[i] = 0.0;
}

// The off-line blocks of every cell (slots r-1, r+1, theta-1, theta+1: 144 values, contiguous in
// Jc) transposed to Jo[(q * 36 + r * 6 + c) * n + X], so that a half-sweep's right-hand side reads
// them with consecutive cells in consecutive lanes.  OT cells per block through LDS; once per
// preconditioner set-up.
constexpr int OT = 32;
template <typename TO>
__global__ void __launch_bounds__(256) k_gs_offline(int n, const double* __restrict__ Jc, TO* __restrict__ Jo) {
  __shared__ double t[144][OT + 1];
  const int X0 = blockIdx.x * OT;
  const int nc = min(OT, n - X0);
  for (int e = threadIdx.x; e < nc * 144; e += blockDim.x) {
    const int x = e / 144, v = e - x * 144;
    t[v][x] = Jc[((size_t)(X0 + x) * NSLOT + S_RM1) * 36 + v];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 144 * OT; e += blockDim.x) {
    const int v = e / OT, x = e - v * OT;
    if (x < nc) Jo[(size_t)v * n + X0 + x] = (TO)t[v][x];
  }
}

// MODE 0: one wavefront, Fac the cell factors (F); they are staged in segments while the sweeps
//         run cell by cell (zline_sweeps).
// MODE 1: GSW threads, Fac the pair factors (P); the line's pairs are staged into LDS at kernel
//         start (stage_pairs), spread over eight wavefronts so that no wavefront stalls on its
//         count of outstanding copies, the right-hand side is gathered by all of them, and the
//         first wavefront runs the pair recurrences between data-parallel passes.
// MODE 2: as MODE 1 with only G and H staged (2/3 of the LDS); the data-parallel passes read
//         L1, Dinv and ^U1 from global memory.  For lines whose whole factor does not fit, or
//         when staging it would leave fewer blocks resident per CU than a launch has lines.
// MODE 3: as MODE 2 with G staged at the start and H copied over it after the forward
//         recurrence (1/3 of the LDS), the copies in flight during the two passes between the
//         recurrences.
// MODE 4: as MODE 3 with FP32 matrices widened to FP64 in LDS.
// All solve the same line systems; the roundings differ.  MODE 1 .. 4 with tw_m >= 0: the twisted
// pair form (twist pair tw_m), both halves of the line at once.
// The apply's outputs, on a line's last visit: out = U / post for the line's unknowns, and the
// rms of out / red_s over all unknowns (the sum of squares per line into partial[L]; the last
// line to finish sums the lines' partials in line order).  Returns this thread's sum of squares.
// The scaled right-hand side b~ = (pre * scale * in) / s0 of unknown idx: computed on the line's
// first visit (which also stores in * scale to store_in, and keeps b~ for the later visits), read
// back afterwards.  store_in may alias in: each unknown is read and written by one thread, once.
__device__ inline double gs_b(const ApplyArgs& a, const double* __restrict__ s0, double* __restrict__ bs, int idx,
                              int first_visit) {
  if (!first_visit) return bs[idx];
  const double sc = a.in_scale ? *a.in_scale : 1.0;
  double x = a.in[idx] * sc;
  if (a.store_in) a.store_in[idx] = x;
  if (a.pre) x = x * a.pre[idx];
  const double b = x / s0[idx];
  bs[idx] = b;
  return b;
}

__device__ inline double gs_out(const ApplyArgs& a, int idx, double u) {
  const double o = a.post ? u / a.post[idx] : u;
  a.out[idx] = o;
  if (a.out_raw) a.out_raw[idx] = u;
  const double q = a.red_s ? o / a.red_s[idx] : o;
  return q * q;
}

__device__ inline void gs_out_norm(const ApplyArgs& a, double local, int L, int nlines, int N,
                                   double* __restrict__ partial, unsigned* counter, double* sh, bool* sflag) {
  const double bs = wave_block_sum(local, sh);
  if (threadIdx.x == 0) store_partial(partial + L, bs);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned tk = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *sflag = (tk == (unsigned)nlines - 1);
  }
  __syncthreads();
  if (!*sflag || threadIdx.x >= 64) return;
  const double tot = wave_sum_strided(partial, nlines);
  if (threadIdx.x != 0) return;
  ticket_reset(counter);
  const double rms = sqrt(tot / (double)N);
  if (a.red_out) *a.red_out = rms;
  if (a.gs) {
    a.gs->unorm = rms;
    a.gs->zero_u = (rms == 0.0) ? 1 : 0;
    a.gs->sigma = rms == 0.0 ? 0.0 : (1.4901161193847656e-08 * fmax(1.0, *a.zs_rms)) / rms;
  }
}

template <typename T, int MODE>
__global__ void __launch_bounds__(MODE == 0 ? PW : GSW)
    k_gs_half(int n, int Nr, int Nth, int Nz, int ncol, int nw, const T* __restrict__ Fac, GsArgs g, int cbeg,
              const GmresScalars* __restrict__ gs, int respect_done, ApplyArgs a, int first_visit, int final_visit,
              int nlines, int N, double* __restrict__ partial, unsigned* counter, int tw_m) {
  extern __shared__ __align__(16) double lds[];
  __shared__ double sh_norm[(MODE == 0 ? PW : GSW) / 64];
  __shared__ bool s_last;
  double local = 0.0;   // sum of squares of this thread's outputs (last visit only)
  const int L = g.order[cbeg + blockIdx.x];   // loaded together with the flag below
  if (respect_done && gs->done) return;
  constexpr int NT = MODE == 0 ? PW : GSW;
  const int lane = threadIdx.x;
  double* xs = lds;
  const double dh = g.dh;
  if (L < ncol) {
    // cell z-line (i, j)
    const int i = L / Nth, j = L - (L / Nth) * Nth;
    const int X0 = L * Nz;
    const int NP = npairs(Nz);
    double* zs = lds + pair_vec_doubles(Nz);                                  // MODE 1 .. 4
    T* pb = reinterpret_cast<T*>(lds + 2 * pair_vec_doubles(Nz));             // MODE 1 .. 3
    double* pbw = lds + 2 * pair_vec_doubles(Nz);                              // MODE 4
    T* fb0 = reinterpret_cast<T*>(lds + (size_t)Nz * 6 + 8);                  // MODE 0
    T* fb1 = reinterpret_cast<T*>(reinterpret_cast<char*>(fb0) + zline_buf_bytes<T>());
    if (MODE == 1) stage_pairs<T>(Fac, L, NP, pb);
    else if (MODE == 2) stage_pairs_part<T, PO_G, PGH>(Fac, L, NP, pb);
    else if (MODE == 3) stage_pairs_part<T, PO_G, 144>(Fac, L, NP, pb);
    else if (MODE == 0) zline_stage_first<T>(Fac, X0, Nz, fb0);
    if constexpr (MODE == 4) stage_narrow<PO_G>(Fac, L, NP, Nz, pbw);
    // twisted: the twist record, one value per thread, loaded now and put into zs after the gather
    const bool twisted = MODE >= 1 && tw_m >= 0;
    T twv = T(0);
    if (twisted && lane < TW_XCH) twv = Fac[((size_t)L * NP + tw_m) * PP + lane];
    const int jm = j == 0 ? Nth - 1 : j - 1, jp = j + 1 == Nth ? 0 : j + 1;
    // neighbour z-lines' first cells for the off-line slots r-1, r+1, theta-1, theta+1
    const int nb[4] = {i > 0 ? ((i - 1) * Nth + j) * Nz : -1, i + 1 < Nr ? ((i + 1) * Nth + j) * Nz : -1,
                       Nth > 1 ? (i * Nth + jm) * Nz : -1, Nth > 1 ? (i * Nth + jp) * Nz : -1};
    const bool wall = nw > 0 && i == Nr - 1;
    // right-hand side of element e = (row r, cell k): b~ + (dh / s0) (off-line J) U.  Branch-free,
    // so that all of an element's loads are in flight together: a missing neighbour reads its own
    // line's cell and contributes an exact zero.  Accumulation order q = 0..3, c = 0..5, wall.
    auto rhs = [&](int e) -> double {
      const int r = e / Nz, k = e - r * Nz;
      const int X = X0 + k;
      double acc = 0.0;
      if (!g.first) {
        double jv[4][6], uv[4][6];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const T* Jr = reinterpret_cast<const T*>(g.Jo) + (size_t)(q * 36 + r * 6) * n + X;   // coalesced over k
          const int Y = (nb[q] >= 0 ? nb[q] : X0) + k;
#pragma unroll
          for (int c = 0; c < 6; ++c) { jv[q][c] = (double)Jr[(size_t)c * n]; uv[q][c] = g.U[c * n + Y]; }
        }
        double jw = 0.0, uw = 0.0;
        if (nw > 0) {
          const int W = j * Nz + k;
          jw = g.Jcw[W * 4 + (r <= 2 ? r : 3)];
          uw = g.U[6 * n + W];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const bool on = nb[q] >= 0;
#pragma unroll
          for (int c = 0; c < 6; ++c) acc += (on ? jv[q][c] : 0.0) * uv[q][c];
        }
        acc += (wall && (r <= 2 || r == 5) ? jw : 0.0) * uw;
      }
      const int idx = r * n + X;
      return gs_b(a, g.s0, g.bs, idx, first_visit) + (dh / g.s0[idx]) * acc;
    };
    if constexpr (MODE == 4) {
      // the off-line products by (slot q, cell k), each neighbour's 6 unknowns loaded once for
      // its 6 rows: 4 Nz tasks of 42 loads instead of 6 Nz of 48, one round of loads per thread.
      // The partial sums, [4][6][Nz], sit where the widened G goes later; an element adds its
      // four in slot order.
      double* part = reinterpret_cast<double*>(reinterpret_cast<char*>(pbw) + widen_part_off(Nz));
      // this thread's right-hand-side elements e = lane + u NT (6 Nz <= 2 NT whenever the widened
      // matrices fit, Nz <= 112): their global inputs are loaded before the off-line products,
      // so that the gather waits for one round of loads; the arithmetic is gs_b's
      constexpr int EPT = 2;
      double pin[EPT], ppre[EPT], ps0[EPT], pjw[EPT], puw[EPT];
      const double sc = first_visit && a.in_scale ? *a.in_scale : 1.0;
#pragma unroll
      for (int u = 0; u < EPT; ++u) {
        const int e = lane + u * NT;
        const int ec = e < 6 * Nz ? e : 0;
        const int r = ec / Nz, k = ec - r * Nz;
        const int idx = r * n + X0 + k;
        ps0[u] = g.s0[idx];
        pin[u] = first_visit ? a.in[idx] : g.bs[idx];
        ppre[u] = first_visit && a.pre ? a.pre[idx] : 1.0;
        pjw[u] = 0.0;
        puw[u] = 0.0;
        if (!g.first && nw > 0) {
          const int W = j * Nz + k;
          pjw[u] = g.Jcw[W * 4 + (r <= 2 ? r : 3)];
          puw[u] = g.U[6 * n + W];
        }
      }
      if (!g.first) {
        for (int t = lane; t < 4 * Nz; t += NT) {
          const int q = t / Nz, k = t - q * Nz;
          const int X = X0 + k;
          const bool on = nb[q] >= 0;
          const int Y = (on ? nb[q] : X0) + k;
          const T* Jq = reinterpret_cast<const T*>(g.Jo) + (size_t)(q * 36) * n + X;   // coalesced over k
          double uv[6], jv[36];
#pragma unroll
          for (int c = 0; c < 6; ++c) uv[c] = g.U[c * n + Y];
#pragma unroll
          for (int rc = 0; rc < 36; ++rc) jv[rc] = (double)Jq[(size_t)rc * n];
#pragma unroll
          for (int r = 0; r < 6; ++r) {
            double p = 0.0;
#pragma unroll
            for (int c = 0; c < 6; ++c) p += (on ? jv[r * 6 + c] : 0.0) * uv[c];
            part[(q * 6 + r) * Nz + k] = p;
          }
        }
        lds_barrier();
      }
#pragma unroll
      for (int u = 0; u < EPT; ++u) {
        const int e = lane + u * NT;
        if (e >= 6 * Nz) continue;
        const int r = e / Nz, k = e - r * Nz;
        double acc = 0.0;
        if (!g.first) {
          acc = ((part[r * Nz + k] + part[(6 + r) * Nz + k]) + part[(12 + r) * Nz + k]) + part[(18 + r) * Nz + k];
          acc += (wall && (r <= 2 || r == 5) ? pjw[u] : 0.0) * puw[u];
        }
        const int idx = r * n + X0 + k;
        double b = pin[u];
        if (first_visit) {
          double x = pin[u] * sc;
          if (a.store_in) a.store_in[idx] = x;
          if (a.pre) x = x * ppre[u];
          b = x / ps0[u];
          g.bs[idx] = b;
        }
        xs[k * 6 + r] = b + (dh / ps0[u]) * acc;
      }
    } else {
      for (int e = lane; e < 6 * Nz; e += NT) xs[(e % Nz) * 6 + e / Nz] = rhs(e);
    }
    if (MODE >= 1) {
      if (lane < 6 * (2 * NP - Nz)) xs[Nz * 6 + lane] = 0.0;   // phantom cell of an odd line
      // the data-parallel passes read L1, Dinv and ^U1 from the staged line (MODE 1) or from
      // global memory (MODE 2: G and H staged; MODE 3: G, then H in its place; MODE 4: the same
      // widened to FP64); the recurrences read G and H from LDS
      const T* pd = MODE == 1 ? pb : Fac + (size_t)L * NP * PP;
      stage_wait();
      if (twisted && lane < TW_XCH) zs[lane] = (double)twv;
      if constexpr (MODE == 4) widen_in_place(NP, Nz, pbw);
      // b'_{2K+1} = b_{2K+1} - L1_{2K+1} b_{2K} (twisted, below the twist pair: the cells in
      // reverse order; none for the twist pair)
      for (int e = lane; e < NP * 6; e += NT) {
        const int K = e / 6, r = e - K * 6;
        if (twisted && K == tw_m) continue;
        const bool rv = twisted && K > tw_m;
        const int e0 = K * 12 + (rv ? 6 : 0), e1 = K * 12 + (rv ? 0 : 6);
        const T* C = pd + (size_t)K * PP + PO_C + r * 6;
        double v = xs[e1 + r];
#pragma unroll
        for (int c = 0; c < 6; ++c) v = fma(-(double)C[c], xs[e0 + c], v);
        xs[e1 + r] = v;
      }
      lds_barrier();
      constexpr int SG = MODE == 1 ? PP : MODE == 2 ? PGH : 144;   // LDS stride of a pair's matrices
      if (twisted) {
        if (lane < 32 && (lane & 15) < 12) {
          if constexpr (MODE == 4) pair_recurrence_tw<double, true, 144>(pbw, NP, tw_m, 0, xs, xs, 0.0);
          else pair_recurrence_tw<T, true, SG>(pb, NP, tw_m, PO_G, xs, xs, 0.0);
          twist_pair(tw_m, xs, zs);
        }
      } else if (lane < 12) {
        if constexpr (MODE == 4) pair_recurrence<double, true, 144>(pbw, NP, 0, xs, xs);
        else pair_recurrence<T, true, SG>(pb, NP, PO_G, xs, xs);
      }
      lds_barrier();
      // MODE 3, 4: H replaces G in the buffer; the copies run under the two passes below
      if (MODE == 3) stage_pairs_part<T, PO_H, 144>(Fac, L, NP, pb);
      if constexpr (MODE == 4) stage_narrow<PO_H>(Fac, L, NP, Nz, pbw);
      // h_k = Dinv_k z_k
      for (int e = lane; e < NP * 12; e += NT) {
        const int K = e / 12, rr = e - K * 12, hh = rr / 6;
        if (twisted && K == tw_m) continue;
        const int cell = twisted && K > tw_m ? 1 - hh : hh;
        const T* D = pd + (size_t)K * PP + PO_D + rr * 6;
        double v = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) v = fma((double)D[c], xs[K * 12 + cell * 6 + c], v);
        zs[K * 12 + cell * 6 + (rr - hh * 6)] = v;
      }
      lds_barrier();
      // z'_{2K} = h_{2K} - ^U1_{2K} h_{2K+1}
      for (int e = lane; e < NP * 6; e += NT) {
        const int K = e / 6, r = e - K * 6;
        if (twisted && K == tw_m) continue;
        const bool rv = twisted && K > tw_m;
        const int e0 = K * 12 + (rv ? 6 : 0), e1 = K * 12 + (rv ? 0 : 6);
        const T* U = pd + (size_t)K * PP + PO_U + r * 6;
        double v = zs[e0 + r];
#pragma unroll
        for (int c = 0; c < 6; ++c) v = fma(-(double)U[c], zs[e1 + c], v);
        zs[e0 + r] = v;
      }
      if (MODE == 3 || MODE == 4) stage_wait();
      if constexpr (MODE == 4) widen_in_place(NP, Nz, pbw);
      lds_barrier();
      if (twisted) {
        if (lane < 32 && (lane & 15) < 12) {
          const double xm = xs[tw_m * 12 + (lane >= 16 ? tw_perm(lane & 15) : lane)];
          if constexpr (MODE == 4) pair_recurrence_tw<double, false, 144>(pbw, NP, tw_m, 0, zs, xs, xm);
          else pair_recurrence_tw<T, false, SG>(pb, NP, tw_m, MODE == 3 ? 0 : PO_H, zs, xs, xm);
        }
      } else if (lane < 12) {
        if constexpr (MODE == 4) pair_recurrence<double, false, 144>(pbw, NP, 0, zs, xs);
        else pair_recurrence<T, false, SG>(pb, NP, MODE == 3 ? 0 : PO_H, zs, xs);
      }
      lds_barrier();
    } else {
      zline_sweeps<T>(Fac, X0, Nz, xs, fb0, fb1);
    }
    for (int e = lane; e < 6 * Nz; e += NT) {
      const int r = e / Nz, k = e - r * Nz;
      const int idx = r * n + X0 + k;
      const double u = g.s0[idx] * xs[k * 6 + r];
      g.U[idx] = u;
      if (final_visit) local += gs_out(a, idx, u);
    }
  } else {
    // wall line j: tridiagonal along z (axial conduction), couplings to the outer cells and to
    // the neighbouring wall lines moved to the right-hand side.  Factored at set-up
    // (k_gs_wall_factor); the forward and backward substitutions are first-order linear
    // recurrences, solved 64 cells at a time by a scan over the first wavefront.
    const int j = L - ncol;
    const int jm = j == 0 ? Nth - 1 : j - 1, jp = j + 1 == Nth ? 0 : j + 1;
    double* rhs = xs;          // [Nz]
    for (int k = lane; k < Nz; k += NT) {
      const int W = j * Nz + k;
      const int X = ((Nr - 1) * Nth + j) * Nz + k;
      double acc = 0.0;
      if (!g.first) {
        if (Nth > 1) acc += g.Jw[W * 5 + 3] * g.U[6 * n + jm * Nz + k] + g.Jw[W * 5 + 4] * g.U[6 * n + jp * Nz + k];
#pragma unroll
        for (int v = 0; v < 6; ++v) acc += g.Jwc[W * 6 + v] * g.U[v * n + X];
      }
      rhs[k] = gs_b(a, g.s0, g.bs, 6 * n + W, first_visit) + (dh / g.s0[6 * n + W]) * acc;
    }
    lds_barrier();
    if (lane < 64) {
      const double* T3 = g.Tw + (size_t)j * Nz * 3;
      // forward: d_k = (rhs_k - lo_k d_{k-1}) / den_k
      double carry = 0.0;
      for (int k0 = 0; k0 < Nz; k0 += 64) {
        const int k = k0 + lane;
        double a = 0.0, b = 0.0;
        if (k < Nz) {
          const double rd = T3[k * 3 + 1];
          a = rhs[k] * rd;
          b = -T3[k * 3 + 0] * rd;
        }
        if (lane == 0) {
          a = fma(b, carry, a);
          b = 0.0;
        }
        scan_linear(a, b, lane);
        if (k < Nz) rhs[k] = a;
        carry = __shfl(a, min(63, Nz - 1 - k0), 64);
      }
      // backward: x_k = d_k - c_k x_{k+1}, lanes running down from k1
      carry = 0.0;
      for (int k1 = Nz - 1; k1 >= 0; k1 -= 64) {
        const int k = k1 - lane;
        double a = 0.0, b = 0.0;
        if (k >= 0) {
          a = rhs[k];
          b = -T3[k * 3 + 2];
        }
        if (lane == 0) {
          a = fma(b, carry, a);
          b = 0.0;
        }
        scan_linear(a, b, lane);
        if (k >= 0) rhs[k] = a;
        carry = __shfl(a, min(63, k1), 64);
      }
    }
    lds_barrier();
    for (int k = lane; k < Nz; k += NT) {
      const int idx = 6 * n + j * Nz + k;
      const double u = g.s0[idx] * rhs[k];
      g.U[idx] = u;
      if (final_visit) local += gs_out(a, idx, u);
    }
  }
  if (final_visit) gs_out_norm(a, local, L, nlines, N, partial, counter, sh_norm, &s_last);
}

// ------------------------------------------------------------------ host side (preconditioner)
// Greedy colouring of the line graph (cell z-lines (i, j) couple to (i +- 1, j) and (i, j +- 1);
// wall line j couples to cell line (Nr - 1, j) and wall lines j +- 1).  Lines are visited in
// index order, so an even Ntheta gives the red/black pattern (i + j) mod 2.
void gs_colour_lines(int Nr, int Nth, int nwall_lines, std::vector<int>& order, std::vector<int>& cstart) {
  const int ncol = Nr * Nth;
  const int nl = ncol + nwall_lines;
  std::vector<int> colour(nl, -1);
  auto neighbours = [&](int L, std::vector<int>& out) {
    out.clear();
    if (L < ncol) {
      const int i = L / Nth, j = L % Nth;
      if (i > 0) out.push_back(L - Nth);
      if (i + 1 < Nr) out.push_back(L + Nth);
      if (Nth > 1) {
        out.push_back(i * Nth + (j == 0 ? Nth - 1 : j - 1));
        out.push_back(i * Nth + (j + 1 == Nth ? 0 : j + 1));
      }
      if (nwall_lines > 0 && i == Nr - 1) out.push_back(ncol + j);
    } else {
      const int j = L - ncol;
      out.push_back((Nr - 1) * Nth + j);
      if (Nth > 1) {
        out.push_back(ncol + (j == 0 ? Nth - 1 : j - 1));
        out.push_back(ncol + (j + 1 == Nth ? 0 : j + 1));
      }
    }
  };
  std::vector<int> nbr;
  int ncolour = 0;
  for (int L = 0; L < nl; ++L) {
    neighbours(L, nbr);
    for (int c = 0;; ++c) {
      bool used = false;
      for (int M : nbr)
        if (M != L && colour[M] == c) used = true;
      if (!used) { colour[L] = c; if (c + 1 > ncolour) ncolour = c + 1; break; }
    }
  }
  order.clear();
  cstart.assign(ncolour + 1, 0);
  for (int c = 0; c < ncolour; ++c) {
    cstart[c] = (int)order.size();
    for (int L = 0; L < nl; ++L)
      if (colour[L] == c) order.push_back(L);
  }
  cstart[ncolour] = (int)order.size();
}

template <typename T>
static int gs_pick_mode(Handle& h, int o);

void launch_prec_setup(Handle& h, double dh, const double* scale, int kind, int prec, int sweeps, hipStream_t s) {
  (void)hipMemcpyAsync(h.s0, scale, sizeof(double) * h.N, hipMemcpyDeviceToDevice, s);
  const int fkind = kind == VTSA_PREC_BLOCK_JACOBI ? VTSA_PREC_BLOCK_JACOBI : VTSA_PREC_ZLINE;
  const int nb = (h.ncol + PCOLS - 1) / PCOLS;
  // twisted pair form (line Gauss-Seidel, an even number of cells, at least three pairs, a pair
  // form granted both for the staging option and automatically): each line factored from its
  // first cell down and from its last cell up, both only as far as the twist pair, in one launch
  const int Nz = h.desc.Nz, NP = npairs(Nz);
  bool twist = kind == VTSA_PREC_ZLINE_GS && h.gs_twist && Nz % 2 == 0 && NP >= 3;
  if (twist)
    twist = prec == VTSA_FP32 ? gs_pick_mode<float>(h, h.gs_full) >= 1 && gs_pick_mode<float>(h, 1) >= 1
                              : gs_pick_mode<double>(h, h.gs_full) >= 1 && gs_pick_mode<double>(h, 1) >= 1;
  const int tw_m = twist ? (NP - 1) / 2 : -1;
  const int nbr = tw_m >= 0 ? nb : 0;
  const int kend_top = tw_m >= 0 ? 2 * tw_m + 2 : Nz;
  const int kend_rev = tw_m >= 0 ? 2 * (NP - 1 - tw_m) + 2 : 0;
  if (prec == VTSA_FP32)
    hipLaunchKernelGGL(k_prec_factor<float>, dim3(nb + nbr), dim3(PW), 0, s, h.n, Nz, h.ncol, fkind, h.Jc, h.s0, dh,
                       (float*)h.F, (float*)h.Fr, nb, kend_top, kend_rev);
  else
    hipLaunchKernelGGL(k_prec_factor<double>, dim3(nb + nbr), dim3(PW), 0, s, h.n, Nz, h.ncol, fkind, h.Jc, h.s0, dh,
                       (double*)h.F, (double*)h.Fr, nb, kend_top, kend_rev);
  if (h.nw > 0) hipLaunchKernelGGL(k_prec_wall, dim3((h.nw + 255) / 256), dim3(256), 0, s, h.nw, h.Jw, dh, h.Fw);
  if (kind == VTSA_PREC_ZLINE_GS) {
    // the off-line blocks are stored in the precision of the factors (the FP32 form halves the
    // half-sweep's right-hand-side traffic)
    if (prec == VTSA_FP32)
      hipLaunchKernelGGL(k_gs_offline<float>, dim3((h.n + OT - 1) / OT), dim3(256), 0, s, h.n, h.Jc, (float*)h.Jo);
    else
      hipLaunchKernelGGL(k_gs_offline<double>, dim3((h.n + OT - 1) / OT), dim3(256), 0, s, h.n, h.Jc, h.Jo);
    if (h.nw > 0)
      hipLaunchKernelGGL(k_gs_wall_factor, dim3((h.desc.Nth + 63) / 64), dim3(64), 0, s, h.desc.Nth, h.desc.Nz, h.n,
                         h.Jw, h.s0, dh, h.Tw);
    const int nt = h.ncol * NP * 6;
    if (prec == VTSA_FP32)
      hipLaunchKernelGGL(k_gs_pairs<float>, dim3((nt + PPB * 6 - 1) / (PPB * 6)), dim3(PPB * 6), 0, s, Nz, h.ncol,
                         (const float*)h.F, (float*)h.P, (const float*)h.Fr, tw_m);
    else
      hipLaunchKernelGGL(k_gs_pairs<double>, dim3((nt + PPB * 6 - 1) / (PPB * 6)), dim3(PPB * 6), 0, s, Nz, h.ncol,
                         (const double*)h.F, (double*)h.P, (const double*)h.Fr, tw_m);
    if (tw_m >= 0) {
      if (prec == VTSA_FP32)
        hipLaunchKernelGGL(k_gs_twist<float>, dim3(h.ncol), dim3(64), 0, s, h.n, Nz, h.Jc, h.s0, dh,
                           (const float*)h.F, (const float*)h.Fr, tw_m, (float*)h.P);
      else
        hipLaunchKernelGGL(k_gs_twist<double>, dim3(h.ncol), dim3(64), 0, s, h.n, Nz, h.Jc, h.s0, dh,
                           (const double*)h.F, (const double*)h.Fr, tw_m, (double*)h.P);
    }
  }
  h.gs_tw_m = kind == VTSA_PREC_ZLINE_GS ? tw_m : -1;
  h.prec_kind = kind;
  h.prec_prec = prec;
  h.prec_dh = dh;
  h.gs_sweeps = sweeps;
  h.prec_ready = 1;
}

// The apply's LDS (FP64 factors, the larger case) must fit the 64 KiB a block may use.
bool zline_fits(int Nz) { return zline_lds_bytes<double>(Nz) <= 65536; }

size_t partial_size(int ncol, int nw) {
  size_t a = (size_t)RMAXB * (MAX_RESTART + 2 + DEFL_MAX);   // GMRES: K dots + |w|^2 + correction dots
  size_t pg = (size_t)ncol + nw + 1;   // apply grid: a block per column; line Gauss-Seidel: a partial per line
  if (pg < (size_t)RMAXB) pg = RMAXB;                            // line Gauss-Seidel final reduction
  size_t b = (size_t)RMAXB * 8 + pg;
  return a > b ? a : b;
}

// Algorithmic HBM bytes of one half-sweep over colour c, each datum counted once: per cell line
// its factor (Nz x 180 values cell by cell, npairs x PP in pair form), per cell b~, s0 and the U
// written (6 each) and, unless it is the first
// half-sweep (U = 0), the four off-line 6x6 J blocks and the neighbours' U; per outer-ring cell
// Jcw (4) and the wall's U; per wall cell Jw (self, z +- 1), b~, s0, U written and, unless first,
// Jw (theta +- 1), Jwc (6), the theta-neighbours' U (2) and the outer cell's U (6).  On a line's
// first visit b~ is formed instead of read (in, pre, and store_in and b~ written); on its last
// visit out (and post, red_s) per unknown as well.
static double gs_bytes(const Handle& h, int c, bool first, size_t fsz, int mode, bool fv, bool fin,
                       const ApplyArgs& a) {
  const double Nz = h.desc.Nz;
  const double cells = h.gs_ncell[c] * Nz, outer = h.gs_nouter[c] * Nz, walls = h.gs_nwall[c] * Nz;
  const double per_line = (mode >= 1 ? (double)npairs(h.desc.Nz) * PP : Nz * 180.0) * fsz;
  const double per_unknown = (fv ? 8.0 * (1 + (a.pre ? 1 : 0) + (a.store_in ? 1 : 0)) : 0.0) +
                             (fin ? 8.0 * (1 + (a.post ? 1 : 0) + (a.red_s ? 1 : 0)) : 0.0);
  const double per_cell = 8.0 * 18 + (first ? 0.0 : 144.0 * fsz + 8.0 * 24) + 6.0 * per_unknown;
  const double per_outer = first ? 0.0 : 8.0 * 5;
  const double per_wall = 8.0 * 6 + (first ? 0.0 : 8.0 * 16) + per_unknown;
  return h.gs_ncell[c] * per_line + cells * per_cell + outer * per_outer + walls * per_wall;
}

static hipEvent_t prof_event(Handle& h) {
  if (!h.prof_pool.empty()) {
    hipEvent_t e = h.prof_pool.back();
    h.prof_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  // no system-scope fence: the bracket should cost as little as possible around the kernel
  if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
  return e;
}

// Read the pending (start, stop) pairs into the totals (waits for them) and recycle the events.
int prof_resolve(Handle& h) {
  int rc = 0;
  for (size_t q = 0; q + 1 < h.prof_pending.size(); q += 2) {
    hipEvent_t e0 = h.prof_pending[q], e1 = h.prof_pending[q + 1];
    float ms = 0.f;
    if (!rc) rc = check(hipEventSynchronize(e1), "hipEventSynchronize");
    if (!rc) rc = check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    h.prof_ms += ms;
    h.prof_bytes += h.prof_pending_bytes[q / 2];
    h.prof_launches += 1;
    h.prof_pool.push_back(e0);
    h.prof_pool.push_back(e1);
  }
  h.prof_pending.clear();
  h.prof_pending_bytes.clear();
  return rc;
}

template <typename T, int MODE>
static size_t gs_lds_bytes(int Nz) {
  return MODE == 1 ? gs_full_lds_bytes<T>(Nz) : MODE == 2 ? gs_gh_lds_bytes<T>(Nz)
                   : MODE == 3 ? gs_g_lds_bytes<T>(Nz) : MODE == 4 ? gs_w_lds_bytes(Nz)
                   : zline_lds_bytes<T>(Nz);
}

// Whether half-sweep variant MODE can run: its LDS fits the device's limit for one block and the
// kernel accepts that much dynamic LDS (asked for once per variant and size).
template <typename T, int MODE>
static bool gs_mode_ok(const Handle& h) {
  if (MODE == 4 && !widen_fits<T>(h.desc.Nz)) return false;
  static std::mutex mu;
  static int granted = -1;   // -1 not asked, 0 refused, 1 granted (for the largest size asked)
  static size_t granted_bytes = 0;
  const size_t need = gs_lds_bytes<T, MODE>(h.desc.Nz);
  if (need > (size_t)h.lds_max) return false;
  std::lock_guard<std::mutex> lock(mu);
  if (granted == 1 && need <= granted_bytes) return true;
  if (granted == 0 && need >= granted_bytes) return false;
  const bool ok = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_gs_half<T, MODE>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)need) == hipSuccess;
  (void)hipGetLastError();
  granted = ok ? 1 : 0;
  granted_bytes = need;
  return ok;
}

// Half-sweep variant for staging option o: whole-line staging when the option allows it and the
// device grants the LDS: the whole pair factor (MODE 1) while the blocks a CU can hold at that LDS
// size cover the widest colour in one round, else G and H only (MODE 2), else the segmented
// cell-by-cell form (MODE 0).
template <typename T>
static int gs_pick_mode(Handle& h, int o) {
  const int Nz = h.desc.Nz;
  const size_t lds1 = gs_lds_bytes<T, 1>(Nz), lds2 = gs_lds_bytes<T, 2>(Nz), lds3 = gs_lds_bytes<T, 3>(Nz);
  const size_t lds4 = gs_lds_bytes<T, 4>(Nz);
  constexpr bool narrow = sizeof(T) == sizeof(float);
  int widest = 0;
  for (size_t c = 0; c + 1 < h.gs_cstart.size(); ++c) widest = std::max(widest, h.gs_cstart[c + 1] - h.gs_cstart[c]);
  auto per_cu = [&](size_t lds) { return std::max<long long>(1, (long long)h.lds_per_cu / (long long)lds); };
  // option VTSA_OPT_GS_WHOLE_LINES: 0 segments, 1 automatic, 2 G and H, 3 the whole factor,
  // 4 G then H, 5 G then H widened to FP64 (FP32 factors; FP64 ones: as 4); automatic takes the
  // first of G then H widened (FP32 factors: FP64 operands on the recurrences), the whole
  // factor, G and H, G then H whose blocks all fit on the CUs at once, else the first that fits
  bool ok4 = false;
  if constexpr (narrow) ok4 = (o == 1 || o == 5) && gs_mode_ok<T, 4>(h);
  const bool ok1 = (o == 1 || o == 3) && gs_mode_ok<T, 1>(h);
  const bool ok2 = (o == 1 || o == 2) && gs_mode_ok<T, 2>(h);
  const bool ok3 = (o == 1 || o == 4 || (o == 5 && !narrow)) && gs_mode_ok<T, 3>(h);
  auto one_round = [&](size_t lds) { return (long long)widest <= (long long)h.n_cu * per_cu(lds); };
  int mode = 0;
  if (ok4 && one_round(lds4)) mode = 4;
  else if (ok1 && one_round(lds1)) mode = 1;
  else if (ok2 && one_round(lds2)) mode = 2;
  else if (ok3 && one_round(lds3)) mode = 3;
  else if (ok4) mode = 4;
  else if (ok1) mode = 1;
  else if (ok2) mode = 2;
  else if (ok3) mode = 3;
  return mode;
}

template <typename T>
static void gs_launch(Handle& h, const ApplyArgs& a, hipStream_t s) {
  const int N = h.N, Nz = h.desc.Nz;
  const size_t lds1 = gs_lds_bytes<T, 1>(Nz), lds2 = gs_lds_bytes<T, 2>(Nz), lds3 = gs_lds_bytes<T, 3>(Nz);
  const size_t lds0 = gs_lds_bytes<T, 0>(Nz), lds4 = gs_lds_bytes<T, 4>(Nz);
  constexpr bool narrow = sizeof(T) == sizeof(float);
  // a twisted set-up keeps the cell factors only down to the twist pair, so it runs a pair form
  // whatever staging option was set after it (one is granted: the set-up checked)
  int mode = gs_pick_mode<T>(h, h.gs_full);
  if (mode == 0 && h.gs_tw_m >= 0) mode = gs_pick_mode<T>(h, 1);
  GsArgs g{h.Jo, h.Jcw, h.Jw, h.Jwc, h.s0, h.gsb, h.gsu, h.gs_order, h.prec_dh, 1, h.Tw};
  const int C = (int)h.gs_cstart.size() - 1;
  // With two colours no line's first visit reads a line not yet visited, so the right-hand side
  // is formed on each line's first visit; with more, a separate pass forms it and zeroes U first.
  const bool fold = C <= 2;
  if (!fold) hipLaunchKernelGGL(k_gs_gather, dim3((N + 255) / 256), dim3(256), 0, s, N, a, h.s0, h.gsb, h.gsu);
  std::vector<int> seq;
  if (h.gs_forward) {
    // forward sweeps only: colours 0 .. C-1, gs_sweeps times
    for (int sw = 0; sw < h.gs_sweeps; ++sw)
      for (int c = 0; c < C; ++c) seq.push_back(c);
  } else {
    for (int c = 0; c < C; ++c) seq.push_back(c);
    for (int c = C - 2; c >= 0; --c) seq.push_back(c);
    for (int sw = 1; sw < h.gs_sweeps; ++sw) {
      for (int c = 1; c < C; ++c) seq.push_back(c);
      for (int c = C - 2; c >= 0; --c) seq.push_back(c);
    }
  }
  // a colour's last visit writes the apply's outputs for its lines
  std::vector<int> last(C, -1), first(C, -1);
  for (size_t q = 0; q < seq.size(); ++q) {
    last[seq[q]] = (int)q;
    if (first[seq[q]] < 0) first[seq[q]] = (int)q;
  }
  const int nlines = h.gs_cstart[C];
  double* partial = h.red.partial + RMAXB * 8;
  unsigned* counter = h.red.counter + 1;
  for (size_t q = 0; q < seq.size(); ++q) {
    const int c = seq[q];
    const int fin = last[c] == (int)q ? 1 : 0;
    const int fv = fold && first[c] == (int)q ? 1 : 0;
    g.first = q == 0 ? 1 : 0;
    const int cnt = h.gs_cstart[c + 1] - h.gs_cstart[c];
    if (cnt == 0) continue;
    const bool sample = h.prof_every > 0 && (h.prof_seen++ % h.prof_every) == 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (sample) {
      e0 = prof_event(h);
      e1 = prof_event(h);
      if (e0 && e1) (void)hipEventRecord(e0, s);
    }
    if (mode == 4) {
      if constexpr (narrow)
        hipLaunchKernelGGL((k_gs_half<T, 4>), dim3(cnt), dim3(GSW), lds4, s, h.n, h.desc.Nr, h.desc.Nth, Nz, h.ncol,
                           h.nw, (const T*)h.P, g, h.gs_cstart[c], a.gs, a.respect_done, a, fv, fin, nlines, N,
                           partial, counter, h.gs_tw_m);
    } else if (mode == 1)
      hipLaunchKernelGGL((k_gs_half<T, 1>), dim3(cnt), dim3(GSW), lds1, s, h.n, h.desc.Nr, h.desc.Nth, Nz, h.ncol,
                         h.nw, (const T*)h.P, g, h.gs_cstart[c], a.gs, a.respect_done, a, fv, fin, nlines, N,
                         partial, counter, h.gs_tw_m);
    else if (mode == 2)
      hipLaunchKernelGGL((k_gs_half<T, 2>), dim3(cnt), dim3(GSW), lds2, s, h.n, h.desc.Nr, h.desc.Nth, Nz, h.ncol,
                         h.nw, (const T*)h.P, g, h.gs_cstart[c], a.gs, a.respect_done, a, fv, fin, nlines, N,
                         partial, counter, h.gs_tw_m);
    else if (mode == 3)
      hipLaunchKernelGGL((k_gs_half<T, 3>), dim3(cnt), dim3(GSW), lds3, s, h.n, h.desc.Nr, h.desc.Nth, Nz, h.ncol,
                         h.nw, (const T*)h.P, g, h.gs_cstart[c], a.gs, a.respect_done, a, fv, fin, nlines, N,
                         partial, counter, h.gs_tw_m);
    else
      hipLaunchKernelGGL((k_gs_half<T, 0>), dim3(cnt), dim3(PW), lds0, s, h.n, h.desc.Nr, h.desc.Nth, Nz, h.ncol,
                         h.nw, (const T*)h.F, g, h.gs_cstart[c], a.gs, a.respect_done, a, fv, fin, nlines, N,
                         partial, counter, h.gs_tw_m);
    if (sample && e0 && e1) {
      (void)hipEventRecord(e1, s);
      h.prof_pending.push_back(e0);
      h.prof_pending.push_back(e1);
      h.prof_pending_bytes.push_back(gs_bytes(h, c, g.first != 0, sizeof(T), mode, fv != 0, fin != 0, a));
      if (h.prof_pending_bytes.size() >= 512) (void)prof_resolve(h);
    }
  }
}

void apply_launch(Handle& h, const ApplyArgs& a, hipStream_t s) {
  const int Nz = h.desc.Nz;
  if (h.prec_kind == VTSA_PREC_ZLINE_GS) {
    if (h.prec_prec == VTSA_FP32) gs_launch<float>(h, a, s);
    else gs_launch<double>(h, a, s);
    return;
  }
  const int nwb = h.nw > 0 ? (h.nw + PW * 4 - 1) / (PW * 4) : 0;
  const int grid = h.ncol + nwb;
  if (h.prec_prec == VTSA_FP32)
    hipLaunchKernelGGL(k_prec_apply<float>, dim3(grid), dim3(PW), zline_lds_bytes<float>(Nz), s, h.n, Nz, h.ncol,
                       h.nw, h.prec_kind, (const float*)h.F, h.Fw, h.s0, a, h.red.partial + RMAXB * 8,
                       h.red.counter + 1);
  else
    hipLaunchKernelGGL(k_prec_apply<double>, dim3(grid), dim3(PW), zline_lds_bytes<double>(Nz), s, h.n, Nz, h.ncol,
                       h.nw, h.prec_kind, (const double*)h.F, h.Fw, h.s0, a, h.red.partial + RMAXB * 8,
                       h.red.counter + 1);
}

void launch_prec_apply(Handle& h, const double* in, const double* in_scale, double* store_in, const double* pre,
                       const double* post, double* out, bool sigma_for_matvec, hipStream_t s) {
  ApplyArgs a{in, in_scale, store_in, pre, post, out, nullptr, nullptr, sigma_for_matvec ? h.gs : nullptr,
              sigma_for_matvec ? &h.gs->zs_rms : nullptr,
              sigma_for_matvec ? 1 : 0};
  apply_launch(h, a, s);
}

}  // namespace vtsa

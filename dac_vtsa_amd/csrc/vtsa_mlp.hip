// Learned stage guess: a per-cell MLP on the gfx950 matrix cores, trained on the device.
//
// Network: 32 inputs -> 128 -> 128 -> 128 -> 16 outputs (6 used), ReLU after each hidden layer.
// Weights and activations are BF16; every product accumulates in FP32 on the matrix cores
// (v_mfma_f32_16x16x32_bf16).  Master weights, biases and the Adam moments are FP32.
// One wavefront owns 16 rows (cells); a row is one cell of one implicit stage.
//
// Operand layout of the 16x16x32 BF16 MFMA (lane l of 64, i = 0..7, j = 0..3):
//   A[m][k] : lane l holds A[l % 16][8 (l / 16) + i]
//   B[k][n] : lane l holds B[8 (l / 16) + i][l % 16]
//   D[m][n] : lane l holds D[4 (l / 16) + j][l % 16]
//
// Parameters, FP32, flat: W1[128][32] b1[128] W2[128][128] b2[128] W3[128][128] b3[128]
// W4[16][128] b4[16] (W[out][in]).  BF16 copies: wb in the same layout (B operand of the
// forward pass) and wt with every W transposed to [in][out] (B operand of the backward pass).
//
// Training keeps transposed activations XT[32][rows], H_l^T[128][rows] and gradients
// G4^T[16][rows], G_l^T[128][rows] so that the weight-gradient products dW = G^T H read both
// operands as contiguous 16-byte runs.
#include <cmath>
#include <cstring>
#include <vector>

#include "vtsa_internal.h"
#include "vtsa_reduce.h"

namespace vtsa {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int MI = 32, MH = 128, MO = 16, MO_USED = 6;
constexpr int MROWS = 16;    // rows per wavefront
constexpr int MWAVES = 4;    // wavefronts per block of the row passes
constexpr int MCHUNK = 256;  // rows per weight-gradient partial (rows are padded to a multiple)
constexpr int OW1 = 0, OB1 = OW1 + MH * MI, OW2 = OB1 + MH, OB2 = OW2 + MH * MH, OW3 = OB2 + MH,
              OB3 = OW3 + MH * MH, OW4 = OB3 + MH, OB4 = OW4 + MO * MH, NPARAM = OB4 + MO;
constexpr int NNORM = 2 * MI + 2 * MO;   // input mean, input 1/std, target mean, target std
constexpr int NTILE_W = (MH / 16) * (MI / 16) + 2 * (MH / 16) * (MH / 16) + (MO / 16) * (MH / 16);
constexpr int LDH = MH + 8;   // LDS row stride (elements) of a hidden tile
constexpr int LDG = MO + 8;   // LDS row stride of the output-gradient tile

struct Mlp {
  int max_rows = 0, max_pad = 0;
  float *p = nullptr, *m1 = nullptr, *m2 = nullptr, *norm = nullptr, *part = nullptr, *loss = nullptr;
  float* grad = nullptr;
  __bf16 *wb = nullptr, *wt = nullptr, *xt = nullptr, *ht = nullptr, *gt = nullptr;
  float* host_loss = nullptr;   // pinned
  int step = 0;
};

__device__ __forceinline__ void wave_lds_order() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront"); }

__device__ __forceinline__ f32x4 mfma(const bf16x8& a, const bf16x8& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 ld8(const __bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }

__device__ __forceinline__ bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int i = 0; i < 8; ++i) z[i] = (__bf16)0.0f;
  return z;
}

// Four consecutive rows of one column of a transposed [cols][rows] array (8-byte store).
__device__ __forceinline__ void st4(__bf16* p, float a, float b, float c, float d) {
  union { __bf16 h[4]; uint2 u; } v;
  v.h[0] = (__bf16)a; v.h[1] = (__bf16)b; v.h[2] = (__bf16)c; v.h[3] = (__bf16)d;
  *reinterpret_cast<uint2*>(p) = v.u;
}

// One dense layer for the 16 rows of a wavefront: acc[t] (t = 16-column tile of the outputs)
// = A (16 x 32*KS, fragments in registers) times W^T, W = w[off .. ] in [out][in] layout.
template <int KS, int NT>
__device__ __forceinline__ void layer(const bf16x8 (&a)[KS], const __bf16* __restrict__ w, int in_dim, int lr, int lg,
                                      f32x4 (&acc)[NT]) {
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const __bf16* row = w + (size_t)(16 * t + lr) * in_dim + 8 * lg;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) acc[t] = mfma(a[ks], ld8(row + 32 * ks), acc[t]);
  }
}

// Forward pass of rows [0, rows) and, when `target` is given, the loss and the backward data
// pass.  Rows in [rows, rows_pad) are computed from zero inputs and carry zero gradients.
__global__ __launch_bounds__(64 * MWAVES) void k_mlp_pass(const __bf16* __restrict__ wb, const __bf16* __restrict__ wt,
                                                          const float* __restrict__ p, const float* __restrict__ feat,
                                                          int rows, int rows_pad, const float* __restrict__ norm,
                                                          float* __restrict__ out, const float* __restrict__ target,
                                                          float gscale, __bf16* __restrict__ xt,
                                                          __bf16* __restrict__ ht, __bf16* __restrict__ gt,
                                                          float* __restrict__ loss) {
  __shared__ __bf16 lds[MWAVES][3 * 16 * LDH + 16 * LDG];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int tile = blockIdx.x * MWAVES + wv;
  const int r0 = tile * MROWS;
  if (r0 >= rows_pad) return;   // whole wavefronts leave; no block-wide barrier follows
  const int lr = lane & 15, lg = lane >> 4;
  const bool train = target != nullptr;
  __bf16* a1 = lds[wv];
  __bf16* a2 = a1 + 16 * LDH;
  __bf16* a3 = a2 + 16 * LDH;
  __bf16* g4 = a3 + 16 * LDH;
  const size_t rp = (size_t)rows_pad;

  // inputs: lane holds row r0 + lr, columns 8 lg .. 8 lg + 7 (the A layout)
  bf16x8 x[1];
  {
    const int row = r0 + lr;
    float v[8];
    if (row < rows) {
      const float4* fp = reinterpret_cast<const float4*>(feat + (size_t)row * MI + 8 * lg);
      const float4 u0 = fp[0], u1 = fp[1];
      v[0] = u0.x; v[1] = u0.y; v[2] = u0.z; v[3] = u0.w; v[4] = u1.x; v[5] = u1.y; v[6] = u1.z; v[7] = u1.w;
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = (v[i] - norm[8 * lg + i]) * norm[MI + 8 * lg + i];
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = 0.0f;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) x[0][i] = (__bf16)v[i];
    if (train) {
#pragma unroll
      for (int i = 0; i < 8; ++i) xt[(size_t)(8 * lg + i) * rp + row] = x[0][i];
    }
  }

  f32x4 acc[8];
  bf16x8 a[4];
  __bf16* hts[3] = {ht, ht + (size_t)MH * rp, ht + (size_t)2 * MH * rp};
  __bf16* tiles[3] = {a1, a2, a3};
  const int offw[3] = {OW1, OW2, OW3}, offb[3] = {OB1, OB2, OB3};
#pragma unroll
  for (int L = 0; L < 3; ++L) {
    if (L == 0) layer<1, 8>(x, wb + OW1, MI, lr, lg, acc);
    else layer<4, 8>(a, wb + offw[L], MH, lr, lg, acc);
    __bf16* tl = tiles[L];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int col = 16 * t + lr;
      const float b = p[offb[L] + col];
      float h[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        h[j] = fmaxf(acc[t][j] + b, 0.0f);
        tl[(4 * lg + j) * LDH + col] = (__bf16)h[j];
      }
      if (train) st4(hts[L] + (size_t)col * rp + r0 + 4 * lg, h[0], h[1], h[2], h[3]);
    }
    wave_lds_order();
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) a[ks] = ld8(tl + lr * LDH + 32 * ks + 8 * lg);
  }
  f32x4 y[1];
  layer<4, 1>(a, wb + OW4, MH, lr, lg, y);
  float lsum = 0.0f;
  {
    const int col = lr;
    const float b = p[OB4 + col];
    float g[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int row = r0 + 4 * lg + j;
      const float v = y[0][j] + b;
      if (out && row < rows) out[(size_t)row * MO + col] = v;
      g[j] = 0.0f;
      if (train && row < rows && col < MO_USED) {
        const float tn = (target[(size_t)row * 8 + col] - norm[2 * MI + col]) / norm[2 * MI + MO + col];
        const float d = v - tn;
        lsum += d * d;
        g[j] = d * gscale;
      }
    }
    if (!train) return;
#pragma unroll
    for (int j = 0; j < 4; ++j) g4[(4 * lg + j) * LDG + col] = (__bf16)g[j];
    st4(gt + (size_t)col * rp + r0 + 4 * lg, g[0], g[1], g[2], g[3]);
  }
  lsum = wave_sum(lsum);
  if (lane == 0) loss[tile] = lsum;
  wave_lds_order();

  // backward data pass: G_{L} = (G_{L+1} W_{L+1}) * (H_L > 0), L = 3, 2, 1
  bf16x8 ga[4];
  ga[0] = lg < 2 ? ld8(g4 + lr * LDG + 8 * lg) : zero8();
  __bf16* gts[3] = {gt + (size_t)MO * rp, gt + (size_t)MO * rp + (size_t)MH * rp,
                    gt + (size_t)MO * rp + (size_t)2 * MH * rp};
#pragma unroll
  for (int L = 2; L >= 0; --L) {
    // dH_L[r][in] = sum_out G_{L+1}[r][out] W_{L+1}[out][in]; B operand from wt ([in][out])
    if (L == 2) {
      // W4 is [16][128]: K = 16 outputs, padded to 32 with zero lanes
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const bf16x8 bw = lg < 2 ? ld8(wt + OW4 + (size_t)(16 * t + lr) * MO + 8 * lg) : zero8();
        acc[t] = mfma(ga[0], bw, f32x4{0.f, 0.f, 0.f, 0.f});
      }
    } else {
      layer<4, 8>(ga, wt + offw[L + 1], MH, lr, lg, acc);
    }
    __bf16* tl = tiles[L];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int col = 16 * t + lr;
      float g[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float h = (float)tl[(4 * lg + j) * LDH + col];
        g[j] = h > 0.0f ? acc[t][j] : 0.0f;
        if (L > 0) tl[(4 * lg + j) * LDH + col] = (__bf16)g[j];
      }
      st4(gts[L] + (size_t)col * rp + r0 + 4 * lg, g[0], g[1], g[2], g[3]);
    }
    if (L > 0) {
      wave_lds_order();
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) ga[ks] = ld8(tl + lr * LDH + 32 * ks + 8 * lg);
    }
  }
}

// Weight and bias gradients: one wavefront per (16 x 16 tile of one W, chunk of MCHUNK rows);
// partial sums go to part[chunk][NPARAM] and are reduced in a fixed order by k_mlp_adam.
__global__ __launch_bounds__(256) void k_mlp_wgrad(const __bf16* __restrict__ xt, const __bf16* __restrict__ ht,
                                                   const __bf16* __restrict__ gt, int rows_pad,
                                                   float* __restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int id = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (id >= NTILE_W) return;
  const int chunk = blockIdx.y;
  const int lr = lane & 15, lg = lane >> 4;
  const size_t rp = (size_t)rows_pad;
  int L, ot, it;
  if (id < 16) { L = 0; ot = id >> 1; it = id & 1; }
  else if (id < 80) { L = 1; ot = (id - 16) >> 3; it = (id - 16) & 7; }
  else if (id < 144) { L = 2; ot = (id - 80) >> 3; it = (id - 80) & 7; }
  else { L = 3; ot = 0; it = id - 144; }
  const int in_dim = L == 0 ? MI : MH;
  const int ow = L == 0 ? OW1 : L == 1 ? OW2 : L == 2 ? OW3 : OW4;
  const int ob = L == 0 ? OB1 : L == 1 ? OB2 : L == 2 ? OB3 : OB4;
  const __bf16* G = L == 3 ? gt : gt + (size_t)MO * rp + (size_t)L * MH * rp;
  const __bf16* H = L == 0 ? xt : ht + (size_t)(L - 1) * MH * rp;
  const __bf16* ga = G + (size_t)(16 * ot + lr) * rp + 8 * lg;
  const __bf16* hb = H + (size_t)(16 * it + lr) * rp + 8 * lg;
  const int c0 = chunk * MCHUNK;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  float bsum = 0.0f;
#pragma unroll 4
  for (int rr = c0; rr < c0 + MCHUNK; rr += 32) {
    const bf16x8 av = ld8(ga + rr), bv = ld8(hb + rr);
    acc = mfma(av, bv, acc);
    if (it == 0) {
#pragma unroll
      for (int i = 0; i < 8; ++i) bsum += (float)av[i];
    }
  }
  float* pc = part + (size_t)chunk * NPARAM;
#pragma unroll
  for (int j = 0; j < 4; ++j) pc[ow + (16 * ot + 4 * lg + j) * in_dim + 16 * it + lr] = acc[j];
  if (it == 0) {
    bsum += __shfl_xor(bsum, 16, 64);
    bsum += __shfl_xor(bsum, 32, 64);
    if (lg == 0) pc[ob + 16 * ot + lr] = bsum;
  }
}

// Transposed position of parameter index q inside the W blocks (-1 for a bias).
__device__ __forceinline__ int transposed_index(int q) {
  int ow, in_dim, out_dim;
  if (q < OB1) { ow = OW1; in_dim = MI; out_dim = MH; }
  else if (q >= OW2 && q < OB2) { ow = OW2; in_dim = MH; out_dim = MH; }
  else if (q >= OW3 && q < OB3) { ow = OW3; in_dim = MH; out_dim = MH; }
  else if (q >= OW4 && q < OB4) { ow = OW4; in_dim = MH; out_dim = MO; }
  else return -1;
  const int r = q - ow, o = r / in_dim, i = r - o * in_dim;
  return ow + i * out_dim + o;
}

// Reduces the gradient partials (fixed order), applies one Adam step when lr > 0 and refreshes
// the BF16 copies.  nchunk = 0: only the BF16 copies are refreshed (after new parameters).
__global__ void k_mlp_adam(float* __restrict__ p, float* __restrict__ m1, float* __restrict__ m2,
                           const float* __restrict__ part, int nchunk, float* __restrict__ grad, float lr, float b1,
                           float b2, float eps, float c1, float c2, __bf16* __restrict__ wb, __bf16* __restrict__ wt) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= NPARAM) return;
  float w = p[q];
  if (nchunk > 0) {
    float g = 0.0f;
    for (int c = 0; c < nchunk; ++c) g += part[(size_t)c * NPARAM + q];
    grad[q] = g;
    if (lr > 0.0f) {
      const float a = b1 * m1[q] + (1.0f - b1) * g;
      const float v = b2 * m2[q] + (1.0f - b2) * g * g;
      m1[q] = a;
      m2[q] = v;
      w -= lr * (a * c1) / (sqrtf(v * c2) + eps);
      p[q] = w;
    }
  }
  wb[q] = (__bf16)w;
  const int tq = transposed_index(q);
  if (tq >= 0) wt[tq] = (__bf16)w;
}

// Per-cell input rows of the network for one implicit stage (32 FP32 values per cell):
//   [0, 6)   y of the cell (c_CO2, c_H2O, c_N2, q_CO2, q_H2O, e) at the start of the step
//   [6, 12)  the same at the z - 1 neighbour (the cell itself at the column end)
//   [12, 18) the same at the z + 1 neighbour
//   [18, 24) (z_pred - y) / s   predictor increment in tolerance units
//   [24, 30) r / s              r = M^-1 R(z_pred), the preconditioned stage residual at the
//                               predictor (minus the first Newton step)
//   30: log(h), 31: stage index (0: first, 1: second)
__global__ void k_mlp_features(int n, int Nz, const double* __restrict__ y, const double* __restrict__ zp,
                               const double* __restrict__ s, const double* __restrict__ R, float log_h, float stage,
                               float* __restrict__ feat) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int k = c % Nz;
  const int km = k > 0 ? c - 1 : c, kp = k < Nz - 1 ? c + 1 : c;
  float v[32];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const size_t o = (size_t)q * n;
    const double yc = y[o + c], sc = s[o + c];
    v[q] = (float)yc;
    v[6 + q] = (float)y[o + km];
    v[12 + q] = (float)y[o + kp];
    v[18 + q] = (float)((zp[o + c] - yc) / sc);
    v[24 + q] = (float)(R[o + c] / sc);
  }
  v[30] = log_h;
  v[31] = stage;
  float4* out = reinterpret_cast<float4*>(feat + (size_t)c * MI);
#pragma unroll
  for (int q = 0; q < 8; ++q) out[q] = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
}

// Training targets: (z_converged - z_base) / s for the six cell variables (two zero pads), with
// z_base = z_pred - r (r = M^-1 R(z_pred): the base is the predictor after one preconditioned
// Newton step) or z_base = z_pred when r is null.
__global__ void k_mlp_targets(int n, const double* __restrict__ zc, const double* __restrict__ zp,
                              const double* __restrict__ r, const double* __restrict__ s,
                              float* __restrict__ target) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  float v[8];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const size_t o = (size_t)q * n + c;
    const double base = r ? zp[o] - r[o] : zp[o];
    v[q] = (float)((zc[o] - base) / s[o]);
  }
  v[6] = v[7] = 0.0f;
  float4* out = reinterpret_cast<float4*>(target + (size_t)c * 8);
  out[0] = make_float4(v[0], v[1], v[2], v[3]);
  out[1] = make_float4(v[4], v[5], v[6], v[7]);
}

// Guess: z = z_base + s * (out * std + mean) on the six cell variables (z_base as for the
// targets), with the integrator's positivity truncation on c and q (at least 0.1 z_pred,
// core/solver.py:99-108); the wall temperatures take z_base.
__global__ void k_mlp_apply(int n, int N, const float* __restrict__ out, const float* __restrict__ norm,
                            const double* __restrict__ zp, const double* __restrict__ r,
                            const double* __restrict__ s, double* __restrict__ zg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double zo = zp[i];
  const double base = r ? zo - r[i] : zo;
  if (i >= 6 * n) { zg[i] = base; return; }
  const int q = i / n, c = i - q * n;
  const double d = (double)out[(size_t)c * MO + q] * (double)norm[2 * MI + MO + q] + (double)norm[2 * MI + q];
  double v = base + s[i] * d;
  if (q < 5) {
    const double lo = 0.1 * (zo > 0.0 ? zo : 0.0);
    v = v > lo ? v : lo;
  }
  zg[i] = v;
}

static int pad_rows(int rows) { return (rows + MCHUNK - 1) / MCHUNK * MCHUNK; }

static void free_mlp(Mlp& m) {
  void* bufs[] = {m.p, m.m1, m.m2, m.norm, m.part, m.loss, m.grad, m.wb, m.wt, m.xt, m.ht, m.gt};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (m.host_loss) (void)hipHostFree(m.host_loss);
}

}  // namespace vtsa

using namespace vtsa;

struct vtsa_mlp {
  Mlp m;
};

extern "C" {

int vtsa_mlp_sizes(int32_t* n_param, int32_t* n_in, int32_t* n_hidden, int32_t* n_out, int32_t* n_norm) {
  if (n_param) *n_param = NPARAM;
  if (n_in) *n_in = MI;
  if (n_hidden) *n_hidden = MH;
  if (n_out) *n_out = MO;
  if (n_norm) *n_norm = NNORM;
  return 0;
}

int vtsa_mlp_create(int32_t max_rows, vtsa_mlp** out) {
  if (!out || max_rows < 1) { set_error("vtsa_mlp_create: bad argument"); return -1; }
  *out = nullptr;
  vtsa_mlp* M = new vtsa_mlp();
  Mlp& m = M->m;
  m.max_rows = max_rows;
  m.max_pad = pad_rows(max_rows);
  const size_t rp = (size_t)m.max_pad;
  const size_t nchunk = rp / MCHUNK;
  bool bad = check(hipMalloc((void**)&m.p, NPARAM * sizeof(float)), "hipMalloc") ||
             check(hipMalloc((void**)&m.m1, NPARAM * sizeof(float)), "hipMalloc") ||
             check(hipMalloc((void**)&m.m2, NPARAM * sizeof(float)), "hipMalloc") ||
             check(hipMalloc((void**)&m.grad, NPARAM * sizeof(float)), "hipMalloc") ||
             check(hipMalloc((void**)&m.norm, NNORM * sizeof(float)), "hipMalloc") ||
             check(hipMalloc((void**)&m.part, nchunk * NPARAM * sizeof(float)), "hipMalloc") ||
             check(hipMalloc((void**)&m.loss, rp / MROWS * sizeof(float)), "hipMalloc") ||
             check(hipMalloc((void**)&m.wb, NPARAM * sizeof(__bf16)), "hipMalloc") ||
             check(hipMalloc((void**)&m.wt, NPARAM * sizeof(__bf16)), "hipMalloc") ||
             check(hipMalloc((void**)&m.xt, MI * rp * sizeof(__bf16)), "hipMalloc") ||
             check(hipMalloc((void**)&m.ht, 3 * MH * rp * sizeof(__bf16)), "hipMalloc") ||
             check(hipMalloc((void**)&m.gt, (MO + 3 * MH) * rp * sizeof(__bf16)), "hipMalloc") ||
             check(hipHostMalloc((void**)&m.host_loss, rp / MROWS * sizeof(float), hipHostMallocDefault),
                   "hipHostMalloc");
  if (!bad) {
    std::vector<float> ones(NNORM, 0.0f);
    for (int i = 0; i < MI; ++i) ones[MI + i] = 1.0f;
    for (int i = 0; i < MO; ++i) ones[2 * MI + MO + i] = 1.0f;
    bad = check(hipMemset(m.p, 0, NPARAM * sizeof(float)), "hipMemset") ||
          check(hipMemset(m.m1, 0, NPARAM * sizeof(float)), "hipMemset") ||
          check(hipMemset(m.m2, 0, NPARAM * sizeof(float)), "hipMemset") ||
          check(hipMemset(m.wb, 0, NPARAM * sizeof(__bf16)), "hipMemset") ||
          check(hipMemset(m.wt, 0, NPARAM * sizeof(__bf16)), "hipMemset") ||
          check(hipMemcpy(m.norm, ones.data(), NNORM * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy") ||
          check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");   // null-stream fills done
  }
  if (bad) {
    free_mlp(m);
    delete M;
    return -1;
  }
  *out = M;
  return 0;
}

int vtsa_mlp_destroy(vtsa_mlp* M) {
  if (!M) return 0;
  (void)hipDeviceSynchronize();
  free_mlp(M->m);
  delete M;
  return 0;
}

int vtsa_mlp_set_params(vtsa_mlp* M, const float* params, vtsa_stream st) {
  if (!M || !params) { set_error("vtsa_mlp_set_params: null argument"); return -1; }
  Mlp& m = M->m;
  hipStream_t s = (hipStream_t)st;
  if (check(hipMemcpyAsync(m.p, params, NPARAM * sizeof(float), hipMemcpyHostToDevice, s), "hipMemcpyAsync") ||
      check(hipMemsetAsync(m.m1, 0, NPARAM * sizeof(float), s), "hipMemsetAsync") ||
      check(hipMemsetAsync(m.m2, 0, NPARAM * sizeof(float), s), "hipMemsetAsync"))
    return -1;
  m.step = 0;
  k_mlp_adam<<<(NPARAM + 255) / 256, 256, 0, s>>>(m.p, m.m1, m.m2, m.part, 0, m.grad, 0.0f, 0.f, 0.f, 0.f, 0.f, 0.f,
                                                  m.wb, m.wt);
  if (check(hipGetLastError(), "k_mlp_adam")) return -1;
  return check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

int vtsa_mlp_get_params(vtsa_mlp* M, float* params, vtsa_stream st) {
  if (!M || !params) { set_error("vtsa_mlp_get_params: null argument"); return -1; }
  hipStream_t s = (hipStream_t)st;
  if (check(hipMemcpyAsync(params, M->m.p, NPARAM * sizeof(float), hipMemcpyDeviceToHost, s), "hipMemcpyAsync"))
    return -1;
  return check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

int vtsa_mlp_set_norm(vtsa_mlp* M, const float* norm, vtsa_stream st) {
  if (!M || !norm) { set_error("vtsa_mlp_set_norm: null argument"); return -1; }
  for (int i = 0; i < NNORM; ++i)
    if (!std::isfinite(norm[i])) { set_error("vtsa_mlp_set_norm: non-finite value"); return -2; }
  for (int i = 0; i < MO; ++i)
    if (!(norm[2 * MI + MO + i] > 0.0f)) { set_error("vtsa_mlp_set_norm: target std must be > 0"); return -2; }
  hipStream_t s = (hipStream_t)st;
  if (check(hipMemcpyAsync(M->m.norm, norm, NNORM * sizeof(float), hipMemcpyHostToDevice, s), "hipMemcpyAsync"))
    return -1;
  return check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

int vtsa_mlp_forward(vtsa_mlp* M, const float* feat, int32_t rows, float* out, vtsa_stream st) {
  if (!M || !feat || !out) { set_error("vtsa_mlp_forward: null argument"); return -1; }
  Mlp& m = M->m;
  if (rows < 1 || rows > m.max_rows) { set_error("vtsa_mlp_forward: rows out of range"); return -2; }
  const int rp = pad_rows(rows);
  const int tiles = rp / MROWS;
  k_mlp_pass<<<(tiles + MWAVES - 1) / MWAVES, 64 * MWAVES, 0, (hipStream_t)st>>>(
      m.wb, m.wt, m.p, feat, rows, rp, m.norm, out, nullptr, 0.0f, nullptr, nullptr, nullptr, nullptr);
  return check(hipGetLastError(), "k_mlp_pass");
}

int vtsa_mlp_train_step(vtsa_mlp* M, const float* feat, const float* target, int32_t rows, float lr, float* loss,
                        float* grad, vtsa_stream st) {
  if (!M || !feat || !target) { set_error("vtsa_mlp_train_step: null argument"); return -1; }
  Mlp& m = M->m;
  if (rows < 1 || rows > m.max_rows) { set_error("vtsa_mlp_train_step: rows out of range"); return -2; }
  if (!(lr >= 0.0f) || !std::isfinite(lr)) { set_error("vtsa_mlp_train_step: bad learning rate"); return -2; }
  hipStream_t s = (hipStream_t)st;
  const int rp = pad_rows(rows);
  const int tiles = rp / MROWS, nchunk = rp / MCHUNK;
  const float gscale = 2.0f / ((float)rows * (float)MO_USED);
  k_mlp_pass<<<(tiles + MWAVES - 1) / MWAVES, 64 * MWAVES, 0, s>>>(m.wb, m.wt, m.p, feat, rows, rp, m.norm, nullptr,
                                                                   target, gscale, m.xt, m.ht, m.gt, m.loss);
  if (check(hipGetLastError(), "k_mlp_pass")) return -1;
  k_mlp_wgrad<<<dim3((NTILE_W + 3) / 4, nchunk), 256, 0, s>>>(m.xt, m.ht, m.gt, rp, m.part);
  if (check(hipGetLastError(), "k_mlp_wgrad")) return -1;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float c1 = 1.0f, c2 = 1.0f;
  if (lr > 0.0f) {
    m.step += 1;
    c1 = (float)(1.0 / (1.0 - std::pow(0.9, m.step)));
    c2 = (float)(1.0 / (1.0 - std::pow(0.999, m.step)));
  }
  k_mlp_adam<<<(NPARAM + 255) / 256, 256, 0, s>>>(m.p, m.m1, m.m2, m.part, nchunk, m.grad, lr, b1, b2, eps, c1, c2,
                                                  m.wb, m.wt);
  if (check(hipGetLastError(), "k_mlp_adam")) return -1;
  if (loss) {
    if (check(hipMemcpyAsync(m.host_loss, m.loss, tiles * sizeof(float), hipMemcpyDeviceToHost, s), "hipMemcpyAsync"))
      return -1;
  }
  if (grad) {
    if (check(hipMemcpyAsync(grad, m.grad, NPARAM * sizeof(float), hipMemcpyDeviceToHost, s), "hipMemcpyAsync"))
      return -1;
  }
  if (loss || grad) {
    if (check(hipStreamSynchronize(s), "hipStreamSynchronize")) return -1;
  }
  if (loss) {
    double acc = 0.0;
    for (int i = 0; i < tiles; ++i) acc += m.host_loss[i];
    *loss = (float)(acc / ((double)rows * MO_USED));
  }
  return 0;
}

int vtsa_mlp_features(vtsa_handle* H, const double* y, const double* zp, const double* s, const double* R,
                      double log_h, double stage, float* feat, vtsa_stream st) {
  if (!H || !y || !zp || !s || !R || !feat) { set_error("vtsa_mlp_features: null argument"); return -1; }
  Handle& h = H->h;
  k_mlp_features<<<(h.n + 255) / 256, 256, 0, (hipStream_t)st>>>(h.n, h.desc.Nz, y, zp, s, R, (float)log_h,
                                                                  (float)stage, feat);
  return check(hipGetLastError(), "k_mlp_features");
}

int vtsa_mlp_targets(vtsa_handle* H, const double* zc, const double* zp, const double* r, const double* s,
                     float* target, vtsa_stream st) {
  if (!H || !zc || !zp || !s || !target) { set_error("vtsa_mlp_targets: null argument"); return -1; }
  Handle& h = H->h;
  k_mlp_targets<<<(h.n + 255) / 256, 256, 0, (hipStream_t)st>>>(h.n, zc, zp, r, s, target);
  return check(hipGetLastError(), "k_mlp_targets");
}

int vtsa_mlp_apply(vtsa_mlp* M, vtsa_handle* H, const float* out, const double* zp, const double* r,
                   const double* s, double* zg, vtsa_stream st) {
  if (!M || !H || !out || !zp || !s || !zg) { set_error("vtsa_mlp_apply: null argument"); return -1; }
  Handle& h = H->h;
  if (h.n > M->m.max_rows) { set_error("vtsa_mlp_apply: more cells than the network's row capacity"); return -2; }
  k_mlp_apply<<<(h.N + 255) / 256, 256, 0, (hipStream_t)st>>>(h.n, h.N, out, M->m.norm, zp, r, s, zg);
  return check(hipGetLastError(), "k_mlp_apply");
}

}  // extern "C"

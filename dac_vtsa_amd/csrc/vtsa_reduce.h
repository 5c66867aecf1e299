// Device-side reduction helpers shared by the kernels of libvtsa: wavefront butterflies,
// block sums, and one-launch cross-block reductions (per-block partials, last block sums).
#pragma once
#include <hip/hip_runtime.h>

namespace vtsa {

constexpr int RBLK = 256;        // reduction block size
constexpr int RMAXB = 512;       // max blocks of a reduction grid
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

// Sums of partial rows q = q0, q0 + qs, .. (R of them, those below nval; rows of stride RMAXB)
// by one wavefront: per row, lane l adds partials l, l + 64, .. in order, then a butterfly;
// deterministic.  The loads of RMAXB partials of every row are issued before their adds.  Lane 0
// stores row q's sum to out[q].
template <int R>
__device__ inline void wave_sum_rows(const double* __restrict__ partial, int cnt, int q0, int qs, int nval,
                                     double* out) {
  constexpr int CM = RMAXB / 64;
  const int lane = threadIdx.x & 63;
  double acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = 0.0;
  for (int base = 0; base < cnt; base += 64 * CM) {
    double v[R][CM];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int c = 0; c < CM; ++c) {
        const int q = q0 + r * qs, b = base + lane + 64 * c;
        v[r][c] = (q < nval && b < cnt) ? load_partial(partial + (size_t)q * RMAXB + b) : 0.0;
      }
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int c = 0; c < CM; ++c)
        if (base + lane + 64 * c < cnt) acc[r] += v[r][c];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] += __shfl_xor(acc[r], off, 64);
  if (lane == 0)
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (q0 + r * qs < nval) out[q0 + r * qs] = acc[r];
}

// Sum of partial[0 .. cnt) by one wavefront, in the order of wave_sum_rows; every lane returns it.
__device__ inline double wave_sum_strided(const double* __restrict__ partial, int cnt) {
  constexpr int CM = RMAXB / 64;
  const int lane = threadIdx.x & 63;
  double acc = 0.0;
  for (int base = 0; base < cnt; base += 64 * CM) {
    double v[CM];
#pragma unroll
    for (int c = 0; c < CM; ++c) {
      const int b = base + lane + 64 * c;
      v[c] = b < cnt ? load_partial(partial + b) : 0.0;
    }
#pragma unroll
    for (int c = 0; c < CM; ++c)
      if (base + lane + 64 * c < cnt) acc += v[c];
  }
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
  for (int q0 = wv; q0 < nval; q0 += 4 * nwv) wave_sum_rows<4>(partial, gridDim.x, q0, nwv, nval, out);
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

}  // namespace vtsa

// ohm_reduce.h — wave64 / workgroup reductions, the deterministic "last
// workgroup finishes the reduction" pattern (one-off reductions), and
// deferred reductions whose consumer re-sums the producer's partials. Partial sums are combined in a
// fixed order, so results are bitwise reproducible run to run.
#pragma once

#include "ohm_internal.h"

namespace ohm {

// Compensated accumulation: products via TwoProd (fma), sums via TwoSum
// (Ogita, Rump & Oishi, "Accurate sum and dot product", SIAM J. Sci. Comput.
// 26(6), 2005, Algorithm Dot2). FP64 accumulates directly.
template <typename T>
struct Acc;

template <>
struct Acc<double> {
  double s = 0.0;
  __device__ __forceinline__ void add_prod(double a, double b) { s = fma(a, b, s); }
  __device__ __forceinline__ double get() const { return s; }
};

template <>
struct Acc<float> {
  float hi = 0.f, lo = 0.f;
  __device__ __forceinline__ void add_prod(float a, float b) {
    float p = a * b;
    float pe = fmaf(a, b, -p);
    float s = hi + p;
    float bb = s - hi;
    float se = (hi - (s - bb)) + (p - bb);
    hi = s;
    lo += pe + se;
  }
  __device__ __forceinline__ double get() const { return (double)hi + (double)lo; }
};

// op(v, the value a DPP lane permutation brings to this lane), moved as the
// two 32-bit halves of the double; lanes of rows outside kRowMask combine
// with `ident` (the identity of op) instead.
template <int kCtrl, int kRowMask, typename Op>
__device__ __forceinline__ double dpp_op(double v, double ident, Op op) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v), e = __builtin_bit_cast(uint64_t, ident);
  const int lo = __builtin_amdgcn_update_dpp((int)(uint32_t)e, (int)(uint32_t)u, kCtrl, kRowMask,
                                             0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(e >> 32), (int)(uint32_t)(u >> 32),
                                             kCtrl, kRowMask, 0xf, false);
  return op(v, __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// op over the 64 lanes of a wave (every lane active) with DPP lane
// permutations, no LDS traffic: pairs and quads (quad_perm), 8-lane halves
// (row_half_mirror), 16-lane rows (row_mirror), then row 0 into row 1 and
// row 2 into row 3 (row_bcast:15) and rows 0-1 into rows 2-3
// (row_bcast:31). Lane 63 ends with the total, which every lane gets back.
// The combination order is fixed, so sums are reproducible.
template <typename Op>
__device__ __forceinline__ double wave_reduce(double v, double ident, Op op) {
  v = dpp_op<0xb1, 0xf>(v, ident, op);   // quad_perm [1,0,3,2]
  v = dpp_op<0x4e, 0xf>(v, ident, op);   // quad_perm [2,3,0,1]
  v = dpp_op<0x141, 0xf>(v, ident, op);  // row_half_mirror
  v = dpp_op<0x140, 0xf>(v, ident, op);  // row_mirror
  v = dpp_op<0x142, 0xa>(v, ident, op);  // row_bcast:15 into rows 1 and 3
  v = dpp_op<0x143, 0xc>(v, ident, op);  // row_bcast:31 into rows 2 and 3
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)u, 63);
  const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), 63);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ double wave_sum(double v) {
  return wave_reduce(v, 0.0, [](double a, double b) { return a + b; });
}

// fmax over the wave (a NaN lane is ignored, as fmax does)
__device__ __forceinline__ double wave_max(double v) {
  return wave_reduce(v, -INFINITY, [](double a, double b) { return fmax(a, b); });
}

// Sum K values over the workgroup; totals land in thread 0's v[].
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K]) {
  __shared__ double sm[K][kWaves];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; k++) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) sm[k][w] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < kWaves; i++) s += sm[k][i];
      v[k] = s;
    }
  }
  __syncthreads();
}

// Each workgroup deposits its K partials; the last one to arrive (ticket ==
// gridDim.x-1) sums all partials in workgroup order and returns true with
// the grand totals in thread 0's tot[]. The ticket counter is reset for the
// next kernel by that last workgroup.
template <int K>
__device__ __forceinline__ bool grid_sum(double (&v)[K], double *partials, unsigned *counter,
                                         double (&tot)[K]) {
  block_sum<K>(v);
  __shared__ int s_last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++)
      __hip_atomic_store(&partials[(size_t)blockIdx.x * K + k], v[k], __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    // The partials are agent-coherent atomic stores; waiting for them to
    // complete orders them before the ticket without an agent-scope fence
    // (which would write back the whole L2 of this XCD).
    __builtin_amdgcn_s_waitcnt(0);
    unsigned t = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!s_last) return false;
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; k++) acc[k] = 0.0;
  for (unsigned i = threadIdx.x; i < gridDim.x; i += kBlock) {
#pragma unroll
    for (int k = 0; k < K; k++)
      acc[k] += __hip_atomic_load(&partials[(size_t)i * K + k], __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
  }
  block_sum<K>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) tot[k] = acc[k];
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return true;
}

// Producer side of a deferred reduction: each workgroup stores its K block
// sums at partials[blockIdx.x * K ...] with plain stores and the kernel
// boundary publishes them. The consuming kernel sums them itself
// (gather_sum) or, where every consumer workgroup would read too many of
// them, a one-workgroup fold kernel in between reduces them to
// partials[0..K) and the consumer gathers that single partial.
template <int K>
__device__ __forceinline__ void deposit(double (&v)[K], double *partials) {
  block_sum<K>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) partials[(size_t)blockIdx.x * K + k] = v[k];
  }
}

// Consumer side: sums np partials in the same order grid_sum uses, so every
// workgroup of the consuming kernel holds bitwise-identical totals (in every
// thread) and takes the same scalar decisions.
template <int K>
__device__ __forceinline__ void gather_sum(const double *__restrict__ partials, int np,
                                           double (&tot)[K]) {
  __shared__ double bc[K];
#pragma unroll
  for (int k = 0; k < K; k++) tot[k] = 0.0;
  for (int i = threadIdx.x; i < np; i += kBlock) {
#pragma unroll
    for (int k = 0; k < K; k++) tot[k] += partials[(size_t)i * K + k];
  }
  block_sum<K>(tot);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) bc[k] = tot[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; k++) tot[k] = bc[k];
}

// Dense shadow residual value in [-1, 1): SplitMix64 finaliser of (i+1)*gamma.
__host__ __device__ __forceinline__ double shadow_value(uint64_t i) {
  uint64_t z = (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

// Workgroups for an n-element vector kernel: one element per thread (so a
// 65k-entry vector still spreads over every CU) up to kMaxReduceBlocks
// workgroups, grid-striding beyond that. `per_thread` is accepted for call
// sites that state their unroll but does not shrink small grids.
__host__ __device__ __forceinline__ int grid_for(int64_t n, int per_thread = 1) {
  (void)per_thread;
  int64_t g = (n + (int64_t)kBlock - 1) / (int64_t)kBlock;
  if (g < 1) g = 1;
  if (g > kMaxReduceBlocks) g = kMaxReduceBlocks;
  return (int)g;
}

}  // namespace ohm

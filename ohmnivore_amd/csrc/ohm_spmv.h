// ohm_spmv.h — the CSR-stream SpMV workgroup body, shared by every kernel
// that fuses an epilogue (dot products, residuals) onto y = A x.
#pragma once

#include "ohm_reduce.h"

namespace ohm {

template <typename T>
__device__ __forceinline__ T fmac(T a, T b, T c);
template <>
__device__ __forceinline__ double fmac<double>(double a, double b, double c) {
  return fma(a, b, c);
}
template <>
__device__ __forceinline__ float fmac<float>(float a, float b, float c) {
  return fmaf(a, b, c);
}

// Row values y_r for the rows of this workgroup's block; calls f(r, y_r) for
// each row owned by this thread (at most one) — long rows resolve in thread 0.
// The operand is read through xf(j), so a kernel can form x_j on the fly from
// other vectors (a fused vector update) instead of reading a stored x.
template <typename T, typename X, typename F>
__device__ __forceinline__ void spmv_block_g(const uint32_t *__restrict__ rowptr,
                                             const uint32_t *__restrict__ col,
                                             const T *__restrict__ val,
                                             const uint32_t *__restrict__ blocks, X xf, F f) {
  __shared__ T prod[kSpmvNnzPerBlock];
  const uint32_t r0 = blocks[blockIdx.x], r1 = blocks[blockIdx.x + 1];
  const uint32_t k0 = rowptr[r0], k1 = rowptr[r1];
  const int tid = threadIdx.x;
  if (k1 - k0 > (uint32_t)kSpmvNnzPerBlock) {
    // four independent chains keep four gathers in flight per thread
    T a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
    uint32_t k = k0 + tid;
    for (; k + 3 * kBlock < k1; k += 4 * kBlock) {
      a0 = fmac<T>(val[k], xf(col[k]), a0);
      a1 = fmac<T>(val[k + kBlock], xf(col[k + kBlock]), a1);
      a2 = fmac<T>(val[k + 2 * kBlock], xf(col[k + 2 * kBlock]), a2);
      a3 = fmac<T>(val[k + 3 * kBlock], xf(col[k + 3 * kBlock]), a3);
    }
    for (; k < k1; k += kBlock) a0 = fmac<T>(val[k], xf(col[k]), a0);
    double v[1] = {(double)((a0 + a1) + (a2 + a3))};
    block_sum<1>(v);
    if (tid == 0) f(r0, (T)v[0]);
  } else {
    const uint32_t cnt = k1 - k0;
    for (uint32_t k = tid; k < cnt; k += kBlock) prod[k] = val[k0 + k] * xf(col[k0 + k]);
    __syncthreads();
    const uint32_t r = r0 + tid;
    if (r < r1) {
      const uint32_t a = rowptr[r] - k0, e = rowptr[r + 1] - k0;
      T s = T(0);
      for (uint32_t k = a; k < e; k++) s += prod[k];
      f(r, s);
    }
  }
}

template <typename T, typename F>
__device__ __forceinline__ void spmv_block(const uint32_t *__restrict__ rowptr,
                                           const uint32_t *__restrict__ col,
                                           const T *__restrict__ val,
                                           const uint32_t *__restrict__ blocks,
                                           const T *__restrict__ x, F f) {
  spmv_block_g<T>(rowptr, col, val, blocks, [&](uint32_t j) { return x[j]; }, f);
}

__device__ __forceinline__ bool solve_stopped(const SolveState *st, int skip) {
  if (!st || !skip) return false;
  return st->status != ST_RUNNING || (skip == 2 && st->s_conv);
}

}  // namespace ohm

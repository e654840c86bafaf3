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

// Operand values gathered at one column: NV raw vector entries that a
// kernel combines into x_j once its head has produced the scalars.
template <typename T, int NV>
struct Opnd {
  T v[NV];
};

// Row values y_r for the rows of this workgroup's block; calls f(r, y_r) for
// each row owned by this thread (at most one) — long rows resolve in thread
// 0. The operand at column j is gather(j) (NV raw loads) combined by
// comb(); head() runs once every matrix entry, column index, row pointer and
// operand load of this thread has been issued, so its own loads and
// workgroup synchronisation overlap them. head() returning false exits the
// workgroup (every thread gets the same answer); it may synchronise.
// Returns whether the block ran.
template <typename T, int NV, int PER, typename G, typename H, typename C, typename F>
__device__ __forceinline__ bool spmv_block_h(const uint32_t *__restrict__ rowptr,
                                             const uint32_t *__restrict__ col,
                                             const T *__restrict__ val,
                                             const uint32_t *__restrict__ blocks, G gather,
                                             H head, C comb, F f) {
  constexpr int kCap = PER * kBlock;  // entries a block stages (its matrix's nnz cap)
  __shared__ T prod[kCap];
  const uint32_t r0 = blocks[2 * blockIdx.x], k0 = blocks[2 * blockIdx.x + 1];
  const uint32_t r1 = blocks[2 * blockIdx.x + 2], k1 = blocks[2 * blockIdx.x + 3];
  const int tid = threadIdx.x;
  const uint32_t cnt = k1 - k0;
  if (cnt > (uint32_t)kCap) {
    if (!head()) return false;
    // four independent chains keep four gathers in flight per thread
    T a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
    uint32_t k = k0 + tid;
    for (; k + 3 * kBlock < k1; k += 4 * kBlock) {
      a0 = fmac<T>(val[k], comb(gather(col[k])), a0);
      a1 = fmac<T>(val[k + kBlock], comb(gather(col[k + kBlock])), a1);
      a2 = fmac<T>(val[k + 2 * kBlock], comb(gather(col[k + 2 * kBlock])), a2);
      a3 = fmac<T>(val[k + 3 * kBlock], comb(gather(col[k + 3 * kBlock])), a3);
    }
    for (; k < k1; k += kBlock) a0 = fmac<T>(val[k], comb(gather(col[k])), a0);
    double v[1] = {(double)((a0 + a1) + (a2 + a3))};
    block_sum<1>(v);
    if (tid == 0) f(r0, (T)v[0]);
    return true;
  }
  // Clamped indices keep every load valid and unconditional (so they issue
  // back to back); entries past cnt are discarded below. An empty block
  // (cnt == 0: rows without entries) loads nothing.
  const uint32_t last = cnt ? cnt - 1 : 0;
  T av[PER];
  Opnd<T, NV> gv[PER];
  if (cnt) {
#pragma unroll
    for (int j = 0; j < PER; j++) av[j] = val[k0 + min((uint32_t)(tid + j * kBlock), last)];
    uint32_t cj[PER];
#pragma unroll
    for (int j = 0; j < PER; j++) cj[j] = col[k0 + min((uint32_t)(tid + j * kBlock), last)];
#pragma unroll
    for (int j = 0; j < PER; j++) gv[j] = gather(cj[j]);
  }
  const uint32_t r = r0 + tid;
  uint32_t ra = 0, re = 0;
  if (r < r1) {
    ra = rowptr[r] - k0;
    re = rowptr[r + 1] - k0;
  }
  if (!head()) return false;
  if (cnt) {
#pragma unroll
    for (int j = 0; j < PER; j++) {
      const uint32_t k = tid + j * kBlock;
      if (k < cnt) prod[k] = av[j] * comb(gv[j]);
    }
  }
  __syncthreads();
  if (r < r1) {
    T s = T(0);
    for (uint32_t k = ra; k < re; k++) s += prod[k];
    f(r, s);
  }
  return true;
}

// y = A x body without a head.
template <typename T, int PER, typename F>
__device__ __forceinline__ void spmv_block(const uint32_t *__restrict__ rowptr,
                                           const uint32_t *__restrict__ col,
                                           const T *__restrict__ val,
                                           const uint32_t *__restrict__ blocks,
                                           const T *__restrict__ x, F f) {
  spmv_block_h<T, 1, PER>(
      rowptr, col, val, blocks,
      [&](uint32_t j) {
        Opnd<T, 1> o;
        o.v[0] = x[j];
        return o;
      },
      [] { return true; }, [](const Opnd<T, 1> &o) { return o.v[0]; }, f);
}

__device__ __forceinline__ bool solve_stopped(const SolveState *st, int skip) {
  if (!st || !skip) return false;
  return st->status != ST_RUNNING || (skip == 2 && st->s_conv);
}

}  // namespace ohm

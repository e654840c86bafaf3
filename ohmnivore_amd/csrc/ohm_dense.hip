// ohm_dense.hip — dense LU with partial pivoting on the GPU for small
// systems: the direct solver behind a Newton run whose iterative sub-solve
// cannot converge (the reference falls back to a CPU sparse LU there,
// src/analysis/dc.rs:112-131; NewtonLinearMode::CpuSparseLu,
// src/solver/newton.rs:22-30,286-289).
//
// Layout: the CSR matrix is scattered into a row-major n x n FP64 array in
// HBM (134 MB at the 4096-row limit, kept on the backend for later solves).
// Elimination step k is two launches: a one-workgroup pivot search + row
// swap (matrix and right-hand side), then
// a trailing update over rows k+1..n-1 and columns k+1..n-1 (plus the
// right-hand side as column n) with the multiplier a_ik / a_kk recomputed
// per element, so column k is only read during the update and no multiplier
// array is needed. Back substitution runs in one workgroup.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "ohm_internal.h"
#include "ohm_reduce.h"

namespace ohm {

namespace {

constexpr int kDenseWg = 1024;

__global__ void k_dense_zero(int64_t m, double *__restrict__ d) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * kBlock)
    d[i] = 0.0;
}

__global__ void k_dense_scatter(int n, const uint32_t *__restrict__ rowptr,
                                const uint32_t *__restrict__ col, const double *__restrict__ val,
                                const double *__restrict__ rhs, double *__restrict__ d,
                                double *__restrict__ y) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= n) return;
  for (uint32_t k = rowptr[r]; k < rowptr[r + 1]; k++) d[(int64_t)r * n + col[k]] += val[k];
  y[r] = rhs[r];
}

// argmax_{i >= k} |d[i][k]| (lowest index on ties), then swap rows k and p
// of d and y; flag[0] = 1 if the pivot is zero
__global__ __launch_bounds__(kDenseWg) void k_dense_pivot(int n, int k, double *__restrict__ d,
                                                          double *__restrict__ y,
                                                          int *__restrict__ flag) {
  __shared__ double sv[kDenseWg / 64];
  __shared__ int si[kDenseWg / 64];
  __shared__ int s_p;
  double best = -1.0;
  int bi = n;
  for (int i = k + threadIdx.x; i < n; i += kDenseWg) {
    const double a = fabs(d[(int64_t)i * n + k]);
    if (a > best) {
      best = a;
      bi = i;
    }
  }
  // wave64 reduction of (value, index)
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(best, off, 64);
    const int oi = __shfl_xor(bi, off, 64);
    if (ov > best || (ov == best && oi < bi)) {
      best = ov;
      bi = oi;
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    sv[w] = best;
    si[w] = bi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double bv = sv[0];
    int bx = si[0];
    for (int q = 1; q < kDenseWg / 64; q++)
      if (sv[q] > bv || (sv[q] == bv && si[q] < bx)) {
        bv = sv[q];
        bx = si[q];
      }
    if (!(bv > 1e-300) || bx >= n) {
      flag[0] = 1;
      bx = k;
    }
    s_p = bx;
  }
  __syncthreads();
  const int p = s_p;
  if (p != k) {
    for (int j = threadIdx.x; j < n; j += kDenseWg) {
      const double t = d[(int64_t)k * n + j];
      d[(int64_t)k * n + j] = d[(int64_t)p * n + j];
      d[(int64_t)p * n + j] = t;
    }
    if (threadIdx.x == 0) {
      const double t = y[k];
      y[k] = y[p];
      y[p] = t;
    }
  }
}

// rows i in (k, n), columns j in (k, n] (j == n: the right-hand side)
__global__ void k_dense_update(int n, int k, double *__restrict__ d, double *__restrict__ y) {
  const int m = n - k - 1;  // rows and matrix columns below / right of the pivot
  const int64_t total = (int64_t)m * (m + 1);
  const double piv = d[(int64_t)k * n + k];
  for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * kBlock) {
    const int i = k + 1 + (int)(e / (m + 1));
    const int jj = (int)(e % (m + 1));
    const double l = d[(int64_t)i * n + k] / piv;
    if (jj == m)
      y[i] -= l * y[k];
    else {
      const int j = k + 1 + jj;
      d[(int64_t)i * n + j] -= l * d[(int64_t)k * n + j];
    }
  }
}

__global__ __launch_bounds__(kDenseWg) void k_dense_backsub(int n, const double *__restrict__ d,
                                                            double *__restrict__ y,
                                                            double *__restrict__ x) {
  __shared__ double xk;
  for (int k = n - 1; k >= 0; k--) {
    if (threadIdx.x == 0) {
      xk = y[k] / d[(int64_t)k * n + k];
      x[k] = xk;
    }
    __syncthreads();
    const double v = xk;
    for (int i = threadIdx.x; i < k; i += kDenseWg) y[i] -= d[(int64_t)i * n + k] * v;
    __syncthreads();
  }
}

}  // namespace

ohm_status dense_lu_solve(ohm_backend *b, const DevCsr &a, const double *rhs, double *x) {
  const int n = a.n;
  if (n <= 0) return OHM_OK;
  if (n > kDenseMaxRows)
    return fail(OHM_ERR_INVALID, "dense LU is limited to %d unknowns (system has %d)",
                kDenseMaxRows, n);
  hipStream_t st = b->stream;
  // workspace kept on the backend between solves (a Newton run in this mode
  // solves once per iteration): n*n matrix, n right-hand side, one flag
  const int64_t nn = (int64_t)n * n;
  const size_t need = sizeof(double) * (size_t)(nn + n) + 256;
  if (b->dense_ws_bytes < need) {
    if (b->dense_ws) OHM_HIP(hipFree(b->dense_ws));
    b->dense_ws = nullptr;
    b->dense_ws_bytes = 0;
    OHM_HIP(hipMalloc(&b->dense_ws, need));
    b->dense_ws_bytes = need;
  }
  double *d = reinterpret_cast<double *>(b->dense_ws), *y = d + nn;
  int *flag = reinterpret_cast<int *>(y + n);
  OHM_HIP(hipMemsetAsync(flag, 0, sizeof(int), st));
  hipLaunchKernelGGL(k_dense_zero, dim3((unsigned)std::min<int64_t>((nn + kBlock - 1) / kBlock, 8192)),
                     dim3(kBlock), 0, st, nn, d);
  hipLaunchKernelGGL(k_dense_scatter, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, n,
                     a.rowptr, a.col, a.val, rhs, d, y);
  for (int k = 0; k < n; k++) {
    hipLaunchKernelGGL(k_dense_pivot, dim3(1), dim3(kDenseWg), 0, st, n, k, d, y, flag);
    const int64_t m = n - k - 1;
    if (m > 0) {
      const int64_t total = m * (m + 1);
      hipLaunchKernelGGL(k_dense_update,
                         dim3((unsigned)std::min<int64_t>((total + kBlock - 1) / kBlock, 4096)),
                         dim3(kBlock), 0, st, n, k, d, y);
    }
  }
  hipLaunchKernelGGL(k_dense_backsub, dim3(1), dim3(kDenseWg), 0, st, n, d, y, x);
  b->dispatches += 2 * n + 3;
  OHM_HIP(hipGetLastError());
  int h = 0;
  OHM_HIP(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, st));
  OHM_HIP(hipStreamSynchronize(st));
  b->readbacks++;
  if (h) return fail(OHM_ERR_SOLVE, "Solve error: matrix is singular (zero pivot in dense LU)");
  return OHM_OK;
}

}  // namespace ohm

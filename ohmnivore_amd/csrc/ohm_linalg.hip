// ohm_linalg.hip — CSR SpMV, BLAS-1 and Jacobi kernels for gfx950.
//
// SpMV (the SolverBackend::spmv of src/solver/backend.rs:56 and the
// spmv_real/spmv_ds kernels of src/solver/gpu_shaders.rs:45,
// src/solver/ds_shaders.rs:144) is CSR-stream: the host splits the rows
// into blocks of at most 256 rows and 2048 nonzeros; a workgroup streams its
// block's values and column indices with fully coalesced loads, multiplies
// by the gathered x and stages the products in LDS, then each thread sums
// one row from LDS in column order. A row longer than the block budget gets
// a workgroup of its own that strides over it and reduces across lanes.
// Every launch is a bandwidth-bound pass: bytes = nnz*(V+4) + 4(n+1) + 2nV.
#include <algorithm>
#include <cstdlib>

#include "ohm_spmv.h"

namespace ohm {

// --------------------------------------------------------------- SpMV ----
void build_spmv_blocks(const uint32_t *rowptr, int n, std::vector<uint32_t> &blocks, int rows_cap,
                       int nnz_cap) {
  rows_cap = std::max(1, std::min(rows_cap, kSpmvRowsPerBlock));
  nnz_cap = std::max(1, std::min(nnz_cap, kSpmvNnzPerBlock));
  blocks.clear();
  blocks.push_back(0);
  int r = 0;
  while (r < n) {
    uint32_t start = rowptr[r];
    int rows = 0;
    if (rowptr[r + 1] - start > (uint32_t)nnz_cap) {
      r++;  // long row: own block
    } else {
      while (r < n && rows < rows_cap && rowptr[r + 1] - start <= (uint32_t)nnz_cap) {
        r++;
        rows++;
      }
    }
    blocks.push_back((uint32_t)r);
  }
}

// Entries per thread of a matrix's default row blocks (of at most 256
// rows): the mean entries per row over rows of at most 256 entries, rounded
// down, so that blocks fill their per * 256 staged entries a little before
// 256 rows instead of leaving slots empty. Fitting per to the blocks matters
// on large matrices: loads past a block's entries take issue slots and
// cache bandwidth like real ones (a 335 MB five-point matrix: 4.95 TB/s with
// 8 per thread, 5.83 with 4, 6.04 with 5).
int spmv_default_per(const uint32_t *rowptr, int n) {
  if (const char *e = getenv("OHM_SPMV_PER")) {  // fixed entries per thread (1-8)
    const int v = atoi(e);
    if (v >= 1 && v <= 8) return spmv_per_for(v * kBlock);
  }
  uint64_t entries = 0, rows = 0;
  for (int r = 0; r < n; r++) {
    const uint32_t c = rowptr[r + 1] - rowptr[r];
    if (c <= (uint32_t)kBlock) {
      entries += c;
      rows++;
    }
  }
  const int mean = rows ? (int)(entries / rows) : 1;
  return spmv_per_for(std::max(mean, 1) * kBlock);
}

// Device block table: (first row, first nonzero) of every block boundary, so
// a workgroup reads its row and nonzero ranges with one 16-byte scalar load.
std::vector<uint32_t> spmv_block_pairs(const uint32_t *rowptr, const std::vector<uint32_t> &bounds) {
  std::vector<uint32_t> t(2 * bounds.size());
  for (size_t i = 0; i < bounds.size(); i++) {
    t[2 * i] = bounds[i];
    t[2 * i + 1] = rowptr[bounds[i]];
  }
  return t;
}

template <typename T, int PER>
__global__ __launch_bounds__(kBlock) void k_spmv(const uint32_t *__restrict__ rowptr,
                                                 const uint32_t *__restrict__ col,
                                                 const T *__restrict__ val,
                                                 const uint32_t *__restrict__ blocks,
                                                 const T *__restrict__ x, T *__restrict__ y,
                                                 const SolveState *st, int skip) {
  if (solve_stopped(st, skip)) return;
  spmv_block<T, PER>(rowptr, col, val, blocks, x, [&](uint32_t r, T v) { y[r] = v; });
}

void launch_spmv(ohm_backend *b, const DevCsr &a, const double *x, double *y,
                 const SolveState *state, int skip) {
  if (a.n == 0) return;
  with_per(a.per, [&](auto P) {
    hipLaunchKernelGGL((k_spmv<double, decltype(P)::value>), dim3(a.n_blocks), dim3(kBlock), 0,
                       b->stream, a.rowptr, a.col, a.val, a.blocks, x, y, state, skip);
  });
  b->dispatches++;
}

void launch_spmv_f32(ohm_backend *b, const DevCsr &a, const float *x, float *y,
                     const SolveState *state, int skip) {
  if (a.n == 0) return;
  with_per(a.per, [&](auto P) {
    hipLaunchKernelGGL((k_spmv<float, decltype(P)::value>), dim3(a.n_blocks), dim3(kBlock), 0,
                       b->stream, a.rowptr, a.col, a.val32, a.blocks, x, y, state, skip);
  });
  b->dispatches++;
}

// y = A^T x: SpMV over the transpose pattern, values gathered through t_src.
__global__ __launch_bounds__(kBlock) void k_spmv_t(const uint32_t *__restrict__ rowptr,
                                                   const uint32_t *__restrict__ col,
                                                   const uint32_t *__restrict__ src,
                                                   const double *__restrict__ val,
                                                   const uint32_t *__restrict__ blocks,
                                                   const double *__restrict__ x,
                                                   double *__restrict__ y) {
  __shared__ double prod[kSpmvNnzPerBlock];
  const uint32_t r0 = blocks[2 * blockIdx.x], k0 = blocks[2 * blockIdx.x + 1];
  const uint32_t r1 = blocks[2 * blockIdx.x + 2], k1 = blocks[2 * blockIdx.x + 3];
  const int tid = threadIdx.x;
  if (k1 - k0 > (uint32_t)kSpmvNnzPerBlock) {
    double acc = 0.0;
    for (uint32_t k = k0 + tid; k < k1; k += kBlock) acc = fma(val[src[k]], x[col[k]], acc);
    double v[1] = {acc};
    block_sum<1>(v);
    if (tid == 0) y[r0] = v[0];
  } else {
    const uint32_t cnt = k1 - k0;
    for (uint32_t k = tid; k < cnt; k += kBlock) prod[k] = val[src[k0 + k]] * x[col[k0 + k]];
    __syncthreads();
    const uint32_t r = r0 + tid;
    if (r < r1) {
      const uint32_t a = rowptr[r] - k0, e = rowptr[r + 1] - k0;
      double s = 0.0;
      for (uint32_t k = a; k < e; k++) s += prod[k];
      y[r] = s;
    }
  }
}

void launch_spmv_t(ohm_backend *b, const DevCsr &a, const double *x, double *y) {
  if (a.n == 0) return;
  hipLaunchKernelGGL(k_spmv_t, dim3(a.t_n_blocks), dim3(kBlock), 0, b->stream, a.t_rowptr,
                     a.t_col, a.t_src, a.val, a.t_blocks, x, y);
  b->dispatches++;
}

// ------------------------------------------------------------- BLAS-1 ----
__global__ void k_axpy(int n, double alpha, const double *__restrict__ x, double *__restrict__ y) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    y[i] = fma(alpha, x[i], y[i]);
}
__global__ void k_scale(int n, double alpha, double *__restrict__ x) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) x[i] *= alpha;
}
__global__ void k_copy(int n, const double *__restrict__ x, double *__restrict__ y) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) y[i] = x[i];
}
__global__ void k_fill(int n, double v, double *__restrict__ x) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) x[i] = v;
}
__global__ void k_mul(int n, const double *__restrict__ d, const double *__restrict__ x,
                      double *__restrict__ y) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) y[i] = d[i] * x[i];
}
__global__ __launch_bounds__(kBlock) void k_dot(int n, const double *__restrict__ x,
                                                const double *__restrict__ y, double *partials,
                                                unsigned *counter, double *out) {
  Acc<double> a;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    a.add_prod(x[i], y[i]);
  double v[1] = {a.get()}, tot[1];
  if (grid_sum<1>(v, partials, counter, tot) && threadIdx.x == 0) *out = tot[0];
}

#define GRID(n) dim3(grid_for(n, 4)), dim3(kBlock), 0, b->stream
void launch_axpy(ohm_backend *b, int n, double alpha, const double *x, double *y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axpy, GRID(n), n, alpha, x, y);
  b->dispatches++;
}
void launch_scale(ohm_backend *b, int n, double alpha, double *x) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_scale, GRID(n), n, alpha, x);
  b->dispatches++;
}
void launch_copy(ohm_backend *b, int n, const double *x, double *y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_copy, GRID(n), n, x, y);
  b->dispatches++;
}
void launch_fill(ohm_backend *b, int n, double v, double *x) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill, GRID(n), n, v, x);
  b->dispatches++;
}
void launch_mul(ohm_backend *b, int n, const double *d, const double *x, double *y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mul, GRID(n), n, d, x, y);
  b->dispatches++;
}
void launch_dot(ohm_backend *b, int n, const double *x, const double *y, int slot) {
  hipLaunchKernelGGL(k_dot, GRID(n), n, x, y, b->d_partials, b->d_counter, b->d_scalars + slot);
  b->dispatches++;
}
#undef GRID

// ------------------------------------------------------------- Jacobi ----
// inv[r] = 1/a_rr when |a_rr| > 1e-30 (src/solver/gpu.rs:178-189). For a
// zero-diagonal row (V-source / inductor branch) inv[r] = 1/S_rr with
// S_rr = -sum_j a_rj a_jr / a_jj the diagonal of the Schur complement over
// the rows that do have a pivot; 1 when that sum is zero too.
__global__ void k_jacobi_setup(int n, const uint32_t *__restrict__ rowptr,
                               const uint32_t *__restrict__ col, const double *__restrict__ val,
                               const int32_t *__restrict__ diag_pos, double *inv, float *inv32) {
  for (int r = blockIdx.x * kBlock + threadIdx.x; r < n; r += gridDim.x * kBlock) {
    int dp = diag_pos[r];
    double d = dp >= 0 ? val[dp] : 0.0;
    double out;
    if (fabs(d) > 1e-30) {
      out = 1.0 / d;
    } else {
      double s = 0.0;
      for (uint32_t k = rowptr[r]; k < rowptr[r + 1]; k++) {
        uint32_t j = col[k];
        if ((int)j == r) continue;
        int dj = diag_pos[j];
        double ajj = dj >= 0 ? val[dj] : 0.0;
        if (fabs(ajj) <= 1e-30) continue;
        double ajr = 0.0;  // a_jr: search row j
        for (uint32_t q = rowptr[j]; q < rowptr[j + 1]; q++)
          if ((int)col[q] == r) {
            ajr = val[q];
            break;
          }
        s -= val[k] * ajr / ajj;
      }
      out = (fabs(s) > 1e-300) ? 1.0 / s : 1.0;
    }
    inv[r] = out;
    if (inv32) inv32[r] = (float)out;
  }
}

void launch_jacobi_setup(ohm_backend *b, const DevCsr &a, double *inv, float *inv32) {
  if (a.n == 0) return;
  hipLaunchKernelGGL(k_jacobi_setup, dim3(grid_for(a.n)), dim3(kBlock), 0, b->stream, a.n,
                     a.rowptr, a.col, a.val, a.diag_pos, inv, inv32);
  b->dispatches++;
}

// ---------------------------------------------------- matrix management --
__global__ void k_d2f(int64_t n, const double *__restrict__ x, float *__restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    y[i] = (float)x[i];
}

ohm_status csr_ensure_f32(ohm_backend *b, DevCsr &m) {
  if (!m.val32) OHM_HIP(hipMalloc(&m.val32, sizeof(float) * std::max<int64_t>(m.nnz, 1)));
  if (m.nnz > 0) {
    hipLaunchKernelGGL(k_d2f, dim3(grid_for(m.nnz, 4)), dim3(kBlock), 0, b->stream, m.nnz, m.val,
                       m.val32);
    b->dispatches++;
  }
  return OHM_OK;
}

ohm_status csr_upload(ohm_backend *b, DevCsr &m, const double *val, const uint32_t *col,
                      const uint32_t *rowptr, int n, int64_t nnz) {
  m.n = n;
  m.nnz = nnz;
  m.h_rowptr.assign(rowptr, rowptr + n + 1);
  m.h_col.assign(col, col + nnz);
  std::vector<uint32_t> bounds;
  m.per = spmv_default_per(rowptr, n);
  m.tuned_nnz = m.per * kBlock;
  build_spmv_blocks(rowptr, n, bounds, kSpmvRowsPerBlock, m.tuned_nnz);
  m.n_blocks = (int)bounds.size() - 1;
  const std::vector<uint32_t> blocks = spmv_block_pairs(rowptr, bounds);
  std::vector<int32_t> diag(n > 0 ? n : 1, -1);
  for (int r = 0; r < n; r++)
    for (uint32_t k = rowptr[r]; k < rowptr[r + 1]; k++)
      if ((int)col[k] == r) {
        diag[r] = (int32_t)k;
        break;
      }
  size_t nz = (size_t)std::max<int64_t>(nnz, 1);
  OHM_HIP(hipMalloc(&m.rowptr, sizeof(uint32_t) * (n + 1)));
  OHM_HIP(hipMalloc(&m.col, sizeof(uint32_t) * nz));
  OHM_HIP(hipMalloc(&m.val, sizeof(double) * nz));
  OHM_HIP(hipMalloc(&m.blocks, sizeof(uint32_t) * blocks.size()));
  OHM_HIP(hipMalloc(&m.diag_pos, sizeof(int32_t) * diag.size()));
  OHM_HIP(hipMemcpyAsync(m.rowptr, rowptr, sizeof(uint32_t) * (n + 1), hipMemcpyHostToDevice,
                         b->stream));
  if (nnz > 0) {
    OHM_HIP(hipMemcpyAsync(m.col, col, sizeof(uint32_t) * nnz, hipMemcpyHostToDevice, b->stream));
    if (val)
      OHM_HIP(hipMemcpyAsync(m.val, val, sizeof(double) * nnz, hipMemcpyHostToDevice, b->stream));
  }
  OHM_HIP(hipMemcpyAsync(m.blocks, blocks.data(), sizeof(uint32_t) * blocks.size(),
                         hipMemcpyHostToDevice, b->stream));
  OHM_HIP(hipMemcpyAsync(m.diag_pos, diag.data(), sizeof(int32_t) * diag.size(),
                         hipMemcpyHostToDevice, b->stream));
  OHM_HIP(hipStreamSynchronize(b->stream));
  return OHM_OK;
}

ohm_status csr_ensure_transpose(ohm_backend *b, DevCsr &m) {
  if (m.t_rowptr) return OHM_OK;
  int n = m.n;
  std::vector<uint32_t> trp(n + 1, 0), tcol(std::max<int64_t>(m.nnz, 1)), tsrc(std::max<int64_t>(m.nnz, 1));
  for (int64_t k = 0; k < m.nnz; k++) trp[m.h_col[k] + 1]++;
  for (int i = 0; i < n; i++) trp[i + 1] += trp[i];
  std::vector<uint32_t> pos(trp.begin(), trp.end() - 1);
  for (int r = 0; r < n; r++)
    for (uint32_t k = m.h_rowptr[r]; k < m.h_rowptr[r + 1]; k++) {
      uint32_t c = m.h_col[k], p = pos[c]++;
      tcol[p] = (uint32_t)r;
      tsrc[p] = (uint32_t)k;
    }
  std::vector<uint32_t> bounds;
  build_spmv_blocks(trp.data(), n, bounds);
  m.t_n_blocks = (int)bounds.size() - 1;
  const std::vector<uint32_t> blocks = spmv_block_pairs(trp.data(), bounds);
  OHM_HIP(hipMalloc(&m.t_rowptr, sizeof(uint32_t) * (n + 1)));
  OHM_HIP(hipMalloc(&m.t_col, sizeof(uint32_t) * tcol.size()));
  OHM_HIP(hipMalloc(&m.t_src, sizeof(uint32_t) * tsrc.size()));
  OHM_HIP(hipMalloc(&m.t_blocks, sizeof(uint32_t) * blocks.size()));
  OHM_HIP(hipMemcpy(m.t_rowptr, trp.data(), sizeof(uint32_t) * (n + 1), hipMemcpyHostToDevice));
  OHM_HIP(hipMemcpy(m.t_col, tcol.data(), sizeof(uint32_t) * tcol.size(), hipMemcpyHostToDevice));
  OHM_HIP(hipMemcpy(m.t_src, tsrc.data(), sizeof(uint32_t) * tsrc.size(), hipMemcpyHostToDevice));
  OHM_HIP(hipMemcpy(m.t_blocks, blocks.data(), sizeof(uint32_t) * blocks.size(),
                    hipMemcpyHostToDevice));
  return OHM_OK;
}

__global__ void k_gather_vals(int64_t nnz, const uint32_t *__restrict__ src,
                              const double *__restrict__ val, double *__restrict__ out) {
  for (int64_t k = blockIdx.x * (int64_t)kBlock + threadIdx.x; k < nnz;
       k += (int64_t)gridDim.x * kBlock)
    out[k] = val[src[k]];
}

ohm_status csr_transpose_view(ohm_backend *b, DevCsr &m, DevCsr &view) {
  OHM_TRY(csr_ensure_transpose(b, m));
  if (!m.t_val) OHM_HIP(hipMalloc(&m.t_val, sizeof(double) * std::max<int64_t>(m.nnz, 1)));
  if (m.nnz > 0) {
    hipLaunchKernelGGL(k_gather_vals, dim3(grid_for(m.nnz)), dim3(kBlock), 0, b->stream, m.nnz,
                       m.t_src, m.val, m.t_val);
    b->dispatches++;
  }
  view = DevCsr();
  view.n = m.n;
  view.nnz = m.nnz;
  view.rowptr = m.t_rowptr;
  view.col = m.t_col;
  view.val = m.t_val;
  view.blocks = m.t_blocks;
  view.n_blocks = m.t_n_blocks;
  return OHM_OK;
}

// Row-block autotuner: every candidate shape (rows per workgroup x LDS
// products per workgroup) is a valid block table for the same kernel, so
// the choice only changes speed. Each candidate is timed with HIP events over
// back-to-back launches on the backend stream and the fastest kept.
ohm_status csr_autotune(ohm_backend *b, DevCsr &m, int *rows_out, int *nnz_out, bool fused) {
  if (m.n < 4096 || m.h_rowptr.empty()) {
    if (rows_out) *rows_out = m.tuned_rows;
    if (nnz_out) *nnz_out = m.tuned_nnz;
    return OHM_OK;
  }
  // a pattern tuned before keeps its shape (no timing)
  uint64_t h = 1469598103934665603ull;  // FNV-1a over the row pointers
  for (uint32_t v : m.h_rowptr) h = (h ^ v) * 1099511628211ull;
  const std::vector<uint64_t> key = {(uint64_t)m.n, (uint64_t)m.nnz, h, (uint64_t)fused};
  for (auto &e : b->tune_cache)
    if (e.first == key) {
      std::vector<uint32_t> bounds;
      build_spmv_blocks(m.h_rowptr.data(), m.n, bounds, e.second.first, e.second.second);
      const std::vector<uint32_t> blocks = spmv_block_pairs(m.h_rowptr.data(), bounds);
      uint32_t *nb = nullptr;
      OHM_HIP(hipMalloc(&nb, sizeof(uint32_t) * blocks.size()));
      OHM_HIP(hipMemcpy(nb, blocks.data(), sizeof(uint32_t) * blocks.size(), hipMemcpyHostToDevice));
      if (m.blocks) OHM_HIP(hipFree(m.blocks));
      m.blocks = nb;
      m.n_blocks = (int)bounds.size() - 1;
      m.tuned_rows = e.second.first;
      m.tuned_nnz = e.second.second;
      m.per = spmv_per_for(m.tuned_nnz);
      if (rows_out) *rows_out = m.tuned_rows;
      if (nnz_out) *nnz_out = m.tuned_nnz;
      return OHM_OK;
    }
  // Matrices applied by the fused BiCGSTAB kernels keep 256-row blocks
  // (their block count sets how many partial sums each consumer's head
  // re-sums); others also try smaller blocks. OHM_TUNE_ALL_ROWS=1 tries
  // every row count for all.
  static const int kRowsAll[] = {32, 64, 128, 256}, kRowsFused[] = {256};
  const char *all_env = getenv("OHM_TUNE_ALL_ROWS");
  const bool all_rows = !fused || (all_env && atoi(all_env) != 0);
  const int *kRows = all_rows ? kRowsAll : kRowsFused;
  const int n_rows = all_rows ? 4 : 1;
  static const int kPer[] = {1, 2, 3, 4, 5, 6, 8};
  double *x = nullptr, *y = nullptr;
  OHM_HIP(hipMalloc(&x, sizeof(double) * m.n));
  OHM_HIP(hipMalloc(&y, sizeof(double) * m.n));
  OHM_HIP(hipMemsetAsync(x, 0, sizeof(double) * m.n, b->stream));
  std::vector<uint32_t> blocks;
  float best = 1e30f;
  int best_r = m.tuned_rows, best_z = m.tuned_nnz;
  std::vector<uint32_t> best_blocks;
  uint32_t *d_blk = nullptr;
  size_t d_cap = 0;
  ohm_status rc = OHM_OK;
  for (int ri = 0; ri < n_rows; ri++)
    for (int per : kPer) {
      const int rows = kRows[ri];
      const int nz = per * kBlock;
      build_spmv_blocks(m.h_rowptr.data(), m.n, blocks, rows, nz);
      blocks = spmv_block_pairs(m.h_rowptr.data(), blocks);
      if (blocks.size() > d_cap) {
        if (d_blk) (void)hipFree(d_blk);
        d_cap = blocks.size();
        if (hipMalloc(&d_blk, sizeof(uint32_t) * d_cap) != hipSuccess) {
          rc = fail(OHM_ERR_SOLVE, "Solve error: out of device memory in the autotuner");
          goto done;
        }
      }
      (void)hipMemcpyAsync(d_blk, blocks.data(), sizeof(uint32_t) * blocks.size(),
                           hipMemcpyHostToDevice, b->stream);
      {
        DevCsr v = m;  // shallow view with the candidate table
        v.blocks = d_blk;
        v.n_blocks = (int)blocks.size() / 2 - 1;
        v.tuned_nnz = nz;
        v.per = per;
        v.h_rowptr.clear();
        v.h_col.clear();
        for (int w = 0; w < 3; w++) launch_spmv(b, v, x, y, nullptr, 0);
        (void)hipEventRecord(b->ev0, b->stream);
        for (int k = 0; k < 20; k++) launch_spmv(b, v, x, y, nullptr, 0);
        (void)hipEventRecord(b->ev1, b->stream);
        (void)hipEventSynchronize(b->ev1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
        if (ms < best) {
          best = ms;
          best_r = rows;
          best_z = nz;
          best_blocks = blocks;
        }
      }
    }
  if (!best_blocks.empty()) {
    uint32_t *nb = nullptr;
    if (hipMalloc(&nb, sizeof(uint32_t) * best_blocks.size()) != hipSuccess) {
      rc = fail(OHM_ERR_SOLVE, "Solve error: out of device memory in the autotuner");
      goto done;
    }
    (void)hipMemcpyAsync(nb, best_blocks.data(), sizeof(uint32_t) * best_blocks.size(),
                         hipMemcpyHostToDevice, b->stream);
    (void)hipStreamSynchronize(b->stream);
    if (m.blocks) (void)hipFree(m.blocks);
    m.blocks = nb;
    m.n_blocks = (int)best_blocks.size() / 2 - 1;
    m.tuned_rows = best_r;
    m.tuned_nnz = best_z;
    m.per = spmv_per_for(best_z);
    graph_cache_clear(b);
    if (b->tune_cache.size() >= 64) b->tune_cache.erase(b->tune_cache.begin());
    b->tune_cache.push_back({key, {best_r, best_z}});
  }
done:
  (void)hipStreamSynchronize(b->stream);
  if (d_blk) (void)hipFree(d_blk);
  (void)hipFree(x);
  (void)hipFree(y);
  b->dispatches += 12 * 23;
  if (rows_out) *rows_out = m.tuned_rows;
  if (nnz_out) *nnz_out = m.tuned_nnz;
  return rc;
}

void csr_free(DevCsr &m) {
  void *ptrs[] = {m.rowptr, m.col,    m.val,   m.val32, m.blocks,  m.diag_pos,
                  m.t_rowptr, m.t_col, m.t_src, m.t_blocks, m.t_val};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  m = DevCsr();
}

// ---------------------------------------------------- HBM stream probes --
// Streaming read (32-byte loads, four in flight per thread over a grid of
// 8,192 workgroups) and copy over a buffer larger than the Infinity Cache:
// the practical HBM ceilings that the SpMV's bandwidth is read against.
constexpr int kStreamGrid = 8192;

__global__ __launch_bounds__(kBlock) void k_stream_read(int64_t n4, const double4 *__restrict__ x,
                                                       double *__restrict__ sink) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  double s = 0.0;
  int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  for (; i + 3 * stride < n4; i += 4 * stride) {
    double4 v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) v[u] = x[i + u * stride];
#pragma unroll
    for (int u = 0; u < 4; u++) s += (v[u].x + v[u].y) + (v[u].z + v[u].w);
  }
  for (; i < n4; i += stride) {
    const double4 v = x[i];
    s += (v.x + v.y) + (v.z + v.w);
  }
  if (s == -1.0) sink[0] = s;  // data-dependent store keeps the loads
}

__global__ __launch_bounds__(kBlock) void k_stream_copy(int64_t n2, const double2 *__restrict__ x,
                                                       double2 *__restrict__ y) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n2; i += stride) y[i] = x[i];
}

ohm_status stream_timed(ohm_backend *b, int64_t bytes, int mode, int reps, double &avg_ms) {
  const int64_t n2 = std::max<int64_t>(bytes / 32 * 2, 2);  // whole 32-byte groups
  double2 *x = nullptr, *y = nullptr;
  OHM_HIP(hipMalloc(&x, sizeof(double2) * n2));
  if (hipMalloc(&y, mode == 1 ? sizeof(double2) * n2 : sizeof(double2)) != hipSuccess) {
    (void)hipFree(x);
    return fail(OHM_ERR_SOLVE, "Solve error: out of device memory for the stream probe");
  }
  const dim3 blk(kBlock);
  auto one = [&]() {
    if (mode == 1)
      hipLaunchKernelGGL(k_stream_copy, dim3(32768), blk, 0, b->stream, n2, x, y);
    else
      hipLaunchKernelGGL(k_stream_read, dim3(kStreamGrid), blk, 0, b->stream, n2 / 2,
                         reinterpret_cast<const double4 *>(x), (double *)y);
  };
  ohm_status rc = OHM_OK;
  if (hipMemsetAsync(x, 0, sizeof(double2) * n2, b->stream) != hipSuccess) rc = fail(OHM_ERR_SOLVE, "memset");
  if (rc == OHM_OK) {
    for (int w = 0; w < 2; w++) one();
    (void)hipEventRecord(b->ev0, b->stream);
    for (int k = 0; k < reps; k++) one();
    (void)hipEventRecord(b->ev1, b->stream);
    (void)hipEventSynchronize(b->ev1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
    avg_ms = (double)ms / reps;
    if (hipGetLastError() != hipSuccess) rc = fail(OHM_ERR_SOLVE, "Solve error: stream probe failed");
  }
  (void)hipFree(x);
  (void)hipFree(y);
  b->dispatches += reps + 2;
  return rc;
}

ohm_status ensure_partials(ohm_backend *b, size_t count) {
  if (count <= b->partials_cap) return OHM_OK;
  if (b->d_partials) OHM_HIP(hipFree(b->d_partials));
  OHM_HIP(hipMalloc(&b->d_partials, sizeof(double) * count));
  b->partials_cap = count;
  return OHM_OK;
}

}  // namespace ohm

// ohm_internal.h — shared internals of the MI355X solve backend.
//
// Layout in HBM: CSR with u32 row pointers / column indices (the reference's
// GPU upload format, src/solver/gpu.rs:124-125), f64 values plus an optional
// f32 mirror for the mixed-precision inner solver, and flat f64 / f32
// vectors. Every solve keeps its scalars (rho, alpha, omega, status) in a
// device-resident SolveState so no iteration needs a host round trip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/ohm_solver.h"

namespace ohm {

// ---------------------------------------------------------------- errors --
void set_error(const std::string &msg);
ohm_status fail(ohm_status code, const char *fmt, ...);
ohm_status hip_fail(hipError_t e, const char *what, const char *file, int line);

#define OHM_HIP(call)                                                    \
  do {                                                                   \
    hipError_t _e = (call);                                              \
    if (_e != hipSuccess) return ohm::hip_fail(_e, #call, __FILE__, __LINE__); \
  } while (0)

#define OHM_TRY(call)                        \
  do {                                       \
    ohm_status _s = (call);                  \
    if (_s != OHM_OK) return _s;             \
  } while (0)

// ------------------------------------------------------ launch constants --
constexpr int kBlock = 256;              // threads per workgroup (4 wave64)
constexpr int kWaves = kBlock / 64;
constexpr int kSpmvNnzPerBlock = 2048;   // LDS-staged products per row block
constexpr int kSpmvRowsPerBlock = kBlock;
constexpr int kMaxReduceBlocks = 2048;   // cap on grid size of reducing vector kernels

// LDS-staged SpMV: each thread loads `per` matrix entries of its row block
// (the block holds at most per * kBlock entries), so per is chosen to fit
// the blocks of a matrix; loads past a block's entries cost as much as the
// entries themselves. Supported values and the launch dispatcher:
inline int spmv_per_for(int nnz_cap) {
  const int need = (std::max(nnz_cap, 1) + kBlock - 1) / kBlock;
  return need <= 6 ? need : 8;
}
template <typename F>
inline void with_per(int per, F &&f) {
  switch (per) {
    case 1: f(std::integral_constant<int, 1>()); break;
    case 2: f(std::integral_constant<int, 2>()); break;
    case 3: f(std::integral_constant<int, 3>()); break;
    case 4: f(std::integral_constant<int, 4>()); break;
    case 5: f(std::integral_constant<int, 5>()); break;
    case 6: f(std::integral_constant<int, 6>()); break;
    default: f(std::integral_constant<int, 8>()); break;
  }
}
constexpr int64_t kDeferMaxReads = int64_t(1) << 21;  // consumer x producer workgroups re-summed

// ----------------------------------------------------------- solve state --
enum SolveStatus : int {
  ST_RUNNING = 0,
  ST_CONVERGED = 1,
  ST_BRK_RHO = 2,
  ST_BRK_RV = 3,
  ST_BRK_TT = 4,
  ST_NAN = 5,
  ST_BRK_OMEGA = 6,
  ST_MAXITER = 7,
  ST_SKIPPED = 8,  // a captured solve masked off before it started
};

struct SolveState {
  double rho, alpha, omega, beta;
  double rho_new;       // r_hat.r of the current iteration (rho once alpha is formed)
  double abs_tol;       // tolerance * ||b||
  double bnorm2;        // ||b||^2
  double rr;            // last ||r||^2
  double ss;            // last ||s||^2
  double rv, ts, tt;
  double rhrh;          // ||r_hat||^2 (scaled breakdown tests)
  int status;
  int s_conv;           // ||s|| test passed: x += alpha*p_hat then stop
  int iters;
  int max_iters;
  unsigned counter;     // last-block-done ticket
  int scaled;           // OHM_BREAKDOWN_SCALED tests
  double r0;            // ||b - A x0|| of the solve (read with the status word)
  double res2;          // squared residual norm at the last completed iteration
  double eta;           // forcing term chosen on the device (inexact Newton)
};

// ------------------------------------------------------ device objects --
struct DevCsr {
  int n = 0;
  int64_t nnz = 0;
  uint32_t *rowptr = nullptr, *col = nullptr;
  double *val = nullptr;
  float *val32 = nullptr;          // f32 mirror (mixed precision)
  uint32_t *blocks = nullptr;      // SpMV block table: (first row, first nonzero) pairs, n_blocks+1
  int n_blocks = 0;
  int32_t *diag_pos = nullptr;     // CSR position of (r, r) or -1
  // transpose pattern (built lazily for A^T x): CSR of A^T + map to A's slots
  uint32_t *t_rowptr = nullptr, *t_col = nullptr, *t_src = nullptr, *t_blocks = nullptr;
  int t_n_blocks = 0;
  double *t_val = nullptr;         // values of A^T in its CSR order (gathered via t_src)
  int tuned_rows = kSpmvRowsPerBlock, tuned_nnz = kSpmvNnzPerBlock;  // row-block shape
  int per = kSpmvNnzPerBlock / kBlock;  // entries per thread: tuned_nnz / kBlock
  std::vector<uint32_t> h_rowptr, h_col; // host copy of the pattern
};

}  // namespace ohm

namespace ohm {
// A captured chunk of solver iterations, replayable while every pointer and
// launch size in `key` is unchanged.
struct GraphEntry {
  std::vector<uintptr_t> key;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};
}  // namespace ohm

struct ohm_backend {
  std::vector<ohm::GraphEntry> graphs;
  bool use_graphs = true;
  bool autotune = true;              // tune row blocks of solver-internal matrices
  // tuned row-block shapes by pattern (n, nnz, row-pointer hash)
  std::vector<std::pair<std::vector<uint64_t>, std::pair<int, int>>> tune_cache;
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cu = 0;
  double *d_partials = nullptr;      // per-block partial sums for reductions
  size_t partials_cap = 0;
  unsigned *d_counter = nullptr;     // ticket for one-shot reductions
  double *d_scalars = nullptr;       // device results of one-shot dots
  double *h_pinned = nullptr;        // pinned host staging for scalar readback
  void *ac_ws = nullptr;             // banded AC solver's chunk workspace (grow-only)
  size_t ac_ws_bytes = 0;
  uint64_t ac_ws_gen = 0;            // bumped by every chunk written into ac_ws
  void *dense_ws = nullptr;          // dense LU: n*n matrix + rhs + flag (grow-only)
  size_t dense_ws_bytes = 0;
  void *scratch = nullptr;           // this device's address of the allocation-free
                                     // solve's static scratch array (never freed)
  int64_t dispatches = 0, readbacks = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

struct ohm_buffer {
  ohm_backend *owner = nullptr;
  size_t n = 0;
  double *d = nullptr;
};

struct ohm_matrix {
  ohm_backend *owner = nullptr;
  ohm::DevCsr csr;
};

namespace ohm {

// ------------------------------------------------------ kernel launchers --
// All launches go on the backend's stream; `state` (may be null) lets solver
// kernels skip work once a solve has finished.
// Row blocks for the CSR-stream SpMV: at most rows_cap rows (<= 256, one
// thread per row) and nnz_cap LDS-staged products (<= 2048) per workgroup; a
// longer row gets a workgroup of its own.
void build_spmv_blocks(const uint32_t *rowptr, int n, std::vector<uint32_t> &blocks,
                       int rows_cap = kSpmvRowsPerBlock, int nnz_cap = kSpmvNnzPerBlock);
int spmv_default_per(const uint32_t *rowptr, int n);
// Pick the fastest (rows_cap, nnz_cap) for m by timing each candidate.
// fused: m is applied by the fused BiCGSTAB kernels (keep 256-row blocks)
ohm_status csr_autotune(ohm_backend *b, DevCsr &m, int *rows_out, int *nnz_out,
                        bool fused = false);
ohm_status stream_timed(ohm_backend *b, int64_t bytes, int mode, int reps, double &avg_ms);
ohm_status csr_upload(ohm_backend *b, DevCsr &m, const double *val, const uint32_t *col,
                      const uint32_t *rowptr, int n, int64_t nnz);
void csr_free(DevCsr &m);
ohm_status csr_ensure_f32(ohm_backend *b, DevCsr &m);
ohm_status csr_ensure_transpose(ohm_backend *b, DevCsr &m);
// Refresh t_val from the current values and return A^T as a DevCsr view
// (shares the pattern arrays of m; never freed on its own).
ohm_status csr_transpose_view(ohm_backend *b, DevCsr &m, DevCsr &view);
ohm_status ensure_partials(ohm_backend *b, size_t count);
// Union of the patterns of G and C (rows sorted by column, duplicates
// summed) with both value arrays scattered onto it.
ohm_status csr_union(int64_t n, const uint32_t *grp, const uint32_t *gcl, const double *gv,
                     const uint32_t *crp, const uint32_t *ccl, const double *cv,
                     std::vector<uint32_t> &rp, std::vector<uint32_t> &cl, std::vector<double> &g,
                     std::vector<double> &c);
// Drop every cached graph (called whenever device objects are freed).
void graph_cache_clear(ohm_backend *b);
// Backends not yet destroyed; objects whose owner is gone must not touch it.
bool backend_alive(const ohm_backend *b);

// y = A x (f64), optionally skipping when state->status != RUNNING (skip=1)
// or also when state->s_conv (skip=2).
void launch_spmv(ohm_backend *b, const DevCsr &a, const double *x, double *y,
                 const SolveState *state, int skip);
void launch_spmv_f32(ohm_backend *b, const DevCsr &a, const float *x, float *y,
                     const SolveState *state, int skip);
// y = A^T x through the transpose pattern
void launch_spmv_t(ohm_backend *b, const DevCsr &a, const double *x, double *y);

// BLAS-1 on the stream (SolverBackend ops)
void launch_axpy(ohm_backend *b, int n, double alpha, const double *x, double *y);
void launch_scale(ohm_backend *b, int n, double alpha, double *x);
void launch_copy(ohm_backend *b, int n, const double *x, double *y);
void launch_fill(ohm_backend *b, int n, double v, double *x);
// one-shot deterministic dot into d_scalars[slot]
void launch_dot(ohm_backend *b, int n, const double *x, const double *y, int slot);

// Jacobi inverse diagonal (with the Schur-complement rule for zero diagonals)
void launch_jacobi_setup(ohm_backend *b, const DevCsr &a, double *inv, float *inv32);
// Dense LU with partial pivoting on the GPU (ohm_dense.hip): x = A^-1 rhs
// for systems of at most kDenseMaxRows unknowns; x and rhs are device arrays.
constexpr int kDenseMaxRows = 4096;
ohm_status dense_lu_solve(ohm_backend *b, const DevCsr &a, const double *rhs, double *x);
void launch_mul(ohm_backend *b, int n, const double *d, const double *x, double *y);

// ---------------------------------------------------------- BiCGSTAB core --
template <typename T>
struct BicgVecs {
  T *x, *r, *rh, *p, *v, *s, *t, *ph, *sh, *tmp;
  T *p2;  // second p buffer of the fused ISAI p update
};

struct PrecondDev;  // defined in ohm_precond.h

enum { PK_NONE = 0, PK_JACOBI = 1, PK_ISAI = 2 };

}  // namespace ohm

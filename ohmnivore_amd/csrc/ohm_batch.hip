// ohm_batch.hip — batched BiCGSTAB over K independent systems
//
//   (G_k + sigma_k C_k) x_k = b_k,     k = 0 .. K-1
//
// real (f64 / f32) or complex (complex128 / complex64). One batch covers the
// three batched workloads of the solve backend:
//   * AC sweeps: one pattern and one (G, C) pair shared by every frequency
//     point, sigma_k = j*omega_k, so A = G + jωC is formed on the fly inside
//     the SpMV instead of re-tripletising per frequency
//     (form_complex_matrix, src/sparse.rs:149-178);
//   * transient companion systems: sigma = alpha/h gives A = G + (alpha/h) C
//     (form_companion_matrix, src/analysis/transient.rs:285-323);
//   * Monte Carlo instances and uneven-size circuits: each system has its
//     own values and/or pattern.
//
// Vectors of all systems are concatenated. Work is described by a table of
// row blocks (at most 256 rows and 2048 nonzeros, long rows alone), each
// tagged with its system; every kernel of an iteration runs over that table,
// so SpMV and vector updates of all systems go out in one launch. Dot
// products reduce per system: each block writes its partials to a fixed
// slot and a fold kernel with one workgroup per system sums that system's
// slots in block order and applies the scalar update, so results are
// bitwise reproducible and no block waits on a contended counter. Every system
// keeps its own scalars and status (its mask); blocks of finished systems
// exit at their first instruction, and the host drops them from the table
// when it polls (compaction), so converged frequency points stop costing
// bandwidth. Systems of at most kPersistRows rows run instead in one
// persistent launch with one workgroup per system iterating to convergence.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "ohm_precond.h"
#include "ohm_reduce.h"
#include "ohm_sources.h"

namespace ohm {

// ------------------------------------------------------------ scalars ----
template <typename R>
struct cx {
  R re, im;
};
template <typename R>
__device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) {
  return {a.re + b.re, a.im + b.im};
}
template <typename R>
__device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) {
  return {a.re - b.re, a.im - b.im};
}
template <typename R>
__device__ __forceinline__ cx<R> operator*(cx<R> a, cx<R> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

template <typename S>
struct Tr;
template <>
struct Tr<double> {
  using R = double;
  static constexpr bool C = false;
};
template <>
struct Tr<float> {
  using R = float;
  static constexpr bool C = false;
};
template <>
struct Tr<cx<double>> {
  using R = double;
  static constexpr bool C = true;
};
template <>
struct Tr<cx<float>> {
  using R = float;
  static constexpr bool C = true;
};

template <typename S>
__device__ __forceinline__ S mk(double re, double im) {
  using R = typename Tr<S>::R;
  if constexpr (Tr<S>::C)
    return S{(R)re, (R)im};
  else
    return (S)re;
}
template <typename S>
__device__ __forceinline__ S zero() {
  return mk<S>(0.0, 0.0);
}
template <typename S>
__device__ __forceinline__ double re_of(S a) {
  if constexpr (Tr<S>::C)
    return (double)a.re;
  else
    return (double)a;
}
template <typename S>
__device__ __forceinline__ double im_of(S a) {
  if constexpr (Tr<S>::C)
    return (double)a.im;
  else
    return 0.0;
}

// a_k = g_k + sigma c_k on the fly
template <typename S>
__device__ __forceinline__ S aval(typename Tr<S>::R g, typename Tr<S>::R c, double sr, double si) {
  using R = typename Tr<S>::R;
  if constexpr (Tr<S>::C)
    return S{g + (R)sr * c, (R)si * c};
  else
    return g + (R)sr * c;
}

// conj(a) . b and ||a||^2 accumulators (compensated in FP32)
template <typename S>
struct DotAcc {
  Acc<typename Tr<S>::R> re, im;
  __device__ __forceinline__ void add(S a, S b) {
    if constexpr (Tr<S>::C) {
      re.add_prod(a.re, b.re);
      re.add_prod(a.im, b.im);
      im.add_prod(a.re, b.im);
      im.add_prod(-a.im, b.re);
    } else {
      re.add_prod(a, b);
    }
  }
};
template <typename S>
struct NrmAcc {
  Acc<typename Tr<S>::R> s;
  __device__ __forceinline__ void add(S a) {
    if constexpr (Tr<S>::C) {
      s.add_prod(a.re, a.re);
      s.add_prod(a.im, a.im);
    } else {
      s.add_prod(a, a);
    }
  }
};

// double-precision complex helpers for the per-system scalar algebra
struct zd {
  double re, im;
};
__device__ __forceinline__ zd zmul(zd a, zd b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ zd zdiv(zd a, zd b) {
  const double d = b.re * b.re + b.im * b.im;
  return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
__device__ __forceinline__ double zabs2(zd a) { return a.re * a.re + a.im * a.im; }

// ---------------------------------------------------------- batch state --
struct BSys {
  double rho[2], alpha[2], omega[2], beta[2];
  double abs_tol, bnorm2, rr, ss, rhrh;
  int status, s_conv, iters, max_iters;
  unsigned counter;
  int scaled;
  int started;  // persistent kernel: scalars below the vectors are live
  int pad;
};

struct BGeo {
  const uint32_t *rowptr, *col;
  const double *g, *c;      // c may be null (then sigma is ignored)
  const float *g32, *c32;
  double sig_re, sig_im;
  int64_t off;              // offset of this system in the concatenated vectors
  int n;
  int first_block, n_blocks;
  int pad;
};

struct BBlock {
  int sys;
  uint32_t r0, r1;
  int slot;  // position in the full block table: partial-sum slot
};

constexpr int kPersistRows = 4096;  // systems this small run in the persistent kernel
constexpr int kBatchPartials = 5;   // most partial sums one reduction writes

template <typename S>
__device__ __forceinline__ const typename Tr<S>::R *gvals(const BGeo &q) {
  if constexpr (sizeof(typename Tr<S>::R) == 8)
    return q.g;
  else
    return q.g32;
}
template <typename S>
__device__ __forceinline__ const typename Tr<S>::R *cvals(const BGeo &q) {
  if constexpr (sizeof(typename Tr<S>::R) == 8)
    return q.c;
  else
    return q.c32;
}

template <typename S>
__device__ __forceinline__ S load_a(const BGeo &q, const typename Tr<S>::R *g,
                                    const typename Tr<S>::R *c, uint32_t k) {
  using R = typename Tr<S>::R;
  return aval<S>(g[k], c ? c[k] : (R)0, q.sig_re, q.sig_im);
}

// Sum K values over the workgroup for one system; the last block of the
// system (ticket on its BSys) returns true with the system totals in
// thread 0's tot[].
template <int K>
__device__ __forceinline__ bool sys_sum(double (&v)[K], const BBlock &bk, const BGeo &q, BSys *st,
                                        double *partials, double (&tot)[K]) {
  block_sum<K>(v);
  __shared__ int s_last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++)
      __hip_atomic_store(&partials[(size_t)bk.slot * kBatchPartials + k], v[k], __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);
    unsigned t = __hip_atomic_fetch_add(&st->counter, 1u, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
    s_last = (t == (unsigned)q.n_blocks - 1);
  }
  __syncthreads();
  if (!s_last) return false;
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; k++) acc[k] = 0.0;
  for (int i = q.first_block + threadIdx.x; i < q.first_block + q.n_blocks; i += kBlock) {
#pragma unroll
    for (int k = 0; k < K; k++)
      acc[k] += __hip_atomic_load(&partials[(size_t)i * kBatchPartials + k], __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
  }
  block_sum<K>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) tot[k] = acc[k];
    __hip_atomic_store(&st->counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return true;
}

// Producer side: this block's K sums go to its partial slot (plain stores;
// the kernel boundary publishes them to the fold kernel).
template <int K>
__device__ __forceinline__ void bstore(double (&v)[K], const BBlock &bk, double *partials) {
  block_sum<K>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) partials[(size_t)bk.slot * kBatchPartials + k] = v[k];
  }
}

// Fold side: one workgroup per system sums the system's slots in block
// order (the order sys_sum's last block uses); totals in thread 0's tot[].
template <int K>
__device__ __forceinline__ void fold_sum(const BGeo &q, const double *__restrict__ partials,
                                         double (&tot)[K]) {
#pragma unroll
  for (int k = 0; k < K; k++) tot[k] = 0.0;
  for (int i = q.first_block + threadIdx.x; i < q.first_block + q.n_blocks; i += kBlock) {
#pragma unroll
    for (int k = 0; k < K; k++) tot[k] += partials[(size_t)i * kBatchPartials + k];
  }
  block_sum<K>(tot);
}

template <typename S>
__device__ __forceinline__ S wg_sum(S v) {
  double t[2] = {re_of(v), im_of(v)};
  block_sum<2>(t);
  return mk<S>(t[0], t[1]);
}

// y rows of this block of system q (x, y are the system's slices); calls
// f(r, y_r) once for each row owned by this thread.
template <typename S, typename F>
__device__ __forceinline__ void bspmv_block(const BGeo &q, const BBlock &bk,
                                            const S *__restrict__ x, F f) {
  __shared__ S prod[kSpmvNnzPerBlock];
  const auto *g = gvals<S>(q);
  const auto *c = cvals<S>(q);
  const uint32_t k0 = q.rowptr[bk.r0], k1 = q.rowptr[bk.r1];
  const int tid = threadIdx.x;
  if (k1 - k0 > (uint32_t)kSpmvNnzPerBlock) {
    S acc = zero<S>();
    for (uint32_t k = k0 + tid; k < k1; k += kBlock) acc = acc + load_a<S>(q, g, c, k) * x[q.col[k]];
    acc = wg_sum<S>(acc);
    if (tid == 0) f(bk.r0, acc);
  } else {
    const uint32_t cnt = k1 - k0;
    for (uint32_t k = tid; k < cnt; k += kBlock)
      prod[k] = load_a<S>(q, g, c, k0 + k) * x[q.col[k0 + k]];
    __syncthreads();
    const uint32_t r = bk.r0 + tid;
    if (r < bk.r1) {
      const uint32_t a = q.rowptr[r] - k0, e = q.rowptr[r + 1] - k0;
      S s = zero<S>();
      for (uint32_t k = a; k < e; k++) s = s + prod[k];
      f(r, s);
    }
  }
}

__device__ __forceinline__ bool brk_small(double num_abs, double scale, int scaled,
                                          double abs_thresh) {
  return scaled ? num_abs <= 1e-30 * scale : num_abs < abs_thresh;
}

// --------------------------------------------------------------- kernels --
__global__ void k_bsys_reset(int K, BSys *sys, const int *__restrict__ active, int max_iters,
                             int scaled) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < K; k += gridDim.x * blockDim.x) {
    BSys z;
    z.rho[0] = z.alpha[0] = z.omega[0] = z.beta[0] = 1.0;
    z.rho[1] = z.alpha[1] = z.omega[1] = z.beta[1] = 0.0;
    z.abs_tol = z.bnorm2 = z.rr = z.ss = z.rhrh = 0.0;
    z.status = (active && !active[k]) ? ST_CONVERGED : ST_RUNNING;
    z.s_conv = 0;
    z.iters = 0;
    z.max_iters = max_iters;
    z.counter = 0u;
    z.scaled = scaled;
    z.started = 0;
    z.pad = 0;
    sys[k] = z;
  }
}

// y = A x for every block of the table (no state test)
template <typename S>
__global__ __launch_bounds__(kBlock) void k_bspmv(const BBlock *__restrict__ blocks,
                                                  const BGeo *__restrict__ geo,
                                                  const S *__restrict__ x, S *__restrict__ y) {
  const BBlock bk = blocks[blockIdx.x];
  const BGeo q = geo[bk.sys];
  const S *xs = x + q.off;
  S *ys = y + q.off;
  bspmv_block<S>(q, bk, xs, [&](uint32_t r, S v) { ys[r] = v; });
}

// Jacobi inverse diagonal of A_k = G_k + sigma_k C_k: 1/a_rr when
// |a_rr| > 1e-30, else the Schur-complement rule 1/(-sum a_rj a_jr / a_jj)
// (1 when that is zero too).
template <typename S>
__device__ __forceinline__ S row_entry(const BGeo &q, uint32_t r, uint32_t cidx) {
  const auto *g = gvals<S>(q);
  const auto *c = cvals<S>(q);
  for (uint32_t k = q.rowptr[r]; k < q.rowptr[r + 1]; k++)
    if (q.col[k] == cidx) return load_a<S>(q, g, c, k);
  return zero<S>();
}
__device__ __forceinline__ zd to_zd(double re, double im) { return {re, im}; }

template <typename S>
__global__ __launch_bounds__(kBlock) void k_bjacobi(const BBlock *__restrict__ blocks,
                                                    const BGeo *__restrict__ geo,
                                                    S *__restrict__ inv) {
  const BBlock bk = blocks[blockIdx.x];
  const BGeo q = geo[bk.sys];
  for (uint32_t r = bk.r0 + threadIdx.x; r < bk.r1; r += kBlock) {
    S d = row_entry<S>(q, r, r);
    zd dz = to_zd(re_of(d), im_of(d));
    zd out;
    if (sqrt(zabs2(dz)) > 1e-30) {
      out = zdiv({1.0, 0.0}, dz);
    } else {
      zd s = {0.0, 0.0};
      const auto *g = gvals<S>(q);
      const auto *c = cvals<S>(q);
      for (uint32_t k = q.rowptr[r]; k < q.rowptr[r + 1]; k++) {
        const uint32_t j = q.col[k];
        if (j == r) continue;
        S ajj = row_entry<S>(q, j, j);
        zd ajjz = to_zd(re_of(ajj), im_of(ajj));
        if (sqrt(zabs2(ajjz)) <= 1e-30) continue;
        S arj = load_a<S>(q, g, c, k), ajr = row_entry<S>(q, j, r);
        zd t = zdiv(zmul(to_zd(re_of(arj), im_of(arj)), to_zd(re_of(ajr), im_of(ajr))), ajjz);
        s.re -= t.re;
        s.im -= t.im;
      }
      out = sqrt(zabs2(s)) > 1e-300 ? zdiv({1.0, 0.0}, s) : zd{1.0, 0.0};
    }
    inv[q.off + r] = mk<S>(out.re, out.im);
  }
}

template <typename S>
__device__ __forceinline__ S shadow_s(uint32_t i) {
  if constexpr (Tr<S>::C)
    return mk<S>(shadow_value(2ull * i), shadow_value(2ull * i + 1));
  else
    return mk<S>(shadow_value(i), 0.0);
}

struct BInitArgs {
  double tol, floor_rel;
  int shadow, warm, tol_mode;
};

// r = b - A x0 (warm, A x0 in ax) or b; r_hat; p = v = 0; per-system
// ||b||^2, ||r||^2, conj(r_hat).r, ||r_hat||^2.
template <typename S>
__global__ __launch_bounds__(kBlock) void k_binit(const BBlock *__restrict__ blocks,
                                                  const BGeo *__restrict__ geo, BSys *sys,
                                                  const S *__restrict__ b, const S *__restrict__ ax,
                                                  S *__restrict__ x, S *__restrict__ r,
                                                  S *__restrict__ rh, S *__restrict__ p,
                                                  S *__restrict__ v, BInitArgs a,
                                                  double *partials) {
  const BBlock bk = blocks[blockIdx.x];
  BSys *st = &sys[bk.sys];
  if (st->status != ST_RUNNING) return;
  const BGeo q = geo[bk.sys];
  NrmAcc<S> bb, rr, hh;
  DotAcc<S> hr;
  const uint32_t i = bk.r0 + threadIdx.x;
  if (i < bk.r1) {
    const int64_t e = q.off + i;
    S bi = b[e];
    S ri = a.warm ? bi - ax[e] : bi;
    if (!a.warm) x[e] = zero<S>();
    r[e] = ri;
    S hi = a.shadow ? shadow_s<S>(i) : ri;
    rh[e] = hi;
    p[e] = zero<S>();
    v[e] = zero<S>();
    bb.add(bi);
    rr.add(ri);
    hh.add(hi);
    hr.add(hi, ri);
  }
  double vals[5] = {bb.s.get(), rr.s.get(), hr.re.get(), hr.im.get(), hh.s.get()};
  bstore<5>(vals, bk, partials);
}

// Per-system folds: one workgroup per system (grid = K) reduces the
// partials of the producing kernel and applies the reference's scalar
// update and tests for that system. Each system's state is written by its
// own workgroup only.
template <typename S>
__global__ __launch_bounds__(kBlock) void k_bfold_init(const BGeo *__restrict__ geo, BSys *sys,
                                                       BInitArgs a,
                                                       const double *__restrict__ partials) {
  BSys *st = &sys[blockIdx.x];
  if (st->status != ST_RUNNING) return;
  double tot[5];
  fold_sum<5>(geo[blockIdx.x], partials, tot);
  if (threadIdx.x == 0) {
    st->bnorm2 = tot[0];
    st->rr = tot[1];
    st->rhrh = tot[4];
    const double bnorm = sqrt(tot[0]);
    if (bnorm < 1e-30) {  // b == 0: x = 0 (src/solver/bicgstab.rs:93-97)
      st->status = ST_CONVERGED;
      return;
    }
    st->abs_tol = a.tol_mode == OHM_TOL_INITIAL ? fmax(a.tol * sqrt(tot[1]), a.floor_rel * bnorm)
                                                : a.tol * bnorm;
    const zd rho = {tot[2], tot[3]};
    if (a.warm && sqrt(tot[1]) < st->abs_tol) {
      st->status = ST_CONVERGED;
    } else if (brk_small(sqrt(zabs2(rho)), sqrt(tot[4]) * sqrt(tot[1]), st->scaled,
                         Tr<S>::C ? 1e-15 : 1e-30)) {
      st->status = ST_BRK_RHO;
    } else {
      st->rho[0] = st->beta[0] = rho.re;
      st->rho[1] = st->beta[1] = rho.im;
    }
  }
}

// K1: p = r + beta (p - omega v); p_hat = M^-1 p
template <typename S, int PK>
__global__ __launch_bounds__(kBlock) void k_bp(const BBlock *__restrict__ blocks,
                                               const BGeo *__restrict__ geo,
                                               const BSys *__restrict__ sys,
                                               const S *__restrict__ r, S *__restrict__ p,
                                               const S *__restrict__ v, S *__restrict__ ph,
                                               const S *__restrict__ inv) {
  const BBlock bk = blocks[blockIdx.x];
  const BSys *st = &sys[bk.sys];
  if (st->status != ST_RUNNING) return;
  const int64_t off = geo[bk.sys].off;
  const S beta = mk<S>(st->beta[0], st->beta[1]), omega = mk<S>(st->omega[0], st->omega[1]);
  const uint32_t i = bk.r0 + threadIdx.x;
  if (i < bk.r1) {
    const int64_t e = off + i;
    S pi = (p[e] - omega * v[e]) * beta + r[e];
    p[e] = pi;
    ph[e] = PK == PK_JACOBI ? inv[e] * pi : pi;
  }
}

// K2: v = A p_hat; alpha = rho / conj(r_hat).v
template <typename S>
__global__ __launch_bounds__(kBlock) void k_bspmv_alpha(const BBlock *__restrict__ blocks,
                                                        const BGeo *__restrict__ geo, BSys *sys,
                                                        const S *__restrict__ ph,
                                                        S *__restrict__ v,
                                                        const S *__restrict__ rh,
                                                        double *partials) {
  const BBlock bk = blocks[blockIdx.x];
  BSys *st = &sys[bk.sys];
  if (st->status != ST_RUNNING) return;
  const BGeo q = geo[bk.sys];
  DotAcc<S> hv;
  NrmAcc<S> vv;
  S *vs = v + q.off;
  const S *hs = rh + q.off;
  bspmv_block<S>(q, bk, ph + q.off, [&](uint32_t r, S y) {
    vs[r] = y;
    hv.add(hs[r], y);
    vv.add(y);
  });
  double vals[3] = {hv.re.get(), hv.im.get(), vv.s.get()};
  bstore<3>(vals, bk, partials);
}

template <typename S>
__global__ __launch_bounds__(kBlock) void k_bfold_alpha(const BGeo *__restrict__ geo, BSys *sys,
                                                        const double *__restrict__ partials) {
  BSys *st = &sys[blockIdx.x];
  if (st->status != ST_RUNNING) return;
  double tot[3];
  fold_sum<3>(geo[blockIdx.x], partials, tot);
  if (threadIdx.x == 0) {
    const zd rv = {tot[0], tot[1]};
    if (brk_small(sqrt(zabs2(rv)), sqrt(st->rhrh) * sqrt(tot[2]), st->scaled,
                  Tr<S>::C ? 1e-15 : 1e-30)) {
      st->status = ST_BRK_RV;
    } else {
      const zd al = zdiv({st->rho[0], st->rho[1]}, rv);
      st->alpha[0] = al.re;
      st->alpha[1] = al.im;
    }
  }
}

// K3: s = r - alpha v; s_hat = M^-1 s; ||s|| early exit
template <typename S, int PK>
__global__ __launch_bounds__(kBlock) void k_bs(const BBlock *__restrict__ blocks,
                                               const BGeo *__restrict__ geo, BSys *sys,
                                               const S *__restrict__ r, const S *__restrict__ v,
                                               S *__restrict__ s, S *__restrict__ sh,
                                               const S *__restrict__ inv, double *partials) {
  const BBlock bk = blocks[blockIdx.x];
  BSys *st = &sys[bk.sys];
  if (st->status != ST_RUNNING) return;
  const BGeo q = geo[bk.sys];
  const S alpha = mk<S>(st->alpha[0], st->alpha[1]);
  NrmAcc<S> ss;
  const uint32_t i = bk.r0 + threadIdx.x;
  if (i < bk.r1) {
    const int64_t e = q.off + i;
    S si = r[e] - alpha * v[e];
    s[e] = si;
    sh[e] = PK == PK_JACOBI ? inv[e] * si : si;
    ss.add(si);
  }
  double vals[1] = {ss.s.get()};
  bstore<1>(vals, bk, partials);
}

__global__ __launch_bounds__(kBlock) void k_bfold_s(const BGeo *__restrict__ geo, BSys *sys,
                                                    const double *__restrict__ partials) {
  BSys *st = &sys[blockIdx.x];
  if (st->status != ST_RUNNING) return;
  double tot[1];
  fold_sum<1>(geo[blockIdx.x], partials, tot);
  if (threadIdx.x == 0) {
    st->ss = tot[0];
    if (sqrt(tot[0]) < st->abs_tol) st->s_conv = 1;
  }
}

// K4: t = A s_hat; omega = conj(t).s / ||t||^2
template <typename S>
__global__ __launch_bounds__(kBlock) void k_bspmv_omega(const BBlock *__restrict__ blocks,
                                                        const BGeo *__restrict__ geo, BSys *sys,
                                                        const S *__restrict__ sh,
                                                        S *__restrict__ t,
                                                        const S *__restrict__ s,
                                                        double *partials) {
  const BBlock bk = blocks[blockIdx.x];
  BSys *st = &sys[bk.sys];
  if (st->status != ST_RUNNING || st->s_conv) return;
  const BGeo q = geo[bk.sys];
  DotAcc<S> ts;
  NrmAcc<S> tt;
  S *tv = t + q.off;
  const S *sv = s + q.off;
  bspmv_block<S>(q, bk, sh + q.off, [&](uint32_t r, S y) {
    tv[r] = y;
    ts.add(y, sv[r]);
    tt.add(y);
  });
  double vals[3] = {ts.re.get(), ts.im.get(), tt.s.get()};
  bstore<3>(vals, bk, partials);
}

template <typename S>
__global__ __launch_bounds__(kBlock) void k_bfold_omega(const BGeo *__restrict__ geo, BSys *sys,
                                                        const double *__restrict__ partials) {
  BSys *st = &sys[blockIdx.x];
  if (st->status != ST_RUNNING || st->s_conv) return;
  double tot[3];
  fold_sum<3>(geo[blockIdx.x], partials, tot);
  if (threadIdx.x == 0) {
    const bool brk = st->scaled ? tot[2] <= 1e-30 * st->ss
                                : (Tr<S>::C ? tot[2] < 1e-15 : fabs(tot[2]) < 1e-30);
    if (brk) {
      st->status = ST_BRK_TT;
    } else {
      st->omega[0] = tot[0] / tot[2];
      st->omega[1] = tot[1] / tot[2];
    }
  }
}

// K5: x += alpha p_hat + omega s_hat; r = s - omega t; tests; beta
template <typename S>
__global__ __launch_bounds__(kBlock) void k_bxr(const BBlock *__restrict__ blocks,
                                                const BGeo *__restrict__ geo, BSys *sys,
                                                S *__restrict__ x, S *__restrict__ r,
                                                const S *__restrict__ rh, const S *__restrict__ ph,
                                                const S *__restrict__ sh, const S *__restrict__ s,
                                                const S *__restrict__ t, double *partials) {
  const BBlock bk = blocks[blockIdx.x];
  BSys *st = &sys[bk.sys];
  if (st->status != ST_RUNNING) return;
  const BGeo q = geo[bk.sys];
  const int sconv = st->s_conv;
  const S alpha = mk<S>(st->alpha[0], st->alpha[1]), omega = mk<S>(st->omega[0], st->omega[1]);
  NrmAcc<S> rr;
  DotAcc<S> hr;
  const uint32_t i = bk.r0 + threadIdx.x;
  if (i < bk.r1) {
    const int64_t e = q.off + i;
    if (sconv) {
      x[e] = x[e] + alpha * ph[e];
    } else {
      x[e] = (x[e] + alpha * ph[e]) + omega * sh[e];
      S ri = s[e] - omega * t[e];
      r[e] = ri;
      rr.add(ri);
      hr.add(rh[e], ri);
    }
  }
  double vals[3] = {rr.s.get(), hr.re.get(), hr.im.get()};
  bstore<3>(vals, bk, partials);
}

template <typename S>
__global__ __launch_bounds__(kBlock) void k_bfold_xr(const BGeo *__restrict__ geo, BSys *sys,
                                                     const double *__restrict__ partials) {
  BSys *st = &sys[blockIdx.x];
  if (st->status != ST_RUNNING) return;
  const int sconv = st->s_conv;
  double tot[3];
  fold_sum<3>(geo[blockIdx.x], partials, tot);
  if (threadIdx.x == 0) {
    st->iters += 1;
    if (sconv) {
      st->status = ST_CONVERGED;
      return;
    }
    st->rr = tot[0];
    const double rnorm = sqrt(tot[0]);
    const zd om = {st->omega[0], st->omega[1]}, rho_new = {tot[1], tot[2]};
    if (std::isnan(rnorm) || std::isinf(rnorm)) {
      st->status = ST_NAN;
    } else if (rnorm < st->abs_tol) {
      st->status = ST_CONVERGED;
    } else if (Tr<S>::C ? zabs2(om) < 1e-30 : fabs(om.re) < 1e-30) {
      st->status = ST_BRK_OMEGA;
    } else if (st->iters >= st->max_iters) {
      st->status = ST_MAXITER;
    } else if (brk_small(sqrt(zabs2(rho_new)), sqrt(st->rhrh) * sqrt(tot[0]), st->scaled,
                         Tr<S>::C ? 1e-15 : 1e-30)) {
      st->status = ST_BRK_RHO;
    } else {
      const zd be = zmul(zdiv(rho_new, {st->rho[0], st->rho[1]}),
                         zdiv({st->alpha[0], st->alpha[1]}, om));
      st->beta[0] = be.re;
      st->beta[1] = be.im;
      st->rho[0] = rho_new.re;
      st->rho[1] = rho_new.im;
    }
  }
}

// ------------------------------------------------- persistent small systems
// One workgroup per system runs the whole solve (the reference's
// single-workgroup bicgstab_fused, src/solver/ds_shaders.rs:340-728, at any
// n <= kPersistRows). Same arithmetic and tests as the multi-kernel path;
// every loop is bounded by max_iters, so every workgroup exits.
template <typename S>
__device__ __forceinline__ void p_spmv(const BGeo &q, const S *__restrict__ x, S *__restrict__ y) {
  const auto *g = gvals<S>(q);
  const auto *c = cvals<S>(q);
  for (int r = threadIdx.x; r < q.n; r += kBlock) {
    S acc = zero<S>();
    for (uint32_t k = q.rowptr[r]; k < q.rowptr[r + 1]; k++)
      acc = acc + load_a<S>(q, g, c, k) * x[q.col[k]];
    y[r] = acc;
  }
}

template <int K>
__device__ __forceinline__ void wg_reduce(double (&v)[K]) {
  block_sum<K>(v);
  __shared__ double bc[K];
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; k++) bc[k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; k++) v[k] = bc[k];
  __syncthreads();
}

template <typename S, int PK>
__global__ __launch_bounds__(kBlock) void k_bpersist(const int *__restrict__ list,
                                                     const BGeo *__restrict__ geo, BSys *sys,
                                                     const S *__restrict__ b, const S *__restrict__ ax,
                                                     S *__restrict__ X, S *__restrict__ R,
                                                     S *__restrict__ RH, S *__restrict__ P,
                                                     S *__restrict__ V, S *__restrict__ SS,
                                                     S *__restrict__ T, S *__restrict__ PH,
                                                     S *__restrict__ SH,
                                                     const S *__restrict__ inv, BInitArgs a,
                                                     int budget) {
  const int k = list[blockIdx.x];
  BSys *st = &sys[k];
  if (st->status != ST_RUNNING) return;
  const BGeo q = geo[k];
  const int n = q.n;
  const int64_t o = q.off;
  S *x = X + o, *r = R + o, *rh = RH + o, *p = P + o, *v = V + o, *s = SS + o, *t = T + o,
    *ph = PH + o, *sh = SH + o;
  const S *iv = inv ? inv + o : nullptr;
  const int scaled = st->scaled, max_iters = st->max_iters;
  const double th = Tr<S>::C ? 1e-15 : 1e-30;
  int status = ST_RUNNING, iters;
  double abs_tol, rr, rhrh, bnorm2;
  zd rho, alpha, omega, beta;
  if (!st->started) {
    NrmAcc<S> nb, nr, nh;
    DotAcc<S> dhr;
    for (int i = threadIdx.x; i < n; i += kBlock) {
      S bi = b[o + i];
      S ri = a.warm ? bi - ax[o + i] : bi;
      if (!a.warm) x[i] = zero<S>();
      r[i] = ri;
      S hi = a.shadow ? shadow_s<S>(i) : ri;
      rh[i] = hi;
      p[i] = zero<S>();
      v[i] = zero<S>();
      nb.add(bi);
      nr.add(ri);
      nh.add(hi);
      dhr.add(hi, ri);
    }
    double iv5[5] = {nb.s.get(), nr.s.get(), dhr.re.get(), dhr.im.get(), nh.s.get()};
    wg_reduce<5>(iv5);
    bnorm2 = iv5[0];
    rr = iv5[1];
    rhrh = iv5[4];
    iters = 0;
    abs_tol = 0.0;
    rho = beta = zd{iv5[2], iv5[3]};
    alpha = omega = zd{1.0, 0.0};
    const double bnorm = sqrt(bnorm2);
    if (bnorm < 1e-30) {
      status = ST_CONVERGED;
    } else {
      abs_tol = a.tol_mode == OHM_TOL_INITIAL ? fmax(a.tol * sqrt(rr), a.floor_rel * bnorm)
                                              : a.tol * bnorm;
      if (a.warm && sqrt(rr) < abs_tol)
        status = ST_CONVERGED;
      else if (brk_small(sqrt(zabs2(rho)), sqrt(rhrh) * sqrt(rr), scaled, th))
        status = ST_BRK_RHO;
    }
  } else {
    iters = st->iters;
    abs_tol = st->abs_tol;
    rr = st->rr;
    rhrh = st->rhrh;
    bnorm2 = st->bnorm2;
    rho = {st->rho[0], st->rho[1]};
    alpha = {st->alpha[0], st->alpha[1]};
    omega = {st->omega[0], st->omega[1]};
    beta = {st->beta[0], st->beta[1]};
  }
  __syncthreads();
  const int stop_at = iters + budget;
  while (status == ST_RUNNING && iters < stop_at) {
    const S sb = mk<S>(beta.re, beta.im), so = mk<S>(omega.re, omega.im);
    for (int i = threadIdx.x; i < n; i += kBlock) {
      S pi = (p[i] - so * v[i]) * sb + r[i];
      p[i] = pi;
      ph[i] = PK == PK_JACOBI ? iv[i] * pi : pi;
    }
    __syncthreads();
    p_spmv<S>(q, ph, v);
    __syncthreads();
    DotAcc<S> hv;
    NrmAcc<S> vv;
    for (int i = threadIdx.x; i < n; i += kBlock) {
      hv.add(rh[i], v[i]);
      vv.add(v[i]);
    }
    double d3[3] = {hv.re.get(), hv.im.get(), vv.s.get()};
    wg_reduce<3>(d3);
    const zd rv = {d3[0], d3[1]};
    if (brk_small(sqrt(zabs2(rv)), sqrt(rhrh) * sqrt(d3[2]), scaled, th)) {
      status = ST_BRK_RV;
      break;
    }
    alpha = zdiv(rho, rv);
    const S sa = mk<S>(alpha.re, alpha.im);
    NrmAcc<S> ns;
    for (int i = threadIdx.x; i < n; i += kBlock) {
      S si = r[i] - sa * v[i];
      s[i] = si;
      sh[i] = PK == PK_JACOBI ? iv[i] * si : si;
      ns.add(si);
    }
    double d1[1] = {ns.s.get()};
    wg_reduce<1>(d1);
    const double ssq = d1[0];
    if (sqrt(ssq) < abs_tol) {
      for (int i = threadIdx.x; i < n; i += kBlock) x[i] = x[i] + sa * ph[i];
      iters++;
      status = ST_CONVERGED;
      break;
    }
    p_spmv<S>(q, sh, t);
    __syncthreads();
    DotAcc<S> ts;
    NrmAcc<S> tt;
    for (int i = threadIdx.x; i < n; i += kBlock) {
      ts.add(t[i], s[i]);
      tt.add(t[i]);
    }
    double e3[3] = {ts.re.get(), ts.im.get(), tt.s.get()};
    wg_reduce<3>(e3);
    const bool tbrk = scaled ? e3[2] <= 1e-30 * ssq : e3[2] < (Tr<S>::C ? 1e-15 : 1e-30);
    if (tbrk) {
      status = ST_BRK_TT;
      break;
    }
    omega = {e3[0] / e3[2], e3[1] / e3[2]};
    const S so2 = mk<S>(omega.re, omega.im);
    NrmAcc<S> nr2;
    DotAcc<S> hr2;
    for (int i = threadIdx.x; i < n; i += kBlock) {
      x[i] = (x[i] + sa * ph[i]) + so2 * sh[i];
      S ri = s[i] - so2 * t[i];
      r[i] = ri;
      nr2.add(ri);
      hr2.add(rh[i], ri);
    }
    double f3[3] = {nr2.s.get(), hr2.re.get(), hr2.im.get()};
    wg_reduce<3>(f3);
    iters++;
    rr = f3[0];
    const double rnorm = sqrt(rr);
    const zd rho_new = {f3[1], f3[2]};
    if (std::isnan(rnorm) || std::isinf(rnorm)) {
      status = ST_NAN;
    } else if (rnorm < abs_tol) {
      status = ST_CONVERGED;
    } else if (Tr<S>::C ? zabs2(omega) < 1e-30 : fabs(omega.re) < 1e-30) {
      status = ST_BRK_OMEGA;
    } else if (iters >= max_iters) {
      status = ST_MAXITER;
    } else if (brk_small(sqrt(zabs2(rho_new)), sqrt(rhrh) * sqrt(rr), scaled, th)) {
      status = ST_BRK_RHO;
    } else {
      beta = zmul(zdiv(rho_new, rho), zdiv(alpha, omega));
      rho = rho_new;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    st->status = status;
    st->iters = iters;
    st->rr = rr;
    st->rhrh = rhrh;
    st->bnorm2 = bnorm2;
    st->abs_tol = abs_tol;
    st->rho[0] = rho.re;
    st->rho[1] = rho.im;
    st->alpha[0] = alpha.re;
    st->alpha[1] = alpha.im;
    st->omega[0] = omega.re;
    st->omega[1] = omega.im;
    st->beta[0] = beta.re;
    st->beta[1] = beta.im;
    st->started = 1;
  }
}

// ------------------------------------------ FP64 residual / refinement ---
// r = b - A x (FP64, S64 = double or cx<double>); per system ||r||^2 and
// ||b||^2 into res[2k], res[2k+1] (last block of each system).
template <typename S>
__global__ __launch_bounds__(kBlock) void k_bresid(const BBlock *__restrict__ blocks,
                                                   const BGeo *__restrict__ geo, BSys *sys,
                                                   const S *__restrict__ x,
                                                   const S *__restrict__ b, S *__restrict__ r,
                                                   double *partials, double *res) {
  const BBlock bk = blocks[blockIdx.x];
  const BGeo q = geo[bk.sys];
  BSys *st = &sys[bk.sys];
  NrmAcc<S> rr, nb;
  const S *bs = b + q.off;
  S *rs = r + q.off;
  bspmv_block<S>(q, bk, x + q.off, [&](uint32_t i, S y) {
    S ri = bs[i] - y;
    rs[i] = ri;
    rr.add(ri);
    nb.add(bs[i]);
  });
  double vals[2] = {rr.s.get(), nb.s.get()}, tot[2];
  if (sys_sum<2>(vals, bk, q, st, partials, tot) && threadIdx.x == 0) {
    res[2 * bk.sys] = tot[0];
    res[2 * bk.sys + 1] = tot[1];
  }
}

// r32 = r64 / scale_k for the systems being refined
template <typename S64, typename S32>
__global__ __launch_bounds__(kBlock) void k_bscale_down(const BBlock *__restrict__ blocks,
                                                        const BGeo *__restrict__ geo,
                                                        const S64 *__restrict__ r,
                                                        const double *__restrict__ scale,
                                                        S32 *__restrict__ r32) {
  const BBlock bk = blocks[blockIdx.x];
  const BGeo q = geo[bk.sys];
  const double inv = 1.0 / scale[bk.sys];
  const uint32_t i = bk.r0 + threadIdx.x;
  if (i < bk.r1) {
    const S64 v = r[q.off + i];
    r32[q.off + i] = mk<S32>(re_of(v) * inv, im_of(v) * inv);
  }
}

// x += scale_k * d32 for systems whose inner solve produced a usable step
template <typename S64, typename S32>
__global__ __launch_bounds__(kBlock) void k_bcorrect(const BBlock *__restrict__ blocks,
                                                     const BGeo *__restrict__ geo,
                                                     const BSys *__restrict__ sys,
                                                     const S32 *__restrict__ d,
                                                     const double *__restrict__ scale,
                                                     S64 *__restrict__ x) {
  const BBlock bk = blocks[blockIdx.x];
  const BSys *st = &sys[bk.sys];
  if (st->status == ST_NAN || st->iters == 0) return;
  const BGeo q = geo[bk.sys];
  const double sc = scale[bk.sys];
  const uint32_t i = bk.r0 + threadIdx.x;
  if (i < bk.r1) {
    const S32 di = d[q.off + i];
    x[q.off + i] = x[q.off + i] + mk<S64>(re_of(di) * sc, im_of(di) * sc);
  }
}

// float copies of value arrays
__global__ void k_form_a(int64_t nnz, const double *__restrict__ g, const double *__restrict__ c,
                         double sigma, double *__restrict__ a) {
  for (int64_t k = blockIdx.x * (int64_t)kBlock + threadIdx.x; k < nnz;
       k += (int64_t)gridDim.x * kBlock)
    a[k] = g[k] + sigma * c[k];
}

__global__ void k_d2f_b(int64_t n, const double *__restrict__ x, float *__restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    y[i] = (float)x[i];
}

}  // namespace ohm

// =================================================================== host ==
using namespace ohm;

struct ohm_batch {
  ohm_backend *owner = nullptr;
  int K = 0;
  bool cplx = false;
  int64_t N = 0;  // total rows
  int max_n = 0;
  std::vector<BGeo> h_geo;
  BGeo *d_geo = nullptr;
  // The systems as created (G and C separate). For real systems with a C
  // matrix the FP64 SpMV reads A_k = G_k + sigma_k C_k formed once per sigma
  // (h_geo's g points there, c is null); FP32 mirrors, the transient RHS
  // (which needs G x and C x) and re-forming read these originals.
  std::vector<BGeo> h_src;
  BGeo *d_src = nullptr;
  std::vector<double *> formed;
  std::vector<BBlock> h_blocks;  // full table, ordered by system
  BBlock *d_full = nullptr, *d_active = nullptr;
  int *d_list = nullptr;         // active system ids (persistent kernel)
  int *d_mask = nullptr;         // per-system run flag for a reset
  BSys *d_sys = nullptr;
  BSys *h_sys = nullptr;         // pinned poll staging
  double *d_partials = nullptr;
  double *d_res = nullptr, *h_res = nullptr;  // per-system residual pairs (pinned host)
  double *d_scale = nullptr;
  std::vector<void *> allocs;    // patterns and values
  std::vector<int64_t> h_nnz;    // per system
  bool has_f32 = false;
  bool jac64_valid = false, jac32_valid = false;  // Jacobi inverses match the current sigmas
  // workspaces: 10 vectors each (x r rh p v s t ph sh inv) + rhs/ax scratch
  void *ws64 = nullptr, *ws32 = nullptr;
  int64_t dispatches = 0;
  int tuned_shape = -1;  // the row-block shape the autotuner picked (index of its candidates; -1: default)
};

namespace {

size_t elem_bytes(bool cplx, bool f32) { return (cplx ? 2 : 1) * (f32 ? 4 : 8); }

template <typename S>
struct BW {  // workspace view
  S *x, *r, *rh, *p, *v, *s, *t, *ph, *sh, *inv, *b, *ax;
};
template <typename S>
BW<S> ws_view(void *base, int64_t N) {
  const int64_t stride = (N + 63) & ~int64_t(63);
  S *q = reinterpret_cast<S *>(base);
  BW<S> w;
  S **f[] = {&w.x, &w.r, &w.rh, &w.p, &w.v, &w.s, &w.t, &w.ph, &w.sh, &w.inv, &w.b, &w.ax};
  for (int i = 0; i < 12; i++) *f[i] = q + i * stride;
  return w;
}
int64_t ws_bytes(const ohm_batch *bt, bool f32) {
  const int64_t stride = (bt->N + 63) & ~int64_t(63);
  return 12 * std::max<int64_t>(stride, 64) * (int64_t)elem_bytes(bt->cplx, f32);
}

// FP64 values of real systems with a C matrix: A_k = G_k + sigma_k C_k,
// formed on the device whenever sigma changes (h_geo then points at A_k with
// no C); complex systems keep forming G + sigma C inside the SpMV.
ohm_status form_values(ohm_batch *bt) {
  for (int k = 0; k < bt->K; k++) {
    const BGeo &o = bt->h_src[k];
    BGeo &q = bt->h_geo[k];
    const int64_t nnz = bt->h_nnz[k];
    if (bt->cplx || !o.c || o.n == 0 || nnz == 0) {
      q.g = o.g;
      q.c = o.c;
      continue;
    }
    if (!bt->formed[k]) {
      OHM_HIP(hipMalloc(&bt->formed[k], sizeof(double) * nnz));
      bt->allocs.push_back(bt->formed[k]);
    }
    hipLaunchKernelGGL(k_form_a, dim3(std::min<int64_t>((nnz + kBlock - 1) / kBlock, 4096)),
                       dim3(kBlock), 0, bt->owner->stream, nnz, o.g, o.c, o.sig_re, bt->formed[k]);
    q.g = bt->formed[k];
    q.c = nullptr;
  }
  OHM_HIP(hipGetLastError());
  return OHM_OK;
}

struct Poll {
  int running = 0;
};

ohm_status upload_tables(ohm_batch *bt, const std::vector<BBlock> &tab) {
  if (tab.empty()) return OHM_OK;
  OHM_HIP(hipMemcpyAsync(bt->d_active, tab.data(), sizeof(BBlock) * tab.size(),
                         hipMemcpyHostToDevice, bt->owner->stream));
  return OHM_OK;
}

// Blocks of the systems whose flag is set.
std::vector<BBlock> active_table(const ohm_batch *bt, const std::vector<char> &run) {
  std::vector<BBlock> out;
  out.reserve(bt->h_blocks.size());
  for (const BBlock &b : bt->h_blocks)
    if (run[b.sys]) out.push_back(b);
  return out;
}

// One iteration over the active block table: each reducing kernel is
// followed by its per-system fold (grid = K systems).
template <typename S, int PK>
void enqueue_iter(ohm_batch *bt, BW<S> &w, int nblk) {
  hipStream_t st = bt->owner->stream;
  const dim3 g(nblk), gk(bt->K), blk(kBlock);
  hipLaunchKernelGGL((k_bp<S, PK>), g, blk, 0, st, bt->d_active, bt->d_geo, bt->d_sys, w.r, w.p,
                     w.v, w.ph, w.inv);
  hipLaunchKernelGGL(k_bspmv_alpha<S>, g, blk, 0, st, bt->d_active, bt->d_geo, bt->d_sys, w.ph,
                     w.v, w.rh, bt->d_partials);
  hipLaunchKernelGGL(k_bfold_alpha<S>, gk, blk, 0, st, bt->d_geo, bt->d_sys, bt->d_partials);
  hipLaunchKernelGGL((k_bs<S, PK>), g, blk, 0, st, bt->d_active, bt->d_geo, bt->d_sys, w.r, w.v,
                     w.s, w.sh, w.inv, bt->d_partials);
  hipLaunchKernelGGL(k_bfold_s, gk, blk, 0, st, bt->d_geo, bt->d_sys, bt->d_partials);
  hipLaunchKernelGGL(k_bspmv_omega<S>, g, blk, 0, st, bt->d_active, bt->d_geo, bt->d_sys, w.sh,
                     w.t, w.s, bt->d_partials);
  hipLaunchKernelGGL(k_bfold_omega<S>, gk, blk, 0, st, bt->d_geo, bt->d_sys, bt->d_partials);
  hipLaunchKernelGGL(k_bxr<S>, g, blk, 0, st, bt->d_active, bt->d_geo, bt->d_sys, w.x, w.r, w.rh,
                     w.ph, w.sh, w.s, w.t, bt->d_partials);
  hipLaunchKernelGGL(k_bfold_xr<S>, gk, blk, 0, st, bt->d_geo, bt->d_sys, bt->d_partials);
  bt->dispatches += 9;
  bt->owner->dispatches += 9;
}

struct BatchOpts {
  int64_t max_iters = 10000;
  double tol = 1e-5;
  int shadow = OHM_SHADOW_HASHED;
  bool warm = false;
  int check_interval = 0;
  bool scaled = true;
  int tol_mode = OHM_TOL_RHS;
  double tol_floor = 1e-15;
  int precond = OHM_PREC_JACOBI;
  // FP64 solves: a copy of every system's starting x (warm; elements of the
  // solve's scalar type). A system whose solve breaks down is solved once
  // more from it with the other shadow vector (the analyses' recovery of
  // the single-system solver).
  const void *x_start = nullptr;
  // iterations the slowest system of an earlier like batch took (0:
  // unknown): the first chunk runs to about 4 short of it with no poll
  // before it, then chunks of 4 until well past it
  int64_t expect_iters = 0;
};

// Solves the systems flagged in `run` on workspace w (rhs in w.b, x0 in w.x
// when warm). Per-system status / iterations are left in d_sys.
template <typename S>
ohm_status core_solve(ohm_batch *bt, BW<S> &w, const std::vector<char> &run, const BatchOpts &o) {
  ohm_backend *b = bt->owner;
  hipStream_t stream = b->stream;
  const int K = bt->K;
  std::vector<int> mask(K);
  int n_run = 0;
  for (int k = 0; k < K; k++) n_run += (mask[k] = run[k] ? 1 : 0);
  OHM_HIP(hipMemcpyAsync(bt->d_mask, mask.data(), sizeof(int) * K, hipMemcpyHostToDevice, stream));
  const int mi = (int)std::min<int64_t>(o.max_iters, 0x7fffffff);
  hipLaunchKernelGGL(k_bsys_reset, dim3((K + 255) / 256), dim3(256), 0, stream, K, bt->d_sys,
                     bt->d_mask, mi, o.scaled ? 1 : 0);
  if (n_run == 0) return OHM_OK;
  std::vector<BBlock> tab = active_table(bt, run);
  OHM_TRY(upload_tables(bt, tab));
  const int pk = o.precond == OHM_PREC_NONE ? PK_NONE : PK_JACOBI;
  bool &jac_ok = sizeof(typename Tr<S>::R) == 8 ? bt->jac64_valid : bt->jac32_valid;
  if (pk == PK_JACOBI && !jac_ok) {  // every system's inverse diagonal, kept until sigma changes
    hipLaunchKernelGGL(k_bjacobi<S>, dim3(bt->h_blocks.size()), dim3(kBlock), 0, stream,
                       bt->d_full, bt->d_geo, w.inv);
    jac_ok = true;
  }
  if (o.warm)
    hipLaunchKernelGGL(k_bspmv<S>, dim3(tab.size()), dim3(kBlock), 0, stream, bt->d_active,
                       bt->d_geo, w.x, w.ax);
  BInitArgs ia{o.tol, o.tol_floor, o.shadow == OHM_SHADOW_HASHED ? 1 : 0, o.warm ? 1 : 0,
               o.tol_mode};
  bt->dispatches += 3;
  b->dispatches += 3;

  // persistent path: every running system small
  int max_n = 0;
  for (int k = 0; k < K; k++)
    if (run[k]) max_n = std::max(max_n, bt->h_geo[k].n);
  if (max_n <= kPersistRows) {
    // one workgroup per system; relaunched in bounded slices of iterations
    // until no system is running
    std::vector<char> live = run;
    const int budget = (int)std::min<int64_t>(1000, std::max<int64_t>(1, o.max_iters));
    for (int64_t launched = 0;; launched += budget) {
      std::vector<int> list;
      for (int k = 0; k < K; k++)
        if (live[k]) list.push_back(k);
      if (list.empty() || launched > o.max_iters + budget) break;
      OHM_HIP(hipMemcpyAsync(bt->d_list, list.data(), sizeof(int) * list.size(),
                             hipMemcpyHostToDevice, stream));
      const dim3 g(list.size()), blk(kBlock);
      if (pk == PK_JACOBI)
        hipLaunchKernelGGL((k_bpersist<S, PK_JACOBI>), g, blk, 0, stream, bt->d_list, bt->d_geo,
                           bt->d_sys, w.b, w.ax, w.x, w.r, w.rh, w.p, w.v, w.s, w.t, w.ph, w.sh,
                           w.inv, ia, budget);
      else
        hipLaunchKernelGGL((k_bpersist<S, PK_NONE>), g, blk, 0, stream, bt->d_list, bt->d_geo,
                           bt->d_sys, w.b, w.ax, w.x, w.r, w.rh, w.p, w.v, w.s, w.t, w.ph, w.sh,
                           nullptr, ia, budget);
      bt->dispatches++;
      b->dispatches++;
      OHM_HIP(hipMemcpyAsync(bt->h_sys, bt->d_sys, sizeof(BSys) * K, hipMemcpyDeviceToHost,
                             stream));
      OHM_HIP(hipStreamSynchronize(stream));
      b->readbacks++;
      for (int k = 0; k < K; k++)
        if (live[k] && bt->h_sys[k].status != ST_RUNNING) live[k] = 0;
    }
    return OHM_OK;
  }

  hipLaunchKernelGGL(k_binit<S>, dim3(tab.size()), dim3(kBlock), 0, stream, bt->d_active,
                     bt->d_geo, bt->d_sys, w.b, w.ax, w.x, w.r, w.rh, w.p, w.v, ia,
                     bt->d_partials);
  hipLaunchKernelGGL(k_bfold_init<S>, dim3(K), dim3(kBlock), 0, stream, bt->d_geo, bt->d_sys, ia,
                     bt->d_partials);
  bt->dispatches += 2;
  b->dispatches += 2;
  std::vector<char> live = run;
  int64_t done = 0;
  for (bool first = true;; first = false) {
    int chunk = 0;
    if (first && o.expect_iters > 0 && o.check_interval <= 0) {
      // at most 64 iterations, so that a long earlier solve cannot queue a
      // long run of no-ops
      chunk = (int)std::min<int64_t>(64, std::max<int64_t>(4, (o.expect_iters - 4) / 8 * 8));
    } else {
      OHM_HIP(hipMemcpyAsync(bt->h_sys, bt->d_sys, sizeof(BSys) * K, hipMemcpyDeviceToHost,
                             stream));
      OHM_HIP(hipStreamSynchronize(stream));
      b->readbacks++;
      int still = 0;
      bool shrank = false;
      int64_t min_iters = INT64_MAX;
      for (int k = 0; k < K; k++) {
        if (!live[k]) continue;
        if (bt->h_sys[k].status != ST_RUNNING) {
          live[k] = 0;
          shrank = true;
        } else {
          still++;
          min_iters = std::min<int64_t>(min_iters, bt->h_sys[k].iters);
        }
      }
      if (still == 0) break;
      if (done > o.max_iters + 64) break;  // every system has hit max_iters by now
      if (shrank) {
        tab = active_table(bt, live);
        OHM_TRY(upload_tables(bt, tab));
      }
      chunk = o.check_interval > 0
                  ? o.check_interval
                  : (int)std::min<int64_t>(16, std::max<int64_t>(4, min_iters / 2));
      if (o.check_interval <= 0 && o.expect_iters > 0) {
        const int64_t e = o.expect_iters;
        if (min_iters < e - 4)
          chunk = (int)std::min<int64_t>(16, std::max<int64_t>(4, (e - 4 - min_iters + 3) / 4 * 4));
        else if (min_iters < e + 8)
          chunk = 4;
      }
    }
    for (int c = 0; c < chunk; c++) {
      if (pk == PK_JACOBI)
        enqueue_iter<S, PK_JACOBI>(bt, w, (int)tab.size());
      else
        enqueue_iter<S, PK_NONE>(bt, w, (int)tab.size());
    }
    done += chunk;
  }
  return OHM_OK;
}

template <typename S>
ohm_status run_spmv(ohm_batch *bt, const S *x, S *y) {
  if (bt->h_blocks.empty()) return OHM_OK;
  OHM_HIP(hipMemcpyAsync(bt->d_active, bt->h_blocks.data(), sizeof(BBlock) * bt->h_blocks.size(),
                         hipMemcpyHostToDevice, bt->owner->stream));
  hipLaunchKernelGGL(k_bspmv<S>, dim3(bt->h_blocks.size()), dim3(kBlock), 0, bt->owner->stream,
                     bt->d_active, bt->d_geo, x, y);
  bt->dispatches++;
  bt->owner->dispatches++;
  return OHM_OK;
}

ohm_status ensure_ws(ohm_batch *bt, bool f32) {
  void **p = f32 ? &bt->ws32 : &bt->ws64;
  if (!*p) OHM_HIP(hipMalloc(p, ws_bytes(bt, f32)));
  if (f32 && !bt->has_f32) {
    // float mirrors of every distinct value array
    std::map<const double *, const float *> done;
    for (int k = 0; k < bt->K; k++) {
      BGeo &q = bt->h_geo[k];
      BGeo &o = bt->h_src[k];
      const int64_t nnz = bt->h_nnz[k];
      for (int which = 0; which < 2; which++) {
        const double *src = which ? o.c : o.g;
        if (!src) continue;
        const float *dst;
        auto it = done.find(src);
        if (it != done.end()) {
          dst = it->second;
        } else {
          float *f = nullptr;
          OHM_HIP(hipMalloc(&f, sizeof(float) * std::max<int64_t>(nnz, 1)));
          bt->allocs.push_back(f);
          if (nnz > 0)
            hipLaunchKernelGGL(k_d2f_b, dim3(std::min<int64_t>((nnz + 255) / 256, 4096)), dim3(256),
                               0, bt->owner->stream, nnz, src, f);
          done[src] = f;
          dst = f;
        }
        if (which)
          q.c32 = o.c32 = dst;
        else
          q.g32 = o.g32 = dst;
      }
    }
    OHM_HIP(hipMemcpyAsync(bt->d_geo, bt->h_geo.data(), sizeof(BGeo) * bt->K, hipMemcpyHostToDevice,
                           bt->owner->stream));
    OHM_HIP(hipMemcpyAsync(bt->d_src, bt->h_src.data(), sizeof(BGeo) * bt->K, hipMemcpyHostToDevice,
                           bt->owner->stream));
    bt->has_f32 = true;
  }
  return OHM_OK;
}

ohm_status fill_results(ohm_batch *bt, const std::vector<int64_t> &iters,
                        const std::vector<int64_t> &refs, const std::vector<int> &status,
                        ohm_batch_result *res, int64_t max_iters) {
  // FP64 true residuals of the final x
  ohm_status first = OHM_OK;
  std::string msg;
  for (int k = 0; k < bt->K; k++) {
    ohm_batch_result &r = res[k];
    const double rr = bt->h_res[2 * k], bb = bt->h_res[2 * k + 1];
    r.iterations = iters[k];
    r.refinements = refs[k];
    r.residual_norm = bb > 0 ? sqrt(rr / bb) : 0.0;
    r.reason = status[k];
    if (status[k] == ST_CONVERGED) {
      r.status = OHM_OK;
    } else {
      ohm_status e = status_to_error(status[k], iters[k], max_iters);
      r.status = e;
      if (first == OHM_OK) {
        first = e;
        msg = std::string(ohm_last_error()) + " (batch system " + std::to_string(k) + ")";
      }
    }
  }
  if (first != OHM_OK) return fail(first, "%s", msg.c_str());
  return OHM_OK;
}

template <typename S64, typename S32>
ohm_status batch_solve_t(ohm_batch *bt, const BatchOpts &o, int precision, double inner_tol,
                         int64_t max_ref, ohm_batch_result *res, bool report = true) {
  ohm_backend *b = bt->owner;
  hipStream_t stream = b->stream;
  const int K = bt->K;
  BW<S64> w = ws_view<S64>(bt->ws64, bt->N);
  std::vector<int64_t> iters(K, 0), refs(K, 0);
  std::vector<int> status(K, ST_RUNNING);
  std::vector<char> run(K, 1);
  if (precision == OHM_PRECISION_FP64) {
    OHM_TRY(core_solve<S64>(bt, w, run, o));
    OHM_HIP(hipMemcpyAsync(bt->h_sys, bt->d_sys, sizeof(BSys) * K, hipMemcpyDeviceToHost, stream));
    OHM_HIP(hipStreamSynchronize(stream));
    bool retry = false;
    for (int k = 0; k < K; k++) {
      iters[k] = bt->h_sys[k].iters;
      status[k] = bt->h_sys[k].status;
      const int s = status[k];
      run[k] = s == ST_BRK_RHO || s == ST_BRK_RV || s == ST_BRK_TT || s == ST_BRK_OMEGA;
      retry |= run[k] != 0;
    }
    if (retry && o.warm && o.x_start) {
      for (int k = 0; k < K; k++)
        if (run[k] && bt->h_geo[k].n > 0)
          OHM_HIP(hipMemcpyAsync(w.x + bt->h_geo[k].off,
                                 static_cast<const S64 *>(o.x_start) + bt->h_geo[k].off,
                                 sizeof(S64) * bt->h_geo[k].n, hipMemcpyDeviceToDevice, stream));
      BatchOpts o2 = o;
      o2.shadow = o.shadow == OHM_SHADOW_HASHED ? OHM_SHADOW_RHS : OHM_SHADOW_HASHED;
      OHM_TRY(core_solve<S64>(bt, w, run, o2));
      OHM_HIP(hipMemcpyAsync(bt->h_sys, bt->d_sys, sizeof(BSys) * K, hipMemcpyDeviceToHost, stream));
      OHM_HIP(hipStreamSynchronize(stream));
      for (int k = 0; k < K; k++)
        if (run[k]) {
          iters[k] += bt->h_sys[k].iters;
          status[k] = bt->h_sys[k].status;
        }
    }
  } else {
    // FP32 inner solves inside FP64 iterative refinement, per system
    OHM_TRY(ensure_ws(bt, true));
    BW<S32> w32 = ws_view<S32>(bt->ws32, bt->N);
    if (!o.warm) OHM_HIP(hipMemsetAsync(w.x, 0, sizeof(S64) * bt->N, stream));
    std::vector<double> prev(K, INFINITY);
    std::vector<char> ref_run(K, 1), need64(K, 0);
    std::vector<double> scale(K, 1.0), target(K, 0.0), rel0(K, 0.0);
    OHM_TRY(upload_tables(bt, bt->h_blocks));
    for (int64_t it = 0;; it++) {
      hipLaunchKernelGGL(k_bresid<S64>, dim3(bt->h_blocks.size()), dim3(kBlock), 0, stream,
                         bt->d_active, bt->d_geo, bt->d_sys, w.x, w.b, w.r, bt->d_partials,
                         bt->d_res);
      OHM_HIP(hipMemcpyAsync(bt->h_res, bt->d_res, sizeof(double) * 2 * K, hipMemcpyDeviceToHost,
                             stream));
      OHM_HIP(hipStreamSynchronize(stream));
      b->readbacks++;
      int left = 0;
      for (int k = 0; k < K; k++) {
        if (!ref_run[k]) continue;
        const double rr = sqrt(bt->h_res[2 * k]), bb = sqrt(bt->h_res[2 * k + 1]);
        if (bb < 1e-30) {  // b == 0 -> x = 0
          status[k] = ST_CONVERGED;
          ref_run[k] = 0;
          continue;
        }
        const double rel = rr / bb;
        if (it == 0) {
          rel0[k] = rel;
          target[k] = o.tol_mode == OHM_TOL_INITIAL ? std::max(o.tol * rel, o.tol_floor) : o.tol;
        }
        if (!std::isfinite(rel)) {
          status[k] = ST_NAN;
          ref_run[k] = 0;
        } else if (rel < target[k]) {
          status[k] = ST_CONVERGED;
          ref_run[k] = 0;
        } else if (it >= max_ref || rel > 0.5 * prev[k] || iters[k] >= o.max_iters) {
          need64[k] = 1;  // refinement stalled: finish in FP64
          ref_run[k] = 0;
        } else {
          prev[k] = rel;
          scale[k] = rr;
          left++;
        }
      }
      if (left == 0) break;
      OHM_HIP(hipMemcpyAsync(bt->d_scale, scale.data(), sizeof(double) * K, hipMemcpyHostToDevice,
                             stream));
      std::vector<BBlock> tab = active_table(bt, ref_run);
      OHM_TRY(upload_tables(bt, tab));
      hipLaunchKernelGGL((k_bscale_down<S64, S32>), dim3(tab.size()), dim3(kBlock), 0, stream,
                         bt->d_active, bt->d_geo, w.r, bt->d_scale, w32.b);
      BatchOpts oi = o;
      oi.warm = false;
      oi.tol_mode = OHM_TOL_RHS;
      oi.tol = std::min(0.5, std::max(inner_tol, 1e-6));
      OHM_TRY(core_solve<S32>(bt, w32, ref_run, oi));
      OHM_TRY(upload_tables(bt, tab));
      hipLaunchKernelGGL((k_bcorrect<S64, S32>), dim3(tab.size()), dim3(kBlock), 0, stream,
                         bt->d_active, bt->d_geo, bt->d_sys, w32.x, bt->d_scale, w.x);
      OHM_HIP(hipMemcpyAsync(bt->h_sys, bt->d_sys, sizeof(BSys) * K, hipMemcpyDeviceToHost, stream));
      OHM_HIP(hipStreamSynchronize(stream));
      for (int k = 0; k < K; k++)
        if (ref_run[k]) {
          iters[k] += bt->h_sys[k].iters;
          refs[k]++;
        }
      OHM_TRY(upload_tables(bt, bt->h_blocks));
    }
    bool any64 = false;
    for (int k = 0; k < K; k++) any64 |= need64[k] != 0;
    if (any64) {
      BatchOpts o64 = o;
      o64.warm = true;
      o64.tol_mode = OHM_TOL_RHS;
      OHM_TRY(core_solve<S64>(bt, w, need64, o64));
      OHM_HIP(hipMemcpyAsync(bt->h_sys, bt->d_sys, sizeof(BSys) * K, hipMemcpyDeviceToHost, stream));
      OHM_HIP(hipStreamSynchronize(stream));
      for (int k = 0; k < K; k++)
        if (need64[k]) {
          iters[k] += bt->h_sys[k].iters;
          status[k] = bt->h_sys[k].status;
        }
    }
  }
  if (!report) {  // per-step solves of a transient: statuses only
    for (int k = 0; k < K; k++) bt->h_res[2 * k] = bt->h_res[2 * k + 1] = 0.0;
    return fill_results(bt, iters, refs, status, res, o.max_iters);
  }
  // final FP64 residuals for the report
  OHM_TRY(upload_tables(bt, bt->h_blocks));
  hipLaunchKernelGGL(k_bresid<S64>, dim3(bt->h_blocks.size()), dim3(kBlock), 0, stream,
                     bt->d_active, bt->d_geo, bt->d_sys, w.x, w.b, w.ax, bt->d_partials, bt->d_res);
  OHM_HIP(hipMemcpyAsync(bt->h_res, bt->d_res, sizeof(double) * 2 * K, hipMemcpyDeviceToHost,
                         stream));
  OHM_HIP(hipStreamSynchronize(stream));
  return fill_results(bt, iters, refs, status, res, o.max_iters);
}

}  // namespace

// ================================================================ C ABI ==
extern "C" {

// Batches of at least kBatchTuneRows rows pick their row-block shape by
// timing the batched SpMV (the body every batched BiCGSTAB kernel shares)
// over each candidate table: 3 launches to warm, 10 timed with HIP events.
// The shape only changes how rows are grouped into workgroups (and so the
// order in which a system's per-block dot partials are added).
constexpr int64_t kBatchTuneRows = 65536;

static ohm_status batch_tune_blocks(ohm_batch *bt, const std::vector<std::vector<BBlock>> &tables) {
  ohm_backend *b = bt->owner;
  hipStream_t st = b->stream;
  const size_t eb = elem_bytes(bt->cplx, false);
  const int64_t stride = (bt->N + 63) & ~int64_t(63);
  char *base = reinterpret_cast<char *>(bt->ws64);
  void *x = base, *y = base + stride * eb;
  float best = 1e30f;
  size_t pick = tables.size() - 1;
  for (size_t c = 0; c < tables.size(); c++) {
    const std::vector<BBlock> &tab = tables[c];
    if (tab.empty()) continue;
    OHM_HIP(hipMemcpyAsync(bt->d_full, tab.data(), sizeof(BBlock) * tab.size(),
                           hipMemcpyHostToDevice, st));
    auto launch = [&]() {
      if (bt->cplx)
        hipLaunchKernelGGL(k_bspmv<cx<double>>, dim3(tab.size()), dim3(kBlock), 0, st,
                           bt->d_full, bt->d_geo, (const cx<double> *)x, (cx<double> *)y);
      else
        hipLaunchKernelGGL(k_bspmv<double>, dim3(tab.size()), dim3(kBlock), 0, st, bt->d_full,
                           bt->d_geo, (const double *)x, (double *)y);
    };
    for (int w = 0; w < 3; w++) launch();
    OHM_HIP(hipEventRecord(b->ev0, st));
    for (int r = 0; r < 10; r++) launch();
    OHM_HIP(hipEventRecord(b->ev1, st));
    OHM_HIP(hipEventSynchronize(b->ev1));
    OHM_HIP(hipGetLastError());
    bt->dispatches += 13;
    float ms = 0.f;
    OHM_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    if (ms < best) {
      best = ms;
      pick = c;
    }
  }
  bt->h_blocks = tables[pick];
  for (int k = 0; k < bt->K; k++) bt->h_geo[k].first_block = bt->h_geo[k].n_blocks = 0;
  for (const BBlock &bk : bt->h_blocks) {
    BGeo &g = bt->h_geo[bk.sys];
    if (g.n_blocks == 0) g.first_block = bk.slot;
    g.n_blocks++;
  }
  for (int k = 0; k < bt->K; k++) {
    bt->h_src[k].first_block = bt->h_geo[k].first_block;
    bt->h_src[k].n_blocks = bt->h_geo[k].n_blocks;
  }
  bt->tuned_shape = (int)pick;
  OHM_HIP(hipMemcpyAsync(bt->d_full, bt->h_blocks.data(), sizeof(BBlock) * bt->h_blocks.size(),
                         hipMemcpyHostToDevice, st));
  OHM_HIP(hipMemcpyAsync(bt->d_geo, bt->h_geo.data(), sizeof(BGeo) * bt->K, hipMemcpyHostToDevice,
                         st));
  OHM_HIP(hipMemcpyAsync(bt->d_src, bt->h_src.data(), sizeof(BGeo) * bt->K, hipMemcpyHostToDevice,
                         st));
  OHM_HIP(hipStreamSynchronize(st));
  return OHM_OK;
}

ohm_status ohm_batch_create(ohm_backend *b, int32_t is_complex, int64_t k,
                            const ohm_batch_system *systems, ohm_batch **out) {
  if (!b || !out || (k > 0 && !systems)) return fail(OHM_ERR_INVALID, "null argument");
  if (k <= 0 || k > (1 << 24)) return fail(OHM_ERR_INVALID, "batch size %lld out of range", (long long)k);
  OHM_HIP(hipSetDevice(b->device));
  ohm_batch *bt = new ohm_batch();
  bt->owner = b;
  bt->K = (int)k;
  bt->cplx = is_complex != 0;
  auto bail = [&](ohm_status s) {
    ohm_batch_free(bt);
    return s;
  };
  // dedupe patterns and value arrays by host pointer
  std::map<std::tuple<const void *, const void *, int64_t>, std::pair<uint32_t *, uint32_t *>> pats;
  std::vector<const uint32_t *> host_rowptr;  // per system, for the block-shape autotuner
  std::map<std::pair<const void *, int64_t>, double *> vals;
  int64_t off = 0;
  for (int64_t s = 0; s < k; s++) {
    const ohm_batch_system &q = systems[s];
    if (q.n < 0 || q.n > 0x7fffffff || (q.n > 0 && (!q.rowptr || !q.g_val)))
      return bail(fail(OHM_ERR_INVALID, "batch system %lld: bad size or null arrays", (long long)s));
    const int n = (int)q.n;
    const int64_t nnz = n > 0 ? q.rowptr[n] : 0;
    if (n > 0) {
      if (q.rowptr[0] != 0) return bail(fail(OHM_ERR_INVALID, "batch system %lld: rowptr[0] != 0", (long long)s));
      for (int r = 0; r < n; r++)
        if (q.rowptr[r + 1] < q.rowptr[r])
          return bail(fail(OHM_ERR_INVALID, "batch system %lld: rowptr not monotone", (long long)s));
      if (nnz > 0 && !q.col) return bail(fail(OHM_ERR_INVALID, "batch system %lld: null col", (long long)s));
      for (int64_t e = 0; e < nnz; e++)
        if (q.col[e] >= (uint32_t)n)
          return bail(fail(OHM_ERR_INVALID, "batch system %lld: column index out of range", (long long)s));
    }
    BGeo g{};
    g.n = n;
    g.off = off;
    g.sig_re = q.sigma_re;
    g.sig_im = q.sigma_im;
    auto pk = std::make_tuple((const void *)q.rowptr, (const void *)q.col, (int64_t)n);
    auto it = pats.find(pk);
    if (it == pats.end()) {
      uint32_t *rp = nullptr, *cl = nullptr;
      if (hipMalloc(&rp, sizeof(uint32_t) * (n + 1)) != hipSuccess ||
          hipMalloc(&cl, sizeof(uint32_t) * std::max<int64_t>(nnz, 1)) != hipSuccess)
        return bail(fail(OHM_ERR_SOLVE, "Solve error: out of device memory for batch patterns"));
      bt->allocs.push_back(rp);
      bt->allocs.push_back(cl);
      static const uint32_t zero_rp = 0;
      OHM_HIP(hipMemcpy(rp, n > 0 ? q.rowptr : &zero_rp, sizeof(uint32_t) * (n > 0 ? n + 1 : 1),
                        hipMemcpyHostToDevice));
      if (nnz > 0) OHM_HIP(hipMemcpy(cl, q.col, sizeof(uint32_t) * nnz, hipMemcpyHostToDevice));
      it = pats.emplace(pk, std::make_pair(rp, cl)).first;
    }
    g.rowptr = it->second.first;
    g.col = it->second.second;
    for (int which = 0; which < 2; which++) {
      const double *src = which ? q.c_val : q.g_val;
      if (!src || n == 0) continue;
      auto vk = std::make_pair((const void *)src, nnz);
      auto vt = vals.find(vk);
      if (vt == vals.end()) {
        double *d = nullptr;
        if (hipMalloc(&d, sizeof(double) * std::max<int64_t>(nnz, 1)) != hipSuccess)
          return bail(fail(OHM_ERR_SOLVE, "Solve error: out of device memory for batch values"));
        bt->allocs.push_back(d);
        if (nnz > 0) OHM_HIP(hipMemcpy(d, src, sizeof(double) * nnz, hipMemcpyHostToDevice));
        vt = vals.emplace(vk, d).first;
      }
      if (which)
        g.c = vt->second;
      else
        g.g = vt->second;
    }
    // row blocks of this system
    g.first_block = (int)bt->h_blocks.size();
    if (n > 0) {
      std::vector<uint32_t> bounds;
      build_spmv_blocks(q.rowptr, n, bounds);
      for (size_t i = 0; i + 1 < bounds.size(); i++) {
        BBlock bk{(int)s, bounds[i], bounds[i + 1], (int)bt->h_blocks.size()};
        bt->h_blocks.push_back(bk);
      }
    }
    g.n_blocks = (int)bt->h_blocks.size() - g.first_block;
    bt->h_geo.push_back(g);
    bt->h_nnz.push_back(nnz);
    host_rowptr.push_back(q.rowptr);
    bt->max_n = std::max(bt->max_n, n);
    off += n;
  }
  bt->N = off;
  // Row-block shapes the autotuner below may pick (rows per workgroup; the
  // LDS stages at most 2,048 entries per block): tables for every system,
  // device arrays sized for the largest.
  std::vector<int> tune_rows;
  if (b->autotune && bt->N >= kBatchTuneRows) tune_rows = {64, 128, kSpmvRowsPerBlock};
  std::vector<std::vector<BBlock>> tables;
  size_t nb = std::max<size_t>(bt->h_blocks.size(), 1);
  for (int rows : tune_rows) {
    std::vector<BBlock> tab;
    for (int64_t s2 = 0; s2 < k; s2++) {
      const int n = bt->h_geo[s2].n;
      if (n <= 0) continue;
      std::vector<uint32_t> bounds;
      build_spmv_blocks(host_rowptr[s2], n, bounds, rows, kSpmvNnzPerBlock);
      for (size_t i = 0; i + 1 < bounds.size(); i++)
        tab.push_back(BBlock{(int)s2, bounds[i], bounds[i + 1], (int)tab.size()});
    }
    nb = std::max(nb, tab.size());
    tables.push_back(std::move(tab));
  }
  bt->h_src = bt->h_geo;
  bt->formed.assign(k, nullptr);
  if (hipMalloc(&bt->d_geo, sizeof(BGeo) * k) != hipSuccess ||
      hipMalloc(&bt->d_src, sizeof(BGeo) * k) != hipSuccess ||
      hipMalloc(&bt->d_full, sizeof(BBlock) * nb) != hipSuccess ||
      hipMalloc(&bt->d_active, sizeof(BBlock) * nb) != hipSuccess ||
      hipMalloc(&bt->d_list, sizeof(int) * k) != hipSuccess ||
      hipMalloc(&bt->d_mask, sizeof(int) * k) != hipSuccess ||
      hipMalloc(&bt->d_sys, sizeof(BSys) * k) != hipSuccess ||
      hipMalloc(&bt->d_partials, sizeof(double) * kBatchPartials * nb) != hipSuccess ||
      hipMalloc(&bt->d_res, sizeof(double) * 2 * k) != hipSuccess ||
      hipMalloc(&bt->d_scale, sizeof(double) * k) != hipSuccess ||
      hipHostMalloc((void **)&bt->h_sys, sizeof(BSys) * k, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void **)&bt->h_res, sizeof(double) * 2 * k, hipHostMallocDefault) != hipSuccess)
    return bail(fail(OHM_ERR_SOLVE, "Solve error: out of device memory for the batch"));
  OHM_HIP(hipMemcpy(bt->d_src, bt->h_src.data(), sizeof(BGeo) * k, hipMemcpyHostToDevice));
  {
    ohm_status fs = form_values(bt);
    if (fs != OHM_OK) return bail(fs);
  }
  OHM_HIP(hipMemcpy(bt->d_geo, bt->h_geo.data(), sizeof(BGeo) * k, hipMemcpyHostToDevice));
  if (!bt->h_blocks.empty())
    OHM_HIP(hipMemcpy(bt->d_full, bt->h_blocks.data(), sizeof(BBlock) * bt->h_blocks.size(),
                      hipMemcpyHostToDevice));
  if (hipMalloc(&bt->ws64, ws_bytes(bt, false)) != hipSuccess)
    return bail(fail(OHM_ERR_SOLVE, "Solve error: out of device memory for batch vectors"));
  hipLaunchKernelGGL(k_bsys_reset, dim3((k + 255) / 256), dim3(256), 0, b->stream, (int)k,
                     bt->d_sys, nullptr, 1, 1);
  OHM_HIP(hipStreamSynchronize(b->stream));
  if (!tables.empty()) {
    ohm_status ts = batch_tune_blocks(bt, tables);
    if (ts != OHM_OK) return bail(ts);
  }
  *out = bt;
  return OHM_OK;
}

void ohm_batch_free(ohm_batch *bt) {
  if (!bt) return;
  if (backend_alive(bt->owner)) {
    (void)hipSetDevice(bt->owner->device);
    (void)hipStreamSynchronize(bt->owner->stream);
    for (void *p : bt->allocs) (void)hipFree(p);
    void *ds[] = {bt->d_geo, bt->d_src, bt->d_full, bt->d_active, bt->d_list, bt->d_mask,
                  bt->d_sys, bt->d_partials, bt->d_res, bt->d_scale, bt->ws64, bt->ws32};
    for (void *p : ds)
      if (p) (void)hipFree(p);
    if (bt->h_sys) (void)hipHostFree(bt->h_sys);
    if (bt->h_res) (void)hipHostFree(bt->h_res);
  }
  delete bt;
}

int64_t ohm_batch_rows(const ohm_batch *bt) { return bt ? bt->N : 0; }

ohm_status ohm_batch_set_sigma(ohm_batch *bt, const double *sigma) {
  if (!bt || !sigma) return fail(OHM_ERR_INVALID, "null argument");
  if (!backend_alive(bt->owner)) return fail(OHM_ERR_INVALID, "backend destroyed");
  for (int k = 0; k < bt->K; k++) {
    bt->h_geo[k].sig_re = bt->h_src[k].sig_re = sigma[2 * k];
    bt->h_geo[k].sig_im = bt->h_src[k].sig_im = sigma[2 * k + 1];
  }
  OHM_TRY(form_values(bt));
  OHM_HIP(hipMemcpyAsync(bt->d_geo, bt->h_geo.data(), sizeof(BGeo) * bt->K, hipMemcpyHostToDevice,
                         bt->owner->stream));
  OHM_HIP(hipMemcpyAsync(bt->d_src, bt->h_src.data(), sizeof(BGeo) * bt->K, hipMemcpyHostToDevice,
                         bt->owner->stream));
  OHM_HIP(hipStreamSynchronize(bt->owner->stream));
  bt->jac64_valid = bt->jac32_valid = false;
  return OHM_OK;
}

static ohm_status batch_opts(const ohm_bicgstab_params *prm, int32_t precond, BatchOpts &o) {
  ohm_bicgstab_params d;
  ohm_bicgstab_default_params(&d);
  if (!prm) prm = &d;
  if (prm->max_iterations <= 0) return fail(OHM_ERR_SOLVE, "Solve error: BiCGSTAB max_iterations must be > 0");
  if (!(prm->tolerance > 0.0) || !std::isfinite(prm->tolerance))
    return fail(OHM_ERR_SOLVE, "Solve error: BiCGSTAB tolerance must be finite and > 0");
  if (precond == OHM_PREC_ISAI)
    return fail(OHM_ERR_INVALID, "ISAI is not available for batched systems (use Jacobi or none)");
  if (precond != OHM_PREC_NONE && precond != OHM_PREC_JACOBI && precond != OHM_PREC_AUTO)
    return fail(OHM_ERR_INVALID, "unknown preconditioner %d", precond);
  o.max_iters = prm->max_iterations;
  o.tol = prm->tolerance;
  o.shadow = prm->shadow;
  o.warm = prm->warm_start != 0;
  o.check_interval = prm->check_interval;
  o.scaled = prm->breakdown != OHM_BREAKDOWN_ABSOLUTE;
  o.tol_mode = prm->tol_mode;
  o.tol_floor = prm->tol_floor > 0 ? prm->tol_floor : 1e-15;
  o.precond = precond == OHM_PREC_NONE ? OHM_PREC_NONE : OHM_PREC_JACOBI;
  return OHM_OK;
}

// Solve on device vectors already in the batch workspace layout.
static ohm_status batch_solve_ws(ohm_batch *bt, const ohm_bicgstab_params *prm, int32_t precond,
                                 ohm_batch_result *res) {
  BatchOpts o;
  OHM_TRY(batch_opts(prm, precond, o));
  ohm_bicgstab_params d;
  ohm_bicgstab_default_params(&d);
  if (!prm) prm = &d;
  const int prec = prm->precision;
  if (prec != OHM_PRECISION_FP64 && prec != OHM_PRECISION_FP32_IR)
    return fail(OHM_ERR_INVALID, "unknown precision %d", prec);
  const double itol = prm->inner_tolerance > 0 ? prm->inner_tolerance : 1e-4;
  const int64_t mref = prm->max_refinements > 0 ? prm->max_refinements : 20;
  if (bt->cplx)
    return batch_solve_t<cx<double>, cx<float>>(bt, o, prec, itol, mref, res);
  return batch_solve_t<double, float>(bt, o, prec, itol, mref, res);
}

ohm_status ohm_batch_solve(ohm_batch *bt, const double *rhs, double *x,
                           const ohm_bicgstab_params *prm, int32_t precond,
                           ohm_batch_result *res) {
  if (!bt || !rhs || !x || !res) return fail(OHM_ERR_INVALID, "null argument");
  if (!backend_alive(bt->owner)) return fail(OHM_ERR_INVALID, "backend destroyed");
  OHM_HIP(hipSetDevice(bt->owner->device));
  hipStream_t stream = bt->owner->stream;
  const size_t eb = elem_bytes(bt->cplx, false);
  const int64_t stride = (bt->N + 63) & ~int64_t(63);
  char *base = reinterpret_cast<char *>(bt->ws64);
  char *dx = base, *db = base + 10 * stride * eb;
  OHM_HIP(hipMemcpyAsync(db, rhs, eb * bt->N, hipMemcpyHostToDevice, stream));
  const bool warm = prm && prm->warm_start;
  if (warm) OHM_HIP(hipMemcpyAsync(dx, x, eb * bt->N, hipMemcpyHostToDevice, stream));
  ohm_status s = batch_solve_ws(bt, prm, precond, res);
  if (s != OHM_OK && s != OHM_ERR_SOLVE) return s;  // hard errors: x untouched
  std::string keep = s == OHM_OK ? "" : ohm_last_error();
  hipError_t e = hipMemcpyAsync(x, dx, eb * bt->N, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync", __FILE__, __LINE__);
  if (s != OHM_OK) return fail(s, "%s", keep.c_str());
  return OHM_OK;
}

ohm_status ohm_batch_spmv(ohm_batch *bt, const double *x, double *y) {
  if (!bt || !x || !y) return fail(OHM_ERR_INVALID, "null argument");
  if (!backend_alive(bt->owner)) return fail(OHM_ERR_INVALID, "backend destroyed");
  OHM_HIP(hipSetDevice(bt->owner->device));
  hipStream_t stream = bt->owner->stream;
  const size_t eb = elem_bytes(bt->cplx, false);
  const int64_t stride = (bt->N + 63) & ~int64_t(63);
  char *base = reinterpret_cast<char *>(bt->ws64);
  char *dxv = base + 3 * stride * eb, *dyv = base + 4 * stride * eb;  // p, v slots
  OHM_HIP(hipMemcpyAsync(dxv, x, eb * bt->N, hipMemcpyHostToDevice, stream));
  if (bt->cplx)
    OHM_TRY(run_spmv<cx<double>>(bt, (const cx<double> *)dxv, (cx<double> *)dyv));
  else
    OHM_TRY(run_spmv<double>(bt, (const double *)dxv, (double *)dyv));
  OHM_HIP(hipMemcpyAsync(y, dyv, eb * bt->N, hipMemcpyDeviceToHost, stream));
  OHM_HIP(hipStreamSynchronize(stream));
  return OHM_OK;
}

ohm_status ohm_solve_complex(ohm_backend *b, size_t n, size_t nnz, const double *values,
                             const uint32_t *col_indices, const uint32_t *row_pointers,
                             const double *rhs, double *x, int precond,
                             const ohm_bicgstab_params *params, ohm_bicgstab_result *result) {
  if (!b) return fail(OHM_ERR_INVALID, "null backend");
  if (n == 0) return OHM_OK;  // src/solver/gpu.rs:221-223
  if (!values || !col_indices || !row_pointers || !rhs || !x)
    return fail(OHM_ERR_INVALID, "null argument");
  if (n > 0x7fffffffu || row_pointers[n] != nnz)
    return fail(OHM_ERR_SOLVE, "Solve error: dimension mismatch: row_pointers[n] = %u, nnz = %zu",
                row_pointers[n], nnz);
  // split A = Re + j Im: g = Re, c = Im, sigma = j
  std::vector<double> re(nnz), im(nnz);
  for (size_t k = 0; k < nnz; k++) {
    re[k] = values[2 * k];
    im[k] = values[2 * k + 1];
  }
  ohm_batch_system sys{(int64_t)n, row_pointers, col_indices, re.data(), im.data(), 0.0, 1.0};
  ohm_batch *bt = nullptr;
  OHM_TRY(ohm_batch_create(b, 1, 1, &sys, &bt));
  ohm_batch_result r{};
  const int64_t d0 = b->dispatches, rb0 = b->readbacks;
  (void)hipEventRecord(b->ev0, b->stream);
  ohm_status s = ohm_batch_solve(bt, rhs, x, params, precond, &r);
  std::string keep = s == OHM_OK ? "" : ohm_last_error();
  (void)hipEventRecord(b->ev1, b->stream);
  (void)hipEventSynchronize(b->ev1);
  if (result) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
    result->iterations = r.iterations;
    result->refinements = r.refinements;
    result->residual_norm = r.residual_norm;
    result->device_ms = ms;
    result->dispatches = b->dispatches - d0;
    result->readbacks = b->readbacks - rb0;
    result->fp64_fallback = 0;
  }
  ohm_batch_free(bt);
  if (s != OHM_OK) return fail(s, "%s", keep.c_str());
  return OHM_OK;
}

}  // extern "C"

// Union of two CSR patterns with G and C values scattered onto it.
ohm_status ohm::csr_union(int64_t n, const uint32_t *grp, const uint32_t *gcl,
                                const double *gv, const uint32_t *crp, const uint32_t *ccl,
                                const double *cv, std::vector<uint32_t> &rp,
                                std::vector<uint32_t> &cl, std::vector<double> &g,
                                std::vector<double> &c) {
  for (int64_t r = 0; r < n; r++) {
    for (uint32_t a = grp[r] + 1; a < grp[r + 1]; a++)
      if (gcl[a] < gcl[a - 1]) return fail(OHM_ERR_INVALID, "row %lld of G is not sorted by column", (long long)r);
    if (crp)
      for (uint32_t a = crp[r] + 1; a < crp[r + 1]; a++)
        if (ccl[a] < ccl[a - 1]) return fail(OHM_ERR_INVALID, "row %lld of C is not sorted by column", (long long)r);
  }
  rp.assign(n + 1, 0);
  cl.clear();
  g.clear();
  c.clear();
  for (int64_t r = 0; r < n; r++) {
    uint32_t a = grp[r], ae = grp[r + 1], q = crp ? crp[r] : 0, qe = crp ? crp[r + 1] : 0;
    while (a < ae || q < qe) {
      uint32_t ca = a < ae ? gcl[a] : UINT32_MAX, cc = q < qe ? ccl[q] : UINT32_MAX;
      uint32_t col = std::min(ca, cc);
      if (col >= (uint32_t)n) return fail(OHM_ERR_INVALID, "column index out of range in row %lld", (long long)r);
      double gval = 0.0, cval = 0.0;
      // duplicates within a row are summed
      while (a < ae && gcl[a] == col) gval += gv[a++];
      while (q < qe && ccl[q] == col) cval += cv[q++];
      cl.push_back(col);
      g.push_back(gval);
      c.push_back(cval);
    }
    rp[r + 1] = (uint32_t)cl.size();
  }
  return OHM_OK;
}

extern "C" {

ohm_status ohm_ac_sweep(ohm_backend *b, int64_t n, const uint32_t *g_rowptr,
                        const uint32_t *g_col, const double *g_val, const uint32_t *c_rowptr,
                        const uint32_t *c_col, const double *c_val, const double *b_ac,
                        int64_t n_freqs, const double *freqs, int64_t batch,
                        const ohm_bicgstab_params *params, int32_t precond, double *x,
                        ohm_batch_result *results) {
  if (!b) return fail(OHM_ERR_INVALID, "null backend");
  if (n < 0 || n_freqs < 0) return fail(OHM_ERR_INVALID, "negative size");
  if (n == 0 || n_freqs == 0) return OHM_OK;
  if (!g_rowptr || !g_val || !b_ac || !freqs || !x || !results || (c_val && !c_rowptr))
    return fail(OHM_ERR_INVALID, "null argument");
  std::vector<uint32_t> rp, cl;
  std::vector<double> g, c;
  OHM_TRY(csr_union(n, g_rowptr, g_col, g_val, c_val ? c_rowptr : nullptr, c_col, c_val, rp,
                    cl, g, c));
  const int64_t per = batch > 0 ? std::min(batch, n_freqs) : n_freqs;
  ohm_status first = OHM_OK;
  std::string msg;
  std::vector<double> rhs;
  for (int64_t f0 = 0; f0 < n_freqs; f0 += per) {
    const int64_t kk = std::min(per, n_freqs - f0);
    std::vector<ohm_batch_system> sys(kk);
    for (int64_t k = 0; k < kk; k++) {
      const double omega = 2.0 * M_PI * freqs[f0 + k];  // ac.rs:57
      sys[k] = ohm_batch_system{n, rp.data(), cl.data(), g.data(), c.data(), 0.0, omega};
    }
    ohm_batch *bt = nullptr;
    OHM_TRY(ohm_batch_create(b, 1, kk, sys.data(), &bt));
    rhs.resize((size_t)kk * 2 * n);
    for (int64_t k = 0; k < kk; k++) std::copy(b_ac, b_ac + 2 * n, rhs.begin() + k * 2 * n);
    ohm_status s = ohm_batch_solve(bt, rhs.data(), x + f0 * 2 * n, params, precond, results + f0);
    if (s != OHM_OK && first == OHM_OK) {
      first = s;
      msg = ohm_last_error();
    }
    ohm_batch_free(bt);
    if (s != OHM_OK && s != OHM_ERR_SOLVE) break;  // hard error
  }
  if (first != OHM_OK) return fail(first, "%s", msg.c_str());
  return OHM_OK;
}

int64_t ohm_generate_frequencies(int32_t sweep, int64_t n_points, double f_start, double f_stop,
                                 double *out, int64_t cap) {
  std::vector<double> f;
  if (sweep == OHM_SWEEP_DEC || sweep == OHM_SWEEP_OCT) {
    const double span = sweep == OHM_SWEEP_DEC ? log10(f_stop / f_start) : log2(f_stop / f_start);
    const double tot = ceil((double)n_points * span);
    const int64_t total = tot > 0 ? (int64_t)tot : 0;  // `as usize` saturates at 0
    const double base = sweep == OHM_SWEEP_DEC ? 10.0 : 2.0;
    for (int64_t i = 0; i < total; i++)
      f.push_back(f_start * pow(base, (double)i / (double)n_points));
  } else if (sweep == OHM_SWEEP_LIN) {
    if (n_points <= 1) {
      f.push_back(f_start);
    } else {
      const double step = (f_stop - f_start) / (double)(n_points - 1);
      for (int64_t i = 0; i < n_points; i++) f.push_back(f_start + step * (double)i);
    }
  } else {
    fail(OHM_ERR_INVALID, "unknown sweep type %d", sweep);
    return -1;
  }
  if (out)
    for (int64_t i = 0; i < (int64_t)f.size() && i < cap; i++) out[i] = f[i];
  return (int64_t)f.size();
}

}  // extern "C"

// ============================================================ transient ==
// transient::run (src/analysis/transient.rs:25-199) on the batch engine.
// The companion matrices A = G + (alpha/h) C (form_companion_matrix,
// :285-323) are the batch's on-the-fly G + sigma C; the companion right-hand
// sides b(t) + (alpha/h) C x - [TRAP] G x + [TRAP] b(t_prev)
// (form_companion_rhs_be/_trap, :326-352) and the source values
// (evaluate_b_at_time, :265-280; transient_source.rs) are one fused kernel
// per step, so a timestep never leaves the GPU: rhs -> warm-started solve
// (x_n is already in the iterate) -> probe gather.
namespace ohm {

// A HIP call's error as an ohm_status (for code that must clean up or
// record the status instead of returning at once).
#define OHM_HIP_STATUS(call) ohm::hip_status_of((call), #call, __FILE__, __LINE__)
static inline ohm_status hip_status_of(hipError_t e, const char *what, const char *file, int line) {
  return e == hipSuccess ? OHM_OK : hip_fail(e, what, file, line);
}

struct TRhsArgs {
  double t_next, t_prev;
  double sig_c;   // alpha / h
  int trap;       // TRAP: subtract G x and add b(t_prev)
};

// Device state of a graph-captured segment of fixed-size backward-Euler
// steps. Each step's closing kernel accepts the step (time, counters, probe
// record) when its solve converged; otherwise it sets `stop`, and every
// later kernel of the segment exits at its head: 1 = the step's solve is
// still running (the host finishes it), 2 = it failed with `status`.
struct SegWord {
  int stop;
  int status;
  int steps;       // steps accepted in this segment
  int n_rec;       // probe records written (the run's running count)
  long long iters; // BiCGSTAB iterations of the accepted steps
  int max_iters;   // most iterations one accepted step took
  int pad;
  double t;        // time of the last accepted step
};

// b_eff for every row of every instance; sources are gathered per row in
// source order (src_rowptr / src_ent / src_slot), the order in which
// evaluate_b_at_time accumulates them.
__global__ __launch_bounds__(kBlock) void k_tran_rhs(const BBlock *__restrict__ blocks,
                                                     const BGeo *__restrict__ geo,
                                                     const double *__restrict__ x,
                                                     double *__restrict__ b,
                                                     const int *__restrict__ srp,
                                                     const int *__restrict__ sent,
                                                     const int *__restrict__ sslot,
                                                     const TSrc *__restrict__ src,
                                                     const double *__restrict__ pwl, TRhsArgs a,
                                                     const SegWord *seg, double h) {
  if (seg) {  // a captured step: t from the segment word, nothing once it stopped
    if (seg->stop) return;
    a.t_prev = seg->t;
    a.t_next = seg->t + h;
  }
  const BBlock bk = blocks[blockIdx.x];
  const BGeo q = geo[bk.sys];
  const double *xs = x + q.off;
  for (uint32_t r = bk.r0 + threadIdx.x; r < bk.r1; r += kBlock) {
    double bn = 0.0, bp = 0.0;
    for (int e = srp[r]; e < srp[r + 1]; e++) {
      const TSrc &s = src[sent[e]];
      const double sg = s.sign[sslot[e]];
      bn += sg * src_value(s, pwl, a.t_next);
      if (a.trap) bp += sg * src_value(s, pwl, a.t_prev);
    }
    double cx = 0.0, gx = 0.0;
    for (uint32_t k = q.rowptr[r]; k < q.rowptr[r + 1]; k++) {
      const double xv = xs[q.col[k]];
      if (q.c) cx += q.c[k] * xv;
      if (a.trap) gx += q.g[k] * xv;
    }
    b[q.off + r] = a.trap ? bn + a.sig_c * cx - gx + bp : bn + a.sig_c * cx;
  }
}

// Start of a fixed BE step's solve, extrapolated from the last solutions:
// x <- x_n + ca (x_n - x_{n-1}) + cb (x_{n-1} - x_{n-2}) (ca = cb = 0 on the
// first step: the plain warm start from x_n), then x_{n-2} <- x_{n-1},
// x_{n-1} <- x_n (p2 may be null: linear extrapolation, cb = 0). Launched
// after k_tran_rhs, which reads x_n. start (if set) receives a copy of the
// new starting x for a second attempt of a broken-down batched solve. In a
// captured step (seg set) it does nothing once the segment has stopped.
__global__ void k_tran_predict(int64_t n, double ca, double cb, double *__restrict__ x,
                               double *__restrict__ p1, double *__restrict__ p2,
                               double *__restrict__ start, const SegWord *seg) {
  if (seg && seg->stop) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double xn = x[i];
    // history not yet written is never read (its coefficient is 0)
    const double x1 = ca != 0.0 || cb != 0.0 ? p1[i] : xn;
    const double x2 = cb != 0.0 ? p2[i] : x1;
    if (p2) p2[i] = x1;
    p1[i] = xn;
    const double x0 = fma(cb, x1 - x2, fma(ca, xn - x1, xn));
    x[i] = x0;
    if (start) start[i] = x0;
  }
}

__global__ void k_tran_probe(int K, int np, const uint32_t *__restrict__ probes,
                             const BGeo *__restrict__ geo, const double *__restrict__ x,
                             double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K * np) return;
  const int k = i / np, j = i % np;
  out[i] = x[geo[k].off + probes[j]];
}

// Closing kernel of a captured step (one wave): accept the step when its
// solve converged (t += h as the host loop adds it, probe record while
// t >= tstart and records remain), else stop the segment.
__global__ __launch_bounds__(64) void k_seg_close(SegWord *seg, const SolveState *st, double h,
                                                  double tstart, int max_records, int np,
                                                  const uint32_t *__restrict__ probes,
                                                  const double *__restrict__ x,
                                                  double *__restrict__ rec) {
  const SegWord w = *seg;  // every lane reads before lane 0 writes (one wave)
  if (w.stop) return;
  const int status = st->status;
  if (status != ST_CONVERGED) {
    if (threadIdx.x == 0) {
      seg->stop = status == ST_RUNNING ? 1 : 2;
      seg->status = status;
    }
    return;
  }
  const double t_next = w.t + h;
  const bool keep = t_next >= tstart && w.n_rec < max_records;
  if (keep)
    for (int j = threadIdx.x; j < np; j += 64) rec[(size_t)w.n_rec * np + j] = x[probes[j]];
  if (threadIdx.x == 0) {
    seg->t = t_next;
    seg->steps = w.steps + 1;
    seg->iters = w.iters + st->iters;
    seg->max_iters = max(w.max_iters, st->iters);
    if (keep) seg->n_rec = w.n_rec + 1;
  }
}

// max |a - b| over n entries into *out (bit pattern of a non-negative
// double, so the integer max is the float max and order-independent)
__global__ void k_maxdiff(int64_t n, const double *__restrict__ a, const double *__restrict__ b,
                          unsigned long long *out) {
  double m = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double d = fabs(a[i] - b[i]);
    m = (d > m || d != d) ? d : m;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(m, off, 64);
    m = (o > m || o != o) ? o : m;
  }
  if ((threadIdx.x & 63) == 0)
    __hip_atomic_fetch_max(out, (unsigned long long)__double_as_longlong(m), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// a = g + sigma c (the single-instance companion matrix values)
__global__ void k_axpby_vals(int64_t nnz, const double *__restrict__ g,
                             const double *__restrict__ c, double sigma, double *__restrict__ a) {
  for (int64_t k = blockIdx.x * (int64_t)kBlock + threadIdx.x; k < nnz;
       k += (int64_t)gridDim.x * kBlock)
    a[k] = c ? g[k] + sigma * c[k] : g[k];
}

}  // namespace ohm

struct ohm_transient {
  ohm_backend *owner = nullptr;
  ohm_batch *bt = nullptr;
  int K = 0;
  int64_t n = 0;  // rows per instance (all instances share the source table)
  int *d_srp = nullptr, *d_sent = nullptr, *d_sslot = nullptr;
  TSrc *d_src = nullptr;
  double *d_pwl = nullptr;
  uint32_t *d_probes = nullptr;
  int probes_cap = 0;
  double *d_rec = nullptr;      // recorded probe values
  int64_t rec_cap = 0;
  // adaptive mode: x_n and x_trap; fixed BE: d_save holds x_{n-1} for the
  // extrapolated start of the next solve
  double *d_save = nullptr, *d_trap = nullptr;
  double *d_prev2 = nullptr;  // fixed BE: x_{n-2} (second-order extrapolation)
  unsigned long long *d_max = nullptr;
  // one instance: the companion matrix as a CSR of its own, solved by the
  // single-system BiCGSTAB (ISAI / Jacobi, hipGraph chunks, FP32-IR); its
  // preconditioner is built once per step size
  std::vector<uint32_t> h_rowptr, h_col;
  DevCsr A1;
  PrecondDev P1;
  int p1_kind = -1;
  double a1_sigma = NAN;
  // iterations of the last step's FP64 solve, or of its first inner FP32
  // solve under FP32_IR (0: none)
  int64_t last_iters = 0;
  // graph-captured segments of fixed BE steps (one instance, FP64): the
  // solve's own workspace, the segment word and the captured segments of
  // this run, by (steps, iterations per step)
  CapturedSolve *cs = nullptr;
  SegWord *d_seg = nullptr;
  struct SegGraph {
    int steps, iters;
    hipGraph_t graph;
    hipGraphExec_t exec;
  };
  std::vector<SegGraph> segs;
};

namespace {

ohm_status tran_rhs(ohm_transient *tr, double t_next, double t_prev, double sig_c, bool trap) {
  ohm_batch *bt = tr->bt;
  const int64_t stride = (bt->N + 63) & ~int64_t(63);
  double *x = reinterpret_cast<double *>(bt->ws64);
  double *b = x + 10 * stride;
  TRhsArgs a{t_next, t_prev, sig_c, trap ? 1 : 0};
  hipLaunchKernelGGL(k_tran_rhs, dim3(bt->h_blocks.size()), dim3(kBlock), 0, bt->owner->stream,
                     bt->d_full, bt->d_src, x, b, tr->d_srp, tr->d_sent, tr->d_sslot, tr->d_src,
                     tr->d_pwl, a, (const SegWord *)nullptr, 0.0);
  bt->owner->dispatches++;
  return OHM_OK;
}

// One instance: refresh A1 = G + sig C (and its preconditioner) when the
// step size changes, then solve A1 x = b_eff warm-started from x_n.
ohm_status tran_single_solve(ohm_transient *tr, double sig, int precond, const BicgOpts &o,
                             int precision, double inner_tol, int64_t max_ref, int64_t &iters) {
  ohm_backend *b = tr->owner;
  ohm_batch *bt = tr->bt;
  const BGeo &q = bt->h_src[0];
  const int64_t nnz = bt->h_nnz[0];
  if (!tr->A1.rowptr)
    OHM_TRY(csr_upload(b, tr->A1, nullptr, tr->h_col.data(), tr->h_rowptr.data(), (int)tr->n, nnz));
  if (!(tr->a1_sigma == sig)) {
    if (nnz > 0)
      hipLaunchKernelGGL(k_axpby_vals, dim3(grid_for(nnz)), dim3(kBlock), 0, b->stream, nnz, q.g,
                         q.c, sig, tr->A1.val);
    if (tr->p1_kind < 0) {
      OHM_TRY(precond_build(b, tr->A1, precond, 1, tr->P1));
      tr->p1_kind = precond;
    } else {
      OHM_TRY(precond_refresh(b, tr->A1, tr->P1));
    }
    if (precision == OHM_PRECISION_FP32_IR) OHM_TRY(csr_ensure_f32(b, tr->A1));
    tr->a1_sigma = sig;
  }
  const int64_t stride = (bt->N + 63) & ~int64_t(63);
  double *x = reinterpret_cast<double *>(bt->ws64);
  double *rhs = x + 10 * stride;
  BicgOut out;
  ohm_status s;
  if (precision == OHM_PRECISION_FP32_IR) {
    // the last step's first inner solve sizes this one's first chunks
    int64_t refs = 0;
    BicgOpts oe = o;
    oe.expect_iters = tr->last_iters;
    s = bicg_solve_ir(b, tr->A1, tr->P1, rhs, x, oe, inner_tol, max_ref, out, refs);
    tr->last_iters = s == OHM_OK && !out.fp64_fallback ? out.first_inner : 0;
  } else {
    // the last step's solve length sizes this one's chunks
    BicgOpts oe = o;
    oe.expect_iters = tr->last_iters;
    s = bicg_solve_f64(b, tr->A1, tr->P1, rhs, x, oe, out);
    tr->last_iters = s == OHM_OK ? out.iters : 0;
  }
  iters += out.iters;
  return s;
}

ohm_status tran_set_sigma(ohm_transient *tr, double sig) {
  if (tr->K == 1 && tr->n > 0) return OHM_OK;  // the single path forms A1 itself
  std::vector<double> s(2 * tr->K, 0.0);
  for (int k = 0; k < tr->K; k++) s[2 * k] = sig;
  bool same = true;
  for (int k = 0; k < tr->K; k++) same &= tr->bt->h_geo[k].sig_re == sig;
  if (same) return OHM_OK;
  return ohm_batch_set_sigma(tr->bt, s.data());
}

}  // namespace

extern "C" {

ohm_status ohm_transient_create(ohm_backend *b, int64_t k, const ohm_batch_system *instances,
                                int64_t n_sources, const ohm_tran_source *sources,
                                int64_t n_pwl, const double *pwl_pairs, ohm_transient **out) {
  if (!b || !out || !instances || (n_sources > 0 && !sources)) return fail(OHM_ERR_INVALID, "null argument");
  if (k <= 0) return fail(OHM_ERR_INVALID, "need at least one instance");
  const int64_t n = instances[0].n;
  for (int64_t i = 1; i < k; i++)
    if (instances[i].n != n) return fail(OHM_ERR_INVALID, "instances must have the same size");
  std::vector<ohm_batch_system> sys(instances, instances + k);
  for (auto &s : sys) s.sigma_re = s.sigma_im = 0.0;
  ohm_transient *tr = new ohm_transient();
  tr->owner = b;
  tr->K = (int)k;
  tr->n = n;
  auto bail = [&](ohm_status s) {
    ohm_transient_free(tr);
    return s;
  };
  ohm_status s = ohm_batch_create(b, 0, k, sys.data(), &tr->bt);
  if (s != OHM_OK) return bail(s);
  if (k == 1 && n > 0) {
    tr->h_rowptr.assign(instances[0].rowptr, instances[0].rowptr + n + 1);
    tr->h_col.assign(instances[0].col, instances[0].col + tr->h_rowptr[n]);
  }
  // per-row source gather lists in source order
  std::vector<std::vector<std::pair<int, int>>> rows(n);
  std::vector<TSrc> src(n_sources);
  for (int64_t i = 0; i < n_sources; i++) {
    const ohm_tran_source &q = sources[i];
    if (q.n_idx < 0 || q.n_idx > 2) return bail(fail(OHM_ERR_INVALID, "source %lld: n_idx must be 0..2", (long long)i));
    if (q.kind < OHM_SRC_DC || q.kind > OHM_SRC_PWL) return bail(fail(OHM_ERR_INVALID, "source %lld: unknown kind", (long long)i));
    if (q.kind == OHM_SRC_PWL && (q.pwl_offset < 0 || q.pwl_count < 0 || q.pwl_offset + q.pwl_count > n_pwl))
      return bail(fail(OHM_ERR_INVALID, "source %lld: PWL pairs out of range", (long long)i));
    TSrc t{};
    t.kind = q.kind;
    t.n_idx = q.n_idx;
    for (int j = 0; j < 2; j++) {
      t.idx[j] = q.idx[j];
      t.sign[j] = q.sign[j];
    }
    t.dc = q.dc;
    for (int j = 0; j < 7; j++) t.p[j] = q.param[j];
    t.pwl_off = (int)q.pwl_offset;
    t.pwl_n = (int)q.pwl_count;
    src[i] = t;
    for (int j = 0; j < q.n_idx; j++) {
      if (q.idx[j] >= (uint64_t)n) return bail(fail(OHM_ERR_INVALID, "source %lld: index out of range", (long long)i));
      rows[q.idx[j]].push_back({(int)i, j});
    }
  }
  std::vector<int> srp(n + 1, 0), sent, sslot;
  for (int64_t r = 0; r < n; r++) {
    for (auto &e : rows[r]) {
      sent.push_back(e.first);
      sslot.push_back(e.second);
    }
    srp[r + 1] = (int)sent.size();
  }
  auto up = [&](void **d, const void *h, size_t bytes) -> ohm_status {
    OHM_HIP(hipMalloc(d, std::max<size_t>(bytes, 16)));
    if (bytes) OHM_HIP(hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice));
    return OHM_OK;
  };
  if ((s = up((void **)&tr->d_srp, srp.data(), sizeof(int) * srp.size())) != OHM_OK ||
      (s = up((void **)&tr->d_sent, sent.data(), sizeof(int) * sent.size())) != OHM_OK ||
      (s = up((void **)&tr->d_sslot, sslot.data(), sizeof(int) * sslot.size())) != OHM_OK ||
      (s = up((void **)&tr->d_src, src.data(), sizeof(TSrc) * src.size())) != OHM_OK ||
      (s = up((void **)&tr->d_pwl, pwl_pairs, sizeof(double) * 2 * std::max<int64_t>(n_pwl, 0))) != OHM_OK)
    return bail(s);
  const size_t nb = sizeof(double) * std::max<int64_t>(tr->bt->N, 1);
  if (hipMalloc(&tr->d_save, nb) != hipSuccess || hipMalloc(&tr->d_trap, nb) != hipSuccess ||
      hipMalloc(&tr->d_prev2, nb) != hipSuccess ||
      hipMalloc(&tr->d_max, sizeof(unsigned long long)) != hipSuccess)
    return bail(fail(OHM_ERR_SOLVE, "Solve error: out of device memory for the transient"));
  *out = tr;
  return OHM_OK;
}

void ohm_transient_free(ohm_transient *tr) {
  if (!tr) return;
  if (tr->bt && backend_alive(tr->owner)) {
    void *ds[] = {tr->d_srp, tr->d_sent, tr->d_sslot, tr->d_src, tr->d_pwl, tr->d_probes,
                  tr->d_rec, tr->d_save, tr->d_trap, tr->d_prev2, tr->d_max};
    (void)hipStreamSynchronize(tr->owner->stream);
    for (void *p : ds)
      if (p) (void)hipFree(p);
    graph_cache_clear(tr->owner);
    for (auto &g : tr->segs) {
      (void)hipGraphExecDestroy(g.exec);
      (void)hipGraphDestroy(g.graph);
    }
    captured_free(tr->cs);
    if (tr->d_seg) (void)hipFree(tr->d_seg);
    precond_free(tr->P1);
    csr_free(tr->A1);
  }
  ohm_batch_free(tr->bt);
  delete tr;
}

void ohm_tran_default_params(ohm_tran_params *p) {
  if (!p) return;
  *p = ohm_tran_params{};
  p->tstep = 1e-12;
  p->tstop = 1e-9;
  p->tstart = 0.0;
  p->method = OHM_TRAN_FIXED_BE;
  p->precond = OHM_PREC_JACOBI;
  ohm_bicgstab_default_params(&p->solver);
  p->solver.tolerance = 1e-10;
}

ohm_status ohm_transient_run(ohm_transient *tr, const double *x0, const ohm_tran_params *prm,
                             int64_t n_probes, const uint32_t *probes, int64_t max_records,
                             double *times, double *waves, double *x_final,
                             ohm_tran_stats *stats) {
  if (!tr || !prm) return fail(OHM_ERR_INVALID, "null argument");
  if (!backend_alive(tr->owner)) return fail(OHM_ERR_INVALID, "backend destroyed");
  if (!(prm->tstep > 0.0) || !(prm->tstop > 0.0)) return fail(OHM_ERR_TIMESTEP, "Timestep error: tstep and tstop must be > 0");
  if (prm->method == OHM_TRAN_ADAPTIVE && tr->K != 1)
    return fail(OHM_ERR_INVALID, "adaptive TRAP/BE stepping runs one instance at a time");
  if (n_probes < 0 || (n_probes > 0 && !probes)) return fail(OHM_ERR_INVALID, "bad probe list");
  for (int64_t j = 0; j < n_probes; j++)
    if (probes[j] >= (uint64_t)tr->n) return fail(OHM_ERR_INVALID, "probe index out of range");
  ohm_backend *b = tr->owner;
  ohm_batch *bt = tr->bt;
  hipStream_t stream = b->stream;
  OHM_HIP(hipSetDevice(b->device));
  const int K = tr->K;
  const int64_t N = bt->N, stride = (N + 63) & ~int64_t(63);
  double *wx = reinterpret_cast<double *>(bt->ws64);
  // probes and record buffer
  if (n_probes > tr->probes_cap) {
    if (tr->d_probes) OHM_HIP(hipFree(tr->d_probes));
    OHM_HIP(hipMalloc(&tr->d_probes, sizeof(uint32_t) * n_probes));
    tr->probes_cap = (int)n_probes;
  }
  if (n_probes) OHM_HIP(hipMemcpyAsync(tr->d_probes, probes, sizeof(uint32_t) * n_probes, hipMemcpyHostToDevice, stream));
  const int64_t rec_need = std::max<int64_t>(max_records, 0) * K * n_probes;
  if (rec_need > tr->rec_cap) {
    if (tr->d_rec) OHM_HIP(hipFree(tr->d_rec));
    OHM_HIP(hipMalloc(&tr->d_rec, sizeof(double) * rec_need));
    tr->rec_cap = rec_need;
  }
  if (x0)
    OHM_HIP(hipMemcpyAsync(wx, x0, sizeof(double) * N, hipMemcpyHostToDevice, stream));
  else
    OHM_HIP(hipMemsetAsync(wx, 0, sizeof(double) * N, stream));
  int64_t n_rec = 0;
  std::vector<double> rec_t;
  auto record = [&](double t) -> ohm_status {
    if (n_rec >= max_records) return OHM_OK;
    if (n_probes > 0) {
      const int tot = (int)(K * n_probes);
      hipLaunchKernelGGL(k_tran_probe, dim3((tot + 255) / 256), dim3(256), 0, stream, K,
                         (int)n_probes, tr->d_probes, bt->d_geo, wx,
                         tr->d_rec + n_rec * K * n_probes);
    }
    rec_t.push_back(t);
    n_rec++;
    return OHM_OK;
  };
  ohm_tran_stats st{};
  tr->last_iters = 0;
  std::vector<ohm_batch_result> res(K);
  ohm_bicgstab_params sp = prm->solver;
  sp.warm_start = 1;
  BatchOpts o;
  const bool single_inst = tr->K == 1 && tr->n > 0;
  OHM_TRY(batch_opts(&sp, single_inst ? OHM_PREC_JACOBI : prm->precond, o));
  const double itol = sp.inner_tolerance > 0 ? sp.inner_tolerance : 1e-4;
  const int64_t mref = sp.max_refinements > 0 ? sp.max_refinements : 20;
  const bool single = K == 1 && tr->n > 0;
  if (!single && (prm->precond == OHM_PREC_ISAI))
    return fail(OHM_ERR_INVALID, "ISAI is not available for several instances (use Jacobi or none)");
  BicgOpts so;
  so.max_iters = sp.max_iterations;
  so.tol = sp.tolerance;
  so.shadow = sp.shadow;
  so.warm = true;
  so.check_interval = sp.check_interval;
  so.scaled_breakdown = sp.breakdown != OHM_BREAKDOWN_ABSOLUTE;
  so.tol_mode = sp.tol_mode;
  so.tol_floor = sp.tol_floor > 0 ? sp.tol_floor : 1e-15;
  so.retry_breakdown = true;
  double cur_sigma = 0.0;
  int64_t last_batch = 0;  // iterations of the last step's slowest batched system
  auto solve = [&]() -> ohm_status {
    st.linear_solves += K;
    if (single)
      return tran_single_solve(tr, cur_sigma, prm->precond, so, sp.precision, itol, mref,
                               st.bicgstab_iterations);
    // the last step's slowest system sizes this step's chunks (FP64)
    o.expect_iters = sp.precision == OHM_PRECISION_FP64 ? last_batch : 0;
    ohm_status s = batch_solve_t<double, float>(bt, o, sp.precision, itol, mref, res.data(), false);
    last_batch = 0;
    for (int k = 0; k < K; k++) {
      st.bicgstab_iterations += res[k].iterations;
      if (s == OHM_OK) last_batch = std::max<int64_t>(last_batch, res[k].iterations);
    }
    return s;
  };
  OHM_HIP(hipEventRecord(b->ev0, stream));
  const double tstep = prm->tstep, tstop = prm->tstop, tstart = prm->tstart;
  const double h_min = tstep / 10000.0, h_max = tstep, lte_tol = 1e-3;  // transient.rs:36-39
  if (tstart <= 0.0) OHM_TRY(record(0.0));
  double t = 0.0, h = std::min(tstep, h_max);
  ohm_status rc = OHM_OK;
  // Graph-captured timesteps (fixed BE, one instance of more than 64
  // unknowns, FP64, graphs on), enabled by OHM_TRAN_SEGMENT_STEPS = S > 0:
  // after the first step has formed A = G + C/h and its preconditioner,
  // runs of up to S full steps are replayed as one graph (per step:
  // right-hand side, warm-started solve with a fixed number of iterations,
  // closing kernel) and the host reads one segment word per run instead of
  // polling every solve. A step whose solve needs more iterations pauses
  // the segment; the host finishes that solve and the next segment starts
  // after it. Off by default: a graph has no conditional nodes here, so the
  // captured iterations a step does not need, and every step after a pause,
  // still launch and exit at their heads (about 2 us per kernel), which costs
  // more than the polls it saves (RC grid 512 x 512: 3.27-4.19 ms per step
  // against 2.95-3.03 ms eager, identical results).
  int seg_steps = 0, seg_fixed = 0;
  if (const char *e = getenv("OHM_TRAN_SEGMENT_STEPS")) seg_steps = std::max(0, atoi(e));
  // OHM_TRAN_SEGMENT_ITERS pins the iterations captured per step (rounded
  // up to a multiple of 8) instead of following the steps' own counts
  if (const char *e = getenv("OHM_TRAN_SEGMENT_ITERS")) seg_fixed = (std::max(0, atoi(e)) + 7) / 8 * 8;
  // Fixed BE steps start their solve from the quadratic through the last
  // three solutions, extrapolated to t_n + h (k_tran_predict;
  // OHM_TRAN_PREDICT=1: linear, 0: from x_n, as the adaptive stepping does)
  int order = 2;  // extrapolation order
  if (const char *e = getenv("OHM_TRAN_PREDICT")) order = std::max(0, std::min(2, atoi(e)));
  const bool predict = prm->method == OHM_TRAN_FIXED_BE && order > 0;
  if (predict && !single) o.x_start = tr->d_trap;  // written by k_tran_predict
  const int pred_grid = (int)std::min<int64_t>((N + kBlock - 1) / kBlock, 2048);
  const bool seg_on = prm->method == OHM_TRAN_FIXED_BE && single && tr->n > 64 &&
                      sp.precision == OHM_PRECISION_FP64 && b->use_graphs && seg_steps > 0 &&
                      so.check_interval <= 0;
  for (auto &g : tr->segs) {  // captured for the previous run's options
    (void)hipGraphExecDestroy(g.exec);
    (void)hipGraphDestroy(g.graph);
  }
  tr->segs.clear();
  int seg_iters = 0;  // iterations captured per step (multiple of 8)
  auto seg_budget = [seg_fixed](int most) {
    if (seg_fixed > 0) return seg_fixed;
    const int want = most + std::max(4, most / 8);
    return std::max(8, (want + 7) / 8 * 8);
  };
  // Runs one segment of the full steps ahead; returns false when the eager
  // loop should take the next step (nothing to capture, or rc set).
  auto run_segment = [&]() -> bool {
    // the full steps ahead, with t accumulated as the eager loop adds it
    std::vector<double> ts;
    for (double tt = t; (int)ts.size() < seg_steps && tt < tstop && !(tt + h > tstop);) {
      tt = tt + h;
      ts.push_back(tt);
    }
    if (ts.size() < 2 || seg_iters > 256) return false;
    const int S = (int)ts.size();
    if (!tr->cs) {
      if ((rc = captured_create(b, tr->A1, tr->P1, wx, &tr->cs)) != OHM_OK) return false;
      if (hipMalloc(&tr->d_seg, sizeof(SegWord)) != hipSuccess) {
        rc = fail(OHM_ERR_INVALID, "hipMalloc of the segment word failed");
        return false;
      }
    }
    hipGraphExec_t exec = nullptr;
    for (auto &g : tr->segs)
      if (g.steps == S && g.iters == seg_iters) exec = g.exec;
    if (!exec) {
      const int64_t saved = b->dispatches;
      ohm_transient::SegGraph g{S, seg_iters, nullptr, nullptr};
      double *rhs = wx + 10 * stride;
      const int np = (int)n_probes;
      if ((rc = OHM_HIP_STATUS(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed))) != OHM_OK)
        return false;
      for (int j = 0; j < S && rc == OHM_OK; j++) {
        hipLaunchKernelGGL(k_tran_rhs, dim3(bt->h_blocks.size()), dim3(kBlock), 0, stream,
                           bt->d_full, bt->d_src, wx, rhs, tr->d_srp, tr->d_sent, tr->d_sslot,
                           tr->d_src, tr->d_pwl, TRhsArgs{0.0, 0.0, 1.0 / h, 0},
                           (const SegWord *)tr->d_seg, h);
        if (predict)  // full steps of one h
          hipLaunchKernelGGL(k_tran_predict, dim3(pred_grid), dim3(kBlock), 0, stream, N,
                             order == 2 ? 2.0 : 1.0, order == 2 ? -1.0 : 0.0, wx, tr->d_save,
                             order == 2 ? tr->d_prev2 : (double *)nullptr, (double *)nullptr,
                             (const SegWord *)tr->d_seg);
        rc = captured_enqueue(b, tr->cs, tr->A1, tr->P1, rhs, so, seg_iters, &tr->d_seg->stop);
        hipLaunchKernelGGL(k_seg_close, dim3(1), dim3(64), 0, stream, tr->d_seg,
                           captured_state(tr->cs), h, tstart, (int)std::max<int64_t>(max_records, 0),
                           np, tr->d_probes, wx + bt->h_geo[0].off, tr->d_rec);
      }
      const ohm_status ce = OHM_HIP_STATUS(hipStreamEndCapture(stream, &g.graph));
      if (rc == OHM_OK) rc = ce;
      if (rc == OHM_OK) rc = OHM_HIP_STATUS(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
      b->dispatches = saved;
      if (rc != OHM_OK) {
        if (g.graph) (void)hipGraphDestroy(g.graph);
        return false;
      }
      tr->segs.push_back(g);
      exec = g.exec;
    }
    SegWord *hw = reinterpret_cast<SegWord *>(b->h_pinned);
    *hw = SegWord{0, 0, 0, (int)n_rec, 0, 0, 0, t};
    if ((rc = OHM_HIP_STATUS(hipMemcpyAsync(tr->d_seg, hw, sizeof(SegWord), hipMemcpyHostToDevice,
                                       stream))) != OHM_OK ||
        (rc = OHM_HIP_STATUS(hipGraphLaunch(exec, stream))) != OHM_OK ||
        (rc = OHM_HIP_STATUS(hipMemcpyAsync(hw, tr->d_seg, sizeof(SegWord), hipMemcpyDeviceToHost,
                                       stream))) != OHM_OK ||
        (rc = OHM_HIP_STATUS(hipStreamSynchronize(stream))) != OHM_OK)
      return false;
    b->readbacks++;
    const SegWord w = *hw;
    for (int j = 0; j < w.steps; j++) {  // accepted steps: the host's own clock
      t = ts[j];
      st.steps_accepted++;
      st.linear_solves++;
      if (t >= tstart && n_rec < max_records) {
        rec_t.push_back(t);
        n_rec++;
      }
    }
    st.bicgstab_iterations += w.iters;
    if (w.n_rec != n_rec || (w.steps > 0 && !(w.t == t))) {
      rc = fail(OHM_ERR_INVALID, "transient segment bookkeeping mismatch");
      return false;
    }
    int most = w.max_iters;
    if (w.stop == 1) {  // the step after the accepted ones needs more iterations
      BicgOut out;
      st.linear_solves++;
      rc = captured_resume(b, tr->cs, tr->A1, tr->P1, so, seg_iters, out);
      st.bicgstab_iterations += out.iters;
      if (rc == OHM_OK) rc = status_to_error(out.status, out.iters, so.max_iters);
      if (rc != OHM_OK) return false;
      st.steps_accepted++;
      t = ts[w.steps];
      if (t >= tstart && (rc = record(t)) != OHM_OK) return false;
      most = std::max<int>(most, (int)out.iters);
    } else if (w.stop == 2) {
      st.linear_solves++;
      rc = status_to_error(w.status, so.max_iters, so.max_iters);
      return false;
    }
    seg_iters = seg_budget(most);
    return true;
  };
  if (prm->method == OHM_TRAN_FIXED_BE) {
    // backward Euler at fixed h (the stepping rule of :84-93 without LTE control)
    double h1 = 0.0, h2 = 0.0;  // the last two steps (0: none)
    while (t < tstop) {
      if (t + h > tstop) h = tstop - t;
      if (h < h_min * 0.5) break;
      if (seg_on && st.steps_accepted > (order == 2 ? 1 : 0) && seg_iters > 0) {
        if (run_segment()) continue;
        if (rc != OHM_OK) break;
      }
      const double t_next = t + h;
      OHM_TRY(tran_set_sigma(tr, 1.0 / h));
      cur_sigma = 1.0 / h;
      OHM_TRY(tran_rhs(tr, t_next, t, 1.0 / h, false));
      if (predict) {
        // Newton form of the polynomial through the last order+1 solutions,
        // in step ratios so that equal steps give exactly 1 (linear) and
        // 2, -1 (quadratic), the coefficients a captured segment uses
        double ca = 0.0, cb = 0.0;
        if (h1 > 0.0) ca = h / h1;
        if (order == 2 && h1 > 0.0 && h2 > 0.0) {
          const double u = h / h1, v = h2 / h1, w = u * (u + 1.0) / (1.0 + v);
          ca = u + w;
          cb = -w / v;
        }
        hipLaunchKernelGGL(k_tran_predict, dim3(pred_grid), dim3(kBlock), 0, stream, N, ca, cb, wx,
                           tr->d_save, order == 2 ? tr->d_prev2 : (double *)nullptr,
                           single ? (double *)nullptr : tr->d_trap, (const SegWord *)nullptr);
        OHM_HIP(hipGetLastError());
        b->dispatches++;
      }
      h2 = h1;
      h1 = h;
      const int64_t it0 = st.bicgstab_iterations;
      rc = solve();
      if (rc != OHM_OK) break;
      if (seg_on) seg_iters = std::max(seg_iters, seg_budget((int)(st.bicgstab_iterations - it0)));
      st.steps_accepted++;
      t = t_next;
      if (t >= tstart) OHM_TRY(record(t));
    }
  } else {
    // adaptive TRAP with a BE companion for the LTE estimate (:84-178)
    bool use_be = true;
    int fails = 0;
    std::vector<double> hb(1);
    while (t < tstop) {
      if (t + h > tstop) h = tstop - t;
      if (h < h_min * 0.5) break;
      const double t_next = t + h;
      double lte = 0.0;
      OHM_HIP(hipMemcpyAsync(tr->d_save, wx, sizeof(double) * N, hipMemcpyDeviceToDevice, stream));
      if (use_be) {
        OHM_TRY(tran_set_sigma(tr, 1.0 / h));
        cur_sigma = 1.0 / h;
        OHM_TRY(tran_rhs(tr, t_next, t, 1.0 / h, false));
        if ((rc = solve()) != OHM_OK) break;
      } else {
        OHM_TRY(tran_set_sigma(tr, 2.0 / h));
        cur_sigma = 2.0 / h;
        OHM_TRY(tran_rhs(tr, t_next, t, 2.0 / h, true));
        if ((rc = solve()) != OHM_OK) break;
        OHM_HIP(hipMemcpyAsync(tr->d_trap, wx, sizeof(double) * N, hipMemcpyDeviceToDevice, stream));
        OHM_HIP(hipMemcpyAsync(wx, tr->d_save, sizeof(double) * N, hipMemcpyDeviceToDevice, stream));
        OHM_TRY(tran_set_sigma(tr, 1.0 / h));
        cur_sigma = 1.0 / h;
        OHM_TRY(tran_rhs(tr, t_next, t, 1.0 / h, false));
        if ((rc = solve()) != OHM_OK) break;
        OHM_HIP(hipMemsetAsync(tr->d_max, 0, sizeof(unsigned long long), stream));
        hipLaunchKernelGGL(k_maxdiff, dim3(std::min<int64_t>((N + 255) / 256, 1024)), dim3(256), 0,
                           stream, N, tr->d_trap, wx, tr->d_max);
        unsigned long long bits = 0;
        OHM_HIP(hipMemcpyAsync(&bits, tr->d_max, sizeof(bits), hipMemcpyDeviceToHost, stream));
        OHM_HIP(hipStreamSynchronize(stream));
        double md;
        std::memcpy(&md, &bits, sizeof(md));
        lte = (2.0 / 3.0) * md;  // compute_lte, :355-362
        if (lte > lte_tol) {
          double h_new = h * std::min(2.0, std::max(0.5, 0.9 * sqrt(lte_tol / lte)));
          h = std::max(h_new, h_min);
          fails++;
          if (fails >= 10 && h <= h_min * 1.01) {
            rc = fail(OHM_ERR_TIMESTEP,
                      "Timestep error: simulation failed: %d consecutive failures at minimum "
                      "timestep h_min=%.2e near t=%.2e", fails, h_min, t);
            break;
          }
          st.steps_rejected++;
          use_be = true;
          OHM_HIP(hipMemcpyAsync(wx, tr->d_save, sizeof(double) * N, hipMemcpyDeviceToDevice, stream));
          continue;
        }
        double h_new = h * std::min(2.0, std::max(0.5, 0.9 * sqrt(lte_tol / lte)));
        OHM_HIP(hipMemcpyAsync(wx, tr->d_trap, sizeof(double) * N, hipMemcpyDeviceToDevice, stream));
        h = std::min(h_max, std::max(h_min, h_new));
      }
      if (use_be && lte == 0.0) use_be = false;  // after the first BE step switch to TRAP
      fails = 0;
      st.steps_accepted++;
      t = t_next;
      if (t >= tstart) OHM_TRY(record(t));
    }
  }
  OHM_HIP(hipEventRecord(b->ev1, stream));
  OHM_HIP(hipEventSynchronize(b->ev1));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
  st.device_ms = ms;
  st.records = n_rec;
  st.t_end = t;
  if (times)
    for (int64_t i = 0; i < n_rec; i++) times[i] = rec_t[i];
  if (waves && n_rec * K * n_probes > 0)
    OHM_HIP(hipMemcpy(waves, tr->d_rec, sizeof(double) * n_rec * K * n_probes, hipMemcpyDeviceToHost));
  if (x_final) OHM_HIP(hipMemcpy(x_final, wx, sizeof(double) * N, hipMemcpyDeviceToHost));
  (void)stride;
  if (stats) *stats = st;
  return rc;
}

}  // extern "C"

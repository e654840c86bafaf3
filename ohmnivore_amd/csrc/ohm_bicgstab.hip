// ohm_bicgstab.hip — right-preconditioned BiCGSTAB (bicgstab_with_params,
// src/solver/bicgstab.rs:56-189) as a five-kernel fused pipeline per
// iteration with every scalar (rho, alpha, omega, status) resident in device
// memory:
//
//   K1  [r.r, r_hat.r -> NaN/convergence/omega/rho tests, beta]
//       p = r + beta (p - omega v);  p_hat = M^-1 p
//   K2  v = A p_hat                         (partials r_hat.v, v.v)
//   K3  [r_hat.v -> alpha, breakdown test]
//       s = r - alpha v;  s_hat = M^-1 s    (partials s.s)
//   K4  [||s|| early-exit test]  t = A s_hat  (partials t.s, t.t)
//   K5  [t.t breakdown test, omega]
//       x += alpha p_hat + omega s_hat; r = s - omega t  (partials r.r, r_hat.r)
//
// Reductions are deferred: a producing kernel leaves one partial per
// workgroup and every workgroup of the consuming kernel re-sums them in a
// fixed order (the bracketed heads), so no kernel waits on an atomic ticket
// and every workgroup derives the same scalars; workgroup 0 publishes them.
// Once the state leaves RUNNING every later kernel exits at its first
// instruction, so the host enqueues iterations in chunks closed by a
// one-workgroup check and polls one status word per chunk instead of syncing
// on each dot product (the reference syncs six times per iteration,
// src/solver/backend.rs:443-446). Jacobi fuses into K1/K3 elementwise; ISAI
// fuses the p and s updates into the first preconditioner SpMV (M = M_U M_L
// when the product is formed, else M_L followed by an M_U launch). The same
// kernels run in FP32 (with compensated dot products) as the inner solver of
// FP64 iterative refinement.
#include <algorithm>
#include <cstddef>
#include <cmath>
#include <mutex>

#include "ohm_precond.h"
#include "ohm_spmv.h"

namespace ohm {

// stop (captured solves): when *stop is set the solve is masked off: the
// status word leaves RUNNING (ST_SKIPPED) and nothing else is written, so
// every kernel of the solve exits at its head and the iteration state of a
// solve paused before it stays as it was (k_unskip reopens it).
__global__ void k_state_reset(SolveState *st, const int *stop) {
  if (threadIdx.x == 0) {
    if (stop && *stop) {
      st->status = ST_SKIPPED;
      return;
    }
    SolveState z;
    z.rho = z.alpha = z.omega = z.beta = z.rho_new = 1.0;
    z.abs_tol = z.bnorm2 = z.rr = z.ss = z.rv = z.ts = z.tt = z.rhrh = 0.0;
    z.status = ST_RUNNING;
    z.s_conv = 0;
    z.iters = 0;
    z.max_iters = 0;
    z.counter = 0u;
    z.scaled = 1;
    z.r0 = 0.0;
    z.res2 = 0.0;
    z.eta = 0.0;
    *st = z;
  }
}

// Breakdown tests: the reference's absolute thresholds, or the same 1e-30
// applied to normalised quantities.
__device__ __forceinline__ bool rho_breakdown(double rho, double hh, double rr, int scaled) {
  return scaled ? fabs(rho) <= 1e-30 * (sqrt(hh) * sqrt(rr)) : fabs(rho) < 1e-30;
}

// Partial-sum regions of one solve (doubles, `stride` workgroups each).
struct Parts {
  double *init;  // k_init's ticket scratch, 4 per workgroup
  double *rv;    // K2: r_hat.v, v.v
  double *ss;    // K3: s.s
  double *ts;    // K4: t.s, t.t
  double *rr;    // K5 (and k_init): r.r, r_hat.r
  int n_rv, n_ss, n_ts, n_rr;  // partials each consumer sums (1 = folded)
  bool f_rv, f_ss, f_ts, f_rr; // a k_fold launch reduces the producer's partials
};

// A workgroup-uniform read of the status word. Workgroup 0 of a consuming
// kernel may publish a stop while later workgroups start; every workgroup
// takes either that status or the same decision from the same partials, so
// reading it once per workgroup keeps all of its waves on one path.
__device__ __forceinline__ bool block_running(const SolveState *st) {
  __shared__ int s_run;
  if (threadIdx.x == 0) s_run = st->status == ST_RUNNING;
  __syncthreads();
  return s_run != 0;
}

// block_running and gather_sum in one pass: the status word (thread 0) and
// the partials are requested together and published by one workgroup sum.
template <int K>
__device__ __forceinline__ bool gather_run(const SolveState *st,
                                           const double *__restrict__ partials, int np,
                                           double (&tot)[K]) {
  __shared__ double bc[K];
  __shared__ int s_run;
  const int run0 = threadIdx.x == 0 ? (st->status == ST_RUNNING) : 0;
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
    s_run = run0;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; k++) tot[k] = bc[k];
  return s_run != 0;
}

// End-of-iteration tests on ||r||^2, r_hat.r (src/solver/bicgstab.rs:158-185
// in the reference's order), taken by the first kernel of the next iteration
// and by the chunk's closing check. Returns the stop status or ST_RUNNING
// with beta for the p update.
__device__ __forceinline__ int decide_rho(const SolveState *st, double rr, double rhr,
                                          double &beta) {
  if (st->s_conv) return ST_CONVERGED;
  const double rnorm = sqrt(rr);
  if (std::isnan(rnorm) || std::isinf(rnorm)) return ST_NAN;
  if (rnorm < st->abs_tol) return ST_CONVERGED;
  if (fabs(st->omega) < 1e-30) return ST_BRK_OMEGA;
  if (st->iters >= st->max_iters) return ST_MAXITER;
  if (rho_breakdown(rhr, st->rhrh, rr, st->scaled)) return ST_BRK_RHO;
  beta = (rhr / st->rho) * (st->alpha / st->omega);
  return ST_RUNNING;
}

// p-update head: every workgroup forms beta from K5's partials; workgroup 0
// publishes rho_new (or the stop status). False means: exit.
__device__ __forceinline__ bool head_rho(SolveState *st, const Parts &pp, double &beta) {
  double tot[2];
  if (!gather_run<2>(st, pp.rr, pp.n_rr, tot)) return false;
  const int stat = decide_rho(st, tot[0], tot[1], beta);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rr = tot[0];
    if (!st->s_conv) st->res2 = tot[0];
    if (stat != ST_RUNNING)
      st->status = stat;
    else
      st->rho_new = tot[1];
  }
  return stat == ST_RUNNING;
}

// s-update head: alpha = rho_new / r_hat.v with the r_hat.v breakdown test.
__device__ __forceinline__ bool head_alpha(SolveState *st, const Parts &pp, double &alpha) {
  double tot[2];
  if (!gather_run<2>(st, pp.rv, pp.n_rv, tot)) return false;
  const bool brk = st->scaled ? fabs(tot[0]) <= 1e-30 * (sqrt(st->rhrh) * sqrt(tot[1]))
                              : fabs(tot[0]) < 1e-30;
  const double rho = st->rho_new;
  alpha = rho / tot[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rv = tot[0];
    if (brk) {
      st->status = ST_BRK_RV;
    } else {
      st->alpha = alpha;
      st->rho = rho;
    }
  }
  return !brk;
}

// Inexact-Newton forcing term for a sub-solve starting at ||F_k|| = f
// (Eisenstat & Walker, SIAM J. Sci. Comput. 17(1), 1996: choice 1 with the
// eta^((1+sqrt 5)/2) safeguard, choice 2 with gamma 0.9, alpha 2), clamped to
// [eta_min, eta_max]; eta_max on the first Newton iteration.
__host__ __device__ inline double forcing_eta(const Forcing &fc, double f) {
  double eta = fc.eta_max;
  if (fc.f_prev > 0.0 && f == f && f < INFINITY) {
    double safe;
    if (fc.kind == OHM_FORCING_EW1) {
      eta = fabs(f - fc.lin_prev) / fc.f_prev;
      safe = pow(fc.eta_prev, 1.6180339887498949);
    } else {
      const double q = f / fc.f_prev;
      eta = 0.9 * q * q;
      safe = 0.9 * fc.eta_prev * fc.eta_prev;
    }
    if (safe > 0.1 || fc.rate_cap) eta = fmax(eta, safe);
    eta = fmin(fc.eta_max, fmax(fc.eta_min, eta));
  }
  return eta;
}

// r = b - A x0 (warm) or b; r_hat = hashed or r; p = v = 0; x = 0 (cold).
// Also leaves the r.r / r_hat.r partials where the first p update reads them.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_init(int n, const T *__restrict__ b,
                                                 const T *__restrict__ ax0, T *__restrict__ x,
                                                 T *__restrict__ r, T *__restrict__ rh,
                                                 T *__restrict__ p, T *__restrict__ v, int shadow,
                                                 int warm, double tol, double floor_rel,
                                                 int tol_mode, int max_iters, int scaled,
                                                 Forcing fc, SolveState *st, Parts pp,
                                                 int masked) {
  if (masked && st->status != ST_RUNNING) return;  // a masked-off captured solve
  Acc<T> bb, rr, rhr, hh;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    T bi = b[i];
    T ri = warm ? bi - ax0[i] : bi;
    if (!warm) x[i] = T(0);
    r[i] = ri;
    T rhi = shadow ? (T)shadow_value((uint64_t)i) : ri;
    rh[i] = rhi;
    p[i] = T(0);
    v[i] = T(0);
    bb.add_prod(bi, bi);
    rr.add_prod(ri, ri);
    rhr.add_prod(rhi, ri);
    hh.add_prod(rhi, rhi);
  }
  double vals[4] = {bb.get(), rr.get(), rhr.get(), hh.get()}, tot[4];
  const bool last = grid_sum<4>(vals, pp.init, &st->counter, tot);
  // grid_sum leaves this workgroup's sums in thread 0's vals[]
  if (!pp.f_rr && threadIdx.x == 0) {
    pp.rr[(size_t)blockIdx.x * 2] = vals[1];
    pp.rr[(size_t)blockIdx.x * 2 + 1] = vals[2];
  }
  if (last && threadIdx.x == 0) {
    if (pp.f_rr) {  // already folded: the totals
      pp.rr[0] = tot[1];
      pp.rr[1] = tot[2];
    }
    st->bnorm2 = tot[0];
    st->rhrh = tot[3];
    st->scaled = scaled;
    st->max_iters = max_iters;
    st->rr = tot[1];
    st->res2 = tot[1];
    st->r0 = sqrt(tot[1]);
    if (fc.kind) {
      tol = forcing_eta(fc, sqrt(tot[1]));
      st->eta = tol;
    }
    double bnorm = sqrt(tot[0]);
    if (bnorm < 1e-30) {  // b == 0: x = 0 (src/solver/bicgstab.rs:93-97)
      st->status = ST_CONVERGED;
      return;
    }
    st->abs_tol = tol_mode == OHM_TOL_INITIAL ? fmax(tol * sqrt(tot[1]), floor_rel * bnorm)
                                              : tol * bnorm;
    if (warm && sqrt(tot[1]) < st->abs_tol) {
      st->status = ST_CONVERGED;
    } else if (rho_breakdown(tot[2], tot[3], tot[1], scaled)) {
      st->status = ST_BRK_RHO;
    }
    // rho = alpha = omega = 1 from k_state_reset: the first beta multiplies
    // p - omega v = 0, so p = r as in the reference's first iteration
  }
}

// p_j = r_j + beta (p_j - omega v_j), one rounding sequence for every use.
template <typename T>
__device__ __forceinline__ T p_next(T r, T p, T v, T beta, T omega) {
  return fmac<T>(fmac<T>(-omega, v, p), beta, r);
}

// Vector kernels prefetch their first element before the head: the index
// is clamped into the vector (n >= 1 for every launch) and the values are
// used only when it is in range.
__device__ __forceinline__ int first_index(int n) {
  return min((int)(blockIdx.x * kBlock + threadIdx.x), n - 1);
}

// K1 (Jacobi / none): p update, p_hat = inv * p (or p).
template <typename T, int PK>
__global__ __launch_bounds__(kBlock) void k_p_update(int n, const T *__restrict__ r,
                                                     T *__restrict__ p, const T *__restrict__ v,
                                                     T *__restrict__ ph,
                                                     const T *__restrict__ inv, SolveState *st,
                                                     Parts pp) {
  const int i0 = blockIdx.x * kBlock + threadIdx.x, ic = first_index(n);
  const T r0 = r[ic], p0 = p[ic], v0 = v[ic], d0 = PK == PK_JACOBI ? inv[ic] : T(1);
  double beta_d;
  if (!head_rho(st, pp, beta_d)) return;
  const T beta = (T)beta_d, omega = (T)st->omega;
  if (i0 < n) {
    const T pi = p_next<T>(r0, p0, v0, beta, omega);
    p[i0] = pi;
    if (PK == PK_JACOBI) ph[i0] = d0 * pi;
    if (PK == PK_NONE) ph[i0] = pi;
  }
  for (int i = i0 + gridDim.x * kBlock; i < n; i += gridDim.x * kBlock) {
    const T pi = p_next<T>(r[i], p[i], v[i], beta, omega);
    p[i] = pi;
    if (PK == PK_JACOBI) ph[i] = inv[i] * pi;
    if (PK == PK_NONE) ph[i] = pi;
  }
}

// K1 (ISAI): the p update fused into the first preconditioner SpMV:
// y = M1 p with p_j = r_j + beta q_j formed on the fly, where
// q = p_prev - omega v_prev was written by the previous K5 (zeroed by k_init,
// so the first p is r). The two roundings are p_next's, split across the
// two kernels; p is written for the rows of this block.
template <typename T, int PER>
__global__ __launch_bounds__(kBlock) void k_p_spmv(const uint32_t *__restrict__ rowptr,
                                                   const uint32_t *__restrict__ col,
                                                   const T *__restrict__ val,
                                                   const uint32_t *__restrict__ blocks,
                                                   const T *__restrict__ r,
                                                   const T *__restrict__ q,
                                                   T *__restrict__ p, T *__restrict__ y,
                                                   SolveState *st, Parts pp) {
  T beta = T(0);
  spmv_block_h<T, 2, PER>(
      rowptr, col, val, blocks,
      [&](uint32_t j) {
        Opnd<T, 2> o;
        o.v[0] = r[j];
        o.v[1] = q[j];
        return o;
      },
      [&] {
        double beta_d;
        if (!head_rho(st, pp, beta_d)) return false;
        beta = (T)beta_d;
        return true;
      },
      [&](const Opnd<T, 2> &o) { return fmac<T>(o.v[1], beta, o.v[0]); },
      [&](uint32_t i, T yi) {
        y[i] = yi;
        p[i] = fmac<T>(q[i], beta, r[i]);
      });
}

// K2: v = A p_hat; partials r_hat.v, v.v.
template <typename T, int PER>
__global__ __launch_bounds__(kBlock) void k_spmv_alpha(const uint32_t *__restrict__ rowptr,
                                                       const uint32_t *__restrict__ col,
                                                       const T *__restrict__ val,
                                                       const uint32_t *__restrict__ blocks,
                                                       const T *__restrict__ ph,
                                                       T *__restrict__ v,
                                                       const T *__restrict__ rh, SolveState *st,
                                                       Parts pp) {
  Acc<T> acc, vv;
  const bool ran = spmv_block_h<T, 1, PER>(
      rowptr, col, val, blocks,
      [&](uint32_t j) {
        Opnd<T, 1> o;
        o.v[0] = ph[j];
        return o;
      },
      [&] { return block_running(st); }, [](const Opnd<T, 1> &o) { return o.v[0]; },
      [&](uint32_t r, T y) {
        v[r] = y;
        acc.add_prod(rh[r], y);
        vv.add_prod(y, y);
      });
  if (!ran) return;
  double vals[2] = {acc.get(), vv.get()};
  deposit<2>(vals, pp.rv);
}

// K3 (Jacobi / none): alpha, s = r - alpha v, s_hat = inv * s; partial s.s.
template <typename T, int PK>
__global__ __launch_bounds__(kBlock) void k_s_update(int n, const T *__restrict__ r,
                                                     const T *__restrict__ v, T *__restrict__ s,
                                                     T *__restrict__ sh,
                                                     const T *__restrict__ inv, SolveState *st,
                                                     Parts pp) {
  const int i0 = blockIdx.x * kBlock + threadIdx.x, ic = first_index(n);
  const T r0 = r[ic], v0 = v[ic], d0 = PK == PK_JACOBI ? inv[ic] : T(1);
  double alpha_d;
  if (!head_alpha(st, pp, alpha_d)) return;
  const T alpha = (T)alpha_d;
  Acc<T> acc;
  if (i0 < n) {
    const T si = fmac<T>(-alpha, v0, r0);
    s[i0] = si;
    if (PK == PK_JACOBI) sh[i0] = d0 * si;
    if (PK == PK_NONE) sh[i0] = si;
    acc.add_prod(si, si);
  }
  for (int i = i0 + gridDim.x * kBlock; i < n; i += gridDim.x * kBlock) {
    const T si = fmac<T>(-alpha, v[i], r[i]);
    s[i] = si;
    if (PK == PK_JACOBI) sh[i] = inv[i] * si;
    if (PK == PK_NONE) sh[i] = si;
    acc.add_prod(si, si);
  }
  double vals[1] = {acc.get()};
  deposit<1>(vals, pp.ss);
}

// K3 (ISAI): the s update fused into the first preconditioner SpMV
// (y = M1 s with s_j = r_j - alpha v_j formed on the fly); partial s.s.
template <typename T, int PER>
__global__ __launch_bounds__(kBlock) void k_s_spmv(const uint32_t *__restrict__ rowptr,
                                                   const uint32_t *__restrict__ col,
                                                   const T *__restrict__ val,
                                                   const uint32_t *__restrict__ blocks,
                                                   const T *__restrict__ r,
                                                   const T *__restrict__ v, T *__restrict__ s,
                                                   T *__restrict__ y, SolveState *st, Parts pp) {
  T alpha = T(0);
  Acc<T> acc;
  const bool ran = spmv_block_h<T, 2, PER>(
      rowptr, col, val, blocks,
      [&](uint32_t j) {
        Opnd<T, 2> o;
        o.v[0] = r[j];
        o.v[1] = v[j];
        return o;
      },
      [&] {
        double alpha_d;
        if (!head_alpha(st, pp, alpha_d)) return false;
        alpha = (T)alpha_d;
        return true;
      },
      [&](const Opnd<T, 2> &o) { return fmac<T>(-alpha, o.v[1], o.v[0]); },
      [&](uint32_t i, T yi) {
        y[i] = yi;
        const T si = fmac<T>(-alpha, v[i], r[i]);
        s[i] = si;
        acc.add_prod(si, si);
      });
  if (!ran) return;
  double vals[1] = {acc.get()};
  deposit<1>(vals, pp.ss);
}

// K4: the ||s|| early-exit test, then t = A s_hat; partials t.s, t.t.
template <typename T, int PER>
__global__ __launch_bounds__(kBlock) void k_spmv_omega(const uint32_t *__restrict__ rowptr,
                                                       const uint32_t *__restrict__ col,
                                                       const T *__restrict__ val,
                                                       const uint32_t *__restrict__ blocks,
                                                       const T *__restrict__ sh,
                                                       T *__restrict__ t,
                                                       const T *__restrict__ s, SolveState *st,
                                                       Parts pp) {
  Acc<T> ts, tt;
  const bool ran = spmv_block_h<T, 1, PER>(
      rowptr, col, val, blocks,
      [&](uint32_t j) {
        Opnd<T, 1> o;
        o.v[0] = sh[j];
        return o;
      },
      [&] {
        double ss[1];
        if (!gather_run<1>(st, pp.ss, pp.n_ss, ss)) return false;
        const bool sconv = sqrt(ss[0]) < st->abs_tol;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
          st->ss = ss[0];
          if (sconv) {
            st->s_conv = 1;
            st->res2 = ss[0];
          }
        }
        return !sconv;
      },
      [](const Opnd<T, 1> &o) { return o.v[0]; },
      [&](uint32_t r, T y) {
        t[r] = y;
        ts.add_prod(y, s[r]);
        tt.add_prod(y, y);
      });
  if (!ran) return;
  double vals[2] = {ts.get(), tt.get()};
  deposit<2>(vals, pp.ts);
}

// K5: omega = t.s / t.t with the t.t breakdown test; x += alpha p_hat +
// omega s_hat, r = s - omega t; partials r.r, r_hat.r for the next p update.
// After the ||s|| exit only x += alpha p_hat runs (src/solver/bicgstab.rs:
// 128-135); the next kernel turns s_conv into CONVERGED.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_xr_update(int n, T *__restrict__ x,
                                                      T *__restrict__ r,
                                                      const T *__restrict__ rh,
                                                      const T *__restrict__ ph,
                                                      const T *__restrict__ sh,
                                                      const T *__restrict__ s,
                                                      const T *__restrict__ t,
                                                      const T *__restrict__ p,
                                                      const T *__restrict__ v,
                                                      T *__restrict__ q, SolveState *st,
                                                      Parts pp) {
  const int i0 = blockIdx.x * kBlock + threadIdx.x, ic = first_index(n);
  const T x0 = x[ic], ph0 = ph[ic], sh0 = sh[ic], s0 = s[ic], t0 = t[ic], rh0 = rh[ic];
  double tot[2];
  if (!gather_run<2>(st, pp.ts, pp.n_ts, tot)) return;
  const T alpha = (T)st->alpha;
  const int stride = gridDim.x * kBlock;
  if (st->s_conv) {
    if (i0 < n) x[i0] = x0 + alpha * ph0;
    for (int i = i0 + stride; i < n; i += stride) x[i] = x[i] + alpha * ph[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) st->iters += 1;
    return;
  }
  const bool brk = st->scaled ? tot[1] <= 1e-30 * st->ss : fabs(tot[1]) < 1e-30;
  const double omega_d = tot[0] / tot[1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->ts = tot[0];
    st->tt = tot[1];
    if (brk) {
      st->status = ST_BRK_TT;
    } else {
      st->omega = omega_d;
      st->iters += 1;
    }
  }
  if (brk) return;
  const T omega = (T)omega_d;
  Acc<T> rr, rhr;
  if (i0 < n) {
    x[i0] = (x0 + alpha * ph0) + omega * sh0;
    const T ri = s0 - omega * t0;
    r[i0] = ri;
    rr.add_prod(ri, ri);
    rhr.add_prod(rh0, ri);
    if (q) q[i0] = fmac<T>(-omega, v[i0], p[i0]);
  }
  for (int i = i0 + stride; i < n; i += stride) {
    x[i] = (x[i] + alpha * ph[i]) + omega * sh[i];
    const T ri = s[i] - omega * t[i];
    r[i] = ri;
    rr.add_prod(ri, ri);
    rhr.add_prod(rh[i], ri);
    if (q) q[i] = fmac<T>(-omega, v[i], p[i]);
  }
  double vals[2] = {rr.get(), rhr.get()};
  deposit<2>(vals, pp.rr);
}

// ------------------------------------------------------ one-wave solve --
// The whole solve of a system of at most 64 unknowns in one 64-lane
// workgroup (the reference's bicgstab_fused, src/solver/ds_shaders.rs:
// 340-728, dispatched for tiny Newton systems at src/solver/newton.rs:
// 216-252): lane r owns row r and keeps its entries of every vector in
// registers; a matrix-vector product gathers from the one vector staged in
// LDS; dot products are DPP wave sums. The recurrences, stopping tests,
// breakdown tests, shadow residual and forcing are those of the multi-kernel
// pipeline above (K1-K5 and k_init), in the same order; the loop is bounded
// by max_iters, so the workgroup always exits.
constexpr int kFusedRows = 64;
constexpr uint32_t kFusedLdsBytes = 60 * 1024;  // staged matrices, beside the vector

struct FusedCsr {
  const uint32_t *rowptr, *col;
  const double *val;
};

// A lane's row of one matrix: its bounds in registers, the column indices
// and values in LDS (STAGE) or global memory.
struct FusedRow {
  uint32_t beg, end;
  const uint32_t *col;
  const double *val;
};

// Bytes of LDS the staged column indices and values of a matrix take.
__host__ __device__ inline uint32_t fused_stage_bytes(uint32_t nnz) {
  return 8 * nnz + ((4 * nnz + 7) & ~7u);
}

// Binds lane r to its row of m; with STAGE the whole matrix's columns and
// values are copied into LDS at lds (advanced past them). The caller
// synchronises before the first use.
template <bool STAGE>
__device__ __forceinline__ FusedRow fused_bind(const FusedCsr &m, int r, int n,
                                               unsigned char *&lds) {
  FusedRow f{0, 0, m.col, m.val};
  if (STAGE) {  // LDS pointers on every path, so the loads compile to ds_read
    f.col = reinterpret_cast<const uint32_t *>(lds);
    f.val = reinterpret_cast<const double *>(lds);
  }
  if (!m.rowptr) return f;
  if (r < n) {
    f.beg = m.rowptr[r];
    f.end = m.rowptr[r + 1];
  }
  if (STAGE) {
    const uint32_t nnz = m.rowptr[n];
    double *v = reinterpret_cast<double *>(lds);
    uint32_t *c = reinterpret_cast<uint32_t *>(lds + 8 * nnz);
    for (uint32_t k = threadIdx.x; k < nnz; k += kFusedRows) {
      v[k] = m.val[k];
      c[k] = m.col[k];
    }
    lds += fused_stage_bytes(nnz);
    f.col = c;
    f.val = v;
  }
  return f;
}

// A lane's row times src: each product rounded, then summed in column order
// from 0, the rounding sequence of the pipeline's row-block SpMV (products
// staged in LDS, summed per row), so the two paths agree to the bit.
__device__ __forceinline__ double fused_row(const FusedRow &m, const double *__restrict__ src) {
#pragma clang fp contract(off)
  double acc = 0.0;
  for (uint32_t k = m.beg; k < m.end; k++) {
    const double pr = m.val[k] * src[m.col[k]];
    acc = acc + pr;
  }
  return acc;
}

// y_r = (A v)_r with v staged in LDS (every lane writes its v_r first)
__device__ __forceinline__ double fused_spmv(const FusedRow &m, double v, double *sv) {
  __syncthreads();
  sv[threadIdx.x] = v;
  __syncthreads();
  return fused_row(m, sv);
}

template <int PK>
__device__ __forceinline__ double fused_prec(const FusedRow &ml, const FusedRow &mu, bool two,
                                             double inv, double v, double *sv) {
  if (PK == PK_JACOBI) return inv * v;
  if (PK == PK_ISAI) {
    const double y = fused_spmv(ml, v, sv);
    return two ? fused_spmv(mu, y, sv) : y;
  }
  return v;
}

template <int PK, bool STAGE>
__global__ __launch_bounds__(kFusedRows) void k_bicg_fused(
    int n, FusedCsr ag, FusedCsr mlg, FusedCsr mug, int two, const double *__restrict__ dinv,
    const double *__restrict__ b, double *__restrict__ x, int shadow, int warm, double tol,
    double floor_rel, int tol_mode, int max_iters, int scaled, Forcing fc, SolveState *st) {
  __shared__ double sv[kFusedRows];
  extern __shared__ __align__(16) unsigned char fused_lds[];
  const int r = threadIdx.x;
  const bool in = r < n;
  unsigned char *lds = fused_lds;
  const FusedRow a = fused_bind<STAGE>(ag, r, n, lds);
  const FusedRow ml = fused_bind<STAGE>(mlg, r, n, lds);
  const FusedRow mu = fused_bind<STAGE>(mug, r, n, lds);
  if (STAGE) __syncthreads();
  const double inv = (PK == PK_JACOBI && in) ? dinv[r] : 1.0;
  // dot products: lanes past n add +0, as threads without a row do in the
  // pipeline's block sums
  const auto sum = [in](double v) { return wave_sum(in ? v : 0.0); };
  // k_init: r = b - A x0 (warm) or b, r_hat, p = v = 0
  const double bi = in ? b[r] : 0.0;
  double xi = (in && warm) ? x[r] : 0.0;
  double ri = bi;
  if (warm) ri = bi - fused_spmv(a, xi, sv);
  if (!in) ri = 0.0;
  const double rh = in ? (shadow ? shadow_value((uint64_t)r) : ri) : 0.0;
  double p = 0.0, v = 0.0;
  const double bb = sum(bi * bi), hh = sum(rh * rh);
  double rr = sum(ri * ri), rhr = sum(rh * ri);
  const double r0 = sqrt(rr);
  double eta = 0.0;
  if (fc.kind) {
    tol = forcing_eta(fc, r0);
    eta = tol;
  }
  const double bnorm = sqrt(bb);
  const double abs_tol = tol_mode == OHM_TOL_INITIAL ? fmax(tol * r0, floor_rel * bnorm) : tol * bnorm;
  double rho = 1.0, alpha = 1.0, omega = 1.0, res2 = rr;
  int status = ST_RUNNING, iters = 0, s_conv = 0;
  if (bnorm < 1e-30) {  // b == 0: x = 0 (src/solver/bicgstab.rs:93-97)
    status = ST_CONVERGED;
  } else if (warm && r0 < abs_tol) {
    status = ST_CONVERGED;
  } else if (rho_breakdown(rhr, hh, rr, scaled)) {
    status = ST_BRK_RHO;
  }
  while (status == ST_RUNNING) {
    // K1 head (decide_rho) and p update
    const double rnorm = sqrt(rr);
    if (s_conv) { status = ST_CONVERGED; break; }
    if (std::isnan(rnorm) || std::isinf(rnorm)) { status = ST_NAN; break; }
    if (rnorm < abs_tol) { status = ST_CONVERGED; break; }
    if (fabs(omega) < 1e-30) { status = ST_BRK_OMEGA; break; }
    if (iters >= max_iters) { status = ST_MAXITER; break; }
    if (rho_breakdown(rhr, hh, rr, scaled)) { status = ST_BRK_RHO; break; }
    const double beta = (rhr / rho) * (alpha / omega);
    const double rho_new = rhr;
    p = p_next<double>(ri, p, v, beta, omega);
    const double ph = fused_prec<PK>(ml, mu, two, inv, p, sv);
    // K2: v = A p_hat; K3 head: alpha
    v = fused_spmv(a, ph, sv);
    if (!in) v = 0.0;
    const double rv = sum(rh * v), vv = sum(v * v);
    const bool brk = scaled ? fabs(rv) <= 1e-30 * (sqrt(hh) * sqrt(vv)) : fabs(rv) < 1e-30;
    if (brk) { status = ST_BRK_RV; break; }
    alpha = rho_new / rv;
    rho = rho_new;
    // K3: s = r - alpha v; K4 head: ||s|| exit
    const double s = fma(-alpha, v, ri);
    const double ss = sum(s * s);
    if (sqrt(ss) < abs_tol) {
      xi = xi + alpha * ph;
      iters++;
      s_conv = 1;
      res2 = ss;
      continue;  // the next head turns s_conv into CONVERGED
    }
    const double sh = fused_prec<PK>(ml, mu, two, inv, s, sv);
    // K4: t = A s_hat; K5: omega, x and r updates
    double t = fused_spmv(a, sh, sv);
    if (!in) t = 0.0;
    const double ts = sum(t * s), tt = sum(t * t);
    if (scaled ? tt <= 1e-30 * ss : fabs(tt) < 1e-30) { status = ST_BRK_TT; break; }
    omega = ts / tt;
    iters++;
    xi = (xi + alpha * ph) + omega * sh;
    ri = s - omega * t;
    rr = sum(ri * ri);
    rhr = sum(rh * ri);
    res2 = rr;
  }
  if (in) x[r] = xi;
  if (r == 0) {
    st->status = status;
    st->s_conv = s_conv;
    st->iters = iters;
    st->max_iters = max_iters;
    st->r0 = r0;
    st->res2 = res2;
    st->eta = eta;
    st->abs_tol = abs_tol;
  }
}

// Reopens a paused captured solve whose status a later masked-off solve on
// the same state set to ST_SKIPPED.
__global__ void k_unskip(SolveState *st) {
  if (threadIdx.x == 0 && st->status == ST_SKIPPED) st->status = ST_RUNNING;
}

// Reduces a producer's np partials (K each) to partials[0..K) for a
// consumer whose workgroups would otherwise each re-read all of them.
template <int K>
__global__ __launch_bounds__(kBlock) void k_fold(double *partials, int np) {
  double tot[K];
  gather_sum<K>(partials, np, tot);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) partials[k] = tot[k];
  }
}

// Closing test of an enqueued chunk (one workgroup): publishes the stop
// status of the chunk's last iteration so the host's poll sees it.
__global__ __launch_bounds__(kBlock) void k_check(SolveState *st, Parts pp) {
  double tot[2], beta;
  if (!gather_run<2>(st, pp.rr, pp.n_rr, tot)) return;
  const int stat = decide_rho(st, tot[0], tot[1], beta);
  if (threadIdx.x == 0) {
    if (!st->s_conv) st->res2 = tot[0];
    if (stat != ST_RUNNING) st->status = stat;
  }
}

// ------------------------------------------------------------ workspace --
template <typename T>
struct Workspace {
  int n = 0;
  T *base = nullptr;
  SolveState *st = nullptr;
  BicgVecs<T> v{};
  double *partials = nullptr;  // null: the backend's partials buffer
};

static Workspace<double> g_ws64;  // one cached workspace per type (single-threaded backend)
static Workspace<float> g_ws32;

constexpr int kWsVectors = 11;  // vectors of one solve

static size_t ws_stride(int n) { return ((size_t)std::max(n, 1) + 63) & ~size_t(63); }

// The eleven solve vectors, `stride` elements apart from q.
template <typename T>
static void ws_views(Workspace<T> &w, T *q, size_t stride) {
  w.v.x = q;
  w.v.r = q + stride;
  w.v.rh = q + 2 * stride;
  w.v.p = q + 3 * stride;
  w.v.v = q + 4 * stride;
  w.v.s = q + 5 * stride;
  w.v.t = q + 6 * stride;
  w.v.ph = q + 7 * stride;
  w.v.sh = q + 8 * stride;
  w.v.tmp = q + 9 * stride;
  w.v.p2 = q + 10 * stride;
}

template <typename T>
static ohm_status ws_get(Workspace<T> &w, int n) {
  if (w.base && w.n >= n) return OHM_OK;
  if (w.base) OHM_HIP(hipFree(w.base));
  if (w.st) OHM_HIP(hipFree(w.st));
  w.base = nullptr;
  w.st = nullptr;
  w.n = 0;
  const size_t stride = ws_stride(n);
  OHM_HIP(hipMalloc(&w.base, sizeof(T) * stride * kWsVectors));
  OHM_HIP(hipMalloc(&w.st, sizeof(SolveState)));
  ws_views(w, w.base, stride);
  w.n = n;
  return OHM_OK;
}

template <typename T>
struct Ops;
template <>
struct Ops<double> {
  static const double *vals(const DevCsr &a) { return a.val; }
  static const double *inv(const PrecondDev &p) { return p.inv; }
  static void spmv(ohm_backend *b, const DevCsr &a, const double *x, double *y,
                   const SolveState *st, int skip) {
    launch_spmv(b, a, x, y, st, skip);
  }
  static void prec(ohm_backend *b, const PrecondDev &p, const double *in, double *out,
                   double *tmp, const SolveState *st, int skip) {
    precond_apply(b, p, in, out, tmp, st, skip);
  }
};
template <>
struct Ops<float> {
  static const float *vals(const DevCsr &a) { return a.val32; }
  static const float *inv(const PrecondDev &p) { return p.inv32; }
  static void spmv(ohm_backend *b, const DevCsr &a, const float *x, float *y,
                   const SolveState *st, int skip) {
    launch_spmv_f32(b, a, x, y, st, skip);
  }
  static void prec(ohm_backend *b, const PrecondDev &p, const float *in, float *out, float *tmp,
                   const SolveState *st, int skip) {
    precond_apply_f32(b, p, in, out, tmp, st, skip);
  }
};

// First SpMV of the ISAI apply: the product M when formed, else M_L (or the
// transpose view's first factor).
static const DevCsr &isai_first(const PrecondDev &p) { return p.mm.rowptr ? p.mm : p.ml; }

// Widest producing grid of one solve (vector kernels, SpMV, first ISAI SpMV)
// and the doubles its partial regions take.
static size_t parts_stride(const DevCsr &a, const PrecondDev &p, int n) {
  const int64_t gv = grid_for(n, 4), ga = std::max(a.n_blocks, 1);
  const int64_t gm = p.kind == PK_ISAI ? std::max(isai_first(p).n_blocks, 1) : gv;
  return (size_t)std::max(std::max(gv, ga), gm);
}
static size_t parts_doubles(const DevCsr &a, const PrecondDev &p, int n) {
  return std::max<size_t>(parts_stride(a, p, n) * 11, (size_t)kMaxReduceBlocks * 4);
}

// Partial regions and consumer widths for one solve, in `base` (at least
// parts_doubles() doubles) or, when base is null, in the backend's partials
// buffer (grown to fit). A consumer re-sums its producer's partials unless
// that means consumer workgroups x producer workgroups beyond
// kDeferMaxReads; then a one-workgroup k_fold launch reduces them first.
static ohm_status parts_for(ohm_backend *b, const DevCsr &a, const PrecondDev &p, int n,
                            Parts &pp, double *base = nullptr) {
  const int64_t gv = grid_for(n, 4), ga = std::max(a.n_blocks, 1);
  const bool isai = p.kind == PK_ISAI;
  const int64_t gm = isai ? std::max(isai_first(p).n_blocks, 1) : gv;
  const int64_t g3 = gm;  // K3 runs on the first ISAI SpMV's grid, else gv
  auto fold = [](int64_t producer, int64_t consumer) {
    return producer * consumer > kDeferMaxReads;
  };
  const size_t S = parts_stride(a, p, n);
  if (!base) OHM_TRY(ensure_partials(b, parts_doubles(a, p, n)));
  double *q = base ? base : b->d_partials;
  pp.init = q;
  pp.rv = q + 4 * S;
  pp.ss = q + 6 * S;
  pp.ts = q + 7 * S;
  pp.rr = q + 9 * S;
  pp.f_rr = fold(gv, gm);  // K5 -> K1 (on gm workgroups: gv, or the ISAI SpMV's)
  pp.f_rv = fold(ga, gm);  // K2 -> K3
  pp.f_ss = fold(g3, ga);  // K3 -> K4
  pp.f_ts = fold(ga, gv);  // K4 -> K5
  pp.n_rv = pp.f_rv ? 1 : (int)ga;
  pp.n_ss = pp.f_ss ? 1 : (int)g3;
  pp.n_ts = pp.f_ts ? 1 : (int)ga;
  pp.n_rr = pp.f_rr ? 1 : (int)gv;
  return OHM_OK;
}

template <int K>
static void fold_launch(ohm_backend *b, double *partials, int np) {
  hipLaunchKernelGGL(k_fold<K>, dim3(1), dim3(kBlock), 0, b->stream, partials, np);
  b->dispatches++;
}

template <typename T>
static void launch_spmv_alpha(ohm_backend *b, const DevCsr &a, const T *av, BicgVecs<T> &v,
                              SolveState *st, const Parts &pp) {
  with_per(a.per, [&](auto P) {
    hipLaunchKernelGGL((k_spmv_alpha<T, decltype(P)::value>), dim3(a.n_blocks), dim3(kBlock), 0,
                       b->stream, a.rowptr, a.col, av, a.blocks, v.ph, v.v, v.rh, st, pp);
  });
}

// One iteration: K1 .. K5 (ISAI: the p and s updates ride on the first
// preconditioner SpMV; a two-factor apply adds its second SpMV after each).
template <typename T, int PK>
static void enqueue_iteration(ohm_backend *b, const DevCsr &a, const PrecondDev &p,
                              BicgVecs<T> &v, SolveState *st, int n, const Parts &pp) {
  const dim3 gv(grid_for(n, 4)), blk(kBlock), ga(a.n_blocks);
  const T *inv = Ops<T>::inv(p);
  const T *av = Ops<T>::vals(a);
  if (PK == PK_ISAI) {
    const DevCsr &m1 = isai_first(p);
    const bool two = !p.mm.rowptr;
    // q (= p - omega v, written by K5) lives in v.p, which k_init zeroes;
    // p itself in v.p2
    T *y1 = two ? v.tmp : v.ph;
    with_per(m1.per, [&](auto P) {
      hipLaunchKernelGGL((k_p_spmv<T, decltype(P)::value>), dim3(m1.n_blocks), blk, 0, b->stream,
                         m1.rowptr, m1.col, Ops<T>::vals(m1), m1.blocks, v.r, v.p, v.p2, y1, st,
                         pp);
    });
    if (two) Ops<T>::spmv(b, p.mu, v.tmp, v.ph, st, 1);
    launch_spmv_alpha<T>(b, a, av, v, st, pp);
    if (pp.f_rv) fold_launch<2>(b, pp.rv, a.n_blocks);
    T *y3 = two ? v.tmp : v.sh;
    with_per(m1.per, [&](auto P) {
      hipLaunchKernelGGL((k_s_spmv<T, decltype(P)::value>), dim3(m1.n_blocks), blk, 0, b->stream,
                         m1.rowptr, m1.col, Ops<T>::vals(m1), m1.blocks, v.r, v.v, v.s, y3, st,
                         pp);
    });
    if (pp.f_ss) fold_launch<1>(b, pp.ss, m1.n_blocks);
    if (two) Ops<T>::spmv(b, p.mu, v.tmp, v.sh, st, 1);
  } else {
    hipLaunchKernelGGL((k_p_update<T, PK>), gv, blk, 0, b->stream, n, v.r, v.p, v.v, v.ph, inv,
                       st, pp);
    launch_spmv_alpha<T>(b, a, av, v, st, pp);
    if (pp.f_rv) fold_launch<2>(b, pp.rv, a.n_blocks);
    hipLaunchKernelGGL((k_s_update<T, PK>), gv, blk, 0, b->stream, n, v.r, v.v, v.s, v.sh, inv,
                       st, pp);
    if (pp.f_ss) fold_launch<1>(b, pp.ss, (int)gv.x);
  }
  with_per(a.per, [&](auto P) {
    hipLaunchKernelGGL((k_spmv_omega<T, decltype(P)::value>), ga, blk, 0, b->stream, a.rowptr,
                       a.col, av, a.blocks, v.sh, v.t, v.s, st, pp);
  });
  if (pp.f_ts) fold_launch<2>(b, pp.ts, a.n_blocks);
  hipLaunchKernelGGL(k_xr_update<T>, gv, blk, 0, b->stream, n, v.x, v.r, v.rh, v.ph, v.sh, v.s,
                     v.t, v.p2, v.v, PK == PK_ISAI ? v.p : nullptr, st, pp);
  if (pp.f_rr) fold_launch<2>(b, pp.rr, (int)gv.x);
  b->dispatches += 5;
}

template <typename T>
static void enqueue_any(ohm_backend *b, const DevCsr &a, const PrecondDev &p, BicgVecs<T> &v,
                        SolveState *st, int n, const Parts &pp) {
  if (p.kind == PK_JACOBI)
    enqueue_iteration<T, PK_JACOBI>(b, a, p, v, st, n, pp);
  else if (p.kind == PK_ISAI)
    enqueue_iteration<T, PK_ISAI>(b, a, p, v, st, n, pp);
  else
    enqueue_iteration<T, PK_NONE>(b, a, p, v, st, n, pp);
}

static void enqueue_check(ohm_backend *b, SolveState *st, const Parts &pp) {
  hipLaunchKernelGGL(k_check, dim3(1), dim3(kBlock), 0, b->stream, st, pp);
  b->dispatches++;
}

// Captured chunk lengths: an inexact-Newton sub-solve (tolerance relative to
// its own starting residual) opens with a short chunk, since most of them end
// within a few iterations and a replay runs to its end even when the solve
// stops inside it; every later chunk, and every other solve, replays long
// chunks.
constexpr int kGraphItersShort = 4, kGraphIters = 8;

void graph_cache_clear(ohm_backend *b) {
  for (auto &g : b->graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
  }
  b->graphs.clear();
}

// Replays `reps` launches of a captured `iters`-iteration chunk. The key
// holds every pointer and launch size baked into the captured kernels.
template <typename T>
static ohm_status replay_chunk(ohm_backend *b, const DevCsr &a, const PrecondDev &p,
                               BicgVecs<T> &v, SolveState *st, int n, const Parts &pp,
                               int iters, int reps) {
  std::vector<uintptr_t> key = {
      sizeof(T), (uintptr_t)p.kind, (uintptr_t)n, (uintptr_t)a.rowptr, (uintptr_t)a.col,
      (uintptr_t)Ops<T>::vals(a), (uintptr_t)a.blocks, (uintptr_t)a.n_blocks,
      (uintptr_t)Ops<T>::inv(p), (uintptr_t)p.ml.rowptr, (uintptr_t)Ops<T>::vals(p.ml),
      (uintptr_t)p.ml.blocks, (uintptr_t)p.ml.n_blocks, (uintptr_t)p.mu.rowptr,
      (uintptr_t)Ops<T>::vals(p.mu), (uintptr_t)p.mu.blocks, (uintptr_t)p.mu.n_blocks,
      (uintptr_t)p.mm.rowptr, (uintptr_t)Ops<T>::vals(p.mm), (uintptr_t)p.mm.blocks,
      (uintptr_t)p.mm.n_blocks,
      (uintptr_t)v.x, (uintptr_t)v.tmp, (uintptr_t)st, (uintptr_t)pp.init,
      (uintptr_t)iters};
  hipGraphExec_t exec = nullptr;
  for (auto &g : b->graphs)
    if (g.key == key) exec = g.exec;
  const int64_t per_iter = 5 + (p.kind == PK_ISAI && !p.mm.rowptr ? 2 : 0) + pp.f_rv + pp.f_ss +
                           pp.f_ts + pp.f_rr;
  if (!exec) {
    int64_t saved = b->dispatches;
    GraphEntry e;
    e.key = key;
    OHM_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
    for (int k = 0; k < iters; k++) enqueue_any<T>(b, a, p, v, st, n, pp);
    enqueue_check(b, st, pp);
    OHM_HIP(hipStreamEndCapture(b->stream, &e.graph));
    OHM_HIP(hipGraphInstantiate(&e.exec, e.graph, nullptr, nullptr, 0));
    b->dispatches = saved;
    if (b->graphs.size() >= 32) graph_cache_clear(b);
    b->graphs.push_back(e);
    exec = e.exec;
  }
  for (int r = 0; r < reps; r++) OHM_HIP(hipGraphLaunch(exec, b->stream));
  b->dispatches += (per_iter * iters + 1) * reps;
  return OHM_OK;
}

struct StatusWord {  // SolveState from `status` on
  int status, s_conv, iters, max_iters;
  unsigned counter;
  int scaled;
  double r0, res2, eta;
};
static_assert(offsetof(SolveState, r0) - offsetof(SolveState, status) ==
                      offsetof(StatusWord, r0) &&
                  offsetof(SolveState, eta) - offsetof(SolveState, status) ==
                      offsetof(StatusWord, eta) &&
                  sizeof(SolveState) - offsetof(SolveState, status) == sizeof(StatusWord),
              "StatusWord mirrors the tail of SolveState");

// Enqueues the start of a solve on the workspace vectors: state reset, the
// warm-start product A x0 and k_init (rhs is a separate device vector; x0 is
// read from w.v.x). With stop set (captured solves) the solve is masked off
// while *stop is non-zero.
template <typename T>
static ohm_status bicg_start(ohm_backend *b, const DevCsr &a, const PrecondDev &p, const T *rhs,
                             Workspace<T> &w, const BicgOpts &o, Parts &pp,
                             const int *stop = nullptr) {
  const int n = a.n;
  SolveState *st = w.st;
  BicgVecs<T> &v = w.v;
  OHM_TRY(parts_for(b, a, p, n, pp, w.partials));
  hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(64), 0, b->stream, st, stop);
  if (o.warm) Ops<T>::spmv(b, a, v.x, v.t, stop ? st : nullptr, stop ? 1 : 0);
  hipLaunchKernelGGL(k_init<T>, dim3(grid_for(n, 4)), dim3(kBlock), 0, b->stream, n, rhs, v.t, v.x,
                     v.r, v.rh, v.p, v.v, o.shadow == OHM_SHADOW_HASHED ? 1 : 0, o.warm ? 1 : 0,
                     o.tol, o.tol_floor, o.tol_mode,
                     (int)std::min<int64_t>(o.max_iters, 0x7fffffff), o.scaled_breakdown ? 1 : 0,
                     o.forcing, st, pp, stop ? 1 : 0);
  b->dispatches += 2;
  return OHM_OK;
}

// Drives a started solve to its end: chunks of iterations (graph replays or
// eager launches), each followed by a poll of the status word. `enqueued`
// counts iterations already enqueued (a resumed captured solve); `poll`
// says whether to read the status before the first chunk.
template <typename T>
static ohm_status bicg_drive(ohm_backend *b, const DevCsr &a, const PrecondDev &p,
                             Workspace<T> &w, const BicgOpts &o, const Parts &pp, bool poll,
                             int64_t enqueued, BicgOut &out) {
  const int n = a.n;
  SolveState *st = w.st;
  BicgVecs<T> &v = w.v;
  StatusWord *h = reinterpret_cast<StatusWord *>(b->h_pinned);
  for (;; poll = true) {
    int64_t done = 0;
    if (poll) {
      OHM_HIP(hipMemcpyAsync(h, &st->status, sizeof(StatusWord), hipMemcpyDeviceToHost,
                             b->stream));
      OHM_HIP(hipStreamSynchronize(b->stream));
      b->readbacks++;
      out.polls++;
      if (h->status != ST_RUNNING) break;
      done = h->iters;
    }
    int chunk = o.check_interval > 0
                    ? o.check_interval
                    : (int)std::min<int64_t>(16, std::max<int64_t>(4, done / 2));
    // With an expected length E: a first chunk ending about 4 short of E
    // (at most 64 iterations, so that a long earlier solve cannot queue a
    // long run of no-ops), then chunks that close on E - 4 (at most 16) and
    // chunks of 4 from there to E + 8; past that the usual growth.
    bool near = false;
    if (o.check_interval <= 0 && o.expect_iters > 0) {
      const int64_t e = o.expect_iters;
      if (enqueued == 0) {
        chunk = (int)std::min<int64_t>(64, std::max<int64_t>(kGraphItersShort,
                                                             (e - 4) / kGraphIters * kGraphIters));
        near = true;
      } else if (done < e - 4) {
        chunk = (int)std::min<int64_t>(16, std::max<int64_t>(kGraphItersShort, (e - 4 - done + 3) / 4 * 4));
        near = true;
      } else if (done < e + 8) {
        chunk = kGraphItersShort;
        near = true;
      }
    }
    chunk = (int)std::min<int64_t>(chunk, std::max<int64_t>(1, o.max_iters - enqueued));
    if (b->use_graphs && !o.eager) {
      // replay captured chunks; surplus iterations are no-ops once the
      // state leaves RUNNING
      const bool newton_sub = o.tol_mode == OHM_TOL_INITIAL || o.forcing.kind != 0;
      const int gi = near ? (chunk % kGraphIters == 0 ? kGraphIters : kGraphItersShort)
                          : (newton_sub && enqueued == 0 && chunk < kGraphIters ? kGraphItersShort
                                                                                 : kGraphIters);
      int reps = (chunk + gi - 1) / gi;
      OHM_TRY(replay_chunk<T>(b, a, p, v, st, n, pp, gi, reps));
      chunk = reps * gi;
    } else {
      for (int k = 0; k < chunk; k++)
        enqueue_any<T>(b, a, p, v, st, n, pp);
      enqueue_check(b, st, pp);
    }
    enqueued += chunk;
    OHM_HIP(hipGetLastError());
    // Every enqueued iteration either advances iters or ends the solve, so a
    // state that is still RUNNING well past max_iters means a lost launch.
    if (enqueued > o.max_iters + 64)
      return fail(OHM_ERR_SOLVE, "Solve error: BiCGSTAB state did not advance (%lld enqueued)",
                  (long long)enqueued);
  }
  out.status = h->status;
  out.iters = h->iters;
  out.r0 = h->r0;
  out.res = sqrt(h->res2);
  out.eta = h->eta;
  OHM_HIP(hipGetLastError());
  return OHM_OK;
}

// Runs BiCGSTAB on the workspace vectors (rhs a separate device vector, x0
// read from w.v.x when warm).
template <typename T>
static ohm_status bicg_core(ohm_backend *b, const DevCsr &a, const PrecondDev &p, const T *rhs,
                            Workspace<T> &w, const BicgOpts &o, BicgOut &out) {
  Parts pp;
  OHM_TRY(bicg_start<T>(b, a, p, rhs, w, o, pp));
  // The first chunk goes out behind k_init without a poll: the rare solve
  // that k_init ends (b = 0, rho breakdown; under OHM_TOL_INITIAL a warm
  // start only once ||r0|| is below the floor) turns every kernel of the
  // chunk into a no-op at its head. A warm start against tol * ||b|| polls
  // first, since k_init ends it whenever x already meets that tolerance.
  // With an expected length the first chunk goes out without that poll too.
  return bicg_drive<T>(b, a, p, w, o, pp,
                       o.warm && o.tol_mode != OHM_TOL_INITIAL && o.expect_iters <= 0, 0, out);
}

// ----------------------------------------------------- captured solves --
// FP64 solves enqueued inside a caller's graph capture: a workspace of their
// own with x aliased to the caller's vector, so that the captured kernels'
// pointers stay valid for every replay.
struct CapturedSolve {
  Workspace<double> w;
  double *vecs = nullptr, *partials = nullptr;
  SolveState *st = nullptr;
  Parts pp{};
  int n = 0;
};

ohm_status captured_create(ohm_backend *b, const DevCsr &a, const PrecondDev &p, double *x,
                           CapturedSolve **out) {
  (void)b;
  CapturedSolve *c = new CapturedSolve();
  const int n = a.n;
  const size_t stride = ws_stride(n);
  hipError_t e = hipMalloc(&c->vecs, sizeof(double) * stride * kWsVectors);
  if (e == hipSuccess) e = hipMalloc(&c->st, sizeof(SolveState));
  if (e == hipSuccess) e = hipMalloc(&c->partials, sizeof(double) * parts_doubles(a, p, n));
  if (e != hipSuccess) {
    captured_free(c);
    return hip_fail(e, "hipMalloc (captured solve workspace)", __FILE__, __LINE__);
  }
  ws_views(c->w, c->vecs, stride);
  c->w.v.x = x;
  c->w.st = c->st;
  c->w.partials = c->partials;
  c->w.n = n;
  c->n = n;
  *out = c;
  return OHM_OK;
}

void captured_free(CapturedSolve *c) {
  if (!c) return;
  if (c->vecs) (void)hipFree(c->vecs);
  if (c->st) (void)hipFree(c->st);
  if (c->partials) (void)hipFree(c->partials);
  delete c;
}

const SolveState *captured_state(const CapturedSolve *c) { return c->st; }

ohm_status captured_enqueue(ohm_backend *b, CapturedSolve *c, const DevCsr &a,
                            const PrecondDev &p, const double *rhs, const BicgOpts &o, int iters,
                            const int *stop) {
  if (a.n != c->n) return fail(OHM_ERR_INVALID, "captured solve: dimension mismatch");
  OHM_TRY(bicg_start<double>(b, a, p, rhs, c->w, o, c->pp, stop));
  for (int k = 0; k < iters; k++) enqueue_any<double>(b, a, p, c->w.v, c->st, c->n, c->pp);
  enqueue_check(b, c->st, c->pp);
  return OHM_OK;
}

ohm_status captured_resume(ohm_backend *b, CapturedSolve *c, const DevCsr &a,
                           const PrecondDev &p, const BicgOpts &o, int64_t enqueued,
                           BicgOut &out) {
  hipLaunchKernelGGL(k_unskip, dim3(1), dim3(64), 0, b->stream, c->st);
  OHM_HIP(hipGetLastError());
  b->dispatches++;
  return bicg_drive<double>(b, a, p, c->w, o, c->pp, true, enqueued, out);
}

ohm_status status_to_error(int status, int64_t iters, int64_t max_iters) {
  switch (status) {
    case ST_CONVERGED:
      return OHM_OK;
    case ST_BRK_RHO:
      return fail(OHM_ERR_SOLVE, "Solve error: BiCGSTAB breakdown: rho ~ 0");
    case ST_BRK_RV:
      return fail(OHM_ERR_SOLVE, "Solve error: BiCGSTAB breakdown: r_hat.v ~ 0");
    case ST_BRK_TT:
      return fail(OHM_ERR_SOLVE, "Solve error: BiCGSTAB breakdown: ||t|| ~ 0");
    case ST_NAN:
      return fail(OHM_ERR_SOLVE, "Solve error: BiCGSTAB diverged: NaN/Inf in residual");
    case ST_BRK_OMEGA:
      return fail(OHM_ERR_SOLVE, "Solve error: BiCGSTAB breakdown: omega ~ 0");
    default:
      return fail(OHM_ERR_SOLVE, "Solve error: BiCGSTAB did not converge within %lld iterations",
                  (long long)max_iters);
  }
}

static ohm_status check_params(const BicgOpts &o) {
  if (o.max_iters <= 0) return fail(OHM_ERR_SOLVE, "Solve error: BiCGSTAB max_iterations must be > 0");
  if (!(o.tol > 0.0) || !std::isfinite(o.tol))
    return fail(OHM_ERR_SOLVE, "Solve error: BiCGSTAB tolerance must be finite and > 0");
  return OHM_OK;
}

// Systems of at most 64 unknowns run the one-wave solve (one launch and one
// status readback per solve) unless OHM_FUSED_SMALL=0.
static bool fused_small(int n) {
  if (n > kFusedRows) return false;
  const char *e = getenv("OHM_FUSED_SMALL");
  return !e || atoi(e) != 0;
}

static ohm_status bicg_solve_fused(ohm_backend *b, const DevCsr &a, const PrecondDev &p,
                                   const double *rhs, double *x, const BicgOpts &o,
                                   BicgOut &out) {
  OHM_TRY(ws_get(g_ws64, a.n));
  SolveState *st = g_ws64.st;
  const FusedCsr A{a.rowptr, a.col, a.val};
  FusedCsr L{nullptr, nullptr, nullptr}, U{nullptr, nullptr, nullptr};
  int two = 0;
  uint32_t lds = fused_stage_bytes((uint32_t)a.nnz);
  if (p.kind == PK_ISAI) {
    const DevCsr &m1 = isai_first(p);
    L = FusedCsr{m1.rowptr, m1.col, m1.val};
    lds += fused_stage_bytes((uint32_t)m1.nnz);
    two = p.mm.rowptr ? 0 : 1;
    if (two) {
      U = FusedCsr{p.mu.rowptr, p.mu.col, p.mu.val};
      lds += fused_stage_bytes((uint32_t)p.mu.nnz);
    }
  }
  // the matrices go to LDS when they fit beside the vector staging
  const bool stage = lds <= kFusedLdsBytes;
  const int shadow = o.shadow == OHM_SHADOW_HASHED ? 1 : 0, warm = o.warm ? 1 : 0;
  const int max_iters = (int)std::min<int64_t>(o.max_iters, 0x7fffffff);
  const int scaled = o.scaled_breakdown ? 1 : 0;
  const dim3 g(1), blk(kFusedRows);
  const unsigned shm = stage ? lds : 0;
  auto launch = [&](auto kern) {
    hipLaunchKernelGGL(kern, g, blk, shm, b->stream, a.n, A, L, U, two, p.inv, rhs, x, shadow,
                       warm, o.tol, o.tol_floor, o.tol_mode, max_iters, scaled, o.forcing, st);
  };
  if (p.kind == PK_JACOBI)
    stage ? launch(k_bicg_fused<PK_JACOBI, true>) : launch(k_bicg_fused<PK_JACOBI, false>);
  else if (p.kind == PK_ISAI)
    stage ? launch(k_bicg_fused<PK_ISAI, true>) : launch(k_bicg_fused<PK_ISAI, false>);
  else
    stage ? launch(k_bicg_fused<PK_NONE, true>) : launch(k_bicg_fused<PK_NONE, false>);
  OHM_HIP(hipGetLastError());
  b->dispatches++;
  StatusWord *h = reinterpret_cast<StatusWord *>(b->h_pinned);
  OHM_HIP(hipMemcpyAsync(h, &st->status, sizeof(StatusWord), hipMemcpyDeviceToHost, b->stream));
  OHM_HIP(hipStreamSynchronize(b->stream));
  b->readbacks++;
  out.polls++;
  out.status = h->status;
  out.iters = h->iters;
  out.r0 = h->r0;
  out.res = sqrt(h->res2);
  out.eta = h->eta;
  return status_to_error(out.status, out.iters, o.max_iters);
}

static bool is_breakdown(int status) {
  return status == ST_BRK_RHO || status == ST_BRK_RV || status == ST_BRK_TT ||
         status == ST_BRK_OMEGA;
}

// The retry of a broken-down solve: the other shadow vector; counts add up.
static BicgOpts retry_opts(const BicgOpts &o) {
  BicgOpts r = o;
  r.shadow = o.shadow == OHM_SHADOW_HASHED ? OHM_SHADOW_RHS : OHM_SHADOW_HASHED;
  return r;
}
static void merge_retry(BicgOut &out, const BicgOut &r) {
  out.iters += r.iters;
  out.polls += r.polls;
  out.status = r.status;
  out.res = r.res;
  out.eta = r.eta;
}

ohm_status bicg_solve_f64(ohm_backend *b, const DevCsr &a, const PrecondDev &p, const double *rhs,
                          double *x, const BicgOpts &o, BicgOut &out) {
  OHM_TRY(check_params(o));
  if (a.n == 0) {
    out.status = ST_CONVERGED;
    return OHM_OK;
  }
  const size_t xb = sizeof(double) * a.n;
  if (fused_small(a.n)) {
    if (!o.retry_breakdown) return bicg_solve_fused(b, a, p, rhs, x, o, out);
    OHM_TRY(ws_get(g_ws64, a.n));  // the starting iterate, kept for a retry
    OHM_HIP(hipMemcpyAsync(g_ws64.v.r, x, xb, hipMemcpyDeviceToDevice, b->stream));
    ohm_status s = bicg_solve_fused(b, a, p, rhs, x, o, out);
    if (!is_breakdown(out.status)) return s;
    OHM_HIP(hipMemcpyAsync(x, g_ws64.v.r, xb, hipMemcpyDeviceToDevice, b->stream));
    BicgOut r;
    s = bicg_solve_fused(b, a, p, rhs, x, retry_opts(o), r);
    merge_retry(out, r);
    return s;
  }
  OHM_TRY(ws_get(g_ws64, a.n));
  if (o.warm) OHM_HIP(hipMemcpyAsync(g_ws64.v.x, x, xb, hipMemcpyDeviceToDevice, b->stream));
  OHM_TRY(bicg_core<double>(b, a, p, rhs, g_ws64, o, out));
  if (o.retry_breakdown && is_breakdown(out.status)) {
    // x still holds the starting iterate (copied back only below)
    if (o.warm) OHM_HIP(hipMemcpyAsync(g_ws64.v.x, x, xb, hipMemcpyDeviceToDevice, b->stream));
    BicgOut r;
    OHM_TRY(bicg_core<double>(b, a, p, rhs, g_ws64, retry_opts(o), r));
    merge_retry(out, r);
  }
  OHM_HIP(hipMemcpyAsync(x, g_ws64.v.x, xb, hipMemcpyDeviceToDevice, b->stream));
  OHM_HIP(hipStreamSynchronize(b->stream));
  return status_to_error(out.status, out.iters, o.max_iters);
}

ohm_status bicg_solve_transpose(ohm_backend *b, DevCsr &a, PrecondDev &p, const double *rhs,
                                double *x, const BicgOpts &o, BicgOut &out) {
  DevCsr at;
  OHM_TRY(csr_transpose_view(b, a, at));
  PrecondDev pt;  // view: shares every device array of p
  pt.kind = p.kind;
  pt.level = p.level;
  pt.n = p.n;
  pt.inv = p.inv;  // diag(A^T) = diag(A); the zero-diagonal rule is symmetric too
  if (p.kind == PK_ISAI) {
    DevCsr mlt, mut;
    OHM_TRY(csr_transpose_view(b, p.ml, mlt));
    OHM_TRY(csr_transpose_view(b, p.mu, mut));
    pt.ml = mut;  // applied first: M_U^T
    pt.mu = mlt;  // then M_L^T
  }
  BicgOpts ot = o;
  ot.transpose = false;
  return bicg_solve_f64(b, at, pt, rhs, x, ot, out);
}

// ------------------------------------------------- iterative refinement --
// r = b - A x (FP64) with ||r||^2 and ||b||^2 into d_scalars[0..1]; the
// scaled residual r / ||r|| is written in FP32 by the next kernel.
template <int PER>
__global__ __launch_bounds__(kBlock) void k_residual(const uint32_t *__restrict__ rowptr,
                                                     const uint32_t *__restrict__ col,
                                                     const double *__restrict__ val,
                                                     const uint32_t *__restrict__ blocks,
                                                     const double *__restrict__ x,
                                                     const double *__restrict__ rhs,
                                                     double *__restrict__ r, double *partials,
                                                     unsigned *counter, double *out) {
  Acc<double> rr, bb;
  spmv_block<double, PER>(rowptr, col, val, blocks, x, [&](uint32_t i, double y) {
    double ri = rhs[i] - y;
    r[i] = ri;
    rr.add_prod(ri, ri);
    bb.add_prod(rhs[i], rhs[i]);
  });
  double vals[2] = {rr.get(), bb.get()}, tot[2];
  if (grid_sum<2>(vals, partials, counter, tot) && threadIdx.x == 0) {
    out[0] = tot[0];
    out[1] = tot[1];
  }
}

__global__ void k_scale_to_f32(int n, const double *__restrict__ r, const double *__restrict__ nrm2,
                               float *__restrict__ r32) {
  const double inv = 1.0 / sqrt(nrm2[0]);
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    r32[i] = (float)(r[i] * inv);
}

__global__ void k_correct(int n, const float *__restrict__ d, const double *__restrict__ nrm2,
                          double *__restrict__ x) {
  const double sc = sqrt(nrm2[0]);
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    x[i] += sc * (double)d[i];
}

static double *g_ir_r = nullptr;
static float *g_ir_r32 = nullptr;
static int g_ir_n = 0;
static double *g_ir_x0 = nullptr;  // the caller's starting iterate (FP64 rerun)
static int g_ir_x0_n = 0;

static ohm_status residual_norm(ohm_backend *b, const DevCsr &a, const double *rhs,
                                const double *x, double *r, double *partials, double &rel);

ohm_status true_residual(ohm_backend *b, const DevCsr &a, const double *rhs, const double *x,
                         double &rel) {
  if (a.n == 0) {
    rel = 0.0;
    return OHM_OK;
  }
  if (g_ir_n < a.n) {
    if (g_ir_r) OHM_HIP(hipFree(g_ir_r));
    if (g_ir_r32) OHM_HIP(hipFree(g_ir_r32));
    OHM_HIP(hipMalloc(&g_ir_r, sizeof(double) * a.n));
    OHM_HIP(hipMalloc(&g_ir_r32, sizeof(float) * a.n));
    g_ir_n = a.n;
  }
  OHM_TRY(ensure_partials(b, (size_t)std::max(a.n_blocks, kMaxReduceBlocks) * 4));
  return residual_norm(b, a, rhs, x, g_ir_r, b->d_partials, rel);
}

// ||b - A x|| / ||b|| with r and the partial sums (max(n_blocks, 2048) * 4
// doubles) in the caller's buffers (synchronous).
static ohm_status residual_norm(ohm_backend *b, const DevCsr &a, const double *rhs,
                                const double *x, double *r, double *partials, double &rel) {
  with_per(a.per, [&](auto P) {
    hipLaunchKernelGGL((k_residual<decltype(P)::value>), dim3(a.n_blocks), dim3(kBlock), 0,
                       b->stream, a.rowptr, a.col, a.val, a.blocks, x, rhs, r, partials,
                       b->d_counter, b->d_scalars);
  });
  OHM_HIP(hipGetLastError());
  b->dispatches++;
  OHM_HIP(hipMemcpyAsync(b->h_pinned, b->d_scalars, 2 * sizeof(double), hipMemcpyDeviceToHost,
                         b->stream));
  OHM_HIP(hipStreamSynchronize(b->stream));
  b->readbacks++;
  double rn = sqrt(b->h_pinned[0]), bn = sqrt(b->h_pinned[1]);
  rel = bn > 0 ? rn / bn : rn;
  return OHM_OK;
}

// ------------------------------------------------ allocation-free solve --
// Scratch of the allocation-free solve: the eleven solve vectors, the state
// word and the partial sums of one FP64 solve, carved from one file-scope
// device array that each GPU holds for the life of the process. On the host
// the array's name is a stand-in symbol, not a device address: each device's
// address is looked up once (hipGetSymbolAddress, with that device current)
// and only that address is passed to kernels and copies. Every solve on a
// GPU shares that GPU's array, so solves that overlap on one GPU race.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxScratchDevices = 64;
static void *g_scratch_addr[kMaxScratchDevices] = {};
static std::mutex g_scratch_mu;

ohm_status static_scratch_init(ohm_backend *b) {
  if (b->device < 0 || b->device >= kMaxScratchDevices)
    return fail(OHM_ERR_INVALID, "device %d is beyond the scratch table (%d devices)", b->device,
                kMaxScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  void *&addr = g_scratch_addr[b->device];
  if (!addr) {
    OHM_HIP(hipSetDevice(b->device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    addr = p;
  }
  b->scratch = addr;
  return OHM_OK;
}

static size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

ohm_status bicg_solve_f64_noalloc(ohm_backend *b, const DevCsr &a, const PrecondDev &p,
                                  const double *rhs, double *x, BicgOpts o, BicgOut &out,
                                  double &rel) {
  OHM_TRY(check_params(o));
  if (!b->scratch) return fail(OHM_ERR_INVALID, "the backend's scratch address is not set");
  const int n = a.n;
  rel = 0.0;
  if (n == 0) {
    out.status = ST_CONVERGED;
    return OHM_OK;
  }
  // vectors | state word | partial sums (the solve's regions, then the
  // residual's max(n_blocks, 2048) * 4 doubles, which they cover)
  const size_t stride = ws_stride(n);
  const size_t st_off = align256(sizeof(double) * stride * kWsVectors);
  const size_t part_off = st_off + align256(sizeof(SolveState));
  const size_t need = part_off + sizeof(double) * parts_doubles(a, p, n);
  if (need > kScratchBytes)
    return fail(OHM_ERR_INVALID,
                "ohm_bicgstab_noalloc: %d unknowns need %zu bytes of scratch, more than its "
                "%zu bytes (about 700,000 unknowns)",
                n, need, kScratchBytes);
  char *base = (char *)b->scratch;
  Workspace<double> w;
  ws_views(w, (double *)base, stride);
  w.st = (SolveState *)(base + st_off);
  w.partials = (double *)(base + part_off);
  w.n = n;
  o.eager = true;
  if (o.warm)
    OHM_HIP(hipMemcpyAsync(w.v.x, x, sizeof(double) * n, hipMemcpyDeviceToDevice, b->stream));
  OHM_TRY(bicg_core<double>(b, a, p, rhs, w, o, out));
  OHM_HIP(hipMemcpyAsync(x, w.v.x, sizeof(double) * n, hipMemcpyDeviceToDevice, b->stream));
  // the solve is over: its t vector and partial regions hold the residual
  ohm_status s = residual_norm(b, a, rhs, x, w.v.t, w.partials, rel);
  if (s != OHM_OK) return s;
  return status_to_error(out.status, out.iters, o.max_iters);
}

ohm_status bicg_solve_ir(ohm_backend *b, DevCsr &a, PrecondDev &p, const double *rhs, double *x,
                         const BicgOpts &o, double inner_tol, int64_t max_ref, BicgOut &out,
                         int64_t &refinements) {
  OHM_TRY(check_params(o));
  refinements = 0;
  const int n = a.n;
  if (n == 0) {
    out.status = ST_CONVERGED;
    return OHM_OK;
  }
  OHM_TRY(csr_ensure_f32(b, a));
  OHM_TRY(precond_ensure_f32(b, p));
  OHM_TRY(ws_get(g_ws32, n));
  if (o.warm) {
    if (g_ir_x0_n < n) {
      if (g_ir_x0) OHM_HIP(hipFree(g_ir_x0));
      OHM_HIP(hipMalloc(&g_ir_x0, sizeof(double) * n));
      g_ir_x0_n = n;
    }
    OHM_HIP(hipMemcpyAsync(g_ir_x0, x, sizeof(double) * n, hipMemcpyDeviceToDevice, b->stream));
  } else {
    OHM_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, b->stream));
  }
  double rel = 0.0, prev = INFINITY, target = o.tol;
  int64_t total = 0;
  int last_status = ST_RUNNING, stalled = 0;
  constexpr int64_t kInnerMaxIters = 1000;  // per refinement
  for (int64_t k = 0; k <= max_ref; k++) {
    OHM_TRY(true_residual(b, a, rhs, x, rel));
    double bn = sqrt(b->h_pinned[1]);
    if (k == 0) {
      out.r0 = rel * bn;
      double tol = o.tol;
      if (o.forcing.kind) {
        tol = forcing_eta(o.forcing, rel * bn);
        out.eta = tol;
      }
      if (o.tol_mode == OHM_TOL_INITIAL || o.forcing.kind) target = std::max(tol * rel, o.tol_floor);
    }
    out.res = rel * bn;
    if (bn < 1e-30 || rel < target) {
      out.status = ST_CONVERGED;
      out.iters = total;
      return OHM_OK;
    }
    if (!std::isfinite(rel)) {
      out.status = ST_NAN;
      return status_to_error(ST_NAN, total, o.max_iters);
    }
    // A refinement that does not halve the FP64 residual means the FP32
    // inner solver is beyond its attainable accuracy (kappa * 2^-24 >~ 1).
    if (k > 0 && rel > 0.5 * prev) stalled = 1;
    if (k == max_ref || total >= o.max_iters || stalled) break;
    prev = rel;
    hipLaunchKernelGGL(k_scale_to_f32, dim3(grid_for(n, 4)), dim3(kBlock), 0, b->stream, n, g_ir_r,
                       b->d_scalars, g_ir_r32);
    BicgOpts io = o;
    io.warm = false;
    io.tol_mode = OHM_TOL_RHS;
    io.forcing.kind = 0;
    io.expect_iters = k == 0 ? o.expect_iters : 0;
    io.tol = std::max(inner_tol, 0.5 * target / std::max(rel, 1e-300));
    io.tol = std::min(io.tol, 0.5);
    io.max_iters = std::max<int64_t>(1, std::min<int64_t>(kInnerMaxIters, o.max_iters - total));
    BicgOut iout;
    OHM_TRY(bicg_core<float>(b, a, p, g_ir_r32, g_ws32, io, iout));
    total += iout.iters;
    out.polls += iout.polls;
    if (k == 0) out.first_inner = iout.iters;
    last_status = iout.status;
    refinements++;
    // An FP32 breakdown or iteration cap still leaves a useful correction (the
    // inner residual stalls at its FP32 attainable accuracy); only a NaN
    // iterate is discarded. The FP64 residual decides convergence.
    if (iout.status != ST_NAN && iout.iters > 0) {
      hipLaunchKernelGGL(k_correct, dim3(grid_for(n, 4)), dim3(kBlock), 0, b->stream, n,
                         g_ws32.v.x, b->d_scalars, x);
      b->dispatches += 1;
    }
    b->dispatches += 1;
  }
  // FP32 refinement stalled: the system is beyond FP32's reach (for MNA this
  // is a node grounded only through GMIN or an off junction). Finish in FP64
  // from the current iterate, to the same target.
  OHM_HIP(hipStreamSynchronize(b->stream));
  BicgOpts fo = o;
  fo.warm = true;
  fo.tol_mode = OHM_TOL_RHS;
  fo.tol = target;
  fo.forcing.kind = 0;
  fo.expect_iters = 0;
  fo.max_iters = std::max<int64_t>(1, o.max_iters - total);
  BicgOut f64;
  ohm_status s = bicg_solve_f64(b, a, p, rhs, x, fo, f64);
  out.iters = total + f64.iters;
  out.polls += f64.polls;
  out.status = f64.status;
  out.res = f64.res;
  out.fp64_fallback = true;
  if (s != OHM_OK) {
    // Continuing from the refined iterate failed (its residual is left in
    // the modes BiCGSTAB reduces slowest, and a breakdown or divergence
    // followed): solve the original system in FP64 from the caller's start,
    // so FP32 refinement never fails where the FP64 solver alone succeeds.
    if (o.warm)
      OHM_HIP(hipMemcpyAsync(x, g_ir_x0, sizeof(double) * n, hipMemcpyDeviceToDevice, b->stream));
    else
      OHM_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, b->stream));
    BicgOut re;
    BicgOpts ro = o;
    ro.expect_iters = 0;
    s = bicg_solve_f64(b, a, p, rhs, x, ro, re);
    out.iters += re.iters;
    out.polls += re.polls;
    out.status = re.status;
    out.res = re.res;
    if (o.forcing.kind) out.eta = re.eta;
  }
  (void)last_status;
  return s;
}

}  // namespace ohm

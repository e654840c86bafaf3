// ohm_precond.h — device-side preconditioner objects.
#pragma once

#include "ohm_internal.h"

namespace ohm {

struct IsaiPlan;  // symbolic ILU(0)+ISAI data, defined in ohm_isai.hip

struct PrecondDev {
  int kind = PK_NONE;
  int level = 1;
  int n = 0;
  double *inv = nullptr;   // Jacobi inverse diagonal
  float *inv32 = nullptr;
  DevCsr ml, mu;           // ISAI: z = M_U (M_L r)
  DevCsr mm;               // ISAI: the product M_U M_L when formed (one SpMV per apply)
  IsaiPlan *plan = nullptr;
  int ilu_sweeps = 5;      // fixed-point ILU(0) sweeps per numeric refresh
  bool has_f32 = false;
};

ohm_status precond_build(ohm_backend *b, const DevCsr &a, int kind, int level, PrecondDev &p);
ohm_status precond_refresh(ohm_backend *b, const DevCsr &a, PrecondDev &p);
ohm_status precond_ensure_f32(ohm_backend *b, PrecondDev &p);
void precond_free(PrecondDev &p);
// forget the ILU(0) factors so the next refresh restarts from A
void isai_reset(PrecondDev &p);
// out = M^-1 in (f64), skipping per the solve state like the SpMV kernels
void precond_apply(ohm_backend *b, const PrecondDev &p, const double *in, double *out, double *tmp,
                   const SolveState *st, int skip);
void precond_apply_f32(ohm_backend *b, const PrecondDev &p, const float *in, float *out,
                       float *tmp, const SolveState *st, int skip);
bool matrix_has_zero_diag(ohm_backend *b, const DevCsr &a, const double *h_values);

// ---- BiCGSTAB drivers (ohm_bicgstab.hip) --------------------------------
// Forcing inputs: the previous sub-solve's ||F_k-1||, final linear residual
// and forcing term (f_prev < 0: first Newton iteration, eta = eta_max).
struct Forcing {
  int kind = 0;  // 0: off, OHM_FORCING_EW, OHM_FORCING_EW1
  double f_prev = -1.0, lin_prev = 0.0, eta_prev = 0.0;
  double eta_min = 1e-6, eta_max = 0.1;
  int rate_cap = 0;  // apply the safeguard at every iteration (eta falls at most quadratically)
};

struct BicgOpts {
  int64_t max_iters = 10000;
  double tol = 1e-5;
  int shadow = OHM_SHADOW_HASHED;
  bool warm = false;
  int check_interval = 0;
  bool transpose = false;  // solve A^T x = b (FP64 only)
  bool scaled_breakdown = true;
  int tol_mode = OHM_TOL_RHS;
  double tol_floor = 1e-15;
  // Inexact-Newton forcing evaluated on the device from ||b - A x0|| (the
  // Newton residual ||F_k|| when x0 = x_k) as the solve starts:
  // OHM_FORCING_EW (Eisenstat-Walker choice 2) or OHM_FORCING_EW1
  // (choice 1); tol is then ignored and the forcing term becomes the
  // relative tolerance (OHM_TOL_INITIAL).
  Forcing forcing;
  bool eager = false;  // launch every chunk eagerly even where graphs are on
  // A breakdown (rho, r_hat.v, t.t or omega) is retried once from the
  // solve's starting iterate with the other shadow vector (hashed <-> the
  // initial residual) before it is reported. The analyses (Newton, transient)
  // set it; the plain solve reports breakdowns as the reference does.
  bool retry_breakdown = false;
  // Iterations an earlier solve of a like system took (0: unknown). The
  // first chunk then runs to about 4 short of it without a poll, and the
  // chunks near it are short, so that few iterations run past the stop.
  // Under FP32_IR it is the expected length of the first inner FP32 solve.
  int64_t expect_iters = 0;
};
struct BicgOut {
  bool fp64_fallback = false;  // FP32_IR finished in FP64
  int status = ST_RUNNING;
  int64_t iters = 0;
  int64_t polls = 0;
  double r0 = 0.0;   // ||b - A x0|| (the Newton residual when x0 = x_k)
  double res = 0.0;  // residual norm at exit (recurrence; FP64 true residual under IR)
  double eta = 0.0;  // forcing term the solve used (forcing on)
  int64_t first_inner = 0;  // FP32_IR: iterations of the first inner FP32 solve
};

// FP64 solve on device vectors (rhs, x of length a.n).
ohm_status bicg_solve_f64(ohm_backend *b, const DevCsr &a, const PrecondDev &p, const double *rhs,
                          double *x, const BicgOpts &o, BicgOut &out);
// FP64 solve of A^T x = rhs: explicit transpose of A (gather through the
// transpose map) and the transposed preconditioner (Jacobi: the same
// inverse diagonal; ISAI: (M_U M_L)^T = M_L^T M_U^T from transposes of the
// two factors).
ohm_status bicg_solve_transpose(ohm_backend *b, DevCsr &a, PrecondDev &p, const double *rhs,
                                double *x, const BicgOpts &o, BicgOut &out);
// FP32 inner BiCGSTAB inside FP64 iterative refinement.
ohm_status bicg_solve_ir(ohm_backend *b, DevCsr &a, PrecondDev &p, const double *rhs, double *x,
                         const BicgOpts &o, double inner_tol, int64_t max_ref, BicgOut &out,
                         int64_t &refinements);
// ||b - A x|| / ||b|| in FP64 (synchronous)
ohm_status true_residual(ohm_backend *b, const DevCsr &a, const double *rhs, const double *x,
                         double &rel);
// FP64 solve whose vectors, state word and partial sums live in the
// backend's scratch array (no device allocation; eager launches, no graph);
// rel = ||b - A x|| / ||b|| of the result. Systems whose working set exceeds
// the array are refused with OHM_ERR_INVALID.
ohm_status bicg_solve_f64_noalloc(ohm_backend *b, const DevCsr &a, const PrecondDev &p,
                                  const double *rhs, double *x, BicgOpts o, BicgOut &out,
                                  double &rel);
// Looks up this device's address of the scratch array (once per device,
// with the backend's device current) and keeps it in b->scratch.
ohm_status static_scratch_init(ohm_backend *b);
// FP64 solves for graph capture (timestep and Newton segments): a
// workspace of their own, x aliased to the caller's vector x.
struct CapturedSolve;
ohm_status captured_create(ohm_backend *b, const DevCsr &a, const PrecondDev &p, double *x,
                           CapturedSolve **out);
void captured_free(CapturedSolve *c);
const SolveState *captured_state(const CapturedSolve *c);
// Enqueues (no poll, no host decision) a solve of a x = rhs warm-started
// from x: state reset, A x0, k_init, `iters` iterations and a closing check.
// While *stop is non-zero (stop may be null) the whole solve is a no-op and
// its state reads ST_SKIPPED. A solve still RUNNING after `iters`
// iterations is finished by captured_resume.
ohm_status captured_enqueue(ohm_backend *b, CapturedSolve *c, const DevCsr &a,
                            const PrecondDev &p, const double *rhs, const BicgOpts &o, int iters,
                            const int *stop);
// Drives an enqueued solve that was still RUNNING after its captured
// iterations to its end (polls between chunks, as a plain solve does), also
// after later masked-off solves on the same state; `enqueued` iterations
// are already done.
ohm_status captured_resume(ohm_backend *b, CapturedSolve *c, const DevCsr &a,
                           const PrecondDev &p, const BicgOpts &o, int64_t enqueued,
                           BicgOut &out);
ohm_status status_to_error(int status, int64_t iters, int64_t max_iters);

}  // namespace ohm

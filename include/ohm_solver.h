/*
 * ohm_solver.h — C ABI of the MI355X (gfx950) solve backend for ohmnivore.
 *
 * The entry points replace the reference's solver traits one for one:
 *   LinearSolver      src/solver/mod.rs:28-34
 *   SolverBackend     src/solver/backend.rs:51-96
 *   NonlinearBackend  src/solver/nonlinear.rs:31-150
 *   newton_solve      src/solver/newton.rs:65-78
 *   compile           src/compiler.rs:194
 * plus the mixed-precision, batched and adjoint forms, which have no
 * reference counterpart.
 *
 * Conventions: every call returns an ohm_status (mirroring OhmnivoreError,
 * src/error.rs:4-36); on failure ohm_last_error() returns the message
 * formatted like the reference's Display impl (e.g. "Solve error: BiCGSTAB
 * breakdown: rho ~ 0"). Handles are opaque and owned by the caller until
 * their *_free/_destroy call. Every call is complete on return unless its
 * comment says otherwise (the reference's blocking semantics,
 * src/solver/backend.rs:322-326). A backend is single-threaded, like the
 * reference's (Cell counters, src/solver/backend.rs:205-209).
 */
#ifndef OHM_SOLVER_H
#define OHM_SOLVER_H

#include "ohm_circuit.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ohm_backend ohm_backend; /* one HIP device + stream */
typedef struct ohm_buffer ohm_buffer;   /* device f64 vector        (WgpuBuffer, backend.rs:30-34) */
typedef struct ohm_matrix ohm_matrix;   /* device CSR f64 (+f32)    (GpuCsrMatrix, backend.rs:20-27) */
typedef struct ohm_precond ohm_precond; /* Jacobi or ISAI(level)    */

const char *ohm_last_error(void);
const char *ohm_version(void);

/* ---- MNA compiler: compiler::compile (src/compiler.rs:194-754). The CSR
 * pattern, ordering and values match the reference bit for bit. */
ohm_status ohm_compile(const ohm_circuit *circuit, ohm_mna *out);
void ohm_mna_free(ohm_mna *m);

/* ---- Backend lifecycle: GpuSolver::new / WgpuDsBackend::new
 * (src/solver/gpu.rs:44-49). Fails with OHM_ERR_NO_DEVICE ("no GPU adapter")
 * when no gfx950 device is visible. */
ohm_status ohm_backend_create(int device, ohm_backend **out);
void ohm_backend_destroy(ohm_backend *b);
ohm_status ohm_backend_sync(ohm_backend *b);
/* Replay solver iterations from captured hipGraphs (default on) or launch
 * every kernel from the host. Results are bitwise identical either way. */
ohm_status ohm_backend_set_graphs(ohm_backend *b, int enable);
/* Dispatch / readback counters (src/solver/ds_backend.rs:182-190). */
ohm_status ohm_backend_counters(const ohm_backend *b, int64_t *dispatches, int64_t *readbacks);

/* ---- SolverBackend (src/solver/backend.rs:51-96). Vectors are f64: the
 * reference's DS (hi,lo) pairs become native FP64 on CDNA4. */
ohm_status ohm_buffer_new(ohm_backend *b, size_t n, ohm_buffer **out);       /* new_buffer  :81 */
void ohm_buffer_free(ohm_buffer *buf);
size_t ohm_buffer_len(const ohm_buffer *buf);
ohm_status ohm_upload_vec(ohm_backend *b, const double *data, size_t n, ohm_buffer *buf); /* :93 */
ohm_status ohm_download_vec(ohm_backend *b, const ohm_buffer *buf, double *out, size_t n); /* :96 */
ohm_status ohm_upload_matrix(ohm_backend *b, const double *values, const uint32_t *col_indices,
                             const uint32_t *row_pointers, size_t n, size_t nnz,
                             ohm_matrix **out);                              /* :84-90 */
void ohm_matrix_free(ohm_matrix *m);
/* Replace the values of an uploaded matrix (same pattern). */
ohm_status ohm_matrix_set_values(ohm_backend *b, ohm_matrix *m, const double *values);
ohm_status ohm_spmv(ohm_backend *b, const ohm_matrix *a, const ohm_buffer *x, ohm_buffer *y); /* :56 */
ohm_status ohm_spmv_transpose(ohm_backend *b, const ohm_matrix *a, const ohm_buffer *x,
                              ohm_buffer *y); /* y = A^T x (adjoint; no reference counterpart) */
/* Measurement helper (no reference counterpart): `reps` back-to-back SpMV
 * launches on the backend stream bracketed by HIP events; average ms per
 * launch in *avg_ms. */
ohm_status ohm_spmv_timed(ohm_backend *b, const ohm_matrix *a, const ohm_buffer *x, ohm_buffer *y,
                          int reps, int precision, double *avg_ms);
/* HBM stream probe (no reference counterpart): `reps` launches streaming a
 * `bytes`-sized device buffer (mode 0: read only, 32-byte loads, four in
 * flight per thread; mode 1: copy into a second buffer of the same size),
 * average ms per launch. With a buffer beyond the 256 MB Infinity Cache this
 * is the practical HBM ceiling the SpMV bandwidth is read against. */
ohm_status ohm_stream_timed(ohm_backend *b, int64_t bytes, int mode, int reps, double *avg_ms);
/* SpMV autotuner (no reference counterpart): times the matrix's SpMV under
 * each candidate row-block shape (rows per workgroup 32-256 x staged
 * entries per workgroup 256-2048, i.e. 1-8 entries per thread) and keeps
 * the fastest; the chosen shape is returned. Every matrix starts with a
 * shape fitted to its rows (256-row blocks, entries per thread covering
 * nine in ten of them). The shape changes no result except for rows longer
 * than a block's staged entries, which are summed in four interleaved
 * chains. The preconditioner matrices the library builds itself (ISAI) are
 * tuned automatically unless ohm_backend_set_autotune(b, 0); system matrices
 * are not, because their block table also sets the grid of the fused SpMV +
 * reduction kernels, whose cost grows with the number of workgroups. */
ohm_status ohm_matrix_autotune(ohm_backend *b, ohm_matrix *m, int32_t *rows_per_block,
                               int32_t *nnz_per_block);
ohm_status ohm_backend_set_autotune(ohm_backend *b, int enable);
ohm_status ohm_dot(ohm_backend *b, const ohm_buffer *x, const ohm_buffer *y, double *out); /* :59 */
ohm_status ohm_dot_n(ohm_backend *b, const ohm_buffer *x, const ohm_buffer *y, size_t n,
                     double *out);                                                   /* :66 */
ohm_status ohm_axpy(ohm_backend *b, double alpha, const ohm_buffer *x, ohm_buffer *y); /* :72 */
ohm_status ohm_copy(ohm_backend *b, const ohm_buffer *src, ohm_buffer *dst);           /* :75 */
ohm_status ohm_scale(ohm_backend *b, double alpha, ohm_buffer *x);                     /* :78 */

/* ---- Preconditioners. Jacobi: inv_diag exactly as src/solver/gpu.rs:178-189,
 * except that a zero-diagonal row r takes 1/S_rr, S_rr = -sum_j a_rj a_jr / a_jj
 * (the diagonal of the Schur complement), or 1 if that is zero too.
 * ISAI(level): the reference's approximate inverses M_L ~ L^-1, M_U ~ U^-1 on
 * pattern(L^(level+1)), pattern(U^(level+1)) (src/solver/preconditioner.rs:239-277),
 * with L, U from a GPU ILU(0) of pattern(A) + diagonal. AUTO picks ISAI(1)
 * when any diagonal is zero, else Jacobi (src/solver/gpu.rs:128-139). */
enum { OHM_PREC_NONE = 0, OHM_PREC_JACOBI = 1, OHM_PREC_ISAI = 2, OHM_PREC_AUTO = 3 };
ohm_status ohm_precond_create(ohm_backend *b, const ohm_matrix *a, int kind, int level,
                              ohm_precond **out);
/* Recompute the numeric part after ohm_matrix_set_values (pattern reused). */
ohm_status ohm_precond_update(ohm_backend *b, ohm_precond *p, const ohm_matrix *a);
ohm_status ohm_precond_apply(ohm_backend *b, const ohm_precond *p, const ohm_buffer *in,
                             ohm_buffer *out);
int ohm_precond_kind(const ohm_precond *p);
/* ISAI: fixed-point ILU(0) sweeps per numeric refresh (default 5); takes
 * effect at the next ohm_precond_update, which restarts from A. */
ohm_status ohm_precond_set_ilu_sweeps(ohm_precond *p, int sweeps);
void ohm_precond_free(ohm_precond *p);

/* ---- BiCGSTAB: bicgstab_with_params (src/solver/bicgstab.rs:56-189). */
enum { OHM_SHADOW_RHS = 0, OHM_SHADOW_HASHED = 1 };
enum { OHM_PRECISION_FP64 = 0, OHM_PRECISION_FP32_IR = 1 };
typedef struct {
  int64_t max_iterations;  /* default 10000 (bicgstab.rs:10) */
  double tolerance;        /* default 1e-5  (bicgstab.rs:11); relative to ||b|| */
  int32_t shadow;          /* OHM_SHADOW_RHS = r_hat = r0 (reference); HASHED (default) */
  int32_t warm_start;      /* 1: x holds x0 on entry; 0: x0 = 0 (reference) */
  int32_t precision;       /* FP64, or FP32 inner BiCGSTAB inside FP64 iterative refinement */
  int32_t check_interval;  /* iterations between host status polls (0 = adaptive) */
  double inner_tolerance;  /* FP32_IR: inner relative tolerance (default 1e-4) */
  int64_t max_refinements; /* FP32_IR: outer steps (default 30) */
  int32_t breakdown;       /* OHM_BREAKDOWN_SCALED (default) or OHM_BREAKDOWN_ABSOLUTE */
  int32_t tol_mode;        /* OHM_TOL_RHS: stop at tolerance*||b|| (reference);
                              OHM_TOL_INITIAL: stop at max(tolerance*||r0||, tol_floor*||b||) */
  double tol_floor;        /* OHM_TOL_INITIAL floor, relative to ||b|| (default 1e-15) */
} ohm_bicgstab_params;
enum { OHM_TOL_RHS = 0, OHM_TOL_INITIAL = 1 };
/* Breakdown tests. ABSOLUTE is the reference's |rho|, |r_hat.v|, t.t, |omega|
 * < 1e-30 (src/solver/bicgstab.rs:107,125,151,178), which fires on healthy
 * iterations once MNA residuals (amperes at ~mS scale) fall below ~1e-15.
 * SCALED applies the same 1e-30 to the normalised quantities
 * rho/(|r_hat||r|), r_hat.v/(|r_hat||v|), t.t/s.s and omega. */
enum { OHM_BREAKDOWN_SCALED = 0, OHM_BREAKDOWN_ABSOLUTE = 1 };

typedef struct {
  int64_t iterations;      /* BiCGSTAB iterations (summed over refinements) */
  int64_t refinements;     /* FP32_IR outer steps */
  double residual_norm;    /* FP64 ||b - A x|| / ||b|| at exit */
  double device_ms;        /* device time of the solve (HIP events) */
  int64_t dispatches;      /* kernels enqueued */
  int64_t readbacks;       /* host status polls */
  int64_t fp64_fallback;   /* FP32_IR: 1 when refinement stalled and the solve
                              finished in FP64 (kappa * 2^-24 >~ 1) */
} ohm_bicgstab_result;

void ohm_bicgstab_default_params(ohm_bicgstab_params *p);
ohm_status ohm_bicgstab(ohm_backend *b, const ohm_matrix *a, const ohm_buffer *rhs, ohm_buffer *x,
                        const ohm_precond *p, const ohm_bicgstab_params *params,
                        ohm_bicgstab_result *result);
/* Solve A^T x = rhs (adjoint systems; transposed preconditioner). */
ohm_status ohm_bicgstab_transpose(ohm_backend *b, const ohm_matrix *a, const ohm_buffer *rhs,
                                  ohm_buffer *x, const ohm_precond *p,
                                  const ohm_bicgstab_params *params, ohm_bicgstab_result *result);
/* ohm_bicgstab without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). The caller passes the matrix, preconditioner, rhs and x as for
 * ohm_bicgstab (same results, bit for bit); the solve's own vectors, state
 * word and partial sums live in a 64 MiB static device array that each GPU
 * the library runs on holds until the process ends, so systems up to about
 * 700,000 unknowns fit (a larger one is refused with OHM_ERR_INVALID). FP64
 * only (precision FP32_IR is refused), every chunk launched eagerly. All
 * calls on one GPU share that GPU's array: calls that overlap on one GPU
 * (several host threads, or backends on different streams of the same
 * device) race and must be serialised by the caller. */
ohm_status ohm_bicgstab_noalloc(ohm_backend *b, const ohm_matrix *a, const ohm_buffer *rhs,
                                ohm_buffer *x, const ohm_precond *p,
                                const ohm_bicgstab_params *params, ohm_bicgstab_result *result);

/* ---- LinearSolver::solve_real (src/solver/mod.rs:30): host slices in and
 * out, preconditioner chosen by `precond` (OHM_PREC_AUTO = the reference's
 * rule). params may be NULL for defaults. */
ohm_status ohm_solve_real(ohm_backend *b, size_t n, size_t nnz, const double *values,
                          const uint32_t *col_indices, const uint32_t *row_pointers,
                          const double *rhs, double *x, int precond,
                          const ohm_bicgstab_params *params, ohm_bicgstab_result *result);

/* ---- LinearSolver::solve_complex (src/solver/mod.rs:33; GPU version
 * src/solver/gpu.rs:210-735): complex Jacobi-BiCGSTAB with conjugated dot
 * products. `values`, `rhs` and `x` hold interleaved (re, im) pairs
 * (Complex64 layout). precision OHM_PRECISION_FP32_IR runs complex64 inner
 * solves inside complex128 refinement. Unlike the reference, which returns
 * its iterate as Ok after 10 000 iterations without converging
 * (gpu.rs:729-734), a solve that does not reach the tolerance returns an
 * error (x still holds the last iterate). */
ohm_status ohm_solve_complex(ohm_backend *b, size_t n, size_t nnz, const double *values,
                             const uint32_t *col_indices, const uint32_t *row_pointers,
                             const double *rhs, double *x, int precond,
                             const ohm_bicgstab_params *params, ohm_bicgstab_result *result);

/* ---- Batched systems (no counterpart in the reference, which solves one
 * system per call): K independent systems (G_k + sigma_k C_k) x_k = b_k,
 * real or complex, solved together. Vectors of all systems are concatenated
 * in system order (complex as (re, im) pairs). Systems may share patterns
 * and value arrays: identical host pointers are uploaded once. Each system
 * converges, breaks down or stops on its own (per-system status). */
typedef struct ohm_batch ohm_batch;
typedef struct {
  int64_t n;                     /* rows (= columns) */
  const uint32_t *rowptr, *col;  /* CSR pattern: n+1 / rowptr[n] entries, columns < n */
  const double *g_val;           /* values on the pattern */
  const double *c_val;           /* second value array on the same pattern, or NULL */
  double sigma_re, sigma_im;     /* A = G + sigma C (complex batches use both parts) */
} ohm_batch_system;
typedef struct {
  int64_t iterations;            /* BiCGSTAB iterations (inner ones under FP32_IR) */
  int64_t refinements;           /* FP32_IR refinement steps */
  double residual_norm;          /* ||b - A x|| / ||b|| in FP64 */
  int32_t status;                /* OHM_OK or the error this system's solve ended with */
  int32_t reason;                /* 1 converged, 2-6 breakdown kinds / NaN, 7 max iterations */
} ohm_batch_result;
ohm_status ohm_batch_create(ohm_backend *b, int32_t is_complex, int64_t k,
                            const ohm_batch_system *systems, ohm_batch **out);
void ohm_batch_free(ohm_batch *bt);
int64_t ohm_batch_rows(const ohm_batch *bt); /* total rows over all systems */
/* Replace every sigma_k (2k doubles: re, im) without re-uploading values. */
ohm_status ohm_batch_set_sigma(ohm_batch *bt, const double *sigma);
/* y = A x for every system (host vectors, ohm_batch_rows entries). */
ohm_status ohm_batch_spmv(ohm_batch *bt, const double *x, double *y);
/* Solve all systems. x is the warm start when params->warm_start, and
 * receives every system's final iterate. results: k entries. Returns OHM_OK
 * when every system converged, else the first failing system's error. */
ohm_status ohm_batch_solve(ohm_batch *bt, const double *rhs, double *x,
                           const ohm_bicgstab_params *params, int32_t precond,
                           ohm_batch_result *results);

/* ---- AC sweep (src/analysis/ac.rs:25-77): for every frequency f_i solve
 * (G + j 2 pi f_i C) x_i = b_ac, with A formed on the fly from G and C
 * (form_complex_matrix, src/sparse.rs:149-178) and the frequency points
 * batched `batch` at a time (0 = all). C's pattern may differ from G's (the
 * union is used). b_ac: 2n doubles; x: n_freqs * 2n; results: n_freqs. */
ohm_status ohm_ac_sweep(ohm_backend *b, int64_t n, const uint32_t *g_rowptr,
                        const uint32_t *g_col, const double *g_val, const uint32_t *c_rowptr,
                        const uint32_t *c_col, const double *c_val, const double *b_ac,
                        int64_t n_freqs, const double *freqs, int64_t batch,
                        const ohm_bicgstab_params *params, int32_t precond, double *x,
                        ohm_batch_result *results);

/* ---- AC sweep over many circuits (the reference runs ac::run once per
 * circuit). method OHM_AC_DIRECT: every (circuit, frequency) system by a
 * banded LU with partial pivoting after a reverse Cuthill-McKee ordering
 * (circuits whose half bandwidth is at most 4: ladders, chains), one system
 * per GPU lane; OHM_AC_ITERATIVE: ohm_ac_sweep per circuit; OHM_AC_AUTO:
 * direct where the band allows, iterative elsewhere. Each circuit's x
 * (n_freqs * 2n doubles, may be NULL) receives its solutions; results has
 * n_circuits * n_freqs entries (circuit-major) with the FP64 residual
 * ||b - A x|| / ||b|| of every system. device_ms (may be NULL): kernel time
 * of the direct part. OHM_AC_DIRECT_BUDGET_GB bounds its working memory
 * (default 64): that workspace stays allocated on the backend for later
 * sweeps (re-allocating tens of GB per sweep costs seconds) and is released
 * by ohm_backend_destroy. */
enum { OHM_AC_AUTO = 0, OHM_AC_ITERATIVE = 1, OHM_AC_DIRECT = 2 };
typedef struct {
  int64_t n;
  const uint32_t *g_rowptr, *g_col;
  const double *g_val;
  const uint32_t *c_rowptr, *c_col; /* may differ from G's pattern; NULL c_val = no C */
  const double *c_val;
  const double *b_ac;               /* 2n doubles */
  double *x;                        /* out: n_freqs * 2n doubles, or NULL */
} ohm_ac_circuit;
ohm_status ohm_ac_sweep_circuits(ohm_backend *b, int64_t n_circuits,
                                 const ohm_ac_circuit *circuits, int64_t n_freqs,
                                 const double *freqs, int32_t method,
                                 const ohm_bicgstab_params *params, int32_t precond,
                                 ohm_batch_result *results, double *device_ms);

/* ---- The direct AC sweep split into its symbolic and numeric halves (the
 * reference re-forms and re-solves every (circuit, frequency) system from
 * scratch, src/analysis/ac.rs:56-70; a sparse direct solver's analyse /
 * factor split, as faer's sp_lu has, src/solver/sparse_direct.rs:39-47).
 * ohm_ac_plan_create: per circuit the union pattern of G and C, a reverse
 * Cuthill-McKee ordering and the band rows of G and C in that order, built
 * on host threads and uploaded once (every circuit's half bandwidth after
 * RCM must be at most 4; the circuits' x fields are ignored).
 * ohm_ac_plan_solve: every (circuit, frequency) system by banded LU with
 * partial pivoting on the device (at half bandwidth 1 each system's rows
 * are cut into parts eliminated in parallel, joined through a reduced
 * system of the parts' boundary unknowns); x (may be NULL, else n_circuits pointers,
 * each NULL or n_freqs * 2n doubles) receives solutions; results
 * (n_circuits * n_freqs, circuit-major) the FP64 residual of every system.
 * When the sweep fits the device in one chunk its solutions stay resident:
 * ohm_ac_plan_get_x copies frequencies [f_first, f_first + f_count) of one
 * circuit (f_count * 2n doubles) until the next sweep on the backend. */
typedef struct ohm_ac_plan ohm_ac_plan;
ohm_status ohm_ac_plan_create(ohm_backend *b, int64_t n_circuits, const ohm_ac_circuit *circuits,
                              ohm_ac_plan **out);
ohm_status ohm_ac_plan_solve(ohm_ac_plan *plan, int64_t n_freqs, const double *freqs,
                             double *const *x, ohm_batch_result *results, double *device_ms);
ohm_status ohm_ac_plan_get_x(ohm_ac_plan *plan, int64_t circuit, int64_t f_first,
                             int64_t f_count, double *x);
/* Figures of the last solve (no counterpart in the reference; for
 * benchmarks and tests): phase_ms (6 doubles, HIP events on the backend's
 * stream) = part relations, reduced boundary systems, part solves (the
 * whole one-lane LU when unpartitioned), residuals, one-lane redo of
 * systems whose partitioned residual exceeded 1e-12, unpermute; the rows
 * per part (0: one lane per whole system; picked by the autotuner on the
 * plan's first solve unless OHM_AC_PART_ROWS is set), the systems redone,
 * and the chunks the sweep needed. Any pointer may be NULL. */
ohm_status ohm_ac_plan_stats(const ohm_ac_plan *plan, double *phase_ms, int64_t *rows_per_part,
                             int64_t *fallback_systems, int32_t *chunks);
void ohm_ac_plan_free(ohm_ac_plan *plan);

/* generate_frequencies (src/analysis/ac.rs:80-109). Writes up to `cap`
 * points and returns how many the sweep has. */
enum { OHM_SWEEP_DEC = 0, OHM_SWEEP_OCT = 1, OHM_SWEEP_LIN = 2 };
int64_t ohm_generate_frequencies(int32_t sweep, int64_t n_points, double f_start, double f_stop,
                                 double *out, int64_t cap);

/* ---- Transient (src/analysis/transient.rs:25-199). One or more instances
 * of a linear circuit (Monte Carlo: same pattern and sources, own G and C
 * values) stepped together on the GPU: A = G + (alpha/h) C formed on the
 * fly, companion right-hand side and source values evaluated by one kernel
 * per step, warm-started solve, probe values gathered on the device.
 * Instances are ohm_batch_system entries whose c_val holds C on G's
 * pattern (the compiler zero-fills C positions into G, src/compiler.rs:577-579);
 * sigma is ignored. */
enum { OHM_SRC_DC = 0, OHM_SRC_PULSE = 1, OHM_SRC_SIN = 2, OHM_SRC_EXP = 3, OHM_SRC_PWL = 4 };
/* SourceEntry (transient.rs:200-262): b[idx[j]] += sign[j] * value(t).
 * param: PULSE v1 v2 td tr tf pw per | SIN vo va freq td theta |
 * EXP v1 v2 td1 tau1 td2 tau2 (transient_source.rs:39-116); PWL uses
 * pwl_count (t, v) pairs from the shared pair table at pwl_offset. */
typedef struct {
  int32_t kind;
  int32_t n_idx;
  uint32_t idx[2];
  double sign[2];
  double dc;
  double param[7];
  int64_t pwl_offset, pwl_count;
} ohm_tran_source;
enum { OHM_TRAN_FIXED_BE = 0, OHM_TRAN_ADAPTIVE = 1 };
typedef struct {
  double tstep, tstop, tstart;
  int32_t method;    /* FIXED_BE: backward Euler at h = tstep; ADAPTIVE: the
                        reference's TRAP/BE stepping with LTE control (one instance) */
  int32_t precond;   /* OHM_PREC_JACOBI or OHM_PREC_NONE */
  ohm_bicgstab_params solver;  /* per-step solves (always warm-started) */
} ohm_tran_params;
typedef struct {
  int64_t steps_accepted, steps_rejected;
  int64_t linear_solves, bicgstab_iterations;
  int64_t records;
  double t_end;
  double device_ms;
} ohm_tran_stats;
typedef struct ohm_transient ohm_transient;
void ohm_tran_default_params(ohm_tran_params *p);
ohm_status ohm_transient_create(ohm_backend *b, int64_t k, const ohm_batch_system *instances,
                                int64_t n_sources, const ohm_tran_source *sources,
                                int64_t n_pwl, const double *pwl_pairs, ohm_transient **out);
void ohm_transient_free(ohm_transient *tr);
/* x0: k*n initial state (NULL = zero, i.e. UIC with zero initial
 * conditions). Records t = 0 (when tstart <= 0) and every accepted step
 * with t >= tstart, up to max_records: times[r], waves[r][k][probe].
 * x_final: k*n. */
ohm_status ohm_transient_run(ohm_transient *tr, const double *x0, const ohm_tran_params *params,
                             int64_t n_probes, const uint32_t *probes, int64_t max_records,
                             double *times, double *waves, double *x_final,
                             ohm_tran_stats *stats);

/* ---- Newton-Raphson: NewtonParams (src/solver/newton.rs:34-55) and
 * newton_solve (:65-513). Device evaluation, stamping, linear solve,
 * limiting and the convergence test all run on the GPU; the host polls one
 * status word per Newton iteration. */
typedef struct {
  int64_t max_iterations;          /* default 50 */
  double abs_tol;                  /* default 1e-4 (max |x - x_old|) */
  int64_t bicgstab_max_iterations; /* default 10000 */
  double bicgstab_tolerance;       /* default 1e-5, relative to the Newton residual ||F(x_k)|| */
  int32_t precision;               /* OHM_PRECISION_* for the linear sub-solve */
  int32_t precond;                 /* OHM_PREC_AUTO (reference rule) / JACOBI / ISAI */
  const double *initial_guess;     /* NewtonParams::initial_guess, or NULL */
  double bicgstab_floor;           /* sub-solve floor relative to ||b(x_k)|| (default 1e-15) */
  int32_t forcing;                 /* OHM_FORCING_FIXED (default): every sub-solve to
                                      bicgstab_tolerance; OHM_FORCING_EW: Eisenstat-Walker
                                      choice 2, eta_k = 0.9 (||F_k|| / ||F_k-1||)^2 with the
                                      0.9 eta_k-1^2 safeguard; OHM_FORCING_EW1: choice 1,
                                      eta_k = | ||F_k|| - ||F_k-1 + J_k-1 s_k-1|| | / ||F_k-1||
                                      with the eta_k-1^((1+sqrt 5)/2) safeguard. Both clamped
                                      to [bicgstab_tolerance, forcing_max] and evaluated on
                                      the device from ||F_k|| as each sub-solve starts. */
  int32_t forcing_safeguard;       /* 0 (default): the safeguard applies when it exceeds 0.1
                                      (as published); 1: at every iteration, so eta falls at
                                      most as fast as the safeguard allows */
  double forcing_max;              /* OHM_FORCING_EW / EW1: largest eta (default 0.1) */
  int32_t linear_solver;           /* OHM_LINEAR_BICGSTAB (default): the preconditioned
                                      BiCGSTAB sub-solve above; OHM_LINEAR_DENSE_LU: every
                                      Newton step by dense LU with partial pivoting on the GPU
                                      (systems of at most 4096 unknowns; the reference's
                                      NewtonLinearMode::CpuSparseLu, src/solver/newton.rs:22-30) */
} ohm_newton_params;
enum { OHM_FORCING_FIXED = 0, OHM_FORCING_EW = 1, OHM_FORCING_EW1 = 2 };
enum { OHM_LINEAR_BICGSTAB = 0, OHM_LINEAR_DENSE_LU = 1 };
/* The sub-solve is warm-started from x_k, so its initial residual is the
 * Newton residual F(x_k) = b(x_k) - A(x_k) x_k and the stopping rule is
 * max(bicgstab_tolerance*||F||, bicgstab_floor*||b||) (an inexact Newton
 * step). The reference solves from x = 0 to 1e-5*||b|| (src/solver/newton.rs:407-421),
 * which caps the accuracy of the converged point at about kappa*1e-5. */

typedef struct {
  int64_t iterations;         /* Newton iterations (Stats::newton_iterations) */
  int64_t bicgstab_iterations;/* summed over Newton steps */
  double max_diff;            /* last max |x - x_old| */
  double device_ms;
  int64_t fp64_fallbacks;     /* FP32_IR: Newton steps redone in FP64 after the
                                 refinement stalled (kappa * 2^-24 >~ 1) */
} ohm_newton_result;

void ohm_newton_default_params(ohm_newton_params *p);
/* bicgstab_iters_per_newton (ohm_newton_solve, ohm_newton_run): NULL, or an
 * array of at least params->max_iterations int32 entries; entry k receives
 * the BiCGSTAB iterations of Newton iteration k (Stats, newton.rs:430-435),
 * and entries past result->iterations are left untouched. */
ohm_status ohm_newton_solve(ohm_backend *b, const double *base_values,
                            const uint32_t *col_indices, const uint32_t *row_pointers,
                            const double *base_b, const ohm_diode_desc *diodes, size_t n_diodes,
                            const ohm_bjt_desc *bjts, size_t n_bjts, const ohm_mosfet_desc *mosfets,
                            size_t n_mosfets, size_t system_size, size_t matrix_nnz,
                            const ohm_newton_params *params, double *x_out,
                            ohm_newton_result *result, int32_t *bicgstab_iters_per_newton);

/* Device-resident Newton context: uploads once, solves many times (the
 * timed form: inputs already in HBM). */
typedef struct ohm_newton_ctx ohm_newton_ctx;
ohm_status ohm_newton_create(ohm_backend *b, const double *base_values,
                             const uint32_t *col_indices, const uint32_t *row_pointers,
                             const double *base_b, const ohm_diode_desc *diodes, size_t n_diodes,
                             const ohm_bjt_desc *bjts, size_t n_bjts,
                             const ohm_mosfet_desc *mosfets, size_t n_mosfets,
                             size_t system_size, size_t matrix_nnz, ohm_newton_ctx **out);
ohm_status ohm_newton_run(ohm_newton_ctx *ctx, const ohm_newton_params *params,
                          ohm_newton_result *result, int32_t *bicgstab_iters_per_newton);
/* New base values (matrix_nnz), base RHS (system_size) and device
 * parameters for the context's circuit (no counterpart in the reference,
 * which recompiles per analysis; for loops that re-solve one topology with
 * new parameter values). Any argument may be NULL to keep the current one;
 * descriptor arrays have the counts given at create and must keep every node
 * index and stamp slot (only the device parameters may differ), else
 * OHM_ERR_INVALID. The symbolic preconditioner plan, tuned SpMV shapes and
 * captured graphs of the context stay valid. */
ohm_status ohm_newton_update(ohm_newton_ctx *ctx, const double *base_values,
                             const double *base_b, const ohm_diode_desc *diodes,
                             const ohm_bjt_desc *bjts, const ohm_mosfet_desc *mosfets);
ohm_status ohm_newton_get_x(ohm_newton_ctx *ctx, double *x_out);
/* The assembled A(x), b(x) of the last iteration (for checks and adjoints). */
ohm_status ohm_newton_get_system(ohm_newton_ctx *ctx, double *values_out, double *b_out);
void ohm_newton_free(ohm_newton_ctx *ctx);

/* ---- dc::run (src/analysis/dc.rs:25-52): the DC operating point of a
 * compiled system. Linear systems: one ohm_solve_real (AUTO preconditioner,
 * relative tolerance 1e-12). Nonlinear systems: the continuation cascade of
 * run_nonlinear (:60-293) on one device-resident Newton context whose
 * values and right-hand side are replaced per step (ohm_newton_update):
 *   direct Newton (params->max_iterations, 200 BiCGSTAB iterations);
 *   source stepping when MOSFETs are present (:480-640; sources scaled by
 *     alpha from 0.1, the step doubled on success and halved on failure);
 *   adaptive GMIN stepping (:152-260; extra node-diagonal conductance from
 *     0.1 down by 1000x, subdivided by 10x on failure) and a final true-GMIN
 *     solve; pseudo-transient continuation (:340-470; alpha I on node rows
 *     anchored at b + alpha x_prev) when that final solve fails.
 * Where the reference retries a failed step on the CPU with a sparse LU,
 * the step is retried on the GPU with OHM_LINEAR_DENSE_LU (systems of at
 * most 4096 unknowns) or else in FP64 with a BiCGSTAB budget of
 * max(10 x budget, 20000) (and 15 Newton iterations where the reference's
 * retry uses 15). params: the caller's NewtonParams (abs_tol, sub-solve
 * tolerance and forcing, precision, preconditioner apply to every step; the
 * cascade sets the iteration caps and initial guesses as the reference does);
 * NULL = defaults. x_out: mna->size doubles. stats and trace may be NULL;
 * trace receives the first trace_cap Newton runs, stats->runs counts all. */
enum {
  OHM_DC_LINEAR = 0,
  OHM_DC_DIRECT = 1,
  OHM_DC_SOURCE_STEPPING = 2,
  OHM_DC_GMIN_STEPPING = 3,
  OHM_DC_GMIN_FINAL = 4,
  OHM_DC_PSEUDO_TRANSIENT = 5,
  OHM_DC_PSEUDO_TRANSIENT_FINAL = 6
};
typedef struct {
  int32_t stage;       /* OHM_DC_* */
  int32_t converged;   /* 1 if this run converged */
  double level;        /* source scale, extra GMIN or alpha of the step (0 otherwise) */
  int64_t iterations;  /* Newton iterations of the run */
  int32_t retry;       /* 1 for the FP64 larger-budget retry of a failed run */
  int32_t _pad;
} ohm_dc_step;
typedef struct {
  int32_t stage;               /* OHM_DC_* of the run that produced x */
  int32_t _pad;
  int64_t runs;                /* Newton runs (continuation steps and retries) */
  int64_t newton_iterations;   /* summed over every run (Stats::newton_iterations) */
  int64_t bicgstab_iterations; /* summed over every run */
  double device_ms;
} ohm_dc_stats;
ohm_status ohm_dc_run(ohm_backend *b, const ohm_mna *mna, const ohm_newton_params *params,
                      double *x_out, ohm_dc_stats *stats, ohm_dc_step *trace,
                      int64_t trace_cap);

/* ---- Nonlinear transient (no counterpart in the reference, whose transient
 * steps G and C only, src/analysis/transient.rs:84-178, leaving diode, BJT
 * and MOSFET currents out of the time loop). Every timestep is a Newton
 * solve on the device (ohm_newton_* on one context):
 *   BE:   (G + C/h) x + i(x) = b(t+h) + (C/h) x_n
 *   TRAP: (G + 2C/h) x + i(x) = b(t+h) + (2C/h) x_n + d_n,
 *         d_n = b(t) - G x_n - i(x_n)  (= C dx/dt at t_n)
 * which for a circuit without devices is the reference's companion form
 * (:285-352). Step control is the reference's: FIXED_BE steps at tstep;
 * ADAPTIVE takes a BE first step, then TRAP with a BE companion solve for
 * the LTE estimate (:84-178, same h_min, h_max and tolerance); a Newton
 * failure rejects the step like an LTE failure (h halved, next step BE).
 * Sources: the ohm_tran_source table of ohm_transient_create, evaluated on
 * the host per step (b(t) starts from zero, evaluate_b_at_time :265-280).
 * mna: the compiled system (G, C with C's pattern inside G's, descriptors).
 * params->method and ->tstep/tstop/tstart apply; the linear-solver fields
 * of params are unused (newton->bicgstab_* govern the sub-solves).
 * x0: n doubles or NULL (zero). Records as ohm_transient_run with one
 * instance: times[r], waves[r][probe]. stats->linear_solves counts Newton
 * iterations. */
ohm_status ohm_transient_nonlinear(ohm_backend *b, const ohm_mna *mna, int64_t n_sources,
                                   const ohm_tran_source *sources, int64_t n_pwl,
                                   const double *pwl_pairs, const ohm_tran_params *params,
                                   const ohm_newton_params *newton, const double *x0,
                                   int64_t n_probes, const uint32_t *probes, int64_t max_records,
                                   double *times, double *waves, double *x_final,
                                   ohm_tran_stats *stats);

/* ---- NonlinearBackend kernels (src/solver/nonlinear.rs:31-150) on a
 * Newton context's device buffers: evaluate + stamp into A, b at x. */
ohm_status ohm_newton_assemble_at(ohm_newton_ctx *ctx, const double *x, double *values_out,
                                  double *b_out);
/* Limiting of x_new against x_old (BJT, MOSFET, diode passes). */
ohm_status ohm_newton_limit(ohm_newton_ctx *ctx, const double *x_old, double *x_new);

/* ---- Adjoint DC sensitivity (no counterpart in the reference). At the
 * operating point left in ctx by ohm_newton_run, with F(x, p) = 0 the MNA
 * residual and J = dF/dx the assembled Newton matrix there:
 *   solve J^T lambda = dJdx (transposed BiCGSTAB, transposed preconditioner)
 *   grad_g[r]  = dObj/dg_r  = -lambda^T dF/dg_r  for resistor conductances
 *                g_r between res_nodes[2r], res_nodes[2r+1] (OHM_SENTINEL = ground)
 *   grad_kp[m] = dObj/dkp_m = -lambda^T dF/dkp_m for every MOSFET's kp (= KP W/L)
 * dJdx: n doubles (dObjective/dx). grad_kp (n_mosfets) and lambda_out (n)
 * may be NULL. params NULL = defaults with tolerance 1e-12. */
ohm_status ohm_adjoint_dc(ohm_newton_ctx *ctx, const double *dJdx,
                          const ohm_bicgstab_params *params, int64_t n_res,
                          const uint32_t *res_nodes, double *grad_g, double *grad_kp,
                          double *lambda_out, ohm_bicgstab_result *result);

#ifdef __cplusplus
}
#endif
#endif /* OHM_SOLVER_H */

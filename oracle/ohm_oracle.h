/*
 * ohm_oracle.h — CPU oracle for the ohmnivore solve path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference's
 * algorithms (each function cites the reference file:line it follows). Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, and only as the reference answer or the timed CPU baseline — never as part of
 * the product path.
 *
 * Parity status of each function is recorded in oracle/README.md.
 */
#ifndef OHM_ORACLE_H
#define OHM_ORACLE_H

#include "../include/ohm_circuit.h"
#include "../include/ohm_solver.h" /* ohm_tran_source layout only */

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failed call (formatted like OhmnivoreError's Display). */
const char *oo_last_error(void);

/* CsrMatrix::from_triplets (src/sparse.rs:52-100). Output arrays must hold
 * ntrip entries (nrows+1 for rowptr). Returns nnz. vals/out_vals hold 1
 * double per entry (real) or 2 (complex, interleaved) when `complex_vals`. */
int64_t oo_from_triplets(int64_t nrows, int64_t ntrip, const int64_t *rows, const int64_t *cols,
                         const double *vals, int complex_vals, int64_t *rowptr, int64_t *out_cols,
                         double *out_vals);

/* compiler::compile (src/compiler.rs:194-754). */
int oo_compile(const ohm_circuit *c, ohm_mna *out);
void oo_mna_free(ohm_mna *m);

/* CSR y = A x (src/sparse.rs:116-143); complex when `cplx`. */
void oo_spmv(int64_t n, const uint32_t *rowptr, const uint32_t *col, const double *val,
             const double *x, double *y, int cplx);

/* CpuSolver::solve_real / solve_complex: dense Gaussian elimination with
 * partial pivoting (src/solver/cpu.rs:28-170). */
int oo_dense_solve(int64_t nrows, int64_t ncols, int64_t nb, const uint32_t *rowptr,
                   const uint32_t *col, const double *val, const double *b, double *x, int cplx);

/* Sparse direct LU: the role of faer 0.24 sp_lu (src/solver/sparse_direct.rs:12-61),
 * restated as reverse Cuthill-McKee ordering + LAPACK-style banded LU with
 * partial pivoting. Real or complex. */
int oo_sparse_lu_solve(int64_t n, const uint32_t *rowptr, const uint32_t *col, const double *val,
                       const double *b, double *x, int cplx);

/* ac::run over one circuit (src/analysis/ac.rs:56-70) with oo_sparse_lu_solve's
 * factorisation as the linear solve; A_f = G + j 2 pi f C on the union
 * pattern (form_complex_matrix, src/sparse.rs:149-178), ordering computed
 * once. b_ac: 2n doubles; x: n_freqs * 2n doubles, frequency-major. c_val may
 * be NULL (no capacitive part). */
int oo_ac_sweep(int64_t n, const uint32_t *g_rp, const uint32_t *g_col, const double *g_val,
                const uint32_t *c_rp, const uint32_t *c_col, const double *c_val,
                const double *b_ac, int64_t n_freqs, const double *freqs, double *x);

/* Preconditioner kinds for the oracle BiCGSTAB. */
enum { OO_PREC_NONE = 0, OO_PREC_JACOBI = 1, OO_PREC_ISAI = 2 };

typedef struct {
  int kind;
  const double *inv_diag;                 /* JACOBI */
  const uint32_t *ml_rowptr, *ml_col;     /* ISAI: z = M_U (M_L r) */
  const double *ml_val;
  const uint32_t *mu_rowptr, *mu_col;
  const double *mu_val;
} oo_precond;

/* bicgstab_with_params (src/solver/bicgstab.rs:56-189), right-preconditioned,
 * x0 = 0. shadow = 0 keeps the reference's r_hat = b (:88-90); shadow = 1 uses
 * the dense hashed vector oo_shadow_value. scaled = 0 keeps the reference's
 * absolute 1e-30 breakdown tests; 1 applies them to normalised quantities.
 * Returns OHM_OK and the iteration
 * count in *iters. */
int oo_bicgstab(int64_t n, const uint32_t *rowptr, const uint32_t *col, const double *val,
                const double *b, double *x, const oo_precond *p, int64_t max_iters, double tol,
                int shadow, int scaled, int64_t *iters);
double oo_shadow_value(int64_t i);

/* Jacobi inverse diagonal exactly as src/solver/gpu.rs:178-189 (0 for |d|<=1e-30). */
void oo_jacobi_inv_diag(int64_t n, const uint32_t *rowptr, const uint32_t *col, const double *val,
                        double *inv);

/* Reference ISAI (src/solver/preconditioner.rs:64-428): full LU with partial
 * pivoting and fill ("ilu0"), then level-`level` ISAI. Small matrices only.
 * Results are malloc'd; free with oo_free. */
int oo_isai_reference(int64_t n, const uint32_t *rowptr, const uint32_t *col, const double *val,
                      int level, int64_t *ml_nnz, uint32_t **ml_rowptr, uint32_t **ml_col,
                      double **ml_val, int64_t *mu_nnz, uint32_t **mu_rowptr, uint32_t **mu_col,
                      double **mu_val, int64_t **perm);
/* Same factorisation, returning L, U and the row permutation (PA = LU). */
int oo_lu_reference(int64_t n, const uint32_t *rowptr, const uint32_t *col, const double *val,
                    int64_t *l_nnz, uint32_t **l_rowptr, uint32_t **l_col, double **l_val,
                    int64_t *u_nnz, uint32_t **u_rowptr, uint32_t **u_col, double **u_val,
                    int64_t **perm);

/* The backend's preconditioner, computed exactly: ILU(0) without pivoting on
 * pattern(A) plus the diagonal, then ISAI(level) columns by triangular
 * substitution on pattern(L^(level+1)), pattern(U^(level+1)). */
int oo_ilu0_isai(int64_t n, const uint32_t *rowptr, const uint32_t *col, const double *val,
                 int level, int64_t *l_nnz, uint32_t **l_rowptr, uint32_t **l_col, double **l_val,
                 int64_t *u_nnz, uint32_t **u_rowptr, uint32_t **u_col, double **u_val,
                 int64_t *ml_nnz, uint32_t **ml_rowptr, uint32_t **ml_col, double **ml_val,
                 int64_t *mu_nnz, uint32_t **mu_rowptr, uint32_t **mu_col, double **mu_val);

void oo_free(void *p);

/* ---- Device models (FP64 restatements of ds_nonlinear_shaders.rs) ---- */
void oo_diode_eval(const ohm_diode_desc *d, const double *x, double *i_out, double *g_out);
/* out[6]: ic, ib, dic_dvbe, dic_dvbc, dib_dvbe, dib_dvbc (currents polarity-signed). */
void oo_bjt_eval(const ohm_bjt_desc *d, const double *x, double *out);
/* out[3]: id (polarity-signed), gm, gds. */
void oo_mosfet_eval(const ohm_mosfet_desc *d, const double *x, double *out);
/* Smooth level-1 core on (vgs, vds) in the polarity-normalised frame. */
void oo_mosfet_core(double vgs, double vds, double vto, double kp, double lambda, double polarity,
                    double *id, double *gm, double *gds);
double oo_pnjlim(double vnew, double vold, double vt, double vcrit);

typedef struct {
  int64_t max_iterations;
  double abs_tol;
  int linear_mode; /* 0 = sparse LU (faer stand-in), 1 = dense GE (CpuSolver) */
  const double *initial_guess;
} oo_newton_params;

/* newton_solve (src/solver/newton.rs:65-513) with the linear sub-problem
 * solved by direct LU. On return *iterations holds the Newton count and
 * *max_diff the last max|x-x_old|. */
int oo_newton_solve(int64_t n, int64_t nnz, const uint32_t *rowptr, const uint32_t *col,
                    const double *base_val, const double *base_b, int64_t n_diodes,
                    const ohm_diode_desc *diodes, int64_t n_bjts, const ohm_bjt_desc *bjts,
                    int64_t n_mosfets, const ohm_mosfet_desc *mosfets,
                    const oo_newton_params *params, double *x, int64_t *iterations,
                    double *max_diff);

/* Assemble A(x), b(x) = base + device companion stamps (no solve). */
void oo_newton_assemble(int64_t n, int64_t nnz, const double *base_val, const double *base_b,
                        int64_t n_diodes, const ohm_diode_desc *diodes, int64_t n_bjts,
                        const ohm_bjt_desc *bjts, int64_t n_mosfets,
                        const ohm_mosfet_desc *mosfets, const double *x, double *val, double *b);

/* Voltage limiting passes (BJT, MOSFET, diode, in the reference's order). */
void oo_newton_limit(int64_t n, int64_t n_diodes, const ohm_diode_desc *diodes, int64_t n_bjts,
                     const ohm_bjt_desc *bjts, int64_t n_mosfets, const ohm_mosfet_desc *mosfets,
                     const double *x_old, double *x_new);

/* PULSE source value at time t (src/analysis/transient_source.rs:39-71). */
double oo_eval_pulse(double v1, double v2, double td, double tr, double tf, double pw, double per,
                     double t);

/* Source value at time t: evaluate (src/analysis/transient_source.rs:8-37). */
double oo_source_value(const ohm_tran_source *s, const double *pwl_pairs, double t);

/* transient::run (src/analysis/transient.rs:25-199) with every linear
 * solve done by oo_sparse_lu_solve: method OHM_TRAN_FIXED_BE (backward
 * Euler at h = tstep, the same stepping rule without LTE control) or
 * OHM_TRAN_ADAPTIVE (the reference's TRAP/BE with LTE). Records t = 0 when
 * tstart <= 0 and every accepted step with t >= tstart: times[r],
 * waves[r * n_probes + j]; *n_rec receives the record count. */
int oo_transient(int64_t n, const uint32_t *g_rowptr, const uint32_t *g_col, const double *g_val,
                 const uint32_t *c_rowptr, const uint32_t *c_col, const double *c_val,
                 int64_t n_src, const ohm_tran_source *src, const double *pwl_pairs, int method,
                 double tstep, double tstop, double tstart, const double *x0, int64_t n_probes,
                 const uint32_t *probes, int64_t max_rec, double *times, double *waves,
                 int64_t *n_rec, double *x_final, int64_t *accepted, int64_t *rejected);

#ifdef __cplusplus
}
#endif
#endif

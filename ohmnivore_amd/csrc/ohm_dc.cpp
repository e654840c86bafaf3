// ohm_dc.cpp — dc::run (src/analysis/dc.rs:25-650) over the device Newton.
//
// The cascade is host orchestration around one device-resident Newton
// context: every continuation point replaces the context's base values (G
// plus the step's extra node-diagonal conductance) and right-hand side
// (scaled or anchored b) with ohm_newton_update and runs Newton from the
// previous point's answer. The context keeps its CSR upload, symbolic
// preconditioner plan, tuned SpMV shapes and captured graphs across steps,
// so a step costs one values/RHS upload plus the Newton iterations.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "ohm_internal.h"

namespace {

bool recoverable(ohm_status s) {
  // convergence failures move the cascade on; anything else is returned
  return s == OHM_ERR_SOLVE || s == OHM_ERR_NEWTON_NOT_CONVERGED ||
         s == OHM_ERR_NEWTON_NUMERICAL || s == OHM_ERR_CONVERGENCE;
}

struct Cascade {
  const ohm_mna *m = nullptr;
  ohm_newton_params base{};
  ohm_newton_ctx *ctx = nullptr;
  std::vector<int64_t> diag;  // value positions of the node diagonals
  std::vector<double> vals, x, x_prev;
  ohm_dc_stats st{};
  ohm_dc_step *trace = nullptr;
  int64_t cap = 0;

  ~Cascade() {
    if (ctx) ohm_newton_free(ctx);
  }

  // find_node_diagonal_positions (src/analysis/dc.rs:43-53)
  void find_diagonals() {
    for (int64_t r = 0; r < m->n_nodes; r++)
      for (uint32_t k = m->g_rowptr[r]; k < m->g_rowptr[r + 1]; k++)
        if ((int64_t)m->g_col[k] == r) {
          diag.push_back(k);
          break;
        }
  }

  ohm_status run_once(int stage, double level, const ohm_newton_params &p, int retry) {
    ohm_newton_result res{};
    ohm_status s = ohm_newton_run(ctx, &p, &res, nullptr);
    st.runs++;
    st.newton_iterations += res.iterations;
    st.bicgstab_iterations += res.bicgstab_iterations;
    st.device_ms += res.device_ms;
    if (st.runs <= cap) {
      ohm_dc_step &t = trace[st.runs - 1];
      t.stage = stage;
      t.converged = s == OHM_OK;
      t.level = level;
      t.iterations = res.iterations;
      t.retry = retry;
      t._pad = 0;
    }
    return s;
  }

  // gmin_newton_step (:297-332): Newton on G + extra_diag at the node
  // diagonals with right-hand side b_step, from `guess` (NULL = 0). A failed
  // run is retried in FP64 with a larger sub-solve budget and retry_max_it
  // Newton iterations (the reference's CPU sparse-LU retry; 0 = no retry,
  // as for the final true-GMIN solve). On success x holds the answer.
  ohm_status step(int stage, double level, const double *b_step, double extra_diag,
                  const double *guess, int64_t max_it, int64_t budget, int64_t retry_max_it) {
    std::copy(m->g_val, m->g_val + m->g_nnz, vals.begin());
    for (int64_t p : diag) vals[p] += extra_diag;
    ohm_status s = ohm_newton_update(ctx, vals.data(), b_step, nullptr, nullptr, nullptr);
    if (s != OHM_OK) return s;
    ohm_newton_params p = base;
    p.max_iterations = max_it;
    p.bicgstab_max_iterations = budget;
    p.initial_guess = guess;
    s = run_once(stage, level, p, 0);
    if (s != OHM_OK && recoverable(s) && retry_max_it > 0) {
      // the reference's CPU sparse-LU retry: a direct solve per Newton step
      // where the system fits the dense LU, else a larger FP64 budget
      if (m->size <= ohm::kDenseMaxRows) {
        p.linear_solver = OHM_LINEAR_DENSE_LU;
      } else {
        p.precision = OHM_PRECISION_FP64;
        p.bicgstab_max_iterations = std::max<int64_t>(10 * budget, 20000);
      }
      p.max_iterations = retry_max_it;
      s = run_once(stage, level, p, 1);
    }
    if (s == OHM_OK) {
      s = ohm_newton_get_x(ctx, x.data());
      if (s == OHM_OK) st.stage = stage;
    }
    return s;
  }

  // run_source_stepping (:480-640)
  ohm_status source_stepping() {
    const int64_t n = m->size;
    std::vector<double> bs(n);
    bool have = false;
    double last = 0.0, target = 0.1;
    int sub = 0;
    while (target <= 1.0) {
      for (int64_t i = 0; i < n; i++) bs[i] = target * m->b_dc[i];
      ohm_status s = step(OHM_DC_SOURCE_STEPPING, target, bs.data(), 0.0,
                          have ? x_prev.data() : nullptr, 30, 200, 15);
      if (s == OHM_OK) {
        x_prev = x;
        have = true;
        const double prev_step = target - last;
        last = target;
        sub = 0;
        if (std::fabs(target - 1.0) < 1e-12) return OHM_OK;
        target = std::min(target + prev_step * 2.0, 1.0);
      } else {
        if (!recoverable(s)) return s;
        if (++sub > 20) return s;
        target = last + (target - last) / 2.0;
        if (target - last < 1e-6) return s;
      }
    }
    return ohm::fail(OHM_ERR_SOLVE,
                     "Solve error: Source stepping completed without reaching alpha=1.0");
  }

  // run_pseudo_transient_continuation (:352-470), seeded by x_prev
  ohm_status pseudo_transient() {
    const int64_t n = m->size, nn = m->n_nodes;
    std::vector<double> bs(n);
    double last = 1.0, target = 1e-1;
    int sub = 0;
    while (target > 1e-8) {
      std::copy(m->b_dc, m->b_dc + n, bs.begin());
      for (int64_t i = 0; i < nn; i++) bs[i] += target * x_prev[i];
      ohm_status s = step(OHM_DC_PSEUDO_TRANSIENT, target, bs.data(), target, x_prev.data(),
                          20, 500, 20);
      if (s == OHM_OK) {
        x_prev = x;
        last = target;
        sub = 0;
        target *= 1e-2;
      } else {
        if (!recoverable(s) || ++sub > 7) return s;
        target = last * std::pow(10.0, -(double)sub);
      }
    }
    return step(OHM_DC_PSEUDO_TRANSIENT_FINAL, 0.0, m->b_dc, 0.0, x_prev.data(), 25, 1500, 25);
  }

  // run_nonlinear (:60-293)
  ohm_status run() {
    // the caller's max_iterations caps the direct attempt and its retry
    ohm_status s = step(OHM_DC_DIRECT, 0.0, m->b_dc, 0.0, nullptr, base.max_iterations, 200,
                        std::min<int64_t>(15, base.max_iterations));
    if (s == OHM_OK || !recoverable(s)) return s;
    if (m->n_mosfets > 0) {
      s = source_stepping();
      if (s == OHM_OK || !recoverable(s)) return s;
    }
    bool have = false;
    double last = 1.0, target = 1e-1;
    int sub = 0;
    while (target > 1e-13) {
      s = step(OHM_DC_GMIN_STEPPING, target, m->b_dc, target, have ? x_prev.data() : nullptr, 15,
               300, 15);
      if (s == OHM_OK) {
        x_prev = x;
        have = true;
        last = target;
        sub = 0;
        target *= 1e-3;
      } else {
        if (!recoverable(s) || ++sub > 13) return s;
        target = last * std::pow(10.0, -(double)sub);
      }
    }
    // the final true-GMIN step runs NewtonParams::default()'s iteration cap
    // (OHM_DC_FINAL_MAX_ITERATIONS overrides it: a cap too small to converge
    // sends the cascade on to pseudo-transient continuation)
    ohm_newton_params dflt;
    ohm_newton_default_params(&dflt);
    int64_t final_it = dflt.max_iterations;
    if (const char *env = getenv("OHM_DC_FINAL_MAX_ITERATIONS"))
      final_it = std::max<int64_t>(0, atoll(env));
    s = step(OHM_DC_GMIN_FINAL, 0.0, m->b_dc, 0.0, have ? x_prev.data() : nullptr, final_it,
             dflt.bicgstab_max_iterations, 0);
    if (s == OHM_OK || !recoverable(s) || !have) return s;
    return pseudo_transient();
  }
};

}  // namespace

extern "C" ohm_status ohm_dc_run(ohm_backend *b, const ohm_mna *mna,
                                 const ohm_newton_params *params, double *x_out,
                                 ohm_dc_stats *stats, ohm_dc_step *trace, int64_t trace_cap) {
  if (!b || !mna || !x_out) return ohm::fail(OHM_ERR_INVALID, "null argument");
  if (mna->size < 0 || (mna->size > 0 && (!mna->g_rowptr || !mna->b_dc)) ||
      (mna->g_nnz > 0 && (!mna->g_col || !mna->g_val)) || trace_cap < 0 ||
      (trace_cap > 0 && !trace))
    return ohm::fail(OHM_ERR_INVALID, "invalid system");
  if (stats) *stats = ohm_dc_stats{};
  const int64_t n = mna->size;
  if (n == 0) return OHM_OK;
  const bool nonlinear = mna->n_diodes > 0 || mna->n_bjts > 0 || mna->n_mosfets > 0;
  if (!nonlinear) {
    // dc::run's linear branch: solver.solve_real(&system.g, &system.b_dc)
    ohm_bicgstab_params bp;
    ohm_bicgstab_default_params(&bp);
    bp.tolerance = 1e-12;
    bp.max_iterations = 200000;
    ohm_bicgstab_result r{};
    ohm_status s = ohm_solve_real(b, (size_t)n, (size_t)mna->g_nnz, mna->g_val, mna->g_col,
                                  mna->g_rowptr, mna->b_dc, x_out, OHM_PREC_AUTO, &bp, &r);
    if (stats) {
      stats->stage = OHM_DC_LINEAR;
      stats->bicgstab_iterations = r.iterations;
      stats->device_ms = r.device_ms;
    }
    return s;
  }
  Cascade c;
  c.m = mna;
  if (params)
    c.base = *params;
  else
    ohm_newton_default_params(&c.base);
  c.trace = trace;
  c.cap = trace_cap;
  c.vals.resize(mna->g_nnz);
  c.x.resize(n);
  c.x_prev.assign(n, 0.0);
  c.find_diagonals();
  ohm_status s = ohm_newton_create(b, mna->g_val, mna->g_col, mna->g_rowptr, mna->b_dc,
                                   mna->diodes, (size_t)mna->n_diodes, mna->bjts,
                                   (size_t)mna->n_bjts, mna->mosfets, (size_t)mna->n_mosfets,
                                   (size_t)n, (size_t)mna->g_nnz, &c.ctx);
  if (s == OHM_OK) s = c.run();
  if (stats) *stats = c.st;
  if (s == OHM_OK) std::copy(c.x.begin(), c.x.end(), x_out);
  return s;
}

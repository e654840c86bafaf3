// ohm_trannl.cpp — transient analysis with the device currents in the time
// loop (the reference's transient steps G and C only).
//
// Each timestep is one Newton solve on a device-resident context whose base
// matrix is the companion matrix G + (alpha/h) C and whose base right-hand
// side is the step's constant part; the device evaluation, stamping,
// preconditioned BiCGSTAB and limiting of every Newton iteration stay on the
// GPU. The host carries x_n, evaluates the sources and forms the constant
// part (O(n + nnz(C)) per step). TRAP needs d_n = b(t) - G x_n - i(x_n),
// read from a second context (base G, right-hand side 0) whose assembled
// system at x_n gives G x_n + i(x_n) = A(x_n) x_n - b(x_n).
#include <algorithm>
#include <cmath>
#include <vector>

#include "ohm_internal.h"
#include "ohm_sources.h"

namespace {

bool step_failure(ohm_status s) {
  return s == OHM_ERR_SOLVE || s == OHM_ERR_NEWTON_NOT_CONVERGED ||
         s == OHM_ERR_NEWTON_NUMERICAL || s == OHM_ERR_CONVERGENCE;
}

struct NlTran {
  ohm_backend *b = nullptr;
  const ohm_mna *m = nullptr;
  int64_t n = 0, nnz = 0;
  std::vector<double> cg;  // C on G's pattern
  std::vector<ohm::TSrc> src;
  const double *pwl = nullptr;
  ohm_newton_ctx *ctx = nullptr, *ev = nullptr;
  ohm_newton_params np{};
  std::vector<double> vals, rhs, bsrc, tmp, a_ev, b_ev, zeros;
  double cur_sig = -1.0;
  ohm_tran_stats st{};

  ~NlTran() {
    if (ctx) ohm_newton_free(ctx);
    if (ev) ohm_newton_free(ev);
  }

  // evaluate_b_at_time (transient.rs:265-280): sources only, from zero
  void b_at(double t, double *out) const {
    std::fill(out, out + n, 0.0);
    for (const ohm::TSrc &s : src) {
      const double v = ohm::src_value(s, pwl, t);
      for (int j = 0; j < s.n_idx; j++) out[s.idx[j]] += s.sign[j] * v;
    }
  }

  // out += scale * M x over G's pattern
  void spmv_add(const double *mv, double scale, const double *x, double *out) const {
    for (int64_t r = 0; r < n; r++) {
      double s = 0.0;
      for (uint32_t k = m->g_rowptr[r]; k < m->g_rowptr[r + 1]; k++) s += mv[k] * x[m->g_col[k]];
      out[r] += scale * s;
    }
  }

  // one Newton solve of (G + sig C) x + i(x) = c, from `guess`; x_out on
  // success. If Newton from the previous point fails, the step's system is
  // handed to the DC continuation cascade (source stepping on c, GMIN
  // stepping, pseudo-transient): without capacitors a smaller step does not
  // make the algebraic system any easier, so h-halving alone cannot rescue
  // it (a CMOS chain's first step from zero node voltages is the DC problem).
  ohm_status solve(double sig, const double *c, const double *guess, double *x_out) {
    const double *v = nullptr;
    if (sig != cur_sig) {
      for (int64_t k = 0; k < nnz; k++) vals[k] = m->g_val[k] + sig * cg[k];
      v = vals.data();
    }
    ohm_status s = ohm_newton_update(ctx, v, c, nullptr, nullptr, nullptr);
    if (s != OHM_OK) return s;
    cur_sig = sig;
    ohm_newton_params p = np;
    p.initial_guess = guess;
    ohm_newton_result res{};
    s = ohm_newton_run(ctx, &p, &res, nullptr);
    st.linear_solves += res.iterations;
    st.bicgstab_iterations += res.bicgstab_iterations;
    st.device_ms += res.device_ms;
    if (s == OHM_OK) return ohm_newton_get_x(ctx, x_out);
    if (!step_failure(s)) return s;
    ohm_mna view = *m;
    view.g_val = vals.data();  // holds G + cur_sig C
    view.b_dc = const_cast<double *>(c);  // read only (ohm_dc_run takes a const ohm_mna)
    ohm_dc_stats ds{};
    s = ohm_dc_run(b, &view, &np, x_out, &ds, nullptr, 0);
    st.linear_solves += ds.newton_iterations;
    st.bicgstab_iterations += ds.bicgstab_iterations;
    st.device_ms += ds.device_ms;
    return s;
  }

  // d = b(t) - G x - i(x)  (C dx/dt at a solved point)
  ohm_status deriv(double t, const double *x, double *d) {
    OHM_TRY(ohm_newton_assemble_at(ev, x, a_ev.data(), b_ev.data()));
    b_at(t, d);
    for (int64_t r = 0; r < n; r++) d[r] += b_ev[r];
    spmv_add(a_ev.data(), -1.0, x, d);
    return OHM_OK;
  }
};

}  // namespace

extern "C" ohm_status ohm_transient_nonlinear(
    ohm_backend *b, const ohm_mna *mna, int64_t n_sources, const ohm_tran_source *sources,
    int64_t n_pwl, const double *pwl_pairs, const ohm_tran_params *prm,
    const ohm_newton_params *newton, const double *x0, int64_t n_probes, const uint32_t *probes,
    int64_t max_records, double *times, double *waves, double *x_final, ohm_tran_stats *stats) {
  using ohm::fail;
  if (!b || !mna || !prm) return fail(OHM_ERR_INVALID, "null argument");
  const int64_t n = mna->size;
  if (n <= 0 || !mna->g_rowptr || !mna->g_col || !mna->g_val || !mna->c_rowptr ||
      (mna->c_nnz > 0 && (!mna->c_col || !mna->c_val)))
    return fail(OHM_ERR_INVALID, "invalid system");
  if (n_sources < 0 || (n_sources > 0 && !sources) || n_pwl < 0 || (n_pwl > 0 && !pwl_pairs))
    return fail(OHM_ERR_INVALID, "invalid source table");
  if (n_probes < 0 || (n_probes > 0 && !probes) || max_records < 0 ||
      (max_records > 0 && (!times || (n_probes > 0 && !waves))))
    return fail(OHM_ERR_INVALID, "invalid record buffers");
  for (int64_t i = 0; i < n_probes; i++)
    if (probes[i] >= (uint64_t)n) return fail(OHM_ERR_INVALID, "probe %lld out of range", (long long)i);
  const double tstep = prm->tstep, tstop = prm->tstop, tstart = prm->tstart;
  if (!(tstep > 0.0) || !(tstop >= 0.0))
    return fail(OHM_ERR_INVALID, "tstep must be > 0 and tstop >= 0");
  if (prm->method != OHM_TRAN_FIXED_BE && prm->method != OHM_TRAN_ADAPTIVE)
    return fail(OHM_ERR_INVALID, "unknown transient method %d", prm->method);

  NlTran T;
  T.b = b;
  T.m = mna;
  T.n = n;
  T.nnz = mna->g_nnz;
  T.pwl = pwl_pairs;
  if (newton)
    T.np = *newton;
  else
    ohm_newton_default_params(&T.np);
  // C's values scattered onto G's pattern (the compiler zero-fills every C
  // position into G, src/compiler.rs:577-579)
  T.cg.assign(T.nnz, 0.0);
  for (int64_t r = 0; r < n; r++) {
    const uint32_t *g0 = mna->g_col + mna->g_rowptr[r], *g1 = mna->g_col + mna->g_rowptr[r + 1];
    for (uint32_t q = mna->c_rowptr[r]; q < mna->c_rowptr[r + 1]; q++) {
      const uint32_t *it = std::lower_bound(g0, g1, mna->c_col[q]);
      if (it == g1 || *it != mna->c_col[q])
        return fail(OHM_ERR_INVALID, "C entry (%lld, %u) is not in G's pattern", (long long)r,
                    mna->c_col[q]);
      T.cg[it - mna->g_col] += mna->c_val[q];
    }
  }
  T.src.resize(n_sources);
  for (int64_t i = 0; i < n_sources; i++) {
    const char *why = ohm::to_tsrc(sources[i], n, n_pwl, T.src[i]);
    if (why) return fail(OHM_ERR_INVALID, "source %lld: %s", (long long)i, why);
  }
  T.vals.assign(T.nnz, 0.0);
  T.rhs.assign(n, 0.0);
  T.bsrc.assign(n, 0.0);
  T.tmp.assign(n, 0.0);
  T.a_ev.assign(T.nnz, 0.0);
  T.b_ev.assign(n, 0.0);
  T.zeros.assign(n, 0.0);
  OHM_TRY(ohm_newton_create(b, mna->g_val, mna->g_col, mna->g_rowptr, T.zeros.data(), mna->diodes,
                            (size_t)mna->n_diodes, mna->bjts, (size_t)mna->n_bjts, mna->mosfets,
                            (size_t)mna->n_mosfets, (size_t)n, (size_t)T.nnz, &T.ctx));
  OHM_TRY(ohm_newton_create(b, mna->g_val, mna->g_col, mna->g_rowptr, T.zeros.data(), mna->diodes,
                            (size_t)mna->n_diodes, mna->bjts, (size_t)mna->n_bjts, mna->mosfets,
                            (size_t)mna->n_mosfets, (size_t)n, (size_t)T.nnz, &T.ev));

  std::vector<double> x(n, 0.0), x_be(n), x_tr(n), d(n);
  if (x0) std::copy(x0, x0 + n, x.begin());
  int64_t n_rec = 0;
  auto record = [&](double t) {
    if (n_rec >= max_records) return;
    times[n_rec] = t;
    for (int64_t i = 0; i < n_probes; i++) waves[n_rec * n_probes + i] = x[probes[i]];
    n_rec++;
  };
  // the BE step from x: (G + C/h) y + i(y) = b(t+h) + (C/h) x
  auto be_step = [&](double t_next, double h, double *out) -> ohm_status {
    T.b_at(t_next, T.rhs.data());
    T.spmv_add(T.cg.data(), 1.0 / h, x.data(), T.rhs.data());
    return T.solve(1.0 / h, T.rhs.data(), x.data(), out);
  };
  const double h_min = tstep / 10000.0, h_max = tstep, lte_tol = 1e-3;  // transient.rs:36-39
  if (tstart <= 0.0) record(0.0);
  double t = 0.0, h = std::min(tstep, h_max);
  ohm_status rc = OHM_OK;
  if (prm->method == OHM_TRAN_FIXED_BE) {
    while (t < tstop) {
      if (t + h > tstop) h = tstop - t;
      if (h < h_min * 0.5) break;
      const double t_next = t + h;
      if ((rc = be_step(t_next, h, x_be.data())) != OHM_OK) break;
      x.swap(x_be);
      T.st.steps_accepted++;
      t = t_next;
      if (t >= tstart) record(t);
    }
  } else {
    bool use_be = true;
    int fails = 0;
    while (t < tstop) {
      if (t + h > tstop) h = tstop - t;
      if (h < h_min * 0.5) break;
      const double t_next = t + h;
      double lte = 0.0;
      ohm_status s = OHM_OK;
      if (use_be) {
        s = be_step(t_next, h, x_be.data());
      } else {
        // TRAP: (G + 2C/h) y + i(y) = b(t+h) + (2C/h) x + d(x)
        s = T.deriv(t, x.data(), d.data());
        if (s == OHM_OK) {
          T.b_at(t_next, T.rhs.data());
          T.spmv_add(T.cg.data(), 2.0 / h, x.data(), T.rhs.data());
          for (int64_t r = 0; r < n; r++) T.rhs[r] += d[r];
          s = T.solve(2.0 / h, T.rhs.data(), x.data(), x_tr.data());
        }
        if (s == OHM_OK) s = be_step(t_next, h, x_be.data());
        if (s == OHM_OK) {
          double md = 0.0;
          for (int64_t r = 0; r < n; r++) md = std::max(md, std::fabs(x_tr[r] - x_be[r]));
          lte = (2.0 / 3.0) * md;  // compute_lte, :355-362
        }
      }
      if (s != OHM_OK && !step_failure(s)) {
        rc = s;
        break;
      }
      if (s != OHM_OK || lte > lte_tol) {
        // reject: LTE above tolerance, or Newton failed at this step size
        const double h_new = s != OHM_OK ? 0.5 * h
                                         : h * std::min(2.0, std::max(0.5, 0.9 * std::sqrt(lte_tol / lte)));
        h = std::max(h_new, h_min);
        fails++;
        if (fails >= 10 && h <= h_min * 1.01) {
          rc = fail(OHM_ERR_TIMESTEP,
                    "Timestep error: simulation failed: %d consecutive failures at minimum "
                    "timestep h_min=%.2e near t=%.2e",
                    fails, h_min, t);
          break;
        }
        T.st.steps_rejected++;
        use_be = true;
        continue;
      }
      if (use_be) {
        x.swap(x_be);
        if (lte == 0.0) use_be = false;  // after the first BE step switch to TRAP
      } else {
        const double h_new = h * std::min(2.0, std::max(0.5, 0.9 * std::sqrt(lte_tol / lte)));
        x.swap(x_tr);
        h = std::min(h_max, std::max(h_min, h_new));
      }
      fails = 0;
      T.st.steps_accepted++;
      t = t_next;
      if (t >= tstart) record(t);
    }
  }
  T.st.records = n_rec;
  T.st.t_end = t;
  if (x_final) std::copy(x.begin(), x.end(), x_final);
  if (stats) *stats = T.st;
  return rc;
}

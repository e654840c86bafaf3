// ohm_sources.h — time-varying source values of the transient right-hand
// side (transient_source.rs:8-116), shared by the device RHS kernel and the
// host-driven nonlinear stepper so both evaluate the same expressions.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/ohm_solver.h"

namespace ohm {

struct TSrc {
  int kind;       // OHM_SRC_*
  int n_idx;
  uint32_t idx[2];
  double sign[2];
  double dc;
  double p[7];
  int pwl_off, pwl_n;
};

__host__ __device__ inline double src_value(const TSrc &s, const double *__restrict__ pwl, double t) {
  switch (s.kind) {
    case OHM_SRC_PULSE: {  // eval_pulse, transient_source.rs:39-71
      const double v1 = s.p[0], v2 = s.p[1], td = s.p[2], tr = s.p[3], tf = s.p[4], pw = s.p[5],
                   per = s.p[6];
      if (t < td) return v1;
      const double t_rel = (per > 0.0 && per < 1.7976931348623157e308) ? fmod(t - td, per) : t - td;
      if (t_rel < tr) return tr > 0.0 ? v1 + (v2 - v1) * t_rel / tr : v2;
      if (t_rel < tr + pw) return v2;
      if (t_rel < tr + pw + tf) return tf > 0.0 ? v2 + (v1 - v2) * (t_rel - tr - pw) / tf : v1;
      return v1;
    }
    case OHM_SRC_SIN: {  // eval_sin, :73-83
      const double vo = s.p[0], va = s.p[1], freq = s.p[2], td = s.p[3], theta = s.p[4];
      if (t < td) return vo;
      const double dt = t - td;
      const double env = theta != 0.0 ? exp(-dt * theta) : 1.0;
      return vo + va * sin(2.0 * M_PI * freq * dt) * env;
    }
    case OHM_SRC_EXP: {  // eval_exp, :106-116
      const double v1 = s.p[0], v2 = s.p[1], td1 = s.p[2], tau1 = s.p[3], td2 = s.p[4],
                   tau2 = s.p[5];
      if (t < td1) return v1;
      const double rise = (v2 - v1) * (1.0 - exp(-(t - td1) / tau1));
      if (t < td2) return v1 + rise;
      const double fall = (v1 - v2) * (1.0 - exp(-(t - td2) / tau2));
      return v1 + rise + fall;
    }
    case OHM_SRC_PWL: {  // eval_pwl, :85-104
      const double *q = pwl + 2 * s.pwl_off;
      const int m = s.pwl_n;
      if (m == 0) return 0.0;
      if (t <= q[0]) return q[1];
      if (t >= q[2 * (m - 1)]) return q[2 * (m - 1) + 1];
      for (int i = 1; i < m; i++)
        if (t <= q[2 * i]) {
          const double t0 = q[2 * i - 2], v0 = q[2 * i - 1], t1 = q[2 * i], v1 = q[2 * i + 1];
          return v0 + (v1 - v0) * ((t - t0) / (t1 - t0));
        }
      return q[2 * (m - 1) + 1];
    }
    default:
      return s.dc;
  }
}

}  // namespace ohm

namespace ohm {

// ohm_tran_source -> TSrc with the checks every caller applies; returns
// nullptr or the reason the entry is invalid.
inline const char *to_tsrc(const ohm_tran_source &q, int64_t n, int64_t n_pwl, TSrc &t) {
  if (q.n_idx < 0 || q.n_idx > 2) return "n_idx must be 0..2";
  if (q.kind < OHM_SRC_DC || q.kind > OHM_SRC_PWL) return "unknown kind";
  if (q.kind == OHM_SRC_PWL &&
      (q.pwl_offset < 0 || q.pwl_count < 0 || q.pwl_offset + q.pwl_count > n_pwl))
    return "PWL pairs out of range";
  for (int j = 0; j < q.n_idx; j++)
    if (q.idx[j] >= (uint64_t)n) return "index out of range";
  t = TSrc{};
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
  return nullptr;
}

}  // namespace ohm

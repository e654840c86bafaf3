// ohm_newton.hip — GPU-resident Newton-Raphson (newton_solve,
// src/solver/newton.rs:65-513) and the NonlinearBackend kernels
// (src/solver/nonlinear.rs:31-150; WGSL originals in
// src/solver/ds_nonlinear_shaders.rs), in native FP64.
//
// Per Newton iteration, all on one stream:
//   x_old = x
//   device evaluation  -> per-device stamp values S (conductances and Norton
//                         currents i_eq = i - G v), FP64 junction exponentials
//   assembly (gather)  -> every CSR slot / RHS entry touched by a device is
//                         rewritten as base + sum of its contributions in
//                         device order. Only touched entries are written, and
//                         no two threads write one entry, so assembly is
//                         race-free and deterministic where the reference's
//                         scatter stamps race (ds_nonlinear_shaders.rs:144-163).
//   preconditioner refresh + BiCGSTAB (warm-started from x)
//   limiting           -> BJT pnjlim, MOSFET per-node min step scale, diode pnjlim
//   convergence        -> max |x - x_old| and a NaN flag into one status word
// The host reads that status word once per Newton iteration.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ohm_precond.h"
#include "ohm_spmv.h"

namespace ohm {

constexpr double kVT = 0.02585;          // MOSFET smoothing voltage (ds_nonlinear_shaders.rs:777)
constexpr unsigned kScaleOne = 1000000u; // MVL_SCALE_ONE (:1027)
constexpr double kStepLimit = 0.5;       // MVL_STEP_LIMIT (:1028)

// stamp value layout per device
constexpr int kDiodeS = 2;   // g, i_eq
constexpr int kBjtS = 12;    // G[9], ieq_c, ieq_b, ieq_e
constexpr int kMosS = 8;     // S[6], ieq_d, ieq_s

__device__ __forceinline__ double xv(const double *x, uint32_t i) {
  return i == OHM_SENTINEL ? 0.0 : x[i];
}
__device__ __forceinline__ double clamp_exp(double a) {  // ds_shaders.rs:96-108
  return exp(fmin(fmax(a, -80.0), 80.0));
}

// diode_eval_ds (:41-76) + Norton current (:212-248)
__global__ void k_diode_eval(int nd, const ohm_diode_desc *__restrict__ d,
                             const double *__restrict__ x, double *__restrict__ S) {
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < nd; k += gridDim.x * kBlock) {
    const ohm_diode_desc q = d[k];
    const double is = q.is_val, nvt = q.n_vt;
    const double vd = xv(x, q.anode_idx) - xv(x, q.cathode_idx);
    const double e = clamp_exp(vd / nvt);
    const double i = is * (e - 1.0), g = is * e / nvt;
    S[kDiodeS * k] = g;
    S[kDiodeS * k + 1] = i - g * vd;
  }
}

// bjt_eval_ds (:281-367) + stamps (:394-573)
__global__ void k_bjt_eval(int nq, const ohm_bjt_desc *__restrict__ d, const double *__restrict__ x,
                           double *__restrict__ S) {
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < nq; k += gridDim.x * kBlock) {
    const ohm_bjt_desc &q = d[k];
    const double pol = q.polarity, is = q.is_val, bf = q.bf, br = q.br, nf = q.nf_vt,
                 nr = q.nr_vt;
    const double vc = xv(x, q.collector_idx), vb = xv(x, q.base_idx), ve = xv(x, q.emitter_idx);
    const double vbe = pol * (vb - ve), vbc = pol * (vb - vc);
    const double ef = clamp_exp(vbe / nf), er = clamp_exp(vbc / nr);
    const double i_f = is * (ef - 1.0), i_r = is * (er - 1.0);
    const double bf_frac = bf / (bf + 1.0);
    const double ic = (bf_frac * i_f - i_r / (br + 1.0)) * pol;
    const double ib = (i_f / (bf + 1.0) + i_r / (br + 1.0)) * pol;
    const double gf = is * ef / nf, gr = is * er / nr;
    const double a = bf_frac * gf, b = -(gr / (br + 1.0)), c = gf / (bf + 1.0), dd = gr / (br + 1.0);
    double G[9] = {-b, a + b, -a, -dd, c + dd, -c, b + dd, -(a + b + c + dd), a + c};
    double *s = S + (size_t)kBjtS * k;
#pragma unroll
    for (int j = 0; j < 9; j++) s[j] = G[j];
    const double ie = -(ic + ib);
    s[9] = ic - (G[0] * vc + G[1] * vb + G[2] * ve);
    s[10] = ib - (G[3] * vc + G[4] * vb + G[5] * ve);
    s[11] = ie - (G[6] * vc + G[7] * vb + G[8] * ve);
  }
}

// Smooth level-1 MOSFET (mosfet_eval_ds :747-849; CPU twin gpu_shaders.rs:1452-1493)
__device__ __forceinline__ void mosfet_core(double vgs, double vds, double vto, double kp,
                                            double lambda, double pol, double &id, double &gm,
                                            double &gds) {
  vgs *= pol;
  vds *= pol;
  const double v_ov = vgs - vto, arg = v_ov / kVT;
  double v_ov_eff, sig;
  if (arg > 20.0) {
    v_ov_eff = v_ov;
    sig = 1.0;
  } else if (arg < -20.0) {
    const double e = exp(arg);
    v_ov_eff = kVT * e;
    sig = e;
  } else {
    v_ov_eff = kVT * log1p(exp(arg));
    sig = 1.0 / (1.0 + exp(-arg));
  }
  const double lam = 1.0 + lambda * vds;
  double i;
  if (vds < v_ov_eff) {
    const double inner = v_ov_eff * vds - vds * vds * 0.5;
    i = kp * inner * lam;
    gm = kp * vds * lam * sig;
    gds = kp * (v_ov_eff - vds) * lam + kp * inner * lambda;
  } else {
    const double sat = 0.5 * kp * v_ov_eff * v_ov_eff;
    i = sat * lam;
    gm = kp * v_ov_eff * lam * sig;
    gds = sat * lambda;
  }
  id = pol * i;
}

__global__ void k_mosfet_eval(int nm, const ohm_mosfet_desc *__restrict__ d,
                              const double *__restrict__ x, double *__restrict__ S) {
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < nm; k += gridDim.x * kBlock) {
    const ohm_mosfet_desc &q = d[k];
    const double vd = xv(x, q.drain_idx), vg = xv(x, q.gate_idx), vs = xv(x, q.source_idx);
    double id, gm, gds;
    mosfet_core(vg - vs, vd - vs, q.vto, q.kp, q.lambda, q.polarity, id, gm, gds);
    double *s = S + (size_t)kMosS * k;
    s[0] = gds;
    s[1] = gm;
    s[2] = -(gm + gds);
    s[3] = -gds;
    s[4] = -gm;
    s[5] = gm + gds;
    s[6] = id - (gds * vd + gm * vg - (gm + gds) * vs);
    s[7] = -id - (-gds * vd - gm * vg + (gm + gds) * vs);
  }
}

// Gather assembly: out[target[t]] = base[target[t]] + sum_k sign_k * S[src_k].
// The NaN/Inf test of the assembled values (src/solver/newton.rs:258-263)
// folds into the status flag.
__global__ void k_gather(int nt, const uint32_t *__restrict__ target,
                         const uint32_t *__restrict__ ptr, const uint32_t *__restrict__ src,
                         const double *__restrict__ sign, const double *__restrict__ S,
                         const double *__restrict__ base, double *__restrict__ out,
                         int *__restrict__ nan_flag) {
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < nt; t += gridDim.x * kBlock) {
    const uint32_t pos = target[t];
    double v = base[pos];
    for (uint32_t q = ptr[t]; q < ptr[t + 1]; q++) v += sign[q] * S[src[q]];
    out[pos] = v;
    if (nan_flag && !isfinite(v)) atomicOr(nan_flag, 1);
  }
}

// ------------------------------------------------------------ limiting ----
__device__ __forceinline__ double pnjlim(double vnew, double vold, double vt, double vcrit) {
  if (vnew > vcrit && fabs(vnew - vold) > vt + vt) {  // :1173-1187 as written
    if (vold > 0.0) {
      const double arg = 1.0 + (vnew - vold) / vt;
      if (arg > 0.0) return vold + vt * (2.0 + log(arg - 2.0));
      return vcrit;
    }
    return vt * log(vnew / vt);
  }
  return vnew;
}

__device__ __forceinline__ void scale_node(const double *xo, double *xn, uint32_t t, double s) {
  if (t != OHM_SENTINEL) xn[t] = xo[t] + (xn[t] - xo[t]) * s;
}

__device__ bool bjt_limit_one(const ohm_bjt_desc &q, const double *xo, double *xn) {
  bool lim = false;
  const double pol = q.polarity, is = q.is_val, nf = q.nf_vt, nr = q.nr_vt;
  const double vcf = nf * log(nf / (M_SQRT2 * is)), vcr = nr * log(nr / (M_SQRT2 * is));
  const uint32_t B = q.base_idx, E = q.emitter_idx, C = q.collector_idx;
  double vo = pol * (xv(xo, B) - xv(xo, E)), vn = pol * (xv(xn, B) - xv(xn, E));
  double vl = pnjlim(vn, vo, nf, vcf);
  if (vl != vn && fabs(vn - vo) > 1e-30) {
    const double s = (vl - vo) / (vn - vo);
    if (!q.node_fixed[1]) scale_node(xo, xn, B, s);
    if (!q.node_fixed[2]) scale_node(xo, xn, E, s);
    lim = true;
  }
  vo = pol * (xv(xo, B) - xv(xo, C));
  vn = pol * (xv(xn, B) - xv(xn, C));
  vl = pnjlim(vn, vo, nr, vcr);
  if (vl != vn && fabs(vn - vo) > 1e-30) {
    const double s = (vl - vo) / (vn - vo);
    if (!q.node_fixed[1]) scale_node(xo, xn, B, s);
    if (!q.node_fixed[0]) scale_node(xo, xn, C, s);
    lim = true;
  }
  return lim;
}

__device__ bool diode_limit_one(const ohm_diode_desc &q, const double *xo, double *xn) {
  const double nvt = q.n_vt, is = q.is_val;
  const double vcrit = nvt * log(nvt / (M_SQRT2 * is));
  const double vo = xv(xo, q.anode_idx) - xv(xo, q.cathode_idx);
  const double vn = xv(xn, q.anode_idx) - xv(xn, q.cathode_idx);
  const double vl = pnjlim(vn, vo, nvt, vcrit);
  if (vl != vn && fabs(vn - vo) > 1e-30) {
    const double s = (vl - vo) / (vn - vo);
    scale_node(xo, xn, q.anode_idx, s);
    scale_node(xo, xn, q.cathode_idx, s);
    return true;
  }
  return false;
}

// serial = 1: one thread walks the devices in index order (used when two
// devices of the pass share a node, where a parallel pass would race).
// *limited is set to 1 when any step is cut (every writer stores the same
// value, so the unordered plain stores are benign).
__global__ void k_bjt_limit(int nq, const ohm_bjt_desc *__restrict__ d, const double *xo,
                            double *xn, int serial, int *limited) {
  bool lim = false;
  if (serial) {
    if (blockIdx.x == 0 && threadIdx.x == 0)
      for (int k = 0; k < nq; k++) lim |= bjt_limit_one(d[k], xo, xn);
  } else {
    for (int k = blockIdx.x * kBlock + threadIdx.x; k < nq; k += gridDim.x * kBlock)
      lim |= bjt_limit_one(d[k], xo, xn);
  }
  if (lim) *limited = 1;
}

__global__ void k_diode_limit(int nd, const ohm_diode_desc *__restrict__ d, const double *xo,
                              double *xn, int serial, int *limited) {
  bool lim = false;
  if (serial) {
    if (blockIdx.x == 0 && threadIdx.x == 0)
      for (int k = 0; k < nd; k++) lim |= diode_limit_one(d[k], xo, xn);
  } else {
    for (int k = blockIdx.x * kBlock + threadIdx.x; k < nd; k += gridDim.x * kBlock)
      lim |= diode_limit_one(d[k], xo, xn);
  }
  if (lim) *limited = 1;
}

__global__ void k_fill_u32(int n, unsigned v, unsigned *x) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) x[i] = v;
}

__device__ __forceinline__ unsigned encode_scale(double delta) {
  double e = floor(kStepLimit / fabs(delta) * (double)kScaleOne);
  return e < 1.0 ? 1u : (unsigned)e;
}

// mosfet_voltage_limit_reduce_ds (:1052-1103)
__global__ void k_mos_limit_reduce(int nm, const ohm_mosfet_desc *__restrict__ d,
                                   const double *__restrict__ xo, const double *__restrict__ xn,
                                   unsigned *__restrict__ scale) {
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < nm; k += gridDim.x * kBlock) {
    const ohm_mosfet_desc &q = d[k];
    const uint32_t D = q.drain_idx, G = q.gate_idx, S = q.source_idx, f = q.node_fixed_packed;
    const double dgs = (xv(xn, G) - xv(xn, S)) - (xv(xo, G) - xv(xo, S));
    if (fabs(dgs) > kStepLimit) {
      const unsigned e = encode_scale(dgs);
      if (G != OHM_SENTINEL && !(f & 2u)) atomicMin(&scale[G], e);
      if (S != OHM_SENTINEL && !(f & 4u)) atomicMin(&scale[S], e);
    }
    const double dds = (xv(xn, D) - xv(xn, S)) - (xv(xo, D) - xv(xo, S));
    if (fabs(dds) > kStepLimit) {
      const unsigned e = encode_scale(dds);
      if (D != OHM_SENTINEL && !(f & 1u)) atomicMin(&scale[D], e);
      if (S != OHM_SENTINEL && !(f & 4u)) atomicMin(&scale[S], e);
    }
  }
}

// mosfet_voltage_limit_apply_ds (:1106-1130)
__global__ void k_mos_limit_apply(int n, const unsigned *__restrict__ scale,
                                  const double *__restrict__ xo, double *__restrict__ xn,
                                  int *limited) {
  bool lim = false;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const unsigned s = scale[i];
    if (s < kScaleOne) {
      xn[i] = xo[i] + (xn[i] - xo[i]) * ((double)s / (double)kScaleOne);
      lim = true;
    }
  }
  if (lim) *limited = 1;
}

// convergence_check_ds (:1257-1297): max |x - x_old| and a NaN/Inf flag.
struct NewtonWord {
  double max_diff;
  int nan;
  int done;
  int limited;  // a limiter cut this iteration's step
  int pad;
};

__global__ __launch_bounds__(kBlock) void k_converge(int n, const double *__restrict__ xo,
                                                     const double *__restrict__ xn,
                                                     double *partials, unsigned *counter,
                                                     const int *__restrict__ asm_nan,
                                                     NewtonWord *w) {
  double m = 0.0;
  int bad = 0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double a = xn[i];
    if (!isfinite(a)) bad = 1;
    m = fmax(m, fabs(a - xo[i]));
  }
  __shared__ double sm[kWaves];
  __shared__ int sbad;
  if (threadIdx.x == 0) sbad = 0;
  __syncthreads();
  if (bad) atomicOr(&sbad, 1);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  // reuse grid_sum's ticket protocol with the max in slot 0 and the flag in slot 1
  double v[2] = {0.0, 0.0}, tot[2];
  if (threadIdx.x == 0) {
    double mm = 0.0;
    for (int i = 0; i < kWaves; i++) mm = fmax(mm, sm[i]);
    v[0] = mm;
    v[1] = sbad ? 1.0 : 0.0;
  }
  // grid-wide: the same ticket protocol as grid_sum, combining by max
  __shared__ int s_last;
  if (threadIdx.x == 0) {
    __hip_atomic_store(&partials[2 * blockIdx.x], v[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&partials[2 * blockIdx.x + 1], v[1], __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);
    s_last = (__hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
              gridDim.x - 1);
  }
  __syncthreads();
  if (!s_last) return;
  double mm = 0.0, bb = 0.0;
  for (unsigned i = threadIdx.x; i < gridDim.x; i += kBlock) {
    mm = fmax(mm, __hip_atomic_load(&partials[2 * i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    bb += __hip_atomic_load(&partials[2 * i + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  mm = wave_max(mm);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = mm;
  double bsum[1] = {bb};
  block_sum<1>(bsum);  // contains a barrier, so sm[] is complete below
  if (threadIdx.x == 0) {
    for (int i = 0; i < kWaves; i++) mm = fmax(mm, sm[i]);
    tot[0] = mm;
    tot[1] = bsum[0];
    w->max_diff = tot[0];
    w->nan = (tot[1] > 0.0 || (asm_nan && *asm_nan)) ? 1 : 0;
    w->done = 1;
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__global__ void k_copy_d(int n, const double *__restrict__ x, double *__restrict__ y) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) y[i] = x[i];
}

}  // namespace ohm

using namespace ohm;

// ------------------------------------------------- adjoint scatter ----
// dJ/dp = -lambda^T dF/dp at the operating point x, with F the MNA residual
// G x + i_dev(x) - b (the assembled Newton matrix is dF/dx).
// Resistor conductance g between a and b adds g (x_a - x_b) to F_a and
// subtracts it from F_b.
__global__ void k_adj_conductance(int nr, const uint32_t *__restrict__ nodes,
                                  const double *__restrict__ x, const double *__restrict__ lam,
                                  double *__restrict__ grad) {
  for (int r = blockIdx.x * kBlock + threadIdx.x; r < nr; r += gridDim.x * kBlock) {
    const uint32_t a = nodes[2 * r], c = nodes[2 * r + 1];
    const double v = xv(x, a) - xv(x, c);
    grad[r] = -(xv(lam, a) - xv(lam, c)) * v;
  }
}

// MOSFET transconductance parameter kp (= KP W/L): the drain current is
// linear in kp, i_d = kp * i_unit, entering F at the drain stamp row and
// leaving at the source stamp row (b_idx, the rows the Norton currents go to).
__global__ void k_adj_mosfet_kp(int nm, const ohm_mosfet_desc *__restrict__ d,
                                const double *__restrict__ x, const double *__restrict__ lam,
                                double *__restrict__ grad) {
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < nm; k += gridDim.x * kBlock) {
    const ohm_mosfet_desc &q = d[k];
    const double vd = xv(x, q.drain_idx), vg = xv(x, q.gate_idx), vs = xv(x, q.source_idx);
    double iu, gm, gds;
    mosfet_core(vg - vs, vd - vs, q.vto, 1.0, q.lambda, q.polarity, iu, gm, gds);
    grad[k] = -(xv(lam, q.b_idx[0]) - xv(lam, q.b_idx[1])) * iu;
  }
}

struct ohm_newton_ctx {
  ohm_backend *b = nullptr;
  int n = 0;
  int64_t nnz = 0;
  int nd = 0, nq = 0, nm = 0;
  DevCsr A;                       // pattern + assembled values
  double *base_val = nullptr, *base_b = nullptr, *rhs = nullptr;
  double *x = nullptr, *x_old = nullptr, *S = nullptr;
  ohm_diode_desc *dd = nullptr;
  ohm_bjt_desc *qd = nullptr;
  ohm_mosfet_desc *md = nullptr;
  unsigned *scale = nullptr;
  int *asm_nan = nullptr;
  NewtonWord *word = nullptr;
  // gather lists: matrix slots and RHS entries
  int nt_m = 0, nt_b = 0;
  uint32_t *tm = nullptr, *pm = nullptr, *sm = nullptr, *tb = nullptr, *pb = nullptr, *sb = nullptr;
  double *gm = nullptr, *gb = nullptr;
  bool serial_diode = false, serial_bjt = false;
  PrecondDev P;
  int prec_kind = -1;
  int prec_level = 1;
  double *adj = nullptr;  // adjoint scratch: rhs, lambda, gradients (grown on demand)
  size_t adj_cap = 0;
  int zero_diag = -1;     // base matrix has a zero diagonal (-1: not yet checked)
  // host copies of the descriptors: ohm_newton_update keeps their topology
  std::vector<ohm_diode_desc> h_dd;
  std::vector<ohm_bjt_desc> h_qd;
  std::vector<ohm_mosfet_desc> h_md;
};

namespace {

struct Contrib {
  uint32_t target, src;
  double sign;
};

ohm_status build_gather(const std::vector<Contrib> &cs, int &nt, uint32_t **t, uint32_t **p,
                        uint32_t **s, double **g) {
  std::vector<Contrib> v(cs);
  std::stable_sort(v.begin(), v.end(),
                   [](const Contrib &a, const Contrib &b) { return a.target < b.target; });
  std::vector<uint32_t> tt, pp(1, 0), ss;
  std::vector<double> gg;
  for (size_t i = 0; i < v.size(); i++) {
    if (tt.empty() || tt.back() != v[i].target) {
      if (!tt.empty()) pp.push_back((uint32_t)ss.size());
      tt.push_back(v[i].target);
    }
    ss.push_back(v[i].src);
    gg.push_back(v[i].sign);
  }
  if (!tt.empty()) pp.push_back((uint32_t)ss.size());
  nt = (int)tt.size();
  auto up = [](auto **dst, const auto &vec) -> ohm_status {
    using V = typename std::remove_reference<decltype(vec)>::type::value_type;
    OHM_HIP(hipMalloc(dst, sizeof(V) * std::max<size_t>(vec.size(), 1)));
    if (!vec.empty())
      OHM_HIP(hipMemcpy(*dst, vec.data(), sizeof(V) * vec.size(), hipMemcpyHostToDevice));
    return OHM_OK;
  };
  OHM_TRY(up(t, tt));
  OHM_TRY(up(p, pp));
  OHM_TRY(up(s, ss));
  OHM_TRY(up(g, gg));
  return OHM_OK;
}

template <typename F>
bool shares_nodes(int count, int n, F nodes_of) {
  std::vector<int> owner(n, -1);
  for (int k = 0; k < count; k++) {
    uint32_t ns[3];
    int c = nodes_of(k, ns);
    for (int t = 0; t < c; t++) {
      if (ns[t] == OHM_SENTINEL) continue;
      if (owner[ns[t]] >= 0 && owner[ns[t]] != k) return true;
      owner[ns[t]] = k;
    }
  }
  return false;
}

void enqueue_eval_assemble(ohm_newton_ctx *c, const double *x) {
  ohm_backend *b = c->b;
  const dim3 blk(kBlock);
  if (c->nd)
    hipLaunchKernelGGL(k_diode_eval, dim3(grid_for(c->nd)), blk, 0, b->stream, c->nd, c->dd, x,
                       c->S);
  if (c->nq)
    hipLaunchKernelGGL(k_bjt_eval, dim3(grid_for(c->nq)), blk, 0, b->stream, c->nq, c->qd, x,
                       c->S + (size_t)kDiodeS * c->nd);
  if (c->nm)
    hipLaunchKernelGGL(k_mosfet_eval, dim3(grid_for(c->nm)), blk, 0, b->stream, c->nm, c->md, x,
                       c->S + (size_t)kDiodeS * c->nd + (size_t)kBjtS * c->nq);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, b->stream, 1, 0u, (unsigned *)c->asm_nan);
  if (c->nt_m)
    hipLaunchKernelGGL(k_gather, dim3(grid_for(c->nt_m)), blk, 0, b->stream, c->nt_m, c->tm, c->pm,
                       c->sm, c->gm, c->S, c->base_val, c->A.val, c->asm_nan);
  if (c->nt_b)
    hipLaunchKernelGGL(k_gather, dim3(grid_for(c->nt_b)), blk, 0, b->stream, c->nt_b, c->tb, c->pb,
                       c->sb, c->gb, c->S, c->base_b, c->rhs, c->asm_nan);
  b->dispatches += 6;
}

void enqueue_limit(ohm_newton_ctx *c, const double *xo, double *xn) {
  ohm_backend *b = c->b;
  const dim3 blk(kBlock);
  if (c->nq)
    hipLaunchKernelGGL(k_bjt_limit, dim3(grid_for(c->nq)), blk, 0, b->stream, c->nq, c->qd, xo, xn,
                       c->serial_bjt ? 1 : 0, &c->word->limited);
  if (c->nm) {
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(c->n)), blk, 0, b->stream, c->n, kScaleOne,
                       c->scale);
    hipLaunchKernelGGL(k_mos_limit_reduce, dim3(grid_for(c->nm)), blk, 0, b->stream, c->nm, c->md,
                       xo, xn, c->scale);
    hipLaunchKernelGGL(k_mos_limit_apply, dim3(grid_for(c->n)), blk, 0, b->stream, c->n, c->scale,
                       xo, xn, &c->word->limited);
  }
  if (c->nd)
    hipLaunchKernelGGL(k_diode_limit, dim3(grid_for(c->nd)), blk, 0, b->stream, c->nd, c->dd, xo,
                       xn, c->serial_diode ? 1 : 0, &c->word->limited);
  b->dispatches += 5;
}

bool same_topology(const ohm_diode_desc &a, const ohm_diode_desc &b) {
  return a.anode_idx == b.anode_idx && a.cathode_idx == b.cathode_idx &&
         !std::memcmp(a.g_row_col, b.g_row_col, sizeof a.g_row_col) &&
         !std::memcmp(a.b_idx, b.b_idx, sizeof a.b_idx);
}
bool same_topology(const ohm_bjt_desc &a, const ohm_bjt_desc &b) {
  return a.collector_idx == b.collector_idx && a.base_idx == b.base_idx &&
         a.emitter_idx == b.emitter_idx &&
         !std::memcmp(a.node_fixed, b.node_fixed, sizeof a.node_fixed) &&
         !std::memcmp(a.g_row_col, b.g_row_col, sizeof a.g_row_col) &&
         !std::memcmp(a.b_idx, b.b_idx, sizeof a.b_idx);
}
bool same_topology(const ohm_mosfet_desc &a, const ohm_mosfet_desc &b) {
  return a.drain_idx == b.drain_idx && a.gate_idx == b.gate_idx &&
         a.source_idx == b.source_idx && a.node_fixed_packed == b.node_fixed_packed &&
         !std::memcmp(a.g_row_col, b.g_row_col, sizeof a.g_row_col) &&
         !std::memcmp(a.b_idx, b.b_idx, sizeof a.b_idx);
}
template <typename D>
ohm_status check_desc(const std::vector<D> &host, const D *in, const char *what) {
  if (!in) return OHM_OK;
  for (size_t k = 0; k < host.size(); k++)
    if (!same_topology(host[k], in[k]))
      return fail(OHM_ERR_INVALID,
                  "%s descriptor %zu changes nodes or stamp slots: create a new context", what, k);
  return OHM_OK;
}

// only after check_desc accepted every family, so a rejected update leaves
// the context as it was
template <typename D>
ohm_status replace_desc(std::vector<D> &host, D *dev, const D *in, hipStream_t st) {
  if (!in) return OHM_OK;
  host.assign(in, in + host.size());
  if (!host.empty())
    OHM_HIP(hipMemcpyAsync(dev, host.data(), sizeof(D) * host.size(), hipMemcpyHostToDevice, st));
  return OHM_OK;
}
}  // namespace

extern "C" {

void ohm_newton_default_params(ohm_newton_params *p) {
  if (!p) return;
  p->max_iterations = 50;
  p->abs_tol = 1e-4;
  p->bicgstab_max_iterations = 10000;
  p->bicgstab_tolerance = 1e-5;
  p->precision = OHM_PRECISION_FP64;
  p->precond = OHM_PREC_AUTO;
  p->initial_guess = nullptr;
  p->bicgstab_floor = 1e-15;
  p->forcing = OHM_FORCING_FIXED;
  p->forcing_safeguard = 0;
  p->forcing_max = 0.1;
  p->linear_solver = OHM_LINEAR_BICGSTAB;
}

void ohm_newton_free(ohm_newton_ctx *c) {
  if (c && c->adj && backend_alive(c->b)) (void)hipFree(c->adj);
  if (!c) return;
  if (backend_alive(c->b)) graph_cache_clear(c->b);
  csr_free(c->A);
  precond_free(c->P);
  void *ptrs[] = {c->base_val, c->base_b, c->rhs, c->x,  c->x_old, c->S,  c->dd, c->qd, c->md,
                  c->scale,    c->asm_nan, c->word, c->tm, c->pm,   c->sm, c->tb, c->pb, c->sb,
                  c->gm,       c->gb};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  delete c;
}

ohm_status ohm_newton_create(ohm_backend *b, const double *base_values,
                             const uint32_t *col_indices, const uint32_t *row_pointers,
                             const double *base_b, const ohm_diode_desc *diodes, size_t n_diodes,
                             const ohm_bjt_desc *bjts, size_t n_bjts,
                             const ohm_mosfet_desc *mosfets, size_t n_mosfets,
                             size_t system_size, size_t matrix_nnz, ohm_newton_ctx **out) {
  if (!b || !out || !row_pointers) return fail(OHM_ERR_INVALID, "null handle");
  *out = nullptr;
  const int n = (int)system_size;
  if (row_pointers[n] != matrix_nnz)
    return fail(OHM_ERR_SOLVE, "Solve error: dimension mismatch: row_pointers[n]=%u, nnz=%zu",
                row_pointers[n], matrix_nnz);
  for (size_t k = 0; k < matrix_nnz; k++)
    if (!std::isfinite(base_values[k]))
      return fail(OHM_ERR_NEWTON_NUMERICAL,
                  "Newton solver encountered numerical error at iteration 0");
  auto bad_idx = [&](uint32_t i) { return i != OHM_SENTINEL && i >= (uint32_t)n; };
  auto bad_slot = [&](uint32_t s) { return s != OHM_SENTINEL && s >= (uint32_t)matrix_nnz; };
  for (size_t k = 0; k < n_diodes; k++) {
    const ohm_diode_desc &d = diodes[k];
    bool bad = bad_idx(d.anode_idx) || bad_idx(d.cathode_idx) || bad_idx(d.b_idx[0]) ||
               bad_idx(d.b_idx[1]);
    for (int j = 0; j < 4; j++) bad |= bad_slot(d.g_row_col[j]);
    if (bad) return fail(OHM_ERR_INVALID, "diode descriptor %zu indexes outside the system", k);
  }
  for (size_t k = 0; k < n_bjts; k++) {
    const ohm_bjt_desc &d = bjts[k];
    bool bad = bad_idx(d.collector_idx) || bad_idx(d.base_idx) || bad_idx(d.emitter_idx);
    for (int j = 0; j < 3; j++) bad |= bad_idx(d.b_idx[j]);
    for (int j = 0; j < 9; j++) bad |= bad_slot(d.g_row_col[j]);
    if (bad) return fail(OHM_ERR_INVALID, "BJT descriptor %zu indexes outside the system", k);
  }
  for (size_t k = 0; k < n_mosfets; k++) {
    const ohm_mosfet_desc &d = mosfets[k];
    bool bad = bad_idx(d.drain_idx) || bad_idx(d.gate_idx) || bad_idx(d.source_idx);
    for (int j = 0; j < 2; j++) bad |= bad_idx(d.b_idx[j]);
    for (int j = 0; j < 6; j++) bad |= bad_slot(d.g_row_col[j]);
    if (bad) return fail(OHM_ERR_INVALID, "MOSFET descriptor %zu indexes outside the system", k);
  }
  ohm_newton_ctx *c = new ohm_newton_ctx();
  c->b = b;
  c->n = n;
  c->nnz = (int64_t)matrix_nnz;
  c->nd = (int)n_diodes;
  c->nq = (int)n_bjts;
  c->nm = (int)n_mosfets;
  auto cleanup = [&](ohm_status s) {
    ohm_newton_free(c);
    return s;
  };
  ohm_status s = csr_upload(b, c->A, base_values, col_indices, row_pointers, n, c->nnz);
  if (s != OHM_OK) return cleanup(s);
  size_t ns = (size_t)kDiodeS * c->nd + (size_t)kBjtS * c->nq + (size_t)kMosS * c->nm;
#define ALLOC(p, T, cnt)                                                      \
  do {                                                                        \
    hipError_t e_ = hipMalloc(&(p), sizeof(T) * std::max<size_t>((cnt), 1)); \
    if (e_ != hipSuccess) return cleanup(hip_fail(e_, "hipMalloc", __FILE__, __LINE__)); \
  } while (0)
  ALLOC(c->base_val, double, matrix_nnz);
  ALLOC(c->base_b, double, n);
  ALLOC(c->rhs, double, n);
  ALLOC(c->x, double, n);
  ALLOC(c->x_old, double, n);
  ALLOC(c->S, double, ns);
  ALLOC(c->dd, ohm_diode_desc, n_diodes);
  ALLOC(c->qd, ohm_bjt_desc, n_bjts);
  ALLOC(c->md, ohm_mosfet_desc, n_mosfets);
  ALLOC(c->scale, unsigned, n);
  ALLOC(c->asm_nan, int, 1);
  ALLOC(c->word, NewtonWord, 1);
#undef ALLOC
  hipStream_t st = b->stream;
  if (matrix_nnz)
    (void)hipMemcpyAsync(c->base_val, base_values, sizeof(double) * matrix_nnz,
                         hipMemcpyHostToDevice, st);
  if (n) (void)hipMemcpyAsync(c->base_b, base_b, sizeof(double) * n, hipMemcpyHostToDevice, st);
  if (n) (void)hipMemcpyAsync(c->rhs, base_b, sizeof(double) * n, hipMemcpyHostToDevice, st);
  if (n_diodes)
    (void)hipMemcpyAsync(c->dd, diodes, sizeof(ohm_diode_desc) * n_diodes, hipMemcpyHostToDevice, st);
  if (n_bjts)
    (void)hipMemcpyAsync(c->qd, bjts, sizeof(ohm_bjt_desc) * n_bjts, hipMemcpyHostToDevice, st);
  if (n_mosfets)
    (void)hipMemcpyAsync(c->md, mosfets, sizeof(ohm_mosfet_desc) * n_mosfets,
                         hipMemcpyHostToDevice, st);
  // gather lists (matrix slot / RHS entry <- stamp values, in device order)
  std::vector<Contrib> cm, cb;
  uint32_t so = 0;
  for (size_t k = 0; k < n_diodes; k++, so += kDiodeS) {
    const ohm_diode_desc &d = diodes[k];
    const double sg[4] = {1.0, 1.0, -1.0, -1.0};  // [aa, cc, ac, ca]
    for (int j = 0; j < 4; j++)
      if (d.g_row_col[j] != OHM_SENTINEL) cm.push_back({d.g_row_col[j], so, sg[j]});
    if (d.b_idx[0] != OHM_SENTINEL) cb.push_back({d.b_idx[0], so + 1, -1.0});
    if (d.b_idx[1] != OHM_SENTINEL) cb.push_back({d.b_idx[1], so + 1, 1.0});
  }
  for (size_t k = 0; k < n_bjts; k++, so += kBjtS) {
    const ohm_bjt_desc &d = bjts[k];
    for (int j = 0; j < 9; j++)
      if (d.g_row_col[j] != OHM_SENTINEL) cm.push_back({d.g_row_col[j], so + j, 1.0});
    for (int j = 0; j < 3; j++)
      if (d.b_idx[j] != OHM_SENTINEL) cb.push_back({d.b_idx[j], so + 9 + j, -1.0});
  }
  for (size_t k = 0; k < n_mosfets; k++, so += kMosS) {
    const ohm_mosfet_desc &d = mosfets[k];
    for (int j = 0; j < 6; j++)
      if (d.g_row_col[j] != OHM_SENTINEL) cm.push_back({d.g_row_col[j], so + j, 1.0});
    for (int j = 0; j < 2; j++)
      if (d.b_idx[j] != OHM_SENTINEL) cb.push_back({d.b_idx[j], so + 6 + j, -1.0});
  }
  s = build_gather(cm, c->nt_m, &c->tm, &c->pm, &c->sm, &c->gm);
  if (s == OHM_OK) s = build_gather(cb, c->nt_b, &c->tb, &c->pb, &c->sb, &c->gb);
  if (s != OHM_OK) return cleanup(s);
  c->serial_diode = shares_nodes((int)n_diodes, n, [&](int k, uint32_t *o) {
    o[0] = diodes[k].anode_idx;
    o[1] = diodes[k].cathode_idx;
    return 2;
  });
  c->serial_bjt = shares_nodes((int)n_bjts, n, [&](int k, uint32_t *o) {
    o[0] = bjts[k].collector_idx;
    o[1] = bjts[k].base_idx;
    o[2] = bjts[k].emitter_idx;
    return 3;
  });
  c->h_dd.assign(diodes, diodes + n_diodes);
  c->h_qd.assign(bjts, bjts + n_bjts);
  c->h_md.assign(mosfets, mosfets + n_mosfets);
  hipError_t e = hipStreamSynchronize(st);
  if (e != hipSuccess) return cleanup(hip_fail(e, "newton upload", __FILE__, __LINE__));
  *out = c;
  return OHM_OK;
}

ohm_status ohm_newton_update(ohm_newton_ctx *c, const double *base_values, const double *base_b,
                             const ohm_diode_desc *diodes, const ohm_bjt_desc *bjts,
                             const ohm_mosfet_desc *mosfets) {
  if (!c) return fail(OHM_ERR_INVALID, "null context");
  hipStream_t st = c->b->stream;
  if (base_values) {
    for (int64_t k = 0; k < c->nnz; k++)
      if (!std::isfinite(base_values[k]))
        return fail(OHM_ERR_NEWTON_NUMERICAL,
                    "Newton solver encountered numerical error at iteration 0");
  }
  OHM_TRY(check_desc(c->h_dd, diodes, "diode"));
  OHM_TRY(check_desc(c->h_qd, bjts, "BJT"));
  OHM_TRY(check_desc(c->h_md, mosfets, "MOSFET"));
  OHM_TRY(replace_desc(c->h_dd, c->dd, diodes, st));
  OHM_TRY(replace_desc(c->h_qd, c->qd, bjts, st));
  OHM_TRY(replace_desc(c->h_md, c->md, mosfets, st));
  if (base_values && c->nnz) {
    OHM_HIP(hipMemcpyAsync(c->base_val, base_values, sizeof(double) * c->nnz,
                           hipMemcpyHostToDevice, st));
    c->zero_diag = -1;
  }
  if (base_b && c->n)
    OHM_HIP(hipMemcpyAsync(c->base_b, base_b, sizeof(double) * c->n, hipMemcpyHostToDevice, st));
  // the host arrays are the caller's: finish the copies before returning
  OHM_HIP(hipStreamSynchronize(st));
  return OHM_OK;
}

static ohm_status newton_prepare_precond(ohm_newton_ctx *c, int kind) {
  if (c->prec_kind == kind) return OHM_OK;
  precond_free(c->P);
  c->prec_kind = -1;
  OHM_TRY(precond_build(c->b, c->A, kind, c->prec_level, c->P));
  c->prec_kind = kind;
  return OHM_OK;
}

ohm_status ohm_newton_run(ohm_newton_ctx *c, const ohm_newton_params *params,
                          ohm_newton_result *result, int32_t *iters_per_newton) {
  if (!c) return fail(OHM_ERR_INVALID, "null context");
  ohm_newton_params prm;
  if (params)
    prm = *params;
  else
    ohm_newton_default_params(&prm);
  ohm_backend *b = c->b;
  hipStream_t st = b->stream;
  const int n = c->n;
  if (prm.linear_solver != OHM_LINEAR_BICGSTAB && prm.linear_solver != OHM_LINEAR_DENSE_LU)
    return fail(OHM_ERR_INVALID, "unknown linear_solver %d", prm.linear_solver);
  if (prm.linear_solver == OHM_LINEAR_DENSE_LU && n > kDenseMaxRows)
    return fail(OHM_ERR_INVALID, "dense LU is limited to %d unknowns (system has %d)",
                kDenseMaxRows, n);
  if (prm.initial_guess)
    OHM_HIP(hipMemcpyAsync(c->x, prm.initial_guess, sizeof(double) * n, hipMemcpyHostToDevice, st));
  else
    OHM_HIP(hipMemsetAsync(c->x, 0, sizeof(double) * std::max(n, 1), st));
  // AUTO = the reference's rule (src/solver/newton.rs:340-361): ISAI(1) when
  // the assembled matrix has a zero diagonal, else Jacobi. Device stamps only
  // add to node diagonals, so the base matrix decides.
  int kind = prm.precond;
  // the assembled values and right-hand side start from base: the stamp
  // gathers rewrite only the positions devices touch, so every other entry
  // (voltage-source rows of b among them) must come from the current base,
  // which ohm_newton_update may have replaced since the last run
  OHM_HIP(hipMemcpyAsync(c->A.val, c->base_val, sizeof(double) * c->nnz, hipMemcpyDeviceToDevice, st));
  OHM_HIP(hipMemcpyAsync(c->rhs, c->base_b, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  if (kind == OHM_PREC_AUTO) {
    if (c->zero_diag < 0) c->zero_diag = matrix_has_zero_diag(b, c->A, nullptr) ? 1 : 0;
    kind = c->zero_diag ? OHM_PREC_ISAI : OHM_PREC_JACOBI;
  }
  OHM_TRY(newton_prepare_precond(c, kind));
  isai_reset(c->P);  // restart the ILU fixed point from the matrix for a fresh solve
  OHM_HIP(hipEventRecord(b->ev0, st));
  BicgOpts o;
  o.max_iters = prm.bicgstab_max_iterations;
  o.tol = prm.bicgstab_tolerance;
  o.warm = true;
  o.tol_mode = OHM_TOL_INITIAL;  // inexact Newton: relative to ||F(x_k)||
  o.tol_floor = prm.bicgstab_floor;
  o.retry_breakdown = true;
  NewtonWord *h = reinterpret_cast<NewtonWord *>(b->h_pinned);
  int64_t total_bicg = 0, fallbacks = 0;
  double max_diff = 0.0;
  ohm_status rc = OHM_OK;
  int64_t it = 0;
  bool converged = false;
  // Eisenstat-Walker forcing (choice 1 or 2): loose sub-solves while Newton
  // is far from the solution, tight ones as ||F|| falls. Each sub-solve
  // computes its forcing term on the device from its own starting residual
  // ||F(x_k)||; the host carries ||F_k-1||, the last linear residual and
  // eta_k-1 forward.
  Forcing fc;
  if (prm.forcing == OHM_FORCING_EW || prm.forcing == OHM_FORCING_EW1) {
    fc.kind = prm.forcing;
    fc.eta_min = prm.bicgstab_tolerance;
    fc.eta_max = prm.forcing_max > 0 ? std::max(prm.forcing_max, prm.bicgstab_tolerance) : 0.1;
    fc.rate_cap = prm.forcing_safeguard != 0;
  }
  for (it = 0; it < prm.max_iterations; it++) {
    hipLaunchKernelGGL(k_copy_d, dim3(grid_for(n, 4)), dim3(kBlock), 0, st, n, c->x, c->x_old);
    enqueue_eval_assemble(c, c->x);
    OHM_TRY(precond_refresh(b, c->A, c->P));
    o.tol = prm.bicgstab_tolerance;
    o.forcing = fc;
    BicgOut bo;
    ohm_status ls;
    if (prm.linear_solver == OHM_LINEAR_DENSE_LU) {
      // direct step: no forcing bookkeeping (r0 is not formed)
      bo.r0 = NAN;
      ls = dense_lu_solve(b, c->A, c->rhs, c->x);
    } else if (prm.precision == OHM_PRECISION_FP32_IR) {
      int64_t refs = 0;
      // FP32 refinement hands over to FP64 when kappa * 2^-24 >~ 1 (diodes
      // still off, mesh grounded only through GMIN)
      ls = bicg_solve_ir(b, c->A, c->P, c->rhs, c->x, o, 1e-4, 30, bo, refs);
      if (bo.fp64_fallback) fallbacks++;
    } else {
      ls = bicg_solve_f64(b, c->A, c->P, c->rhs, c->x, o, bo);
    }
    total_bicg += bo.iters;
    if (iters_per_newton) iters_per_newton[it] = (int32_t)bo.iters;
    if (fc.kind && std::isfinite(bo.r0)) {
      fc.f_prev = bo.r0;  // ||F(x_k)||
      fc.lin_prev = bo.res;
      fc.eta_prev = bo.eta;
    }
    // the assembled-value NaN flag (src/solver/newton.rs:258-263): k_converge
    // folds it into the status word read below; a failed sub-solve reads it
    // here, since a NaN in the assembly is then the error to report
    if (ls != OHM_OK) {
      OHM_HIP(hipMemcpyAsync(&h->nan, c->asm_nan, sizeof(int), hipMemcpyDeviceToHost, st));
      OHM_HIP(hipStreamSynchronize(st));
      rc = h->nan ? fail(OHM_ERR_NEWTON_NUMERICAL,
                         "Newton solver encountered numerical error at iteration %lld",
                         (long long)it)
                  : ls;
      break;
    }
    OHM_HIP(hipMemsetAsync(&c->word->limited, 0, sizeof(int), st));
    enqueue_limit(c, c->x_old, c->x);
    hipLaunchKernelGGL(k_converge, dim3(grid_for(n, 4)), dim3(kBlock), 0, st, n, c->x_old, c->x,
                       b->d_partials, b->d_counter, c->asm_nan, c->word);
    b->dispatches++;
    OHM_HIP(hipMemcpyAsync(h, c->word, sizeof(NewtonWord), hipMemcpyDeviceToHost, st));
    OHM_HIP(hipStreamSynchronize(st));
    b->readbacks++;
    max_diff = h->max_diff;
    // A limited step is not the Newton step the forcing theory assumes: the
    // residual change reflects the limiter, so the next sub-solve starts from
    // eta_max again instead of tightening on it.
    if (fc.kind && h->limited) fc.f_prev = -1.0;
    if (h->nan) {
      rc = fail(OHM_ERR_NEWTON_NUMERICAL, "Newton solver encountered numerical error at iteration %lld",
                (long long)it);
      break;
    }
    if (max_diff < prm.abs_tol) {
      converged = true;
      it++;
      break;
    }
  }
  OHM_HIP(hipEventRecord(b->ev1, st));
  OHM_HIP(hipEventSynchronize(b->ev1));
  if (result) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
    result->iterations = converged ? it : (rc == OHM_OK ? prm.max_iterations : it + 1);
    result->bicgstab_iterations = total_bicg;
    result->max_diff = max_diff;
    result->device_ms = ms;
    result->fp64_fallbacks = fallbacks;
  }
  if (rc != OHM_OK) return rc;
  if (!converged)
    return fail(OHM_ERR_NEWTON_NOT_CONVERGED,
                "Newton solver did not converge after %lld iterations (max residual: %.2e)",
                (long long)prm.max_iterations, max_diff);
  return OHM_OK;
}

ohm_status ohm_newton_get_x(ohm_newton_ctx *c, double *x_out) {
  if (!c || !x_out) return fail(OHM_ERR_INVALID, "null handle");
  OHM_HIP(hipMemcpyAsync(x_out, c->x, sizeof(double) * c->n, hipMemcpyDeviceToHost, c->b->stream));
  OHM_HIP(hipStreamSynchronize(c->b->stream));
  return OHM_OK;
}

ohm_status ohm_newton_get_system(ohm_newton_ctx *c, double *values_out, double *b_out) {
  if (!c) return fail(OHM_ERR_INVALID, "null handle");
  if (values_out)
    OHM_HIP(hipMemcpyAsync(values_out, c->A.val, sizeof(double) * c->nnz, hipMemcpyDeviceToHost,
                           c->b->stream));
  if (b_out)
    OHM_HIP(hipMemcpyAsync(b_out, c->rhs, sizeof(double) * c->n, hipMemcpyDeviceToHost,
                           c->b->stream));
  OHM_HIP(hipStreamSynchronize(c->b->stream));
  return OHM_OK;
}

ohm_status ohm_newton_assemble_at(ohm_newton_ctx *c, const double *x, double *values_out,
                                  double *b_out) {
  if (!c || !x) return fail(OHM_ERR_INVALID, "null handle");
  hipStream_t st = c->b->stream;
  OHM_HIP(hipMemcpyAsync(c->A.val, c->base_val, sizeof(double) * c->nnz, hipMemcpyDeviceToDevice, st));
  OHM_HIP(hipMemcpyAsync(c->rhs, c->base_b, sizeof(double) * c->n, hipMemcpyDeviceToDevice, st));
  OHM_HIP(hipMemcpyAsync(c->x_old, x, sizeof(double) * c->n, hipMemcpyHostToDevice, st));
  enqueue_eval_assemble(c, c->x_old);
  OHM_HIP(hipGetLastError());
  return ohm_newton_get_system(c, values_out, b_out);
}

ohm_status ohm_newton_limit(ohm_newton_ctx *c, const double *x_old, double *x_new) {
  if (!c || !x_old || !x_new) return fail(OHM_ERR_INVALID, "null handle");
  hipStream_t st = c->b->stream;
  OHM_HIP(hipMemcpyAsync(c->x_old, x_old, sizeof(double) * c->n, hipMemcpyHostToDevice, st));
  OHM_HIP(hipMemcpyAsync(c->x, x_new, sizeof(double) * c->n, hipMemcpyHostToDevice, st));
  enqueue_limit(c, c->x_old, c->x);
  OHM_HIP(hipGetLastError());
  OHM_HIP(hipMemcpyAsync(x_new, c->x, sizeof(double) * c->n, hipMemcpyDeviceToHost, st));
  OHM_HIP(hipStreamSynchronize(st));
  return OHM_OK;
}

ohm_status ohm_newton_solve(ohm_backend *b, const double *base_values,
                            const uint32_t *col_indices, const uint32_t *row_pointers,
                            const double *base_b, const ohm_diode_desc *diodes, size_t n_diodes,
                            const ohm_bjt_desc *bjts, size_t n_bjts, const ohm_mosfet_desc *mosfets,
                            size_t n_mosfets, size_t system_size, size_t matrix_nnz,
                            const ohm_newton_params *params, double *x_out,
                            ohm_newton_result *result, int32_t *iters_per_newton) {
  ohm_newton_ctx *c = nullptr;
  OHM_TRY(ohm_newton_create(b, base_values, col_indices, row_pointers, base_b, diodes, n_diodes,
                            bjts, n_bjts, mosfets, n_mosfets, system_size, matrix_nnz, &c));
  ohm_status s = ohm_newton_run(c, params, result, iters_per_newton);
  if (s == OHM_OK && x_out) s = ohm_newton_get_x(c, x_out);
  ohm_newton_free(c);
  return s;
}

ohm_status ohm_adjoint_dc(ohm_newton_ctx *c, const double *dJdx, const ohm_bicgstab_params *params,
                          int64_t n_res, const uint32_t *res_nodes, double *grad_g,
                          double *grad_kp, double *lambda_out, ohm_bicgstab_result *result) {
  if (!c || !dJdx) return fail(OHM_ERR_INVALID, "null argument");
  if (n_res < 0 || (n_res > 0 && (!res_nodes || !grad_g)))
    return fail(OHM_ERR_INVALID, "bad resistor parameter list");
  for (int64_t r = 0; r < 2 * n_res; r++)
    if (res_nodes[r] != OHM_SENTINEL && res_nodes[r] >= (uint32_t)c->n)
      return fail(OHM_ERR_INVALID, "resistor node index out of range");
  ohm_backend *b = c->b;
  hipStream_t st = b->stream;
  const int n = c->n;
  const size_t need = 2 * (size_t)n + (size_t)n_res + (size_t)c->nm + 2;
  if (need > c->adj_cap) {
    if (c->adj) OHM_HIP(hipFree(c->adj));
    OHM_HIP(hipMalloc(&c->adj, sizeof(double) * need));
    c->adj_cap = need;
  }
  double *rhs = c->adj, *lam = rhs + n, *gg = lam + n, *gk = gg + n_res;
  uint32_t *d_nodes = nullptr;
  if (n_res > 0) OHM_HIP(hipMalloc(&d_nodes, sizeof(uint32_t) * 2 * n_res));
  struct Guard {
    uint32_t *p;
    ~Guard() {
      if (p) (void)hipFree(p);
    }
  } guard{d_nodes};
  if (n_res > 0)
    OHM_HIP(hipMemcpyAsync(d_nodes, res_nodes, sizeof(uint32_t) * 2 * n_res, hipMemcpyHostToDevice, st));
  // dF/dx at the operating point in c->x
  OHM_HIP(hipMemcpyAsync(c->A.val, c->base_val, sizeof(double) * c->nnz, hipMemcpyDeviceToDevice, st));
  OHM_HIP(hipMemcpyAsync(c->rhs, c->base_b, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  enqueue_eval_assemble(c, c->x);
  if (c->prec_kind < 0) {
    int kind = matrix_has_zero_diag(b, c->A, nullptr) ? OHM_PREC_ISAI : OHM_PREC_JACOBI;
    OHM_TRY(newton_prepare_precond(c, kind));
  }
  OHM_TRY(precond_refresh(b, c->A, c->P));
  OHM_HIP(hipMemcpyAsync(rhs, dJdx, sizeof(double) * n, hipMemcpyHostToDevice, st));
  OHM_HIP(hipMemsetAsync(lam, 0, sizeof(double) * n, st));
  ohm_bicgstab_params prm;
  if (params)
    prm = *params;
  else {
    ohm_bicgstab_default_params(&prm);
    prm.tolerance = 1e-12;
  }
  BicgOpts o;
  o.max_iters = prm.max_iterations;
  o.tol = prm.tolerance;
  o.shadow = prm.shadow;
  o.check_interval = prm.check_interval;
  o.scaled_breakdown = prm.breakdown != OHM_BREAKDOWN_ABSOLUTE;
  BicgOut out;
  OHM_HIP(hipEventRecord(b->ev0, st));
  ohm_status s = bicg_solve_transpose(b, c->A, c->P, rhs, lam, o, out);
  if (s == OHM_OK) {
    if (n_res > 0)
      hipLaunchKernelGGL(k_adj_conductance, dim3(grid_for(n_res)), dim3(kBlock), 0, st, (int)n_res,
                         d_nodes, c->x, lam, gg);
    if (c->nm > 0)
      hipLaunchKernelGGL(k_adj_mosfet_kp, dim3(grid_for(c->nm)), dim3(kBlock), 0, st, c->nm, c->md,
                         c->x, lam, gk);
    b->dispatches += 2;
    if (n_res > 0)
      OHM_HIP(hipMemcpyAsync(grad_g, gg, sizeof(double) * n_res, hipMemcpyDeviceToHost, st));
    if (grad_kp && c->nm > 0)
      OHM_HIP(hipMemcpyAsync(grad_kp, gk, sizeof(double) * c->nm, hipMemcpyDeviceToHost, st));
    if (lambda_out)
      OHM_HIP(hipMemcpyAsync(lambda_out, lam, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  }
  OHM_HIP(hipEventRecord(b->ev1, st));
  OHM_HIP(hipEventSynchronize(b->ev1));
  if (result) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
    result->iterations = out.iters;
    result->refinements = 0;
    result->fp64_fallback = 0;
    result->device_ms = ms;
    result->dispatches = 0;
    result->readbacks = out.polls;
    result->residual_norm = NAN;
  }
  return s;
}

}  // extern "C"

// ohm_isai.hip — Jacobi and ISAI(level) preconditioners, set up on the GPU.
//
// The reference builds ISAI on the CPU every Newton iteration from a full
// pivoted LU with fill whose pivot search is O(n^2)
// (src/solver/preconditioner.rs:64-153,239-277; called from
// src/solver/newton.rs:369). Here the pattern work is done once on the host
// per sparsity pattern and the numeric work runs on the device:
//
//  1. ILU(0) on F = pattern(A) + diagonal by synchronous fixed-point sweeps
//     (Chow & Patel, "Fine-grained parallel incomplete LU factorization",
//     SIAM J. Sci. Comput. 37(2), 2015): every factor entry f_ij is updated
//     in parallel from  a_ij - sum_{k<min(i,j)} l_ik u_kj  (divided by u_jj
//     below the diagonal). The sweeps warm-start from the previous factors,
//     so after a Newton step changes only device slots they converge fast.
//  2. ISAI columns exactly as the reference defines them (M_L column k solves
//     L[J,J] m = e_k on J = pattern(L^(level+1)) column k, likewise M_U with
//     U), each local triangular solve done by level-synchronous substitution
//     sweeps over precomputed dependency lists, all columns in parallel.
//
// Apply is z = M r with the product M = M_U M_L formed on the device after
// every numeric refresh (pattern and the (M_U, M_L) entry pairs of each
// product entry found once on the host, products summed in a fixed order):
// one CSR-stream SpMV per apply instead of two dependent launches. When the
// product would be much denser than the two factors the apply stays
// z = M_U (M_L r).
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <unordered_map>

#include "ohm_precond.h"
#include "ohm_spmv.h"

namespace ohm {

struct IsaiPlan {
  int n = 0;
  int64_t nnz_f = 0;
  // factor pattern F (device)
  uint32_t *f_rowptr = nullptr, *f_col = nullptr;
  int32_t *f_from_a = nullptr;  // A position feeding each F entry, or -1
  int32_t *f_div = nullptr;     // F position of u_jj for strictly lower entries, else -1
  uint32_t *cp_ptr = nullptr;   // Chow-Patel pair list offsets per F entry
  uint32_t *cp_pairs = nullptr; // (l position, u position) pairs
  double *f_val = nullptr, *f_tmp = nullptr;
  bool f_init = false;
  // ISAI columns, one set per factor (0 = M_L, 1 = M_U)
  struct Side {
    int64_t m = 0;               // elements (= nnz of M)
    uint32_t *dep_ptr = nullptr; // per element: dependency offsets
    uint32_t *dep = nullptr;     // (factor position, element index) pairs
    int32_t *diag = nullptr;     // factor position of the pivot (M_U), or -1 (unit)
    uint8_t *is_k = nullptr;     // element is the column's own row (rhs = 1)
    uint32_t *out_pos = nullptr; // CSR position in M
    double *m0 = nullptr, *m1 = nullptr;
    int sweeps = 1;
  } side[2];
  // product M = M_U M_L: per product entry, (M_U position, M_L position) pairs
  uint32_t *prod_ptr = nullptr, *prod_pairs = nullptr;
  int64_t prod_nnz = 0;
  uint32_t *prod_long = nullptr;  // entries with more than kProdPairsPerThread pairs
  int64_t n_long = 0;
};

// A product entry whose pair list is longer than this (a dense row of M_U
// meeting a dense column of M_L, e.g. a supply node that touches every
// stage) is summed by a whole workgroup instead of one thread.
constexpr uint32_t kProdPairsPerThread = 64;

// m[e] = sum over the entry's pairs of mu[q] * ml[t], in pair order, for
// entries with at most kProdPairsPerThread pairs (one thread each).
__global__ void k_isai_product(int64_t nm, const uint32_t *__restrict__ ptr,
                               const uint32_t *__restrict__ pairs, const double *__restrict__ mu,
                               const double *__restrict__ ml, double *__restrict__ m,
                               float *__restrict__ m32) {
  for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < nm;
       e += (int64_t)gridDim.x * kBlock) {
    const uint32_t k0 = ptr[e], k1 = ptr[e + 1];
    if (k1 - k0 > kProdPairsPerThread) continue;
    double acc = 0.0;
    for (uint32_t k = k0; k < k1; k++) acc = fma(mu[pairs[2 * k]], ml[pairs[2 * k + 1]], acc);
    m[e] = acc;
    if (m32) m32[e] = (float)acc;
  }
}

// The long entries: one workgroup each, pairs strided over the threads and
// combined by the fixed-order workgroup sum.
__global__ __launch_bounds__(kBlock) void k_isai_product_long(
    const uint32_t *__restrict__ longs, const uint32_t *__restrict__ ptr,
    const uint32_t *__restrict__ pairs, const double *__restrict__ mu,
    const double *__restrict__ ml, double *__restrict__ m, float *__restrict__ m32) {
  const uint32_t e = longs[blockIdx.x];
  double acc[1] = {0.0};
  for (uint32_t k = ptr[e] + threadIdx.x; k < ptr[e + 1]; k += kBlock)
    acc[0] = fma(mu[pairs[2 * k]], ml[pairs[2 * k + 1]], acc[0]);
  block_sum<1>(acc);
  if (threadIdx.x == 0) {
    m[e] = acc[0];
    if (m32) m32[e] = (float)acc[0];
  }
}

// ------------------------------------------------------------- kernels ----
__global__ void k_ilu_init(int64_t nf, const int32_t *__restrict__ from_a,
                           const int32_t *__restrict__ div, const double *__restrict__ aval,
                           double *__restrict__ f) {
  for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < nf;
       e += (int64_t)gridDim.x * kBlock) {
    int32_t s = from_a[e];
    double v = s >= 0 ? aval[s] : 0.0;
    int32_t d = div[e];
    if (d >= 0) {
      int32_t ad = from_a[d];
      double piv = ad >= 0 ? aval[ad] : 0.0;
      v = fabs(piv) < 1e-30 ? 0.0 : v / piv;
    }
    f[e] = v;
  }
}

__global__ void k_ilu_sweep(int64_t nf, const int32_t *__restrict__ from_a,
                            const int32_t *__restrict__ div, const uint32_t *__restrict__ ptr,
                            const uint32_t *__restrict__ pairs, const double *__restrict__ aval,
                            const double *__restrict__ fin, double *__restrict__ fout) {
  for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < nf;
       e += (int64_t)gridDim.x * kBlock) {
    int32_t s = from_a[e];
    double v = s >= 0 ? aval[s] : 0.0;
    for (uint32_t q = ptr[e]; q < ptr[e + 1]; q++) v -= fin[pairs[2 * q]] * fin[pairs[2 * q + 1]];
    int32_t d = div[e];
    if (d >= 0) {
      double piv = fin[d];
      v = fabs(piv) < 1e-30 ? 0.0 : v / piv;
    }
    fout[e] = v;
  }
}

// One substitution sweep for every ISAI element of one side:
// m_new[e] = (delta - sum f[pos] * m_old[src]) / pivot.
__global__ void k_isai_sweep(int64_t m, const uint32_t *__restrict__ ptr,
                             const uint32_t *__restrict__ dep, const int32_t *__restrict__ diag,
                             const uint8_t *__restrict__ is_k, const double *__restrict__ f,
                             const double *__restrict__ min, double *__restrict__ mout) {
  for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < m;
       e += (int64_t)gridDim.x * kBlock) {
    double v = is_k[e] ? 1.0 : 0.0;
    for (uint32_t q = ptr[e]; q < ptr[e + 1]; q++) v -= f[dep[2 * q]] * min[dep[2 * q + 1]];
    int32_t d = diag[e];
    if (d >= 0) {
      double piv = f[d];
      v = fabs(piv) < 1e-30 ? 0.0 : v / piv;
    }
    mout[e] = v;
  }
}

// Small patterns: every sweep of one side in a single 1024-thread
// workgroup, a barrier between sweeps, so a pattern whose dependency depth
// asks for hundreds of sweeps (a long inverter chain) costs one launch, not
// hundreds. Same arithmetic per sweep as the kernels above; the result ends
// in the buffer the sweep-per-launch loop would leave it in.
constexpr int kSweepWg = 1024;
// Entries up to which one workgroup sweeps: about 8 per thread, where a
// sweep inside the workgroup costs less than a launch; beyond that the
// sweep-per-launch grid is faster (an 8,194-unknown amplifier array's
// factors took 16 ms per operating point in one workgroup against 9 ms).
constexpr int64_t kSweepWgMax = 8192;

__global__ __launch_bounds__(kSweepWg) void k_ilu_sweeps_wg(
    int64_t nf, const int32_t *__restrict__ from_a, const int32_t *__restrict__ div,
    const uint32_t *__restrict__ ptr, const uint32_t *__restrict__ pairs,
    const double *__restrict__ aval, double *f0, double *f1, int sweeps) {
  double *fin = f0, *fout = f1;
  for (int s = 0; s < sweeps; s++) {
    for (int64_t e = threadIdx.x; e < nf; e += kSweepWg) {
      int32_t src = from_a[e];
      double v = src >= 0 ? aval[src] : 0.0;
      for (uint32_t q = ptr[e]; q < ptr[e + 1]; q++) v -= fin[pairs[2 * q]] * fin[pairs[2 * q + 1]];
      int32_t d = div[e];
      if (d >= 0) {
        double piv = fin[d];
        v = fabs(piv) < 1e-30 ? 0.0 : v / piv;
      }
      fout[e] = v;
    }
    __syncthreads();
    double *t = fin;
    fin = fout;
    fout = t;
  }
}

__global__ __launch_bounds__(kSweepWg) void k_isai_sweeps_wg(
    int64_t m, const uint32_t *__restrict__ ptr, const uint32_t *__restrict__ dep,
    const int32_t *__restrict__ diag, const uint8_t *__restrict__ is_k,
    const double *__restrict__ f, double *m0, double *m1, int sweeps) {
  double *min = m0, *mout = m1;
  for (int s = 0; s < sweeps; s++) {
    for (int64_t e = threadIdx.x; e < m; e += kSweepWg) {
      double v = is_k[e] ? 1.0 : 0.0;
      for (uint32_t q = ptr[e]; q < ptr[e + 1]; q++) v -= f[dep[2 * q]] * min[dep[2 * q + 1]];
      int32_t d = diag[e];
      if (d >= 0) {
        double piv = f[d];
        v = fabs(piv) < 1e-30 ? 0.0 : v / piv;
      }
      mout[e] = v;
    }
    __syncthreads();
    double *t = min;
    min = mout;
    mout = t;
  }
}

__global__ void k_scatter(int64_t m, const uint32_t *__restrict__ pos, const double *__restrict__ src,
                          double *__restrict__ dst, float *__restrict__ dst32) {
  for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < m;
       e += (int64_t)gridDim.x * kBlock) {
    double v = src[e];
    dst[pos[e]] = v;
    if (dst32) dst32[pos[e]] = (float)v;
  }
}

__global__ void k_zero_diag_flag(int n, const int32_t *__restrict__ diag_pos,
                                 const double *__restrict__ val, int *flag) {
  for (int r = blockIdx.x * kBlock + threadIdx.x; r < n; r += gridDim.x * kBlock) {
    int d = diag_pos[r];
    if (d < 0 || fabs(val[d]) < 1e-30) atomicOr(flag, 1);
  }
}

__global__ void k_mul_f32(int n, const float *__restrict__ d, const float *__restrict__ x,
                          float *__restrict__ y, const SolveState *st, int skip) {
  if (solve_stopped(st, skip)) return;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) y[i] = d[i] * x[i];
}
__global__ void k_mul_f64(int n, const double *__restrict__ d, const double *__restrict__ x,
                          double *__restrict__ y, const SolveState *st, int skip) {
  if (solve_stopped(st, skip)) return;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) y[i] = d[i] * x[i];
}
template <typename T>
__global__ void k_copy_t(int n, const T *__restrict__ x, T *__restrict__ y, const SolveState *st,
                         int skip) {
  if (solve_stopped(st, skip)) return;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) y[i] = x[i];
}

// ------------------------------------------------------------ symbolic ----
template <typename V>
static ohm_status upload(V **dst, const std::vector<V> &src) {
  size_t cnt = std::max<size_t>(src.size(), 1);
  OHM_HIP(hipMalloc(dst, sizeof(V) * cnt));
  if (!src.empty()) OHM_HIP(hipMemcpy(*dst, src.data(), sizeof(V) * src.size(), hipMemcpyHostToDevice));
  return OHM_OK;
}

// Row-sorted pattern product pattern(P*P) (src/solver/preconditioner.rs:185-201).
// Sparsity pattern of the product A B (row-sorted column lists).
static void pattern_mul(int n, const std::vector<uint32_t> &ra, const std::vector<uint32_t> &ca,
                        const std::vector<uint32_t> &rb, const std::vector<uint32_t> &cb,
                        std::vector<uint32_t> &orp, std::vector<uint32_t> &ocl) {
  orp.assign(n + 1, 0);
  ocl.clear();
  std::vector<int> mark(n, -1);
  for (int i = 0; i < n; i++) {
    size_t start = ocl.size();
    for (uint32_t a = ra[i]; a < ra[i + 1]; a++) {
      uint32_t k = ca[a];
      for (uint32_t b = rb[k]; b < rb[k + 1]; b++) {
        uint32_t j = cb[b];
        if (mark[j] != i) {
          mark[j] = i;
          ocl.push_back(j);
        }
      }
    }
    std::sort(ocl.begin() + start, ocl.end());
    orp[i + 1] = (uint32_t)ocl.size();
  }
}

static ohm_status build_side(int n, int level, bool lower, const std::vector<uint32_t> &frp,
                             const std::vector<uint32_t> &fcl, const std::vector<int32_t> &fdiag,
                             IsaiPlan::Side &sd, DevCsr &M, ohm_backend *b) {
  // factor pattern P (L: strict lower + diag; U: diag + strict upper), row-sorted
  std::vector<uint32_t> prp(n + 1, 0), pcl, ppos;  // ppos: F position (diag of L -> unit)
  for (int r = 0; r < n; r++) {
    for (uint32_t q = frp[r]; q < frp[r + 1]; q++) {
      uint32_t c = fcl[q];
      bool keep = lower ? (c <= (uint32_t)r) : (c >= (uint32_t)r);
      if (keep) {
        pcl.push_back(c);
        ppos.push_back(q);
      }
    }
    prp[r + 1] = (uint32_t)pcl.size();
  }
  std::vector<uint32_t> mrp = prp, mcl = pcl;  // pattern of M = pattern(P^(level+1))
  for (int l = 0; l < level; l++) {
    std::vector<uint32_t> trp, tcl;
    pattern_mul(n, mrp, mcl, prp, pcl, trp, tcl);
    mrp.swap(trp);
    mcl.swap(tcl);
  }
  // column view of M: for each column k, rows J ascending
  std::vector<uint32_t> ccp(n + 1, 0), crow(mcl.size());
  for (uint32_t c : mcl) ccp[c + 1]++;
  for (int i = 0; i < n; i++) ccp[i + 1] += ccp[i];
  {
    std::vector<uint32_t> pos(ccp.begin(), ccp.end() - 1);
    for (int r = 0; r < n; r++)
      for (uint32_t q = mrp[r]; q < mrp[r + 1]; q++) crow[pos[mcl[q]]++] = (uint32_t)r;
  }
  // element e = ccp[k] + t  <->  (row J[t], col k); CSR position of (r, k) in M
  const int64_t m = (int64_t)mcl.size();
  std::vector<uint32_t> out_pos(m), dep_ptr(m + 1, 0), dep;
  std::vector<int32_t> diag(m, -1);
  std::vector<uint8_t> is_k(m, 0);
  std::vector<int> depth(m, 0);
  std::vector<int64_t> local(n, -1);  // row -> element index within the current column
  int max_depth = 0;
  for (int k = 0; k < n; k++) {
    const uint32_t c0 = ccp[k], c1 = ccp[k + 1];
    for (uint32_t e = c0; e < c1; e++) local[crow[e]] = e;
    // process in substitution order: ascending rows for L, descending for U
    for (uint32_t q = 0; q < c1 - c0; q++) {
      uint32_t e = lower ? c0 + q : c1 - 1 - q;
      uint32_t r = crow[e];
      // CSR position of (r, k) in M
      uint32_t lo = mrp[r], hi = mrp[r + 1];
      out_pos[e] = (uint32_t)(std::lower_bound(mcl.begin() + lo, mcl.begin() + hi, (uint32_t)k) -
                              mcl.begin());
      is_k[e] = (r == (uint32_t)k);
      int d = 0;
      for (uint32_t p = prp[r]; p < prp[r + 1]; p++) {
        uint32_t cc = pcl[p];
        if (cc == r) {
          if (!lower) diag[e] = (int32_t)ppos[p];
          continue;
        }
        int64_t se = local[cc];
        if (se < 0) continue;
        dep.push_back(ppos[p]);
        dep.push_back((uint32_t)se);
        d = std::max(d, depth[se] + 1);
      }
      dep_ptr[e + 1] = (uint32_t)(dep.size() / 2);
      depth[e] = d;
      max_depth = std::max(max_depth, d);
    }
    for (uint32_t e = c0; e < c1; e++) local[crow[e]] = -1;
  }
  sd.m = m;
  sd.sweeps = max_depth + 1;
  // dependencies were appended in substitution order (descending rows for U);
  // regroup them into element order
  {
    std::vector<uint32_t> start(m), len(m);
    size_t cursor = 0;
    for (int k = 0; k < n; k++) {
      const uint32_t c0 = ccp[k], c1 = ccp[k + 1];
      for (uint32_t t = 0; t < c1 - c0; t++) {
        uint32_t e = lower ? c0 + t : c1 - 1 - t;
        uint32_t end = dep_ptr[e + 1];
        start[e] = (uint32_t)cursor;
        len[e] = end - (uint32_t)cursor;
        cursor = end;
      }
    }
    std::vector<uint32_t> eptr(m + 1, 0), edep;
    edep.reserve(dep.size());
    for (int64_t e = 0; e < m; e++) {
      for (uint32_t j = 0; j < len[e]; j++) {
        edep.push_back(dep[2 * (start[e] + j)]);
        edep.push_back(dep[2 * (start[e] + j) + 1]);
      }
      eptr[e + 1] = (uint32_t)(edep.size() / 2);
    }
    OHM_TRY(upload(&sd.dep_ptr, eptr));
    OHM_TRY(upload(&sd.dep, edep));
  }
  OHM_TRY(upload(&sd.diag, diag));
  OHM_TRY(upload(&sd.is_k, is_k));
  OHM_TRY(upload(&sd.out_pos, out_pos));
  OHM_HIP(hipMalloc(&sd.m0, sizeof(double) * std::max<int64_t>(m, 1)));
  OHM_HIP(hipMalloc(&sd.m1, sizeof(double) * std::max<int64_t>(m, 1)));
  // M as a CSR matrix for the apply SpMV
  OHM_TRY(csr_upload(b, M, nullptr, mcl.data(), mrp.data(), n, m));
  return OHM_OK;
}

// Pattern of M_U M_L and its entry pairs; skipped (two-SpMV apply) when the
// product would hold more than 3x the nonzeros of the two factors.
static ohm_status build_product(ohm_backend *b, PrecondDev &p) {
  IsaiPlan *pl = p.plan;
  const int n = p.n;
  const DevCsr &U = p.mu, &L = p.ml;
  if (U.h_rowptr.empty() || L.h_rowptr.empty()) return OHM_OK;
  std::vector<uint32_t> rp(n + 1, 0), cl, ptr(1, 0), pairs;
  std::vector<int64_t> mark(n, -1);
  std::vector<std::vector<std::pair<uint32_t, uint32_t>>> acc;
  std::vector<uint32_t> cols;
  const int64_t cap = 3 * (U.nnz + L.nnz);
  for (int i = 0; i < n; i++) {
    cols.clear();
    for (uint32_t q = U.h_rowptr[i]; q < U.h_rowptr[i + 1]; q++) {
      const uint32_t k = U.h_col[q];
      for (uint32_t t = L.h_rowptr[k]; t < L.h_rowptr[k + 1]; t++) {
        const uint32_t j = L.h_col[t];
        if (mark[j] < 0) {
          mark[j] = (int64_t)acc.size();
          acc.emplace_back();
          cols.push_back(j);
        }
        acc[mark[j]].push_back({q, t});
      }
    }
    std::sort(cols.begin(), cols.end());
    for (uint32_t j : cols) {
      cl.push_back(j);
      for (auto &pr : acc[mark[j]]) {
        pairs.push_back(pr.first);
        pairs.push_back(pr.second);
      }
      ptr.push_back((uint32_t)(pairs.size() / 2));
      mark[j] = -1;
    }
    acc.clear();
    rp[i + 1] = (uint32_t)cl.size();
    if ((int64_t)cl.size() > cap) return OHM_OK;  // too dense: keep the two-SpMV apply
  }
  pl->prod_nnz = (int64_t)cl.size();
  std::vector<uint32_t> longs;
  for (size_t e = 0; e + 1 < ptr.size(); e++)
    if (ptr[e + 1] - ptr[e] > kProdPairsPerThread) longs.push_back((uint32_t)e);
  pl->n_long = (int64_t)longs.size();
  OHM_TRY(csr_upload(b, p.mm, nullptr, cl.data(), rp.data(), n, pl->prod_nnz));
  OHM_TRY(upload(&pl->prod_ptr, ptr));
  OHM_TRY(upload(&pl->prod_pairs, pairs));
  if (!longs.empty()) OHM_TRY(upload(&pl->prod_long, longs));
  return OHM_OK;
}

static ohm_status isai_plan_build(ohm_backend *b, const DevCsr &a, int level, PrecondDev &p) {
  const int n = a.n;
  IsaiPlan *pl = new IsaiPlan();
  p.plan = pl;
  pl->n = n;
  std::vector<uint32_t> frp(n + 1, 0), fcl;
  std::vector<int32_t> from_a, fdiag(n, -1);
  for (int r = 0; r < n; r++) {
    bool have = false;
    for (uint32_t k = a.h_rowptr[r]; k < a.h_rowptr[r + 1]; k++) {
      uint32_t c = a.h_col[k];
      if (!have && c > (uint32_t)r) {
        fdiag[r] = (int32_t)fcl.size();
        fcl.push_back((uint32_t)r);
        from_a.push_back(-1);
        have = true;
      }
      if (c == (uint32_t)r) {
        have = true;
        fdiag[r] = (int32_t)fcl.size();
      }
      fcl.push_back(c);
      from_a.push_back((int32_t)k);
    }
    if (!have) {
      fdiag[r] = (int32_t)fcl.size();
      fcl.push_back((uint32_t)r);
      from_a.push_back(-1);
    }
    frp[r + 1] = (uint32_t)fcl.size();
  }
  const int64_t nf = (int64_t)fcl.size();
  pl->nnz_f = nf;
  // Chow-Patel pairs: for entry (i, j), all k < min(i, j) with (i,k) and (k,j) in F
  std::vector<int32_t> div(nf, -1);
  std::vector<uint32_t> cptr(nf + 1, 0), pairs;
  for (int i = 0; i < n; i++) {
    for (uint32_t e = frp[i]; e < frp[i + 1]; e++) {
      uint32_t j = fcl[e];
      uint32_t mn = std::min<uint32_t>(i, j);
      for (uint32_t q = frp[i]; q < frp[i + 1] && fcl[q] < mn; q++) {
        uint32_t k = fcl[q];
        auto it = std::lower_bound(fcl.begin() + frp[k], fcl.begin() + frp[k + 1], j);
        if (it != fcl.begin() + frp[k + 1] && *it == j) {
          pairs.push_back(q);
          pairs.push_back((uint32_t)(it - fcl.begin()));
        }
      }
      cptr[e + 1] = (uint32_t)(pairs.size() / 2);
      if (j < (uint32_t)i) div[e] = fdiag[j];
    }
  }
  OHM_TRY(upload(&pl->f_rowptr, frp));
  OHM_TRY(upload(&pl->f_col, fcl));
  OHM_TRY(upload(&pl->f_from_a, from_a));
  OHM_TRY(upload(&pl->f_div, div));
  OHM_TRY(upload(&pl->cp_ptr, cptr));
  OHM_TRY(upload(&pl->cp_pairs, pairs));
  OHM_HIP(hipMalloc(&pl->f_val, sizeof(double) * std::max<int64_t>(nf, 1)));
  OHM_HIP(hipMalloc(&pl->f_tmp, sizeof(double) * std::max<int64_t>(nf, 1)));
  OHM_TRY(build_side(n, level, true, frp, fcl, fdiag, pl->side[0], p.ml, b));
  OHM_TRY(build_side(n, level, false, frp, fcl, fdiag, pl->side[1], p.mu, b));
  // The apply is two SpMVs (M_L, then M_U), the first carrying BiCGSTAB's
  // fused vector update. OHM_ISAI_PRODUCT=1 forms M = M_U M_L instead (one
  // SpMV per apply): fewer launches, but on stencil patterns M has ~1.6x the
  // nonzeros of the two factors and the fused update gathers three vectors
  // per nonzero, so the two-SpMV form is faster there.
  const char *env = getenv("OHM_ISAI_PRODUCT");
  if (env && atoi(env) != 0) OHM_TRY(build_product(b, p));
  return OHM_OK;
}

static ohm_status isai_numeric(ohm_backend *b, const DevCsr &a, PrecondDev &p) {
  IsaiPlan *pl = p.plan;
  const int64_t nf = pl->nnz_f;
  const dim3 gf(grid_for(nf, 4)), blk(kBlock);
  if (!pl->f_init) {
    hipLaunchKernelGGL(k_ilu_init, gf, blk, 0, b->stream, nf, pl->f_from_a, pl->f_div, a.val,
                       pl->f_val);
    b->dispatches++;
    pl->f_init = true;
  }
  if (nf <= kSweepWgMax && p.ilu_sweeps > 1) {
    hipLaunchKernelGGL(k_ilu_sweeps_wg, dim3(1), dim3(kSweepWg), 0, b->stream, nf, pl->f_from_a,
                       pl->f_div, pl->cp_ptr, pl->cp_pairs, a.val, pl->f_val, pl->f_tmp,
                       p.ilu_sweeps);
    if (p.ilu_sweeps & 1) std::swap(pl->f_val, pl->f_tmp);
    b->dispatches++;
  } else {
    for (int s = 0; s < p.ilu_sweeps; s++) {
      hipLaunchKernelGGL(k_ilu_sweep, gf, blk, 0, b->stream, nf, pl->f_from_a, pl->f_div,
                         pl->cp_ptr, pl->cp_pairs, a.val, pl->f_val, pl->f_tmp);
      std::swap(pl->f_val, pl->f_tmp);
      b->dispatches++;
    }
  }
  for (int sdx = 0; sdx < 2; sdx++) {
    IsaiPlan::Side &sd = pl->side[sdx];
    DevCsr &M = sdx == 0 ? p.ml : p.mu;
    if (sd.m == 0) continue;
    const dim3 gm(grid_for(sd.m, 4));
    double *src = sd.m0, *dst = sd.m1;
    OHM_HIP(hipMemsetAsync(src, 0, sizeof(double) * sd.m, b->stream));
    if (sd.m <= kSweepWgMax && sd.sweeps > 1) {
      hipLaunchKernelGGL(k_isai_sweeps_wg, dim3(1), dim3(kSweepWg), 0, b->stream, sd.m, sd.dep_ptr,
                         sd.dep, sd.diag, sd.is_k, pl->f_val, src, dst, sd.sweeps);
      if (sd.sweeps & 1) std::swap(src, dst);
      b->dispatches++;
    } else {
      for (int s = 0; s < sd.sweeps; s++) {
        hipLaunchKernelGGL(k_isai_sweep, gm, blk, 0, b->stream, sd.m, sd.dep_ptr, sd.dep, sd.diag,
                           sd.is_k, pl->f_val, src, dst);
        std::swap(src, dst);
        b->dispatches++;
      }
    }
    hipLaunchKernelGGL(k_scatter, gm, blk, 0, b->stream, sd.m, sd.out_pos, src, M.val,
                       p.has_f32 ? M.val32 : nullptr);
    b->dispatches++;
  }
  if (p.mm.rowptr && pl->prod_nnz > 0) {
    hipLaunchKernelGGL(k_isai_product, dim3(grid_for(pl->prod_nnz)), blk, 0, b->stream,
                       pl->prod_nnz, pl->prod_ptr, pl->prod_pairs, p.mu.val, p.ml.val, p.mm.val,
                       p.has_f32 ? p.mm.val32 : nullptr);
    b->dispatches++;
    if (pl->n_long > 0) {
      hipLaunchKernelGGL(k_isai_product_long, dim3((unsigned)pl->n_long), blk, 0, b->stream,
                         pl->prod_long, pl->prod_ptr, pl->prod_pairs, p.mu.val, p.ml.val, p.mm.val,
                         p.has_f32 ? p.mm.val32 : nullptr);
      b->dispatches++;
    }
  }
  OHM_HIP(hipGetLastError());
  return OHM_OK;
}

bool matrix_has_zero_diag(ohm_backend *b, const DevCsr &a, const double *) {
  int *d_flag = nullptr;
  if (hipMalloc(&d_flag, sizeof(int)) != hipSuccess) return true;
  (void)hipMemsetAsync(d_flag, 0, sizeof(int), b->stream);
  hipLaunchKernelGGL(k_zero_diag_flag, dim3(grid_for(a.n)), dim3(kBlock), 0, b->stream, a.n,
                     a.diag_pos, a.val, d_flag);
  int h = 1;
  (void)hipMemcpyAsync(&h, d_flag, sizeof(int), hipMemcpyDeviceToHost, b->stream);
  (void)hipStreamSynchronize(b->stream);
  (void)hipFree(d_flag);
  b->readbacks++;
  return h != 0;
}

// ------------------------------------------------------------- public ----
ohm_status precond_build(ohm_backend *b, const DevCsr &a, int kind, int level, PrecondDev &p) {
  p = PrecondDev();
  p.n = a.n;
  p.level = level;
  if (kind == OHM_PREC_AUTO) kind = matrix_has_zero_diag(b, a, nullptr) ? OHM_PREC_ISAI : OHM_PREC_JACOBI;
  if (kind == OHM_PREC_JACOBI) {
    p.kind = PK_JACOBI;
    OHM_HIP(hipMalloc(&p.inv, sizeof(double) * std::max(a.n, 1)));
  } else if (kind == OHM_PREC_ISAI) {
    if (level < 0 || level > 1) return fail(OHM_ERR_INVALID, "ISAI level must be 0 or 1");
    p.kind = PK_ISAI;
    OHM_TRY(isai_plan_build(b, a, level, p));
  } else if (kind == OHM_PREC_NONE) {
    p.kind = PK_NONE;
  } else {
    return fail(OHM_ERR_INVALID, "unknown preconditioner kind %d", kind);
  }
  OHM_TRY(precond_refresh(b, a, p));
  if (b->autotune && p.kind == PK_ISAI) {
    if (p.mm.rowptr) {
      OHM_TRY(csr_autotune(b, p.mm, nullptr, nullptr, true));
    } else {
      OHM_TRY(csr_autotune(b, p.ml, nullptr, nullptr, true));
      OHM_TRY(csr_autotune(b, p.mu, nullptr, nullptr, true));
    }
  }
  return OHM_OK;
}

ohm_status precond_refresh(ohm_backend *b, const DevCsr &a, PrecondDev &p) {
  if (p.kind == PK_JACOBI) {
    launch_jacobi_setup(b, a, p.inv, p.has_f32 ? p.inv32 : nullptr);
  } else if (p.kind == PK_ISAI) {
    OHM_TRY(isai_numeric(b, a, p));
  }
  OHM_HIP(hipGetLastError());
  return OHM_OK;
}

__global__ void k_d2f_vec(int64_t n, const double *__restrict__ x, float *__restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    y[i] = (float)x[i];
}

ohm_status precond_ensure_f32(ohm_backend *b, PrecondDev &p) {
  if (p.has_f32) return OHM_OK;
  if (p.kind == PK_JACOBI) {
    OHM_HIP(hipMalloc(&p.inv32, sizeof(float) * std::max(p.n, 1)));
    hipLaunchKernelGGL(k_d2f_vec, dim3(grid_for(p.n, 4)), dim3(kBlock), 0, b->stream,
                       (int64_t)p.n, p.inv, p.inv32);
  } else if (p.kind == PK_ISAI) {
    OHM_TRY(csr_ensure_f32(b, p.ml));
    OHM_TRY(csr_ensure_f32(b, p.mu));
    if (p.mm.rowptr) OHM_TRY(csr_ensure_f32(b, p.mm));
  }
  p.has_f32 = true;
  return OHM_OK;
}

void precond_apply(ohm_backend *b, const PrecondDev &p, const double *in, double *out, double *tmp,
                   const SolveState *st, int skip) {
  const int n = p.n;
  if (p.kind == PK_JACOBI) {
    hipLaunchKernelGGL(k_mul_f64, dim3(grid_for(n, 4)), dim3(kBlock), 0, b->stream, n, p.inv, in,
                       out, st, skip);
    b->dispatches++;
  } else if (p.kind == PK_ISAI && p.mm.rowptr) {
    launch_spmv(b, p.mm, in, out, st, skip);
  } else if (p.kind == PK_ISAI) {
    launch_spmv(b, p.ml, in, tmp, st, skip);
    launch_spmv(b, p.mu, tmp, out, st, skip);
  } else {
    hipLaunchKernelGGL(k_copy_t<double>, dim3(grid_for(n, 4)), dim3(kBlock), 0, b->stream, n, in,
                       out, st, skip);
    b->dispatches++;
  }
}

void precond_apply_f32(ohm_backend *b, const PrecondDev &p, const float *in, float *out,
                       float *tmp, const SolveState *st, int skip) {
  const int n = p.n;
  if (p.kind == PK_JACOBI) {
    hipLaunchKernelGGL(k_mul_f32, dim3(grid_for(n, 4)), dim3(kBlock), 0, b->stream, n, p.inv32, in,
                       out, st, skip);
    b->dispatches++;
  } else if (p.kind == PK_ISAI && p.mm.rowptr) {
    launch_spmv_f32(b, p.mm, in, out, st, skip);
  } else if (p.kind == PK_ISAI) {
    launch_spmv_f32(b, p.ml, in, tmp, st, skip);
    launch_spmv_f32(b, p.mu, tmp, out, st, skip);
  } else {
    hipLaunchKernelGGL(k_copy_t<float>, dim3(grid_for(n, 4)), dim3(kBlock), 0, b->stream, n, in,
                       out, st, skip);
    b->dispatches++;
  }
}

void isai_reset(PrecondDev &p) {
  if (p.plan) p.plan->f_init = false;
}

void precond_free(PrecondDev &p) {
  if (p.inv) (void)hipFree(p.inv);
  if (p.inv32) (void)hipFree(p.inv32);
  csr_free(p.ml);
  csr_free(p.mu);
  csr_free(p.mm);
  if (p.plan) {
    IsaiPlan *pl = p.plan;
    void *ptrs[] = {pl->f_rowptr, pl->f_col, pl->f_from_a, pl->f_div, pl->cp_ptr,
                    pl->cp_pairs, pl->f_val, pl->f_tmp, pl->prod_ptr, pl->prod_pairs,
                    pl->prod_long};
    for (void *q : ptrs)
      if (q) (void)hipFree(q);
    for (auto &sd : pl->side) {
      void *qs[] = {sd.dep_ptr, sd.dep, sd.diag, sd.is_k, sd.out_pos, sd.m0, sd.m1};
      for (void *q : qs)
        if (q) (void)hipFree(q);
    }
    delete pl;
  }
  p = PrecondDev();
}

}  // namespace ohm

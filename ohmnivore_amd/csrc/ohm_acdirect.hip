// ohm_acdirect.hip — batched banded direct solve of AC systems
//
//   (G_c + j 2 pi f C_c) x = b_ac,c     for every circuit c and frequency f
//
// Long chain-like circuits (the RLC ladders of the AC workload) have
// condition numbers that keep Jacobi-BiCGSTAB from converging at low
// frequencies even at 1,000 sections; after a reverse Cuthill-McKee
// ordering their MNA matrix is banded with half-bandwidth 1. Every system
// is then solved exactly by a banded LU with partial pivoting (LAPACK
// xGBTRF/xGBTRS semantics), one system per lane:
//   * the active (B+1) x (2B+1) window of the elimination lives in
//     registers, the band rows of A = G + j omega C are formed on the fly
//     from the circuit's shared G and C band arrays, and the right-hand
//     side is eliminated alongside, so L is never stored;
//   * the rows of U (2B+1 entries with fill) go to HBM interleaved by lane,
//     so each wave's stores and loads are coalesced; back substitution
//     reads them in reverse and overwrites the eliminated right-hand side
//     with x;
//   * a residual pass recomputes ||b - A x|| / ||b|| per system in FP64;
//   * an unpermute pass scatters x back to the circuit's original order.
// Waves hold 64 systems sorted by size, so lanes of a wave run for similar
// lengths.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "ohm_internal.h"

namespace ohm {

struct cz {
  double re, im;
};
__device__ __forceinline__ cz operator+(cz a, cz b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cz operator-(cz a, cz b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cz operator*(cz a, cz b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ double cabs2(cz a) { return a.re * a.re + a.im * a.im; }
__device__ __forceinline__ cz cinv(cz a) {
  const double d = cabs2(a);
  return {a.re / d, -a.im / d};
}

// One wave = up to 64 frequency points of one circuit (lane l: frequency
// f0 + l). The circuit's band rows are shared by every lane, so the wave
// loads them cooperatively, 64 rows at a time, into LDS.
struct BandWave {
  int64_t n;
  const double *gb, *cb;   // band rows of the permuted G and C: [n][2B+1]
  const double *b;         // permuted right-hand side, (re, im) pairs
  const uint32_t *perm;    // permuted row i holds original unknown perm[i]
  const double *omega;     // omega of lane 0 (consecutive lanes follow)
  int64_t store_off;       // complex offset of this wave's U / Y region (per lane-row)
  int64_t out_off;         // complex offset of lane 0's x in the output (lanes n apart)
  int nf;                  // active lanes
  int sys;                 // result slot of lane 0
};

constexpr int kBandTile = 64;

template <int W>
__device__ __forceinline__ void band_tile_load(const BandWave &q, int64_t row0, double (&rg)[W],
                                               double (&rc)[W], double (&rb)[2]) {
  const int64_t i = row0 + threadIdx.x;
  if (i < q.n) {
#pragma unroll
    for (int d = 0; d < W; d++) {
      rg[d] = q.gb[i * W + d];
      rc[d] = q.cb[i * W + d];
    }
    rb[0] = q.b[2 * i];
    rb[1] = q.b[2 * i + 1];
  } else {
#pragma unroll
    for (int d = 0; d < W; d++) rg[d] = rc[d] = 0.0;
    rb[0] = rb[1] = 0.0;
  }
}

// forward elimination + back substitution for the lanes of one wave
template <int B>
__global__ __launch_bounds__(64) void k_band_lu(const BandWave *__restrict__ waves,
                                                cz *__restrict__ U, cz *__restrict__ Y,
                                                int *__restrict__ singular) {
  constexpr int W = 2 * B + 1;  // band width in, U row width with fill
  __shared__ double sg[kBandTile][W], sc[kBandTile][W], sb[kBandTile][2];
  const int lane = threadIdx.x;
  const BandWave q = waves[blockIdx.x];
  const bool live = lane < q.nf;
  const double om = live ? q.omega[lane] : 0.0;
  const int64_t n = q.n;
  cz *Ub = U + q.store_off * W;
  cz *Yb = Y + q.store_off;
  cz win[B + 1][W], y[B + 1];
#pragma unroll
  for (int r = 0; r <= B; r++) {
    y[r] = r < n ? cz{q.b[2 * r], q.b[2 * r + 1]} : cz{0.0, 0.0};
#pragma unroll
    for (int c = 0; c < W; c++) win[r][c] = cz{0.0, 0.0};
    if (r < n) {
#pragma unroll
      for (int d = 0; d < W; d++) {
        const int col = r - B + d;  // window column == global column at j = 0
        if (col >= 0 && col < W && col < n)
          win[r][col] = cz{q.gb[r * W + d], om * q.cb[r * W + d]};
      }
    }
  }
  // rows entering the window: B+1, B+2, ...; tile k holds rows B+1+64k ..
  double rg[W], rc[W], rb[2];
  band_tile_load<W>(q, B + 1, rg, rc, rb);
  int bad = 0;
  for (int64_t t0 = 0; t0 < n; t0 += kBandTile) {
    __syncthreads();
#pragma unroll
    for (int d = 0; d < W; d++) {
      sg[lane][d] = rg[d];
      sc[lane][d] = rc[d];
    }
    sb[lane][0] = rb[0];
    sb[lane][1] = rb[1];
    __syncthreads();
    band_tile_load<W>(q, B + 1 + t0 + kBandTile, rg, rc, rb);  // next tile in flight
    const int64_t t1 = t0 + kBandTile < n ? t0 + kBandTile : n;
    for (int64_t j = t0; j < t1; j++) {
      int p = 0;
      double best = cabs2(win[0][0]);
#pragma unroll
      for (int r = 1; r <= B; r++) {
        const double a = cabs2(win[r][0]);
        if (j + r < n && a > best) {
          best = a;
          p = r;
        }
      }
      if (!(best > 0.0)) bad = 1;
#pragma unroll
      for (int r = 1; r <= B; r++)
        if (p == r) {
#pragma unroll
          for (int c = 0; c < W; c++) {
            const cz t = win[0][c];
            win[0][c] = win[r][c];
            win[r][c] = t;
          }
          const cz t = y[0];
          y[0] = y[r];
          y[r] = t;
        }
      const cz piv_inv = cinv(win[0][0]);
#pragma unroll
      for (int r = 1; r <= B; r++) {
        const cz l = win[r][0] * piv_inv;
#pragma unroll
        for (int c = 1; c < W; c++) win[r][c] = win[r][c] - l * win[0][c];
        y[r] = y[r] - l * y[0];
      }
      if (live) {
#pragma unroll
        for (int c = 0; c < W; c++) Ub[(j * W + c) * 64 + lane] = win[0][c];
        Yb[j * 64 + lane] = y[0];
      }
#pragma unroll
      for (int r = 0; r < B; r++) {
#pragma unroll
        for (int c = 0; c + 1 < W; c++) win[r][c] = win[r + 1][c + 1];
        win[r][W - 1] = cz{0.0, 0.0};
        y[r] = y[r + 1];
      }
      const int64_t i = j + 1 + B;
      const int s = (int)(j - t0);  // row i sits in tile slot s
      if (i < n) {
#pragma unroll
        for (int d = 0; d < W; d++) win[B][d] = cz{sg[s][d], om * sc[s][d]};
        y[B] = cz{sb[s][0], sb[s][1]};
      } else {
#pragma unroll
        for (int c = 0; c < W; c++) win[B][c] = cz{0.0, 0.0};
        y[B] = cz{0.0, 0.0};
      }
    }
  }
  if (!live) return;
  // back substitution, x overwrites Y; rows fetched 8 at a time
  constexpr int G = 8;
  cz xw[W - 1];
#pragma unroll
  for (int k = 0; k < W - 1; k++) xw[k] = cz{0.0, 0.0};
  for (int64_t hi = n - 1; hi >= 0; hi -= G) {
    cz u[G][W], yy[G];
#pragma unroll
    for (int g = 0; g < G; g++) {
      const int64_t j = hi - g;
      if (j >= 0) {
#pragma unroll
        for (int c = 0; c < W; c++) u[g][c] = Ub[(j * W + c) * 64 + lane];
        yy[g] = Yb[j * 64 + lane];
      }
    }
#pragma unroll
    for (int g = 0; g < G; g++) {
      const int64_t j = hi - g;
      if (j >= 0) {
        cz acc = yy[g];
#pragma unroll
        for (int c = 1; c < W; c++) acc = acc - u[g][c] * xw[c - 1];
        const cz xj = acc * cinv(u[g][0]);
        Yb[j * 64 + lane] = xj;
#pragma unroll
        for (int k = W - 2; k > 0; k--) xw[k] = xw[k - 1];
        xw[0] = xj;
      }
    }
  }
  if (bad) singular[q.sys + lane] = 1;
}

// Per-row residual terms, summed per (system, row chunk): fully parallel
// over rows. partial[(wave * 64 + lane) * kResChunks + chunk] = (rr, bb).
constexpr int kResChunks = 64;

template <int B>
__global__ __launch_bounds__(256) void k_band_resid(const BandWave *__restrict__ waves,
                                                    const cz *__restrict__ Y,
                                                    double *__restrict__ partial) {
  constexpr int W = 2 * B + 1;
  const int lane = threadIdx.x & 63, slot = threadIdx.x >> 6;
  const BandWave q = waves[blockIdx.y];
  const bool live = lane < q.nf;
  const double om = live ? q.omega[lane] : 0.0;
  const cz *Yb = Y + q.store_off;
  double rr = 0.0, bb = 0.0;
  if (live)
    for (int64_t i = (int64_t)blockIdx.x * 4 + slot; i < q.n; i += (int64_t)gridDim.x * 4) {
      cz acc = {q.b[2 * i], q.b[2 * i + 1]};
      bb += cabs2(acc);
#pragma unroll
      for (int d = 0; d < W; d++) {
        const int64_t col = i - B + d;
        if (col >= 0 && col < q.n) {
          const cz a = {q.gb[i * W + d], om * q.cb[i * W + d]};
          acc = acc - a * Yb[col * 64 + lane];
        }
      }
      rr += cabs2(acc);
    }
  __shared__ double s_rr[4][64], s_bb[4][64];
  s_rr[slot][lane] = rr;
  s_bb[slot][lane] = bb;
  __syncthreads();
  if (slot == 0) {
    double a = 0.0, c = 0.0;
    for (int k = 0; k < 4; k++) {
      a += s_rr[k][lane];
      c += s_bb[k][lane];
    }
    const size_t o = ((size_t)blockIdx.y * 64 + lane) * kResChunks + blockIdx.x;
    partial[2 * o] = a;
    partial[2 * o + 1] = c;
  }
}

__global__ void k_band_resid_sum(const BandWave *__restrict__ waves, int n_waves,
                                 const double *__restrict__ partial, double *__restrict__ res) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_waves * 64) return;
  const BandWave &q = waves[t / 64];
  const int lane = t % 64;
  if (lane >= q.nf) return;
  double a = 0.0, c = 0.0;
  for (int k = 0; k < kResChunks; k++) {
    a += partial[2 * ((size_t)t * kResChunks + k)];
    c += partial[2 * ((size_t)t * kResChunks + k) + 1];
  }
  res[2 * (q.sys + lane)] = a;
  res[2 * (q.sys + lane) + 1] = c;
}

// x back to the original unknown order: out[out_off + lane n + perm[i]] = x_i
__global__ void k_band_unpermute(const BandWave *__restrict__ waves, const cz *__restrict__ Y,
                                 cz *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  const BandWave q = waves[blockIdx.y];
  if (lane >= q.nf) return;
  const cz *Yb = Y + q.store_off;
  const int rows_per = blockDim.x / 64;
  for (int64_t i = blockIdx.x * rows_per + (threadIdx.x >> 6); i < q.n;
       i += (int64_t)gridDim.x * rows_per)
    out[q.out_off + (int64_t)lane * q.n + q.perm[i]] = Yb[i * 64 + lane];
}

}  // namespace ohm

using namespace ohm;

namespace {

// Reverse Cuthill-McKee on pattern(A + A^T) (George & Liu, "Computer
// Solution of Large Sparse Positive Definite Systems", 1981): BFS from a
// pseudo-peripheral node of each component, neighbours by ascending degree,
// order reversed. Returns perm (new position -> old row).
std::vector<uint32_t> rcm(int64_t n, const std::vector<uint32_t> &rp,
                          const std::vector<uint32_t> &cl) {
  std::vector<std::vector<uint32_t>> adj(n);
  for (int64_t r = 0; r < n; r++)
    for (uint32_t k = rp[r]; k < rp[r + 1]; k++)
      if ((int64_t)cl[k] != r) {
        adj[r].push_back(cl[k]);
        adj[cl[k]].push_back((uint32_t)r);
      }
  for (auto &a : adj) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  std::vector<char> seen(n, 0);
  std::vector<uint32_t> order;
  order.reserve(n);
  std::vector<int64_t> lvl(n, -1);
  auto bfs_far = [&](uint32_t s, int64_t &depth) {
    std::vector<uint32_t> q{s};
    lvl[s] = 0;
    for (size_t h = 0; h < q.size(); h++)
      for (uint32_t v : adj[q[h]])
        if (lvl[v] < 0) {
          lvl[v] = lvl[q[h]] + 1;
          q.push_back(v);
        }
    depth = lvl[q.back()];
    uint32_t best = q.back();
    for (uint32_t u : q)
      if (lvl[u] == depth && adj[u].size() < adj[best].size()) best = u;
    for (uint32_t u : q) lvl[u] = -1;
    return best;
  };
  std::vector<uint32_t> by_deg(n);
  std::iota(by_deg.begin(), by_deg.end(), 0u);
  std::stable_sort(by_deg.begin(), by_deg.end(),
                   [&](uint32_t a, uint32_t b) { return adj[a].size() < adj[b].size(); });
  for (uint32_t s0 : by_deg) {
    if (seen[s0]) continue;
    uint32_t s = s0;
    int64_t depth = -1;
    for (int it = 0; it < 6; it++) {
      int64_t d;
      uint32_t far = bfs_far(s, d);
      if (d <= depth) break;
      depth = d;
      s = far;
    }
    size_t head = order.size();
    order.push_back(s);
    seen[s] = 1;
    std::vector<uint32_t> nb;
    while (head < order.size()) {
      const uint32_t u = order[head++];
      nb.clear();
      for (uint32_t v : adj[u])
        if (!seen[v]) nb.push_back(v);
      std::sort(nb.begin(), nb.end(), [&](uint32_t a, uint32_t b) {
        return adj[a].size() != adj[b].size() ? adj[a].size() < adj[b].size() : a < b;
      });
      for (uint32_t v : nb) {
        seen[v] = 1;
        order.push_back(v);
      }
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

// f(0) .. f(n-1) on up to 16 host threads (independent items; f must not
// throw).
template <typename F>
void parallel_for(int64_t n, F f) {
  const int64_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const int64_t T = std::min<int64_t>(n, hw);
  if (T <= 1) {
    for (int64_t i = 0; i < n; i++) f(i);
    return;
  }
  std::atomic<int64_t> next{0};
  std::vector<std::thread> pool;
  for (int64_t t = 0; t < T; t++)
    pool.emplace_back([&] {
      for (int64_t i; (i = next.fetch_add(1)) < n;) f(i);
    });
  for (auto &th : pool) th.join();
}

struct Prepared {
  int64_t n = 0;
  int band = -1;  // half bandwidth after RCM (max of kl, ku)
  std::vector<uint32_t> rp, cl, perm, iperm;
  std::vector<double> g, c;  // union-pattern values
};

}  // namespace

extern "C" {

ohm_status ohm_ac_sweep_circuits(ohm_backend *b, int64_t n_circuits, const ohm_ac_circuit *circ,
                                 int64_t n_freqs, const double *freqs, int32_t method,
                                 const ohm_bicgstab_params *params, int32_t precond,
                                 ohm_batch_result *results, double *device_ms) {
  if (!b || (n_circuits > 0 && !circ) || (n_freqs > 0 && !freqs) || !results)
    return fail(OHM_ERR_INVALID, "null argument");
  if (n_circuits < 0 || n_freqs < 0) return fail(OHM_ERR_INVALID, "negative size");
  if (method < OHM_AC_AUTO || method > OHM_AC_DIRECT) return fail(OHM_ERR_INVALID, "unknown AC method %d", method);
  OHM_HIP(hipSetDevice(b->device));
  hipStream_t st = b->stream;
  // symbolic preparation per circuit (pattern union, RCM, bandwidth), on
  // host threads; inputs are checked first so errors are deterministic
  std::vector<Prepared> prep(n_circuits);
  int band = 0;
  std::vector<char> direct(n_circuits, 0);
  for (int64_t c = 0; c < n_circuits; c++) {
    const ohm_ac_circuit &q = circ[c];
    if (q.n < 0 || q.n > 0x7fffffff) return fail(OHM_ERR_INVALID, "circuit %lld: bad size", (long long)c);
    if (q.n == 0) continue;
    if (!q.g_rowptr || !q.g_val || !q.b_ac || (q.c_val && !q.c_rowptr))
      return fail(OHM_ERR_INVALID, "circuit %lld: null arrays", (long long)c);
  }
  std::vector<ohm_status> pst(n_circuits, OHM_OK);
  std::vector<std::string> pmsg(n_circuits);
  parallel_for(n_circuits, [&](int64_t c) {
    const ohm_ac_circuit &q = circ[c];
    Prepared &p = prep[c];
    p.n = q.n;
    if (q.n == 0) return;
    pst[c] = csr_union(q.n, q.g_rowptr, q.g_col, q.g_val, q.c_val ? q.c_rowptr : nullptr, q.c_col,
                       q.c_val, p.rp, p.cl, p.g, p.c);
    if (pst[c] != OHM_OK) {
      pmsg[c] = ohm_last_error();  // this thread's message
      return;
    }
    p.perm = rcm(q.n, p.rp, p.cl);
    p.iperm.assign(q.n, 0);
    for (int64_t i = 0; i < q.n; i++) p.iperm[p.perm[i]] = (uint32_t)i;
    int64_t bw = 0;
    for (int64_t r = 0; r < q.n; r++)
      for (uint32_t k = p.rp[r]; k < p.rp[r + 1]; k++)
        bw = std::max<int64_t>(bw, std::llabs((int64_t)p.iperm[r] - (int64_t)p.iperm[p.cl[k]]));
    p.band = (int)std::min<int64_t>(bw, 1 << 20);
  });
  for (int64_t c = 0; c < n_circuits; c++) {
    if (pst[c] != OHM_OK) {
      set_error(pmsg[c]);
      return pst[c];
    }
    if (circ[c].n == 0) continue;
    const Prepared &p = prep[c];
    direct[c] = method == OHM_AC_ITERATIVE ? 0 : (p.band <= 4);
    if (method == OHM_AC_DIRECT && !direct[c])
      return fail(OHM_ERR_INVALID,
                  "circuit %lld: half bandwidth %d after RCM exceeds the direct solver's 4",
                  (long long)c, p.band);
    if (direct[c]) band = std::max(band, p.band);
  }
  ohm_status first = OHM_OK;
  std::string first_msg;
  auto note = [&](ohm_status s) {
    if (s != OHM_OK && first == OHM_OK) {
      first = s;
      first_msg = ohm_last_error();
    }
  };
  // iterative circuits: the batched BiCGSTAB sweep
  for (int64_t c = 0; c < n_circuits; c++) {
    if (direct[c] || circ[c].n == 0) continue;
    const ohm_ac_circuit &q = circ[c];
    std::vector<double> scratch;
    double *x = q.x;
    if (!x) {
      scratch.assign((size_t)n_freqs * 2 * q.n, 0.0);
      x = scratch.data();
    }
    ohm_status s = ohm_ac_sweep(b, q.n, q.g_rowptr, q.g_col, q.g_val, q.c_rowptr, q.c_col,
                                q.c_val, q.b_ac, n_freqs, freqs, 0, params, precond, x,
                                results + c * n_freqs);
    note(s);
    if (s != OHM_OK && s != OHM_ERR_SOLVE) return s;
  }
  const int B = band <= 1 ? 1 : band <= 2 ? 2 : 4;
  const int W = 2 * B + 1;
  // device band data per direct circuit
  std::vector<void *> allocs;
  struct Free {
    std::vector<void *> &v;
    ~Free() {
      for (void *p : v) (void)hipFree(p);
    }
  } guard{allocs};
  auto dmal = [&](void **p, size_t bytes) -> ohm_status {
    OHM_HIP(hipMalloc(p, std::max<size_t>(bytes, 16)));
    allocs.push_back(*p);
    return OHM_OK;
  };
  std::vector<const double *> d_gb(n_circuits), d_cb(n_circuits), d_b(n_circuits);
  std::vector<const uint32_t *> d_perm(n_circuits);
  // band arrays in the RCM order, built on host threads 32 circuits at a
  // time and uploaded from this thread
  std::vector<int64_t> dlist;
  for (int64_t c = 0; c < n_circuits; c++)
    if (direct[c] && circ[c].n > 0) dlist.push_back(c);
  struct BandHost {
    std::vector<double> gb, cb, bp;
  };
  for (size_t c0 = 0; c0 < dlist.size(); c0 += 32) {
    const size_t cn = std::min<size_t>(32, dlist.size() - c0);
    std::vector<BandHost> hb(cn);
    parallel_for((int64_t)cn, [&](int64_t t) {
      const int64_t c = dlist[c0 + t];
      const Prepared &p = prep[c];
      const int64_t n = p.n;
      BandHost &h = hb[t];
      h.gb.assign((size_t)n * W, 0.0);
      h.cb.assign((size_t)n * W, 0.0);
      h.bp.assign(2 * n, 0.0);
      for (int64_t r = 0; r < n; r++) {
        const int64_t i = p.iperm[r];
        for (uint32_t k = p.rp[r]; k < p.rp[r + 1]; k++) {
          const int64_t d = (int64_t)p.iperm[p.cl[k]] - i + B;
          h.gb[i * W + d] += p.g[k];
          h.cb[i * W + d] += p.c[k];
        }
        h.bp[2 * i] = circ[c].b_ac[2 * r];
        h.bp[2 * i + 1] = circ[c].b_ac[2 * r + 1];
      }
    });
    for (size_t t = 0; t < cn; t++) {
      const int64_t c = dlist[c0 + t];
      const int64_t n = prep[c].n;
      const BandHost &h = hb[t];
      void *pg, *pc, *pb, *pp;
      OHM_TRY(dmal(&pg, sizeof(double) * h.gb.size()));
      OHM_TRY(dmal(&pc, sizeof(double) * h.cb.size()));
      OHM_TRY(dmal(&pb, sizeof(double) * h.bp.size()));
      OHM_TRY(dmal(&pp, sizeof(uint32_t) * n));
      OHM_HIP(hipMemcpy(pg, h.gb.data(), sizeof(double) * h.gb.size(), hipMemcpyHostToDevice));
      OHM_HIP(hipMemcpy(pc, h.cb.data(), sizeof(double) * h.cb.size(), hipMemcpyHostToDevice));
      OHM_HIP(hipMemcpy(pb, h.bp.data(), sizeof(double) * h.bp.size(), hipMemcpyHostToDevice));
      OHM_HIP(hipMemcpy(pp, prep[c].perm.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
      d_gb[c] = (const double *)pg;
      d_cb[c] = (const double *)pc;
      d_b[c] = (const double *)pb;
      d_perm[c] = (const uint32_t *)pp;
    }
  }
  // waves: up to 64 frequencies of one circuit; longest circuits first
  struct WaveItem {
    int64_t c, f0;
    int nf;
  };
  std::vector<WaveItem> items;
  for (int64_t c = 0; c < n_circuits; c++)
    if (direct[c] && circ[c].n > 0)
      for (int64_t f0 = 0; f0 < n_freqs; f0 += 64)
        items.push_back({c, f0, (int)std::min<int64_t>(64, n_freqs - f0)});
  std::stable_sort(items.begin(), items.end(), [&](const WaveItem &a, const WaveItem &b2) {
    return circ[a.c].n > circ[b2.c].n;
  });
  std::vector<double> omega(std::max<int64_t>(n_freqs, 1));
  for (int64_t f = 0; f < n_freqs; f++) omega[f] = 2.0 * M_PI * freqs[f];  // ac.rs:57
  double *d_omega = nullptr;
  OHM_TRY(dmal((void **)&d_omega, sizeof(double) * omega.size()));
  OHM_HIP(hipMemcpy(d_omega, omega.data(), sizeof(double) * omega.size(), hipMemcpyHostToDevice));
  // chunks bounded by a device-memory budget for U, Y and the output
  const char *env = getenv("OHM_AC_DIRECT_BUDGET_GB");
  const double budget = (env ? atof(env) : 64.0) * 1e9;
  std::vector<int> singular_h(std::max<int64_t>(n_circuits * n_freqs, 1), 0);
  std::vector<double> res_h(2 * std::max<int64_t>(n_circuits * n_freqs, 1), 0.0);
  int *d_sing = nullptr;
  double *d_res = nullptr;
  OHM_TRY(dmal((void **)&d_sing, sizeof(int) * singular_h.size()));
  OHM_TRY(dmal((void **)&d_res, sizeof(double) * res_h.size()));
  OHM_HIP(hipMemsetAsync(d_sing, 0, sizeof(int) * singular_h.size(), st));
  OHM_HIP(hipMemsetAsync(d_res, 0, sizeof(double) * res_h.size(), st));
  float kernel_ms = 0.f;
  size_t pos = 0;
  bool want_x = false;
  for (int64_t c = 0; c < n_circuits; c++) want_x |= direct[c] && circ[c].x != nullptr;
  while (pos < items.size()) {
    size_t end = pos;
    double bytes = 0.0;
    while (end < items.size()) {
      const double n = (double)circ[items[end].c].n;
      const double wave_b = n * 64.0 * 16.0 * (W + 1) + n * items[end].nf * 16.0 +
                            64.0 * kResChunks * 16.0;
      if (end > pos && bytes + wave_b > budget) break;
      bytes += wave_b;
      end++;
    }
    const int n_w = (int)(end - pos);
    // output in (circuit, frequency) order so each circuit's block is contiguous
    std::vector<size_t> by_cf(n_w);
    std::iota(by_cf.begin(), by_cf.end(), 0);
    std::stable_sort(by_cf.begin(), by_cf.end(), [&](size_t a, size_t b2) {
      const WaveItem &x = items[pos + a], &y = items[pos + b2];
      return x.c != y.c ? x.c < y.c : x.f0 < y.f0;
    });
    std::vector<int64_t> out_off(n_w);
    std::vector<int64_t> f_lo(n_circuits, INT64_MAX), f_hi(n_circuits, -1), x_at(n_circuits, 0);
    int64_t out_total = 0;
    for (size_t k : by_cf) {
      const WaveItem &it = items[pos + k];
      out_off[k] = out_total;
      if (it.f0 < f_lo[it.c]) {
        f_lo[it.c] = it.f0;
        x_at[it.c] = out_total;
      }
      f_hi[it.c] = std::max<int64_t>(f_hi[it.c], it.f0 + it.nf - 1);
      out_total += circ[it.c].n * it.nf;
    }
    std::vector<BandWave> hw(n_w);
    int64_t store = 0;
    for (int k = 0; k < n_w; k++) {
      const WaveItem &it = items[pos + k];
      BandWave w{};
      w.n = circ[it.c].n;
      w.gb = d_gb[it.c];
      w.cb = d_cb[it.c];
      w.b = d_b[it.c];
      w.perm = d_perm[it.c];
      w.omega = d_omega + it.f0;
      w.store_off = store * 64;
      w.out_off = out_off[k];
      w.nf = it.nf;
      w.sys = (int)(it.c * n_freqs + it.f0);
      store += w.n;
      hw[k] = w;
    }
    // chunk buffers carved from the backend's workspace (kept across sweeps:
    // re-allocating tens of GB per chunk costs seconds)
    auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
    const size_t bw_b = al(sizeof(BandWave) * n_w), u_b = al(sizeof(cz) * store * 64 * W),
                 y_b = al(sizeof(cz) * store * 64),
                 x_b = want_x ? al(sizeof(cz) * std::max<int64_t>(out_total, 1)) : 0,
                 p_b = al(sizeof(double) * 2 * (size_t)n_w * 64 * kResChunks);
    const size_t need = bw_b + u_b + y_b + x_b + p_b;
    if (b->ac_ws_bytes < need) {
      if (b->ac_ws) OHM_HIP(hipFree(b->ac_ws));
      b->ac_ws = nullptr;
      b->ac_ws_bytes = 0;
      OHM_HIP(hipMalloc(&b->ac_ws, need));
      b->ac_ws_bytes = need;
    }
    char *wsp = (char *)b->ac_ws;
    BandWave *d_w = (BandWave *)wsp;
    cz *dU = (cz *)(wsp + bw_b);
    cz *dY = (cz *)(wsp + bw_b + u_b);
    cz *dX = want_x ? (cz *)(wsp + bw_b + u_b + y_b) : nullptr;
    double *d_part = (double *)(wsp + bw_b + u_b + y_b + x_b);
    OHM_HIP(hipMemcpyAsync(d_w, hw.data(), sizeof(BandWave) * n_w, hipMemcpyHostToDevice, st));
    OHM_HIP(hipEventRecord(b->ev0, st));
    const dim3 rg(kResChunks, n_w);
    if (B == 1) {
      hipLaunchKernelGGL(k_band_lu<1>, dim3(n_w), dim3(64), 0, st, d_w, dU, dY, d_sing);
      hipLaunchKernelGGL(k_band_resid<1>, rg, dim3(256), 0, st, d_w, dY, d_part);
    } else if (B == 2) {
      hipLaunchKernelGGL(k_band_lu<2>, dim3(n_w), dim3(64), 0, st, d_w, dU, dY, d_sing);
      hipLaunchKernelGGL(k_band_resid<2>, rg, dim3(256), 0, st, d_w, dY, d_part);
    } else {
      hipLaunchKernelGGL(k_band_lu<4>, dim3(n_w), dim3(64), 0, st, d_w, dU, dY, d_sing);
      hipLaunchKernelGGL(k_band_resid<4>, rg, dim3(256), 0, st, d_w, dY, d_part);
    }
    hipLaunchKernelGGL(k_band_resid_sum, dim3((n_w * 64 + 255) / 256), dim3(256), 0, st, d_w, n_w,
                       d_part, d_res);
    if (want_x) {
      hipLaunchKernelGGL(k_band_unpermute, dim3(64, n_w), dim3(256), 0, st, d_w, dY, dX);
      b->dispatches++;
    }
    OHM_HIP(hipEventRecord(b->ev1, st));
    b->dispatches += 3;
    for (int64_t c = 0; c < n_circuits; c++) {
      if (f_hi[c] < 0 || !circ[c].x) continue;
      const int64_t n = circ[c].n;
      OHM_HIP(hipMemcpyAsync(circ[c].x + f_lo[c] * 2 * n, dX + x_at[c],
                             sizeof(cz) * n * (f_hi[c] - f_lo[c] + 1), hipMemcpyDeviceToHost, st));
    }
    OHM_HIP(hipStreamSynchronize(st));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
    kernel_ms += ms;
    b->readbacks++;
    pos = end;
  }
  OHM_HIP(hipMemcpy(singular_h.data(), d_sing, sizeof(int) * singular_h.size(), hipMemcpyDeviceToHost));
  OHM_HIP(hipMemcpy(res_h.data(), d_res, sizeof(double) * res_h.size(), hipMemcpyDeviceToHost));
  for (int64_t c = 0; c < n_circuits; c++) {
    if (!direct[c]) continue;
    for (int64_t f = 0; f < n_freqs; f++) {
      const int64_t k = c * n_freqs + f;
      ohm_batch_result &r = results[k];
      r.iterations = 0;
      r.refinements = 0;
      const double rr = res_h[2 * k], bb = res_h[2 * k + 1];
      r.residual_norm = bb > 0 ? sqrt(rr / bb) : 0.0;
      const bool sing = singular_h[k] != 0 || !std::isfinite(r.residual_norm);
      r.status = sing ? OHM_ERR_SOLVE : OHM_OK;
      r.reason = sing ? 0 : 1;
      if (sing)
        note(fail(OHM_ERR_SOLVE, "Solve error: banded LU: singular matrix (circuit %lld, frequency %lld)",
                  (long long)c, (long long)f));
    }
  }
  if (device_ms) *device_ms = kernel_ms;
  if (first != OHM_OK) return fail(first, "%s", first_msg.c_str());
  return OHM_OK;
}

}  // extern "C"

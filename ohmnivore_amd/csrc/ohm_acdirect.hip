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
//   * each pivot row goes to HBM divided by its pivot (the eliminated
//     right-hand side and the 2B entries right of the diagonal, with fill:
//     2B+1 complex values), interleaved by lane, so each wave's stores and
//     loads are coalesced; back substitution reads them in reverse and
//     writes x;
//   * the back substitution also sums ||b - A x||^2 and ||b||^2 per system
//     in FP64 (a row is complete once the x it touches are known);
//   * an unpermute pass scatters x back to the circuit's original order.
// Waves hold 64 systems sorted by size, so lanes of a wave run for similar
// lengths.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <memory>
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
// a - b c with each part as two FMAs (four instructions where a product
// formed apart and then subtracted takes six)
__device__ __forceinline__ cz cfms(cz a, cz b, cz c) {
  return {fma(-b.re, c.re, fma(b.im, c.im, a.re)), fma(-b.re, c.im, fma(-b.im, c.re, a.im))};
}
// 1 / a = conj(a) / |a|^2 with the reciprocal of |a|^2 from the hardware
// estimate and two Newton-Raphson steps r <- r + r (1 - d r), each of which
// squares the relative error (to within an ulp of the quotient), instead of
// two IEEE divisions (about nine dependent instructions each).
// cinv_d takes d = |a|^2 already formed.
__device__ __forceinline__ cz cinv_d(cz a, double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(fma(-d, r, 1.0), r, r);
  r = fma(fma(-d, r, 1.0), r, r);
  return {a.re * r, -a.im * r};
}
__device__ __forceinline__ cz cinv(cz a) { return cinv_d(a, cabs2(a)); }

// One band wave (segment) = up to 64 frequency points of one circuit (lane
// l: frequency f0 + l). The circuit's band rows are shared by every lane, so
// the one-lane LU's wave loads them cooperatively, 64 rows at a time, into
// LDS; a partitioned wave (PartWave) holds up to two segments.
struct BandWave {
  int64_t n;
  const double *gb, *cb;   // band rows of the permuted G and C: [n][2B+1]
  const double *b;         // permuted right-hand side, (re, im) pairs
  const uint32_t *perm;    // permuted row i holds original unknown perm[i]
  const double *omega;     // omega of lane 0 (consecutive lanes follow)
  int64_t store_off;       // lane-row offset of the U / Y region this wave's lanes sit in
                           // (U: W entries per lane-row)
  int64_t out_off;         // complex offset of lane 0's x in the output (lanes n apart)
  int nf;                  // active lanes
  int sys;                 // result slot of lane 0
  int stride;              // lanes of the region (row i of lane l: i stride + lane0 + l);
  int lane0;               //   nf and 0 unless the wave shares a partitioned wave
  double bnorm2;           // ||b||^2 of the circuit (the same at every frequency)
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

// Band row i of A = G + j omega C (entries for columns i-B .. i+B) and b_i:
// the same address in every lane of a wave, so one fetch serves the wave.
template <int W>
__device__ __forceinline__ void band_row(const BandWave &q, double om, int64_t i, cz (&a)[W],
                                         cz &bi) {
#pragma unroll
  for (int d = 0; d < W; d++) a[d] = cz{q.gb[i * W + d], om * q.cb[i * W + d]};
  bi = cz{q.b[2 * i], q.b[2 * i + 1]};
}

// |b_i - sum_d a_d x_d|^2
template <int W>
__device__ __forceinline__ double resid2(const cz (&a)[W], cz bi, const cz (&xc)[W]) {
  cz acc = bi;
#pragma unroll
  for (int d = 0; d < W; d++) acc = cfms(acc, a[d], xc[d]);
  return cabs2(acc);
}

// |b_i - sum_d a_d x_d|^2 and |b_i|^2 into rr, bb
template <int W>
__device__ __forceinline__ void resid_row(const cz (&a)[W], cz bi, const cz (&xc)[W], double &rr,
                                          double &bb) {
  rr += resid2<W>(a, bi, xc);
  bb += cabs2(bi);
}

// Forward elimination + back substitution for the lanes of one wave, the
// FP64 residual ||b - A x||^2 and ||b||^2 of every lane's system summed in
// the back substitution (row i = j + B is complete once x_j is known) and
// written to res[2 (sys + lane)].
template <int B>
__global__ __launch_bounds__(64) void k_band_lu(const BandWave *__restrict__ waves,
                                                cz *__restrict__ U, cz *__restrict__ X,
                                                int *__restrict__ singular,
                                                double *__restrict__ res) {
  constexpr int W = 2 * B + 1;  // band width in, U row width with fill
  __shared__ double sg[kBandTile][W], sc[kBandTile][W], sb[kBandTile][2];
  const int lane = threadIdx.x;
  const BandWave q = waves[blockIdx.x];
  const bool live = lane < q.nf;
  const double om = live ? q.omega[lane] : 0.0;
  const int64_t n = q.n;
  cz *Ub = U + q.store_off * W + q.lane0;
  cz *Xb = X + q.store_off + q.lane0;
  const int S = q.stride;  // lane stride of the region's rows
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
      if (live) {  // the pivot row divided by its pivot: x_j = e_0 - sum_c e_c x_{j+c}
        Ub[(j * W) * S + lane] = y[0] * piv_inv;
#pragma unroll
        for (int c = 1; c < W; c++) Ub[(j * W + c) * S + lane] = win[0][c] * piv_inv;
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
  // back substitution into X, 64 rows per tile; the band rows of the tile's
  // residual rows (row j+B for row j of the tile) are staged in LDS at slot
  // hi - j
  constexpr int G = B == 1 ? 8 : B == 2 ? 4 : 2;
  cz xw[W - 1];  // x_{j+1} .. x_{j+2B}
#pragma unroll
  for (int k = 0; k < W - 1; k++) xw[k] = cz{0.0, 0.0};
  double rr = 0.0, bb = 0.0;
  for (int64_t hi = n - 1; hi >= 0; hi -= kBandTile) {
    const int64_t lo = hi - kBandTile + 1 > 0 ? hi - kBandTile + 1 : 0;
    const int64_t ir = hi + B - lane;
    const bool in = ir < n && ir - B >= lo;
    __syncthreads();
#pragma unroll
    for (int d = 0; d < W; d++) {
      sg[lane][d] = in ? q.gb[ir * W + d] : 0.0;
      sc[lane][d] = in ? q.cb[ir * W + d] : 0.0;
    }
    sb[lane][0] = in ? q.b[2 * ir] : 0.0;
    sb[lane][1] = in ? q.b[2 * ir + 1] : 0.0;
    __syncthreads();
    for (int64_t g0 = hi; g0 >= lo; g0 -= G) {
      cz u[G][W];
#pragma unroll
      for (int g = 0; g < G; g++) {
        const int64_t j = g0 - g;
#pragma unroll
        for (int c = 0; c < W; c++)
          u[g][c] = j >= lo && live ? Ub[(j * W + c) * S + lane] : cz{0.0, 0.0};
      }
#pragma unroll
      for (int g = 0; g < G; g++) {
        const int64_t j = g0 - g;
        if (j >= lo) {
          cz xj = u[g][0];
#pragma unroll
          for (int c = 1; c < W; c++) xj = xj - u[g][c] * xw[c - 1];
          if (live) Xb[j * S + lane] = xj;
          if (j + B < n) {  // row j + B: columns j .. j + 2B
            const int sl = (int)(hi - j);
            cz a[W], xc[W];
#pragma unroll
            for (int d = 0; d < W; d++) a[d] = cz{sg[sl][d], om * sc[sl][d]};
            xc[0] = xj;
#pragma unroll
            for (int d = 1; d < W; d++) xc[d] = xw[d - 1];
            resid_row<W>(a, cz{sb[sl][0], sb[sl][1]}, xc, rr, bb);
          }
#pragma unroll
          for (int k = W - 2; k > 0; k--) xw[k] = xw[k - 1];
          xw[0] = xj;
        }
      }
    }
  }
  if (!live) return;
  // rows 0 .. B-1 (columns below 0 are absent); xw[k] = x_k now
#pragma unroll
  for (int i = 0; i < B; i++) {
    if (i < n) {
      cz a[W], bi, xc[W];
      band_row<W>(q, om, i, a, bi);
#pragma unroll
      for (int d = 0; d < W; d++) {
        const int k = i - B + d;
        xc[d] = k >= 0 && k < W - 1 ? xw[k] : cz{0.0, 0.0};
      }
      resid_row<W>(a, bi, xc, rr, bb);
    }
  }
  res[2 * (q.sys + lane)] = rr;
  res[2 * (q.sys + lane) + 1] = bb;
  if (bad) singular[q.sys + lane] = 1;
}

// x back to the original unknown order: out[out_off + lane n + perm[i]] = x_i
// grid: 64 workgroups of 256 threads per wave
__global__ void k_band_unpermute(const BandWave *__restrict__ waves, const cz *__restrict__ Y,
                                 cz *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x / 64, part = blockIdx.x % 64;
  const BandWave q = waves[wave];
  if (lane >= q.nf) return;
  const cz *Yb = Y + q.store_off + q.lane0;
  const int rows_per = blockDim.x / 64;
  for (int64_t i = part * rows_per + (threadIdx.x >> 6); i < q.n; i += (int64_t)64 * rows_per)
    out[q.out_off + (int64_t)lane * q.n + q.perm[i]] = Yb[i * q.stride + lane];
}

// ---------------------------------------------------------------------------
// Partitioned solve of systems of half bandwidth 1 (the RLC ladders after
// RCM). One lane per system leaves the AC workload with about 400 waves for
// 25,856 systems, each walking up to 190k dependent rows; here the rows of
// every system are cut into parts of about R rows, eliminated independently.
// A part block is solved with its neighbours' boundary unknowns set to zero,
// so each cut is placed, within 16 rows of every multiple of R, where the
// coupling of the two rows it separates is best contained in their
// diagonals (in a ladder: between the two ends of a resistor, so that the
// block sees that resistor grounded, never across an inductor's branch
// current, which would leave it floating at low frequency):
//   phase 1 (k_part_rel): every part eliminated top-down and, in a second
//     wave, bottom-up (LU with partial pivoting of the row-reversed part),
//     its couplings to the neighbouring parts' boundary unknowns carried as
//     two extra right-hand sides. Nothing is stored but, per part and
//     direction, the relation
//         x_end = alpha - k_prev x_prev - k_next x_next
//     for its last (top-down) or first (bottom-up) unknown, where x_prev is
//     the previous part's last unknown and x_next the next part's first.
//   phase 2 (k_part_reduce): per system, the 2P boundary unknowns from those
//     relations: a banded system of half bandwidth 2, LU with partial
//     pivoting, one lane per system.
//   phase 3 (k_part_solve): every part's top-down elimination restarted,
//     chunk by chunk from the last, from states phase 1 stored every kCk
//     steps, with the now known boundary values moved to its right-hand
//     side; each chunk's pivot rows stay in registers and are
//     back-substituted at once, x written where the one-lane solve writes
//     it; the residual of the part's interior rows is summed on the way.
//   k_part_resid_sum: per system, the residual of every part's first and
//     last row from the x written by both neighbours, plus the parts' sums.
// Phase 1 reads only the band rows shared by the wave's lanes and stores the
// checkpoints (64 B per lane every kCk rows); phase 3 reads them and writes
// x, so a (system, row) costs 24 B of HBM traffic where the one-lane solve,
// which stores its U rows and reads them back, moves 80 B or more.
// Systems whose residual is not small (a part block close to singular
// after all) are solved again by k_band_lu.
// A wave of these kernels takes 64 (circuit, frequency) systems in the order
// circuits longest first, frequencies ascending, so it holds up to two
// segments (band waves): the last frequencies of one circuit and the first
// of the next, which is no longer. Both are cut at the same rows (chosen for
// the two circuits together); the shorter continues with identity rows up to
// the wave's row count, which couple to nothing and solve to x = 0. With 101
// frequencies per circuit, one circuit per wave left 27 of every 128 lanes
// idle.
constexpr int kPartW = 3;  // band width at half bandwidth 1
// steps per band tile: the tile holds this many rows of each segment
constexpr int kPartTile = 32;

// 32 band rows of each of the wave's two segments, each row's G, C and b
// entries together (9 doubles, the ninth padding the lanes' stores over the
// LDS banks): the 8 values of a row, and of the next rows, sit at small
// offsets from one address. The second segment's rows start one double
// further on, so that the two addresses a wave reads at once fall in
// different banks.
struct PartTile {
  double r[2][kPartTile * 9 + 1];
  __device__ double &g(int h, int sl, int d) { return r[h][sl * 9 + d]; }
  __device__ double &c(int h, int sl, int d) { return r[h][sl * 9 + kPartW + d]; }
  __device__ double &b(int h, int sl, int k) { return r[h][sl * 9 + 2 * kPartW + k]; }
};

// A wave of the partitioned kernels: lanes [0, split) are the lanes of band
// wave seg_a, lanes [split, nf) those of seg_b (seg_b == seg_a and split ==
// nf for one segment). Its parts are parts[pbase .. pbase + nparts), over n
// rows; x of lane L at row i is X[store_off + i nf + L] (the segments' own
// stride and lane0 name the same places).
struct PartWave {
  int seg_a, seg_b, split, nf;
  int pbase, nparts;
  int64_t n, store_off;
};

// one part: its wave (index into the chunk's partitioned-wave table), rows
// [s, s+m) and the index of its first checkpoint of the top-down elimination
struct PartRef {
  int item, s, m, ck;
};

// result slot of lane L of a partitioned wave
__device__ __forceinline__ int part_sys(const BandWave &qa, const BandWave &qb,
                                        const PartWave &w, int lane) {
  return lane >= w.split ? qb.sys + lane - w.split : qa.sys + lane;
}

// Steps of the top-down elimination between two checkpoints of its state.
constexpr int kCk = 8;

// direction-order row k of part [s, s+m) of segment q (rev: row k is
// s+m-1-k, band entries mirrored) into this lane's registers; rows past the
// segment's own n are identity rows
__device__ __forceinline__ void part_row_load(const BandWave &q, int64_t s, int64_t m, bool rev,
                                              int64_t k, double (&rg)[kPartW],
                                              double (&rc)[kPartW], double (&rb)[2]) {
  const int64_t i = rev ? s + m - 1 - k : s + k;
  if (k < m && i < q.n) {
#pragma unroll
    for (int d = 0; d < kPartW; d++) {
      const int dd = rev ? kPartW - 1 - d : d;
      rg[d] = q.gb[i * kPartW + dd];
      rc[d] = q.cb[i * kPartW + dd];
    }
    rb[0] = q.b[2 * i];
    rb[1] = q.b[2 * i + 1];
  } else if (k < m) {
#pragma unroll
    for (int d = 0; d < kPartW; d++) rg[d] = rc[d] = 0.0;
    rg[1] = 1.0;
    rb[0] = rb[1] = 0.0;
  } else {
#pragma unroll
    for (int d = 0; d < kPartW; d++) rg[d] = rc[d] = 0.0;
    rb[0] = rb[1] = 0.0;
  }
}

// The tile's rows: direction rows k0 .. k0+31 of segment a (lanes below 32)
// and of segment b (the others), loaded into this lane's registers, then
// stored to its tile slot.
__device__ __forceinline__ void part_tile_load(const BandWave &qa, const BandWave &qb, int64_t s,
                                               int64_t m, bool rev, int64_t k0,
                                               double (&rg)[kPartW], double (&rc)[kPartW],
                                               double (&rb)[2]) {
  const int lane = threadIdx.x;
  const int64_t k = k0 + (lane & (kPartTile - 1));
  if (lane < kPartTile)
    part_row_load(qa, s, m, rev, k, rg, rc, rb);
  else
    part_row_load(qb, s, m, rev, k, rg, rc, rb);
}
__device__ __forceinline__ void part_tile_store(PartTile &t, const double (&rg)[kPartW],
                                                const double (&rc)[kPartW],
                                                const double (&rb)[2]) {
  const int h = threadIdx.x / kPartTile, sl = threadIdx.x % kPartTile;
#pragma unroll
  for (int d = 0; d < kPartW; d++) {
    t.g(h, sl, d) = rg[d];
    t.c(h, sl, d) = rc[d];
  }
  t.b(h, sl, 0) = rb[0];
  t.b(h, sl, 1) = rb[1];
}

// x and the checkpoints are written once and read once, far apart (tens of
// GB per sweep, well past the caches): their stores and the checkpoints'
// loads are nontemporal (measured: part solve 12.75 -> 12.35 ms, relations
// 8.3 -> 8.15 ms on the 25,856-system sweep).
typedef double d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void nt_store(cz *p, cz v) {
  __builtin_nontemporal_store(d2v{v.re, v.im}, (d2v *)p);
}
__device__ __forceinline__ cz nt_load(const cz *p) {
  const d2v v = __builtin_nontemporal_load((const d2v *)p);
  return cz{v.x, v.y};
}

// Marker kept in place of a pivot row when the pivot of step j is row j+1
// as given (|N_0| > |R_0|; the back substitution then reads that row from
// the band rows): a quiet NaN real part.
__device__ __forceinline__ cz pivot_marker() { return cz{__builtin_nan(""), 0.0}; }
__device__ __forceinline__ bool is_pivot_marker(cz e) { return e.re != e.re; }

// Eliminates the part [s, s+m) in direction order for one lane, carrying
// one remainder row forward. Step j combines the remainder R (row j after
// the earlier steps: entries at columns j, j+1) with row j+1 as given (N:
// columns j, j+1, j+2) into
//     R' = (N_0 R - R_0 N) / max(|Re|, |Im| of R_0 and N_0),
// which has no entry at column j. Up to a scalar of modulus between 1 and 2
// this is the remainder partial pivoting leaves (N - (N_0/R_0) R when
// |R_0| >= |N_0|, else R - (R_0/N_0) N), so no row exchange is needed, and
// the carried row keeps the size partial pivoting gives it: the part solve
// restarts from the stored states and chooses its pivots by comparing
// |N_0| with |R_0|, which needs that size (a carried row rescaled to an
// arbitrary size, by a power of two at each checkpoint, measured 100x
// larger residuals on 31 of the 25,856 AC-sweep systems).
// The right-hand sides are f, the start coupling and the end coupling (the
// entries of direction rows 0 and m-1 outside the part); ly and lu are the
// last remainder's right-hand sides and diagonal. With ck set, the state
// before every kCk-th step (R_0, R_1, f, start coupling; the end coupling is
// still zero there) is stored at ck[(j / kCk * 4 + k) * 64]. The lane's
// system is in segment qb when h is 1, else in qa.
__device__ __forceinline__ void part_eliminate(const BandWave &qa, const BandWave &qb, int h,
                                               int64_t s, int64_t m, bool rev, double om,
                                               PartTile &t, cz (&ly)[3], cz &lu, int &bad,
                                               cz *ck) {
  constexpr int W = kPartW;
  double rg[W], rc[W], rb[2];
  // direction row 0 is the first remainder: columns -1 (left of the part),
  // 0 and 1 (right of the part when m == 1)
  if (h)
    part_row_load(qb, s, m, rev, 0, rg, rc, rb);
  else
    part_row_load(qa, s, m, rev, 0, rg, rc, rb);
  cz r0, r1, y0 = {rb[0], rb[1]}, y1, y2;
  {
    cz a[W];
#pragma unroll
    for (int d = 0; d < W; d++) a[d] = cz{rg[d], om * rc[d]};
    r0 = a[1];
    r1 = m > 1 ? a[2] : cz{0.0, 0.0};
    y1 = a[0];
    y2 = m > 1 ? cz{0.0, 0.0} : a[2];
  }
  // tiles of direction rows 1 + t0 .. 32 + t0; step j reads row j+1
  part_tile_load(qa, qb, s, m, rev, 1, rg, rc, rb);
  for (int64_t t0 = 0; t0 < m - 1; t0 += kPartTile) {
    __syncthreads();
    part_tile_store(t, rg, rc, rb);
    __syncthreads();
    part_tile_load(qa, qb, s, m, rev, 1 + t0 + kPartTile, rg, rc, rb);  // next tile in flight
    const int64_t t1 = t0 + kPartTile < m - 1 ? t0 + kPartTile : m - 1;
    // one elimination step on direction row j + 1 (tile slot sl); LAST: the
    // part's last row, whose column-m entry is the end coupling
    auto step = [&](int sl, auto last) {
      const cz n0 = {t.g(h, sl, 0), om * t.c(h, sl, 0)},
               n1 = {t.g(h, sl, 1), om * t.c(h, sl, 1)};
      cz n2 = {t.g(h, sl, 2), om * t.c(h, sl, 2)};
      const cz f = {t.b(h, sl, 0), t.b(h, sl, 1)};
      cz e2 = {0.0, 0.0};
      if (decltype(last)::value) {
        e2 = n2;
        n2 = cz{0.0, 0.0};
      }
      const double mx = fmax(fmax(fabs(r0.re), fabs(r0.im)), fmax(fabs(n0.re), fabs(n0.im)));
      if (!(mx > 0.0)) bad = 1;
      // (N_0 R - R_0 N) sc = c1 R - cr N
      const double sc = __builtin_amdgcn_rcp(mx);
      const cz c1 = {n0.re * sc, n0.im * sc}, cr = {r0.re * sc, r0.im * sc};
      const cz nr0 = cfms(c1 * r1, cr, n1);
      const cz crn2 = cr * n2;
      y0 = cfms(c1 * y0, cr, f);
      y1 = c1 * y1;
      // y2 (the end coupling's column) is zero until the last row of a part
      // of m > 1 rows brings e2, so it is formed there alone
      if (decltype(last)::value) y2 = cfms(c1 * y2, cr, e2);
      r0 = nr0;
      r1 = cz{-crn2.re, -crn2.im};
    };
    auto checkpoint = [&](int64_t j) {  // state before step j (j a multiple of kCk)
      if (ck) {
        cz *o = ck + (size_t)(j / kCk) * 4 * 64;
        nt_store(&o[0], r0);
        nt_store(&o[64], r1);
        nt_store(&o[128], y0);
        nt_store(&o[192], y1);
      }
    };
    int64_t j = t0;
    // groups of kCk steps that stop short of the last row: no per-step tests
    for (; j + kCk <= t1 && j + kCk < m - 1; j += kCk) {
      checkpoint(j);
#pragma unroll
      for (int k = 0; k < kCk; k++) step((int)(j - t0) + k, std::false_type());
    }
    for (; j < t1; j++) {
      if ((j & (kCk - 1)) == 0) checkpoint(j);
      if (j + 2 >= m)
        step((int)(j - t0), std::true_type());
      else
        step((int)(j - t0), std::false_type());
    }
  }
  lu = r0;
  ly[0] = y0;
  ly[1] = y1;
  ly[2] = y2;
}

// Phase 1: blockIdx.x = 2 ip + direction; relation of (part ip, direction)
// to rel[((ip * 2 + dir) * 3 + k) * 64 + lane], k = alpha, k_prev, k_next.
// The top-down wave also leaves, every kCk steps, the state (R_0, R_1, and
// the right-hand sides f and start coupling) at
// ckpt[((ck + c) * 4 + k) * 64 + lane] for the solve to restart from.
__global__ __launch_bounds__(64) void k_part_rel(const BandWave *__restrict__ waves,
                                                 const PartWave *__restrict__ pwaves,
                                                 const PartRef *__restrict__ parts,
                                                 cz *__restrict__ rel, int *__restrict__ singular,
                                                 cz *__restrict__ ckpt) {
  __shared__ PartTile t;
  const int lane = threadIdx.x;
  const int ip = blockIdx.x >> 1, dir = blockIdx.x & 1;
  const PartRef pr = parts[ip];
  const PartWave w = pwaves[pr.item];
  const BandWave qa = waves[w.seg_a], qb = waves[w.seg_b];
  const int64_t s = pr.s, m = pr.m;
  const int h = lane >= w.split ? 1 : 0;
  const bool live = lane < w.nf;
  const double om = live ? (h ? qb.omega[lane - w.split] : qa.omega[lane]) : 0.0;
  cz ly[3], lu = {1.0, 0.0};
  int bad = 0;
  cz *ck = ckpt && dir == 0 && live ? ckpt + (size_t)pr.ck * 4 * 64 + lane : nullptr;
  part_eliminate(qa, qb, h, s, m, dir == 1, om, t, ly, lu, bad, ck);
  if (!live) return;
  if (bad || !(cabs2(lu) > 0.0)) {
    singular[part_sys(qa, qb, w, lane)] = 2;
    lu = cz{1.0, 0.0};
  }
  const cz iu = cinv(lu);
  // top-down: start = prev, end = next; bottom-up: start = next, end = prev
  const cz alpha = ly[0] * iu, ks = ly[1] * iu, ke = ly[2] * iu;
  cz *o = rel + (size_t)(ip * 2 + dir) * 3 * 64 + lane;
  o[0] = alpha;
  o[64] = dir == 0 ? ks : ke;
  o[128] = dir == 0 ? ke : ks;
}

// Phase 2: one lane per system. Boundary unknowns F_p = x(first row of part
// p) at index 2p and L_p = x(last row) at 2p+1; their equations
//   F_p + k_prev L_{p-1} + k_next F_{p+1} = alpha   (bottom-up relation)
//   L_p + k_prev L_{p-1} + k_next F_{p+1} = alpha   (top-down relation)
// form a band of half bandwidth 2; solved by LU with partial pivoting (U rows
// of 5 entries and y to red, x to bnd[(2 pbase + r) * 64 + lane]).
// blockIdx.x: the partitioned wave.
__global__ __launch_bounds__(64) void k_part_reduce(const BandWave *__restrict__ waves,
                                                    const PartWave *__restrict__ pwaves,
                                                    const cz *__restrict__ rel,
                                                    cz *__restrict__ red, cz *__restrict__ bnd,
                                                    int *__restrict__ singular) {
  constexpr int KB = 2, W = 2 * KB + 1;
  const int lane = threadIdx.x;
  const PartWave q = pwaves[blockIdx.x];
  if (lane >= q.nf) return;
  const int P = q.nparts, n = 2 * P;
  const size_t rb = (size_t)2 * q.pbase;
  // band row r of the reduced system: entries for columns r-2 .. r+2
  auto row = [&](int r, cz (&a)[W], cz &f) {
#pragma unroll
    for (int d = 0; d < W; d++) a[d] = cz{0.0, 0.0};
    f = cz{0.0, 0.0};
    if (r >= n) return;
    const int p = r >> 1, dir = (r & 1) ? 0 : 1;
    const cz *o = rel + (size_t)((q.pbase + p) * 2 + dir) * 3 * 64 + lane;
    f = o[0];
    const cz kp = p > 0 ? o[64] : cz{0.0, 0.0}, kn = p + 1 < P ? o[128] : cz{0.0, 0.0};
    if (dir == 1) {  // F_p: L_{p-1} at d 1, F_p at d 2, F_{p+1} at d 4
      a[1] = kp;
      a[2] = cz{1.0, 0.0};
      a[4] = kn;
    } else {         // L_p: L_{p-1} at d 0, L_p at d 2, F_{p+1} at d 3
      a[0] = kp;
      a[2] = cz{1.0, 0.0};
      a[3] = kn;
    }
  };
  cz win[KB + 1][W], y[KB + 1];
#pragma unroll
  for (int r = 0; r <= KB; r++) {
    cz a[W], f;
    row(r, a, f);
    y[r] = f;
#pragma unroll
    for (int c = 0; c < W; c++) win[r][c] = cz{0.0, 0.0};
#pragma unroll
    for (int d = 0; d < W; d++) {
      const int col = r - KB + d;
      if (col >= 0 && col < W) win[r][col] = a[d];
    }
  }
  int bad = 0;
  // the row that enters the window after the next step is loaded one step
  // ahead, so its loads overlap a step's arithmetic
  cz na[W], nf;
  row(KB + 1, na, nf);
  for (int j = 0; j < n; j++) {
    int pv = 0;
    double best = cabs2(win[0][0]);
#pragma unroll
    for (int r = 1; r <= KB; r++) {
      const double v = cabs2(win[r][0]);
      if (j + r < n && v > best) {
        best = v;
        pv = r;
      }
    }
    if (!(best > 0.0)) bad = 1;
#pragma unroll
    for (int r = 1; r <= KB; r++)
      if (pv == r) {
#pragma unroll
        for (int c = 0; c < W; c++) {
          const cz u = win[0][c];
          win[0][c] = win[r][c];
          win[r][c] = u;
        }
        const cz u = y[0];
        y[0] = y[r];
        y[r] = u;
      }
    const cz pinv = best > 0.0 ? cinv(win[0][0]) : cz{0.0, 0.0};
#pragma unroll
    for (int r = 1; r <= KB; r++) {
      const cz l = win[r][0] * pinv;
#pragma unroll
      for (int c = 1; c < W; c++) win[r][c] = win[r][c] - l * win[0][c];
      y[r] = y[r] - l * y[0];
    }
    cz *u = red + ((rb + j) * (W + 1)) * 64 + lane;
#pragma unroll
    for (int c = 0; c < W; c++) u[c * 64] = win[0][c];
    u[W * 64] = y[0];
#pragma unroll
    for (int r = 0; r < KB; r++) {
#pragma unroll
      for (int c = 0; c + 1 < W; c++) win[r][c] = win[r + 1][c + 1];
      win[r][W - 1] = cz{0.0, 0.0};
      y[r] = y[r + 1];
    }
#pragma unroll
    for (int d = 0; d < W; d++) win[KB][d] = na[d];
    y[KB] = nf;
    row(j + 2 + KB, na, nf);
  }
  cz xw[W - 1];
#pragma unroll
  for (int k = 0; k < W - 1; k++) xw[k] = cz{0.0, 0.0};
  // the next (lower) row of U and its right-hand side loaded one step ahead
  cz un[W + 1];
  auto u_load = [&](int j) {
    if (j < 0) return;
    const cz *u = red + ((rb + j) * (W + 1)) * 64 + lane;
#pragma unroll
    for (int c = 0; c <= W; c++) un[c] = u[c * 64];
  };
  u_load(n - 1);
  for (int j = n - 1; j >= 0; j--) {
    cz u[W + 1];
#pragma unroll
    for (int c = 0; c <= W; c++) u[c] = un[c];
    u_load(j - 1);
    cz acc = u[W];
#pragma unroll
    for (int c = 1; c < W; c++) acc = acc - u[c] * xw[c - 1];
    const cz d = u[0];
    const cz xj = cabs2(d) > 0.0 ? acc * cinv(d) : cz{0.0, 0.0};
    bnd[(rb + j) * 64 + lane] = xj;
#pragma unroll
    for (int k = W - 2; k > 0; k--) xw[k] = xw[k - 1];
    xw[0] = xj;
  }
  if (bad) singular[part_sys(waves[q.seg_a], waves[q.seg_b], q, lane)] = 2;
}

// Phase 3: blockIdx.x = ip. The part's top-down elimination is recomputed
// chunk by chunk, last chunk first, from the states phase 1 left every kCk
// steps (the start coupling column times the now known xs is taken off the
// right-hand side; the end coupling times xe enters at the last row), and
// each chunk's pivot rows stay in registers and are back-substituted at
// once: row j's pivot row is (e_0, e_1), x_j = e_0 - e_1 x_{j+1}, or the
// marker with e_1 = 1 / a_{j+1,j}: row j+1 as given was the pivot, x_j =
// (f_{j+1} - a_{j+1,j+1} x_{j+1} - a_{j+1,j+2} x_{j+2}) e_1. Per (system,
// row) this reads the checkpoints (64 B per kCk rows) and writes x (16 B),
// where storing the pivot rows and reading them back took 80 B. The part's
// interior rows s+1 .. s+m-2 use only x of this part, so their squared
// residuals are summed here, to pres[ip * 64 + lane].
// A step whose pivot choice is the same in every live lane (on the RLC
// ladders about 86 % of them: every third row as given is the pivot at
// every frequency, and the carried row at most others) eliminates
// with the pivot's reciprocal, which the pivot row needs anyway:
// R' = N - (n_0 / r_0) R, or R' = R - (r_0 / n_0) N, both the remainder of
// partial pivoting; a step where lanes differ uses the scaled combination
// of part_eliminate.
// At most 168 VGPRs: three waves share a SIMD.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_part_solve(
    const BandWave *__restrict__ waves, const PartWave *__restrict__ pwaves,
    const PartRef *__restrict__ parts, const cz *__restrict__ bnd, const cz *__restrict__ ckpt,
    cz *__restrict__ X, int *__restrict__ singular, double *__restrict__ pres) {
  constexpr int W = kPartW;
  __shared__ PartTile t;
  const int lane = threadIdx.x;
  const int ip = blockIdx.x;
  const PartRef pr = parts[ip];
  const PartWave q = pwaves[pr.item];
  const BandWave qa = waves[q.seg_a], qb = waves[q.seg_b];
  const int p = ip - q.pbase;
  const int64_t s = pr.s;
  const int m = pr.m;  // loop bounds in 32 bits keep their tests on the scalar unit
  const int h = lane >= q.split ? 1 : 0;
  const bool live = lane < q.nf;
  const uint64_t live_mask = __ballot(live);
  const double om = live ? (h ? qb.omega[lane - q.split] : qa.omega[lane]) : 0.0;
  const size_t rb = (size_t)2 * q.pbase;
  cz xs = {0.0, 0.0}, xe = {0.0, 0.0};
  if (live) {
    if (p > 0) xs = bnd[(rb + 2 * (p - 1) + 1) * 64 + lane];
    if (p + 1 < q.nparts) xe = bnd[(rb + 2 * (p + 1)) * 64 + lane];
  }
  const int S = q.nf;
  cz *Xb = X + q.store_off;
  double rr = 0.0;
  int bad = 0;
  const int ns = m - 1;  // elimination steps; step j pivots x_{s+j} against row s+j+1
  if (ns == 0) {         // one row: a_0 xs + a_1 x + a_2 xe = b
    double rg[W], rc[W], rbv[2];
    if (h)
      part_row_load(qb, s, 1, false, 0, rg, rc, rbv);
    else
      part_row_load(qa, s, 1, false, 0, rg, rc, rbv);
    cz a[W];
#pragma unroll
    for (int d = 0; d < W; d++) a[d] = cz{rg[d], om * rc[d]};
    const cz bi = {rbv[0], rbv[1]};
    const cz f = bi - a[0] * xs - a[2] * xe;
    if (!(cabs2(a[1]) > 0.0)) bad = 1;
    if (live) Xb[s * S + lane] = f * cinv(a[1]);
  } else {
    const cz *ckb = ckpt + (size_t)pr.ck * 4 * 64 + lane;
    cz xw0 = {0.0, 0.0}, xw1 = {0.0, 0.0};  // x_{s+j+1}, x_{s+j+2}
    // the checkpoint of the next chunk (chunks run last first) is loaded
    // while the current one is worked; lanes past nf load nothing (their
    // results are not stored)
    cz ckn[4] = {};
    if (live) {
#pragma unroll
      for (int k = 0; k < 4; k++) ckn[k] = nt_load(&ckb[((size_t)((ns - 1) / kCk) * 4 + k) * 64]);
    }
    // tiles of 32 steps, last first; step j's row s+j+1 at slot j - t0
    for (int t0 = (ns - 1) / kPartTile * kPartTile; t0 >= 0; t0 -= kPartTile) {
      const int t1 = t0 + kPartTile < ns ? t0 + kPartTile : ns;
      double rg[W], rc[W], rbv[2];
      part_tile_load(qa, qb, s, m, false, 1 + t0, rg, rc, rbv);
      __syncthreads();
      part_tile_store(t, rg, rc, rbv);
      __syncthreads();
      for (int j0 = t0 + (t1 - 1 - t0) / kCk * kCk; j0 >= t0; j0 -= kCk) {
        const int j1 = j0 + kCk < t1 ? j0 + kCk : t1;
        cz ck4[4];
#pragma unroll
        for (int k = 0; k < 4; k++) ck4[k] = ckn[k];
        if (live && j0 >= kCk) {
#pragma unroll
          for (int k = 0; k < 4; k++) ckn[k] = nt_load(&ckb[((size_t)(j0 / kCk - 1) * 4 + k) * 64]);
        }
        cz r0 = ck4[0], r1 = ck4[1];
        cz y = cfms(ck4[2], ck4[3], xs);
        cz e0[kCk], e1[kCk];
#pragma unroll
        for (int k = 0; k < kCk; k++) {
          const int j = j0 + k;
          if (j < j1) {
            const int sl = j - t0;
            const cz n0 = {t.g(h, sl, 0), om * t.c(h, sl, 0)},
                     n1 = {t.g(h, sl, 1), om * t.c(h, sl, 1)};
            cz n2 = {t.g(h, sl, 2), om * t.c(h, sl, 2)}, f = {t.b(h, sl, 0), t.b(h, sl, 1)};
            if (j + 2 >= m) {  // row m-1: its column-m entry times xe
              f = cfms(f, n2, xe);
              n2 = cz{0.0, 0.0};
            }
            const double dn = cabs2(n0), dr = cabs2(r0);
            const bool sw = dn > dr;  // row j+1 as given is the pivot
            const uint64_t vote = __ballot(sw) & live_mask;
            if (vote == 0) {  // the carried row is the pivot in every live lane
              if (!(dr > 0.0)) bad = 1;
              const cz inv = cinv_d(r0, dr);
              const cz l = n0 * inv;
              e0[k] = y * inv;
              e1[k] = r1 * inv;
              r0 = cfms(n1, l, r1);
              r1 = n2;
              y = cfms(f, l, y);
            } else if (vote == live_mask) {  // row j+1 is the pivot in every live lane
              const cz inv = cinv_d(n0, dn);
              const cz l = r0 * inv;
              e0[k] = pivot_marker();
              e1[k] = inv;
              r0 = cfms(r1, l, n1);
              const cz ln2 = l * n2;
              r1 = cz{-ln2.re, -ln2.im};
              y = cfms(y, l, f);
            } else {
              const cz inv = cinv_d(sw ? n0 : r0, sw ? dn : dr);
              e0[k] = sw ? pivot_marker() : y * inv;
              e1[k] = sw ? inv : r1 * inv;
              const double mx =
                  fmax(fmax(fabs(r0.re), fabs(r0.im)), fmax(fabs(n0.re), fabs(n0.im)));
              if (!(mx > 0.0)) bad = 1;
              const double sc = __builtin_amdgcn_rcp(mx);
              // (N_0 R - R_0 N) sc
              const cz c1 = {n0.re * sc, n0.im * sc}, cr = {r0.re * sc, r0.im * sc};
              const cz nr0 = cfms(c1 * r1, cr, n1);
              const cz crn2 = cr * n2;
              r1 = cz{-crn2.re, -crn2.im};
              r0 = nr0;
              y = cfms(c1 * y, cr, f);
            }
          }
        }
        if (j1 == ns) {  // the last row: its pivot is the final remainder
          if (!(cabs2(r0) > 0.0)) bad = 1;
          xw0 = y * cinv(r0);
          if (live) nt_store(&Xb[(s + m - 1) * S + lane], xw0);
        }
#pragma unroll
        for (int k = kCk - 1; k >= 0; k--) {
          const int j = j0 + k;
          if (j < j1) {
            const int sl = j - t0;
            cz a[W];  // row s+j+1 (columns s+j .. s+j+2)
#pragma unroll
            for (int d = 0; d < W; d++) a[d] = cz{t.g(h, sl, d), om * t.c(h, sl, d)};
            const cz bj = {t.b(h, sl, 0), t.b(h, sl, 1)};
            cz xj;
            if (is_pivot_marker(e0[k])) {  // row s+j+1 as given was the pivot
              const bool last = j + 2 >= m;
              const cz f = last ? cfms(bj, a[2], xe) : bj;
              const cz a2 = last ? cz{0.0, 0.0} : a[2];
              xj = cfms(cfms(f, a[1], xw0), a2, xw1) * e1[k];
            } else {
              xj = cfms(e0[k], e1[k], xw0);
            }
            if (live) nt_store(&Xb[(s + j) * S + lane], xj);
            if (j + 1 <= m - 2) {  // interior row s+j+1
              const cz xc[W] = {xj, xw0, xw1};
              rr += resid2<W>(a, bj, xc);
            }
            xw1 = xw0;
            xw0 = xj;
          }
        }
      }
    }
  }
  if (!live) return;
  pres[(size_t)ip * 64 + lane] = rr;
  if (bad) singular[part_sys(qa, qb, q, lane)] = 2;
}

// One thread per (partitioned wave, lane): the residual terms of every
// part's first and last row (columns in the neighbouring parts, read from
// the written x) and the parts' interior sums, in part order, to
// res[2 sys], beside the circuit's ||b||^2. Identity rows past the lane's
// own segment have zero residual (x is zero there).
__global__ void k_part_resid_sum(const BandWave *__restrict__ waves,
                                 const PartWave *__restrict__ pwaves, int n_waves,
                                 const PartRef *__restrict__ parts,
                                 const double *__restrict__ pres, const cz *__restrict__ X,
                                 double *__restrict__ res) {
  constexpr int W = kPartW;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_waves * 64) return;
  const PartWave w = pwaves[t / 64];
  const int lane = t % 64;
  if (lane >= w.nf) return;
  const bool second = lane >= w.split;
  const BandWave q = waves[second ? w.seg_b : w.seg_a];
  const int l = second ? lane - w.split : lane;
  const double om = q.omega[l];
  const cz *Xb = X + w.store_off;
  const int S = w.nf;
  double rr = 0.0;
  for (int ip = w.pbase; ip < w.pbase + w.nparts; ip++) {
    rr += pres[(size_t)ip * 64 + lane];
    const PartRef pr = parts[ip];
    for (int e = 0; e < 2; e++) {
      if (e == 1 && pr.m == 1) break;  // one row: first and last
      const int64_t i = e == 0 ? pr.s : (int64_t)pr.s + pr.m - 1;
      if (i >= q.n) continue;
      cz a[W], bi, xc[W];
      band_row<W>(q, om, i, a, bi);
#pragma unroll
      for (int d = 0; d < W; d++) {
        const int64_t col = i - 1 + d;
        xc[d] = col >= 0 && col < q.n ? Xb[col * S + lane] : cz{0.0, 0.0};
      }
      rr += resid2<W>(a, bi, xc);
    }
  }
  res[2 * (q.sys + l)] = rr;
  res[2 * (q.sys + l) + 1] = q.bnorm2;
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

// Symbolic preparation of one circuit: pattern union of G and C, RCM,
// half bandwidth. Runs on a host thread; the message of a failure is kept.
ohm_status prepare_circuit(const ohm_ac_circuit &q, Prepared &p, std::string &msg) {
  p.n = q.n;
  if (q.n == 0) return OHM_OK;
  ohm_status s = csr_union(q.n, q.g_rowptr, q.g_col, q.g_val, q.c_val ? q.c_rowptr : nullptr,
                           q.c_col, q.c_val, p.rp, p.cl, p.g, p.c);
  if (s != OHM_OK) {
    msg = ohm_last_error();
    return s;
  }
  p.perm = rcm(q.n, p.rp, p.cl);
  p.iperm.assign(q.n, 0);
  for (int64_t i = 0; i < q.n; i++) p.iperm[p.perm[i]] = (uint32_t)i;
  int64_t bw = 0;
  for (int64_t r = 0; r < q.n; r++)
    for (uint32_t k = p.rp[r]; k < p.rp[r + 1]; k++)
      bw = std::max<int64_t>(bw, std::llabs((int64_t)p.iperm[r] - (int64_t)p.iperm[p.cl[k]]));
  p.band = (int)std::min<int64_t>(bw, 1 << 20);
  return OHM_OK;
}

ohm_status check_circuits(int64_t n_circuits, const ohm_ac_circuit *circ) {
  for (int64_t c = 0; c < n_circuits; c++) {
    const ohm_ac_circuit &q = circ[c];
    if (q.n < 0 || q.n > 0x7fffffff) return fail(OHM_ERR_INVALID, "circuit %lld: bad size", (long long)c);
    if (q.n == 0) continue;
    if (!q.g_rowptr || !q.g_col || !q.g_val || !q.b_ac || (q.c_val && (!q.c_rowptr || !q.c_col)))
      return fail(OHM_ERR_INVALID, "circuit %lld: null arrays", (long long)c);
  }
  return OHM_OK;
}

constexpr int kMaxDirectBand = 4;

}  // namespace

// The symbolic half of a banded AC sweep, resident on the device: per
// circuit the band rows of G and C in RCM order (n (2B+1) doubles each), the
// permuted right-hand side and the permutation. The last solve's wave table
// and per-system results stay with it; its solutions stay in the backend's
// AC workspace while ws_gen matches.
struct ohm_ac_plan {
  ohm_backend *b = nullptr;
  int64_t nc = 0;
  int B = 1;
  std::vector<int64_t> n;
  void *d_band = nullptr;
  std::vector<const double *> gb, cb, bp;
  std::vector<double> bnorm2;  // ||b||^2 per circuit
  std::vector<const uint32_t *> perm;
  // half bandwidth 1: per circuit, cut_score[i] rates a part boundary
  // between RCM rows i and i+1 (lower is better)
  std::vector<std::vector<float>> cut_score;
  // part boundaries per circuit for rows-per-part cuts_R (kept between
  // sweeps: computing them starts a pool of host threads)
  int64_t cuts_R = -1;
  std::vector<std::vector<int>> cuts;
  // boundaries of the partitioned waves that hold two circuits, by wave
  // index, for (rows per part, frequencies) pair_key
  std::pair<int64_t, int64_t> pair_key{-1, -1};
  std::vector<std::vector<int>> pair_cuts;
  // host copies of the last sweep's singular flags and residual sums, when a
  // one-chunk sweep redid no system (they are then the final values)
  std::vector<int> h_sing;
  std::vector<double> h_res;
  bool h_final = false;
  // last solve
  int64_t nf = 0;
  double *d_omega = nullptr;
  int *d_sing = nullptr;
  double *d_res = nullptr;
  size_t sys_cap = 0, omega_cap = 0;
  bool resident = false;
  uint64_t ws_gen = 0;
  std::vector<BandWave> waves;  // the resident chunk's wave table
  size_t y_off = 0;             // byte offset of its solutions in the workspace
  int64_t tuned_rows = -1;      // autotuned rows per part (0: one lane per system)
  int64_t part_rows = 0;        // rows per part of the last solve
  int chunks = 0;
  int64_t fallback_systems = 0; // systems of the last solve redone in one lane
  // device ms of the last solve: relations, reduced systems, part solves
  // (or one-lane LU), residuals, one-lane redo, unpermute
  double phase_ms[6] = {0, 0, 0, 0, 0, 0};
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
};

namespace {

// |off| / |diag|: 0 for a missing coupling, infinite for a zero diagonal
float coupling_ratio(double off, double diag) {
  off = fabs(off);
  diag = fabs(diag);
  if (off == 0.0) return 0.0f;
  if (diag == 0.0) return INFINITY;
  return (float)(off / diag);
}

// Rating of a part boundary between rows i and i+1 of a tridiagonal band
// (gb: [n][3] = entries at columns i-1, i, i+1 of the DC matrix G): the
// larger of |a(i,i+1)| / |a(i,i)| and |a(i+1,i)| / |a(i+1,i+1)|. A part is
// solved with the unknown past its boundary set to zero; where that coupling
// also sits in the diagonal (a conductance to the neighbour) the block keeps
// it as a conductance to ground, where it does not (a branch current of an
// inductor or source) the block loses a path and can be left floating.
std::vector<float> cut_scores(const std::vector<double> &gb, int64_t n) {
  std::vector<float> s(std::max<int64_t>(n - 1, 0));
  for (int64_t i = 0; i + 1 < n; i++)
    s[i] = std::max(coupling_ratio(gb[i * 3 + 2], gb[i * 3 + 1]),
                    coupling_ratio(gb[(i + 1) * 3], gb[(i + 1) * 3 + 1]));
  return s;
}

constexpr int64_t kCutWindow = 16;

// Part boundaries [0, e_1, ..., n) of an n-row system for about R rows per
// part: e_k is the best-rated cut within kCutWindow rows of k R, rate(i)
// rating the boundary between rows i and i+1.
template <class Rate>
std::vector<int> part_cuts_by(Rate rate, int64_t n, int64_t R) {
  std::vector<int> cut{0};
  for (int64_t e0 = R; e0 < n; e0 += R) {
    const int64_t lo = std::max<int64_t>(e0 - kCutWindow, cut.back() + 1);
    const int64_t hi = std::min<int64_t>(e0 + kCutWindow, n - 1);
    if (lo > hi) continue;
    int64_t best = lo;
    float bs = rate(lo - 1);
    for (int64_t e = lo + 1; e <= hi; e++) {
      const float v = rate(e - 1);
      if (v < bs) {
        best = e;
        bs = v;
      }
    }
    cut.push_back((int)best);
  }
  cut.push_back((int)n);
  return cut;
}

std::vector<int> part_cuts(const std::vector<float> &score, int64_t n, int64_t R) {
  return part_cuts_by([&](int64_t i) { return score[i]; }, n, R);
}

// Boundaries shared by two circuits of na and nb rows cut at the same rows
// (a partitioned wave holding frequencies of both): a cut is rated by the
// worse of its two ratings; past a circuit's own rows (identity rows, no
// coupling) its rating is zero.
std::vector<int> part_cuts_pair(const std::vector<float> &sa, int64_t na,
                                const std::vector<float> &sb, int64_t nb, int64_t R) {
  return part_cuts_by(
      [&](int64_t i) {
        const float a = i + 1 < na ? sa[i] : 0.0f, b = i + 1 < nb ? sb[i] : 0.0f;
        return std::max(a, b);
      },
      std::max(na, nb), R);
}

void plan_release(ohm_ac_plan *p) {
  if (!p) return;
  if (backend_alive(p->b)) {
    (void)hipSetDevice(p->b->device);
    if (p->d_band) (void)hipFree(p->d_band);
    if (p->d_omega) (void)hipFree(p->d_omega);
    if (p->d_sing) (void)hipFree(p->d_sing);
    if (p->d_res) (void)hipFree(p->d_res);
    for (hipEvent_t e : p->ev)
      if (e) (void)hipEventDestroy(e);
  }
  delete p;
}

ohm_status plan_build(ohm_backend *b, int64_t n_circuits, const ohm_ac_circuit *circ,
                      std::vector<Prepared> &prep, ohm_ac_plan **out) {
  std::unique_ptr<ohm_ac_plan, void (*)(ohm_ac_plan *)> p(new ohm_ac_plan, plan_release);
  p->b = b;
  p->nc = n_circuits;
  p->n.resize(n_circuits);
  int band = 0;
  for (int64_t c = 0; c < n_circuits; c++) {
    p->n[c] = circ[c].n;
    if (circ[c].n == 0) continue;
    if (prep[c].band > kMaxDirectBand)
      return fail(OHM_ERR_INVALID,
                  "circuit %lld: half bandwidth %d after RCM exceeds the direct solver's %d",
                  (long long)c, prep[c].band, kMaxDirectBand);
    band = std::max(band, prep[c].band);
  }
  const int B = band <= 1 ? 1 : band <= 2 ? 2 : 4;
  const int W = 2 * B + 1;
  p->B = B;
  // one device allocation for every circuit's band data
  auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
  std::vector<size_t> off(n_circuits + 1, 0);
  for (int64_t c = 0; c < n_circuits; c++) {
    const size_t n = (size_t)circ[c].n;
    off[c + 1] = off[c] + 2 * al(sizeof(double) * n * W) + al(sizeof(double) * 2 * n) +
                 al(sizeof(uint32_t) * n);
  }
  OHM_HIP(hipMalloc(&p->d_band, std::max<size_t>(off[n_circuits], 256)));
  p->gb.assign(n_circuits, nullptr);
  p->cb.assign(n_circuits, nullptr);
  p->bp.assign(n_circuits, nullptr);
  p->bnorm2.assign(n_circuits, 0.0);
  p->perm.assign(n_circuits, nullptr);
  p->cut_score.assign(n_circuits, {});
  // band arrays in the RCM order, built on host threads 32 circuits at a
  // time and uploaded from this thread
  struct BandHost {
    std::vector<double> gb, cb, bp;
  };
  hipStream_t st = b->stream;
  for (int64_t c0 = 0; c0 < n_circuits; c0 += 32) {
    const int64_t cn = std::min<int64_t>(32, n_circuits - c0);
    std::vector<BandHost> hb(cn);
    parallel_for(cn, [&](int64_t t) {
      const int64_t c = c0 + t;
      const Prepared &pr = prep[c];
      const int64_t n = pr.n;
      BandHost &h = hb[t];
      h.gb.assign((size_t)n * W, 0.0);
      h.cb.assign((size_t)n * W, 0.0);
      h.bp.assign(2 * n, 0.0);
      for (int64_t r = 0; r < n; r++) {
        const int64_t i = pr.iperm[r];
        for (uint32_t k = pr.rp[r]; k < pr.rp[r + 1]; k++) {
          const int64_t d = (int64_t)pr.iperm[pr.cl[k]] - i + B;
          h.gb[i * W + d] += pr.g[k];
          h.cb[i * W + d] += pr.c[k];
        }
        h.bp[2 * i] = circ[c].b_ac[2 * r];
        h.bp[2 * i + 1] = circ[c].b_ac[2 * r + 1];
      }
      double b2 = 0.0;
      for (int64_t i = 0; i < 2 * n; i++) b2 += h.bp[i] * h.bp[i];
      p->bnorm2[c] = b2;
      if (B == 1) p->cut_score[c] = cut_scores(h.gb, n);
    });
    for (int64_t t = 0; t < cn; t++) {
      const int64_t c = c0 + t;
      const size_t n = (size_t)circ[c].n;
      if (n == 0) continue;
      char *base = (char *)p->d_band + off[c];
      const size_t gsz = al(sizeof(double) * n * W);
      double *dg = (double *)base, *dc = (double *)(base + gsz),
             *db = (double *)(base + 2 * gsz);
      uint32_t *dp = (uint32_t *)(base + 2 * gsz + al(sizeof(double) * 2 * n));
      OHM_HIP(hipMemcpyAsync(dg, hb[t].gb.data(), sizeof(double) * n * W, hipMemcpyHostToDevice, st));
      OHM_HIP(hipMemcpyAsync(dc, hb[t].cb.data(), sizeof(double) * n * W, hipMemcpyHostToDevice, st));
      OHM_HIP(hipMemcpyAsync(db, hb[t].bp.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice, st));
      OHM_HIP(hipMemcpyAsync(dp, prep[c].perm.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, st));
      p->gb[c] = dg;
      p->cb[c] = dc;
      p->bp[c] = db;
      p->perm[c] = dp;
    }
    // the host band arrays of this group are freed on return
    OHM_HIP(hipStreamSynchronize(st));
  }
  *out = p.release();
  return OHM_OK;
}


// Device-memory budget of one chunk of waves: OHM_AC_DIRECT_BUDGET_GB, or
// 90 % of what is free (counting the workspace the backend already holds).
double direct_budget(ohm_backend *b) {
  if (const char *env = getenv("OHM_AC_DIRECT_BUDGET_GB")) return atof(env) * 1e9;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 64e9;
  return 0.9 * (double)(fr + b->ac_ws_bytes);
}

// residual ||b - A x|| / ||b|| above which a partitioned system is solved
// again in one lane
constexpr double kPartResidMax = 1e-12;
// rows per part when neither the environment nor the autotuner sets it
constexpr int64_t kDefaultPartRows = 2048;

ohm_status ensure_events(ohm_ac_plan *p) {
  for (auto &e : p->ev)
    if (!e) OHM_HIP(hipEventCreate(&e));
  return OHM_OK;
}

// One numeric sweep with rows-per-part R (0: one lane per whole system).
ohm_status plan_solve_once(ohm_ac_plan *p, int64_t n_freqs, double *const *x, int64_t R) {
  ohm_backend *b = p->b;
  hipStream_t st = b->stream;
  const int64_t nc = p->nc;
  const int B = p->B, W = 2 * B + 1;
  const bool part = B == 1 && R > 0;
  p->resident = false;
  p->h_final = false;
  p->waves.clear();
  p->fallback_systems = 0;
  for (double &t : p->phase_ms) t = 0.0;
  const size_t n_sys = (size_t)std::max<int64_t>(nc * n_freqs, 1);
  OHM_HIP(hipMemsetAsync(p->d_sing, 0, sizeof(int) * n_sys, st));
  OHM_HIP(hipMemsetAsync(p->d_res, 0, sizeof(double) * 2 * n_sys, st));
  // band waves (segments): up to 64 frequencies of one circuit, longest
  // circuits first. One lane per system: a wave each. Partitioned: a wave of
  // 64 lanes takes the next segment of the next circuit once one circuit's
  // frequencies run out, up to two segments (OHM_AC_PACK=0: one), and the
  // circuits take their frequencies upwards and downwards in turn, so that
  // a shared wave holds the same end of the band of both: the part solve's
  // pivot choice, which its lanes share on most steps when their
  // frequencies are close, depends on the frequency.
  struct WaveItem {
    int64_t c, f0;
    int nf, lane0;  // lane0: the segment's first lane in its wave
  };
  struct Pack {
    size_t first;  // its first segment in items
    int count, nf;
  };
  std::vector<int64_t> order;
  for (int64_t c = 0; c < nc; c++)
    if (p->n[c] > 0) order.push_back(c);
  std::stable_sort(order.begin(), order.end(),
                   [&](int64_t a, int64_t b2) { return p->n[a] > p->n[b2]; });
  const char *pack_env = getenv("OHM_AC_PACK");
  const bool pack = part && !(pack_env && atoi(pack_env) == 0);
  std::vector<WaveItem> items;
  std::vector<Pack> packs;
  bool down = false;
  for (int64_t c : order) {
    for (int64_t done = 0; done < n_freqs;) {
      const bool join = pack && !packs.empty() && packs.back().count == 1 && packs.back().nf < 64;
      const int lane0 = join ? packs.back().nf : 0;
      const int take = (int)std::min<int64_t>(64 - lane0, n_freqs - done);
      items.push_back({c, down ? n_freqs - done - take : done, take, lane0});
      if (join) {
        packs.back().count++;
        packs.back().nf += take;
      } else {
        packs.push_back({items.size() - 1, 1, take});
      }
      done += take;
    }
    if (pack) down = !down;
  }
  const double budget = direct_budget(b);
  bool want_x = false;
  if (x)
    for (int64_t c = 0; c < nc; c++) want_x |= x[c] != nullptr;
  size_t pos = 0;
  int chunks = 0;
  auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
  // part boundaries per circuit (the same for all of its frequencies)
  if (part && p->cuts_R != R) {
    p->cuts.assign(nc, {});
    parallel_for(nc, [&](int64_t c) {
      if (p->n[c] > 0) p->cuts[c] = part_cuts(p->cut_score[c], p->n[c], R);
    });
    p->cuts_R = R;
  }
  const std::vector<std::vector<int>> &cuts = p->cuts;
  // boundaries of the waves holding two circuits, cut at the same rows
  if (pack && p->pair_key != std::make_pair(R, n_freqs)) {
    p->pair_cuts.assign(packs.size(), {});
    parallel_for((int64_t)packs.size(), [&](int64_t k) {
      const Pack &pk = packs[k];
      if (pk.count < 2) return;
      const int64_t ca = items[pk.first].c, cb2 = items[pk.first + 1].c;
      p->pair_cuts[k] = part_cuts_pair(p->cut_score[ca], p->n[ca], p->cut_score[cb2], p->n[cb2], R);
    });
    p->pair_key = std::make_pair(R, n_freqs);
  }
  auto cuts_of = [&](size_t k) -> const std::vector<int> & {
    const Pack &pk = packs[k];
    return pk.count == 2 ? p->pair_cuts[k] : cuts[items[pk.first].c];
  };
  auto rows_of = [&](size_t k) {
    const Pack &pk = packs[k];
    int64_t n = 0;
    for (int e = 0; e < pk.count; e++) n = std::max(n, p->n[items[pk.first + e].c]);
    return n;
  };
  while (pos < packs.size()) {
    size_t end = pos;
    double bytes = 0.0;
    while (end < packs.size()) {
      const double n = (double)rows_of(end);
      const double nfw = packs[end].nf;
      const double np = part ? (double)cuts_of(end).size() - 1 : 0.0;
      // U rows and x per lane-row; per part: its reference, relations,
      // reduced rows, boundary values and residual partials per lane
      // (and the checkpoints of the top-down elimination: 4 per kCk rows)
      const double wave_b = n * nfw * 16.0 * (W + 1) + (want_x ? n * nfw * 16.0 : 0.0) +
                            2.0 * packs[end].count * sizeof(BandWave) + sizeof(PartWave) +
                            np * (sizeof(PartRef) + 64.0 * 16.0 * (6.0 + 12.0 + 2.0 + 1.0)) +
                            (part ? (n / kCk + np) * 4.0 * 64.0 * 16.0 : 0.0);
      if (end > pos && bytes + wave_b > budget) break;
      bytes += wave_b;
      end++;
    }
    // this chunk: waves [pos, end), segments [s0, s0 + n_w)
    const size_t s0 = packs[pos].first;
    const int n_w = (int)(packs[end - 1].first + packs[end - 1].count - s0);
    const int n_pw = (int)(end - pos);
    // output in (circuit, frequency) order so each circuit's block is contiguous
    std::vector<size_t> by_cf(n_w);
    std::iota(by_cf.begin(), by_cf.end(), 0);
    std::stable_sort(by_cf.begin(), by_cf.end(), [&](size_t a, size_t b2) {
      const WaveItem &u = items[s0 + a], &v = items[s0 + b2];
      return u.c != v.c ? u.c < v.c : u.f0 < v.f0;
    });
    std::vector<int64_t> out_off(n_w);
    std::vector<int64_t> f_lo(nc, INT64_MAX), f_hi(nc, -1), x_at(nc, 0);
    int64_t out_total = 0;
    for (size_t k : by_cf) {
      const WaveItem &it = items[s0 + k];
      out_off[k] = out_total;
      if (it.f0 < f_lo[it.c]) {
        f_lo[it.c] = it.f0;
        x_at[it.c] = out_total;
      }
      f_hi[it.c] = std::max<int64_t>(f_hi[it.c], it.f0 + it.nf - 1);
      out_total += p->n[it.c] * it.nf;
    }
    std::vector<BandWave> hw(n_w);
    std::vector<PartWave> hpw(part ? n_pw : 0);
    std::vector<PartRef> prefs;
    int64_t store = 0;  // lane-rows
    int64_t n_ck = 0;   // checkpoints of the top-down elimination
    for (int kw = 0; kw < n_pw; kw++) {
      const Pack &pk = packs[pos + kw];
      const int64_t rows = rows_of(pos + kw);
      const int ka = (int)(pk.first - s0);
      for (int e = 0; e < pk.count; e++) {
        const int k = ka + e;
        const WaveItem &it = items[s0 + k];
        BandWave w{};
        w.n = p->n[it.c];
        w.gb = p->gb[it.c];
        w.cb = p->cb[it.c];
        w.b = p->bp[it.c];
        w.perm = p->perm[it.c];
        w.omega = p->d_omega + it.f0;
        w.store_off = store;
        w.out_off = out_off[k];
        w.nf = it.nf;
        w.sys = (int)(it.c * n_freqs + it.f0);
        w.stride = pk.nf;
        w.lane0 = it.lane0;
        w.bnorm2 = p->bnorm2[it.c];
        hw[k] = w;
      }
      if (part) {
        const std::vector<int> &ct = cuts_of(pos + kw);
        PartWave v{};
        v.seg_a = ka;
        v.seg_b = ka + pk.count - 1;
        v.split = pk.count == 2 ? items[pk.first].nf : pk.nf;
        v.nf = pk.nf;
        v.pbase = (int)prefs.size();
        v.nparts = (int)ct.size() - 1;
        v.n = rows;
        v.store_off = store;
        for (int q = 0; q < v.nparts; q++) {
          const int mq = ct[q + 1] - ct[q];
          prefs.push_back({kw, ct[q], mq, (int)n_ck});
          n_ck += (mq - 1 + kCk - 1) / kCk;
        }
        hpw[kw] = v;
      }
      store += rows * pk.nf;
    }
    const size_t n_ip = prefs.size();
    // chunk buffers carved from the backend's workspace (kept across sweeps:
    // re-allocating tens of GB per chunk costs seconds)
    const size_t bw_b = al(sizeof(BandWave) * n_w), u_b = al(sizeof(cz) * store * W),
                 y_b = al(sizeof(cz) * store),
                 x_b = want_x ? al(sizeof(cz) * std::max<int64_t>(out_total, 1)) : 0,
                 ip_b = al(sizeof(PartRef) * std::max<size_t>(n_ip, 1)),
                 rel_b = al(sizeof(cz) * n_ip * 6 * 64),
                 red_b = al(sizeof(cz) * n_ip * 12 * 64),
                 bnd_b = al(sizeof(cz) * n_ip * 2 * 64),
                 pres_b = al(sizeof(double) * n_ip * 64),
                 ck_b = part ? al(sizeof(cz) * std::max<int64_t>(n_ck, 1) * 4 * 64) : 0,
                 pw_b = part ? al(sizeof(PartWave) * n_pw) : 0;
    // (partitioned: a second wave table for the one-lane redo after pres)
    const size_t need = bw_b + u_b + y_b + x_b +
                        (part ? ip_b + rel_b + red_b + bnd_b + pres_b + bw_b + ck_b + pw_b : 0);
    if (b->ac_ws_bytes < need) {
      if (b->ac_ws) OHM_HIP(hipFree(b->ac_ws));
      b->ac_ws = nullptr;
      b->ac_ws_bytes = 0;
      OHM_HIP(hipMalloc(&b->ac_ws, need));
      b->ac_ws_bytes = need;
    }
    b->ac_ws_gen++;
    char *wsp = (char *)b->ac_ws;
    BandWave *d_w = (BandWave *)wsp;
    cz *dU = (cz *)(wsp + bw_b);
    cz *dY = (cz *)(wsp + bw_b + u_b);  // x in RCM order, lanes interleaved
    cz *dX = want_x ? (cz *)(wsp + bw_b + u_b + y_b) : nullptr;
    char *pw = wsp + bw_b + u_b + y_b + x_b;
    PartRef *d_parts = (PartRef *)pw;
    cz *d_rel = (cz *)(pw + ip_b), *d_red = (cz *)(pw + ip_b + rel_b),
       *d_bnd = (cz *)(pw + ip_b + rel_b + red_b);
    double *d_pres = (double *)(pw + ip_b + rel_b + red_b + bnd_b);
    // after the redo wave table (ip_b + ... + pres_b + bw_b)
    cz *d_ck = part ? (cz *)(pw + ip_b + rel_b + red_b + bnd_b + pres_b + bw_b) : nullptr;
    PartWave *d_pw = part ? (PartWave *)(pw + ip_b + rel_b + red_b + bnd_b + pres_b + bw_b + ck_b)
                          : nullptr;
    OHM_HIP(hipMemcpyAsync(d_w, hw.data(), sizeof(BandWave) * n_w, hipMemcpyHostToDevice, st));
    if (part) {
      OHM_HIP(hipMemcpyAsync(d_parts, prefs.data(), sizeof(PartRef) * n_ip,
                             hipMemcpyHostToDevice, st));
      OHM_HIP(hipMemcpyAsync(d_pw, hpw.data(), sizeof(PartWave) * n_pw, hipMemcpyHostToDevice, st));
    }
    auto one_lane = [&](int nw, const BandWave *dw) -> ohm_status {
      if (B == 1)
        hipLaunchKernelGGL(k_band_lu<1>, dim3(nw), dim3(64), 0, st, dw, dU, dY, p->d_sing, p->d_res);
      else if (B == 2)
        hipLaunchKernelGGL(k_band_lu<2>, dim3(nw), dim3(64), 0, st, dw, dU, dY, p->d_sing, p->d_res);
      else
        hipLaunchKernelGGL(k_band_lu<4>, dim3(nw), dim3(64), 0, st, dw, dU, dY, p->d_sing, p->d_res);
      OHM_HIP(hipGetLastError());
      b->dispatches++;
      return OHM_OK;
    };
    hipEvent_t *ev = p->ev;
    OHM_HIP(hipEventRecord(ev[0], st));
    if (part) {
      hipLaunchKernelGGL(k_part_rel, dim3((unsigned)(2 * n_ip)), dim3(64), 0, st, d_w, d_pw,
                         d_parts, d_rel, p->d_sing, d_ck);
      OHM_HIP(hipGetLastError());
      OHM_HIP(hipEventRecord(ev[1], st));
      hipLaunchKernelGGL(k_part_reduce, dim3(n_pw), dim3(64), 0, st, d_w, d_pw, d_rel, d_red,
                         d_bnd, p->d_sing);
      OHM_HIP(hipGetLastError());
      OHM_HIP(hipEventRecord(ev[2], st));
      hipLaunchKernelGGL(k_part_solve, dim3((unsigned)n_ip), dim3(64), 0, st, d_w, d_pw,
                         d_parts, d_bnd, d_ck, dY, p->d_sing, d_pres);
      OHM_HIP(hipGetLastError());
      OHM_HIP(hipEventRecord(ev[3], st));
      hipLaunchKernelGGL(k_part_resid_sum, dim3((unsigned)((n_pw * 64 + 255) / 256)), dim3(256),
                         0, st, d_w, d_pw, n_pw, d_parts, d_pres, dY, p->d_res);
      OHM_HIP(hipGetLastError());
      b->dispatches += 4;
    } else {
      OHM_HIP(hipEventRecord(ev[1], st));
      OHM_HIP(hipEventRecord(ev[2], st));
      OHM_TRY(one_lane(n_w, d_w));
      OHM_HIP(hipEventRecord(ev[3], st));
    }
    OHM_HIP(hipEventRecord(ev[4], st));
    OHM_HIP(hipEventSynchronize(ev[4]));
    float ms[4] = {0.f, 0.f, 0.f, 0.f};
    for (int k = 0; k < 4; k++) OHM_HIP(hipEventElapsedTime(&ms[k], ev[k], ev[k + 1]));
    p->phase_ms[0] += ms[0];
    p->phase_ms[1] += ms[1];
    p->phase_ms[2] += ms[2];
    p->phase_ms[3] += ms[3];
    if (part) {
      // systems whose partitioned residual is not small: their waves again, one lane each
      std::vector<int> &sing = p->h_sing;
      std::vector<double> &res = p->h_res;
      sing.resize(n_sys);
      res.resize(2 * n_sys);
      OHM_HIP(hipMemcpy(sing.data(), p->d_sing, sizeof(int) * n_sys, hipMemcpyDeviceToHost));
      OHM_HIP(hipMemcpy(res.data(), p->d_res, sizeof(double) * 2 * n_sys, hipMemcpyDeviceToHost));
      p->h_final = pos == 0 && end == packs.size();
      std::vector<BandWave> redo;
      for (int k = 0; k < n_w; k++) {
        const BandWave &w = hw[k];
        bool again = false;
        for (int l = 0; l < w.nf; l++) {
          const size_t q = (size_t)w.sys + l;
          const double rr = res[2 * q], bb = res[2 * q + 1];
          const double r = bb > 0 ? sqrt(rr / bb) : 0.0;
          if (sing[q] != 0 || !(r <= kPartResidMax)) again = true;
        }
        if (again) {
          redo.push_back(w);
          for (int l = 0; l < w.nf; l++) sing[(size_t)w.sys + l] = 0;
          p->fallback_systems += w.nf;
        }
      }
      if (!redo.empty()) {
        p->h_final = false;
        OHM_HIP(hipMemcpyAsync(p->d_sing, sing.data(), sizeof(int) * n_sys, hipMemcpyHostToDevice, st));
        // wave table of the redone items in its own region (the chunk's
        // table stays for get_x)
        BandWave *d_redo = (BandWave *)(pw + ip_b + rel_b + red_b + bnd_b + pres_b);
        OHM_HIP(hipMemcpyAsync(d_redo, redo.data(), sizeof(BandWave) * redo.size(),
                               hipMemcpyHostToDevice, st));
        OHM_HIP(hipEventRecord(ev[0], st));
        OHM_TRY(one_lane((int)redo.size(), d_redo));
        OHM_HIP(hipEventRecord(ev[1], st));
        OHM_HIP(hipEventSynchronize(ev[1]));
        float fm = 0.f;
        OHM_HIP(hipEventElapsedTime(&fm, ev[0], ev[1]));
        p->phase_ms[4] += fm;
      }
    }
    if (want_x) {
      OHM_HIP(hipEventRecord(ev[0], st));
      hipLaunchKernelGGL(k_band_unpermute, dim3(n_w * 64), dim3(256), 0, st, d_w, dY, dX);
      OHM_HIP(hipGetLastError());
      OHM_HIP(hipEventRecord(ev[1], st));
      b->dispatches++;
      for (int64_t c = 0; c < nc; c++) {
        if (f_hi[c] < 0 || !x[c]) continue;
        const int64_t n = p->n[c];
        OHM_HIP(hipMemcpyAsync(x[c] + f_lo[c] * 2 * n, dX + x_at[c],
                               sizeof(cz) * n * (f_hi[c] - f_lo[c] + 1), hipMemcpyDeviceToHost, st));
      }
      OHM_HIP(hipStreamSynchronize(st));
      float um = 0.f;
      OHM_HIP(hipEventElapsedTime(&um, ev[0], ev[1]));
      p->phase_ms[5] += um;
    }
    b->readbacks++;
    chunks++;
    if (end == packs.size() && chunks == 1) {
      p->waves = std::move(hw);
      p->y_off = bw_b + u_b;
    }
    pos = end;
  }
  // solutions of a one-chunk sweep stay in the workspace for ohm_ac_plan_get_x
  p->resident = chunks == 1;
  p->ws_gen = b->ac_ws_gen;
  p->chunks = chunks;
  p->part_rows = part ? R : 0;
  return OHM_OK;
}

// rows per part: OHM_AC_PART_ROWS, else the autotuned value of this plan
// (the fastest of the candidates, each timed over two whole sweeps on the
// plan's first solve), else 2048
ohm_status plan_solve(ohm_ac_plan *p, int64_t n_freqs, const double *freqs, double *const *x,
                      ohm_batch_result *results, double *device_ms) {
  ohm_backend *b = p->b;
  OHM_HIP(hipSetDevice(b->device));
  hipStream_t st = b->stream;
  const int64_t nc = p->nc;
  OHM_TRY(ensure_events(p));
  p->nf = n_freqs;
  const size_t n_sys = (size_t)std::max<int64_t>(nc * n_freqs, 1);
  if (p->sys_cap < n_sys) {
    if (p->d_sing) OHM_HIP(hipFree(p->d_sing));
    if (p->d_res) OHM_HIP(hipFree(p->d_res));
    p->d_sing = nullptr;
    p->d_res = nullptr;
    p->sys_cap = 0;
    OHM_HIP(hipMalloc((void **)&p->d_sing, sizeof(int) * n_sys));
    OHM_HIP(hipMalloc((void **)&p->d_res, sizeof(double) * 2 * n_sys));
    p->sys_cap = n_sys;
  }
  const size_t n_om = (size_t)std::max<int64_t>(n_freqs, 1);
  if (p->omega_cap < n_om) {
    if (p->d_omega) OHM_HIP(hipFree(p->d_omega));
    p->d_omega = nullptr;
    p->omega_cap = 0;
    OHM_HIP(hipMalloc((void **)&p->d_omega, sizeof(double) * n_om));
    p->omega_cap = n_om;
  }
  std::vector<double> omega(n_om, 0.0);
  for (int64_t f = 0; f < n_freqs; f++) omega[f] = 2.0 * M_PI * freqs[f];  // ac.rs:57
  OHM_HIP(hipMemcpyAsync(p->d_omega, omega.data(), sizeof(double) * n_om, hipMemcpyHostToDevice, st));
  int64_t R = kDefaultPartRows;
  if (const char *env = getenv("OHM_AC_PART_ROWS")) {
    R = atoll(env);
  } else if (p->tuned_rows >= 0) {
    R = p->tuned_rows;
  } else if (b->autotune && p->B == 1) {
    // each candidate timed twice (the faster of the two counts), so that a
    // sweep slowed by its first use of a workspace size does not decide
    const int64_t cand[] = {0, 1024, 2048, 3072, 4096, 8192};
    double best = 1e300;
    for (int64_t c : cand) {
      double tc = 1e300;
      for (int rep = 0; rep < 2; rep++) {
        OHM_TRY(plan_solve_once(p, n_freqs, nullptr, c));
        double t = 0.0;
        for (double v : p->phase_ms) t += v;
        tc = std::min(tc, t);
      }
      if (tc < best) {
        best = tc;
        R = c;
      }
    }
    p->tuned_rows = R;
  }
  OHM_TRY(plan_solve_once(p, n_freqs, x, R));
  if (!p->h_final) {
    p->h_sing.assign(n_sys, 0);
    p->h_res.assign(2 * n_sys, 0.0);
    OHM_HIP(hipMemcpy(p->h_sing.data(), p->d_sing, sizeof(int) * n_sys, hipMemcpyDeviceToHost));
    OHM_HIP(hipMemcpy(p->h_res.data(), p->d_res, sizeof(double) * 2 * n_sys, hipMemcpyDeviceToHost));
  }
  const std::vector<int> &sing = p->h_sing;
  const std::vector<double> &res = p->h_res;
  ohm_status first = OHM_OK;
  std::string first_msg;
  for (int64_t c = 0; c < nc; c++) {
    for (int64_t f = 0; f < n_freqs; f++) {
      const int64_t k = c * n_freqs + f;
      ohm_batch_result &r = results[k];
      r.iterations = 0;
      r.refinements = 0;
      if (p->n[c] == 0) {
        r.residual_norm = 0.0;
        r.status = OHM_OK;
        r.reason = 1;
        continue;
      }
      const double rr = res[2 * k], bb = res[2 * k + 1];
      r.residual_norm = bb > 0 ? sqrt(rr / bb) : 0.0;
      const bool bad = sing[k] != 0 || !std::isfinite(r.residual_norm);
      r.status = bad ? OHM_ERR_SOLVE : OHM_OK;
      r.reason = bad ? 0 : 1;
      if (bad && first == OHM_OK) {
        first = fail(OHM_ERR_SOLVE,
                     "Solve error: banded LU: singular matrix (circuit %lld, frequency %lld)",
                     (long long)c, (long long)f);
        first_msg = ohm_last_error();
      }
    }
  }
  double total = 0.0;
  for (int k = 0; k < 5; k++) total += p->phase_ms[k];
  if (device_ms) *device_ms = total;
  if (first != OHM_OK) return fail(first, "%s", first_msg.c_str());
  return OHM_OK;
}

// x of lanes [l0, l1) of one wave back to the original unknown order:
// out[(lane - l0) n + perm[i]] = x_i
__global__ void k_band_gather_x(BandWave q, int l0, int l1, const cz *__restrict__ Y,
                                cz *__restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t span = l1 - l0;
  if (t >= q.n * span) return;
  const int64_t i = t / span;
  const int lane = l0 + (int)(t % span);
  out[(int64_t)(lane - l0) * q.n + q.perm[i]] = Y[q.store_off + q.lane0 + i * q.stride + lane];
}

}  // namespace

extern "C" {

ohm_status ohm_ac_plan_create(ohm_backend *b, int64_t n_circuits, const ohm_ac_circuit *circ,
                              ohm_ac_plan **out) {
  if (!b || (n_circuits > 0 && !circ) || !out) return fail(OHM_ERR_INVALID, "null argument");
  if (n_circuits < 0) return fail(OHM_ERR_INVALID, "negative size");
  *out = nullptr;
  OHM_TRY(check_circuits(n_circuits, circ));
  OHM_HIP(hipSetDevice(b->device));
  std::vector<Prepared> prep(n_circuits);
  std::vector<ohm_status> pst(n_circuits, OHM_OK);
  std::vector<std::string> pmsg(n_circuits);
  parallel_for(n_circuits, [&](int64_t c) { pst[c] = prepare_circuit(circ[c], prep[c], pmsg[c]); });
  for (int64_t c = 0; c < n_circuits; c++)
    if (pst[c] != OHM_OK) {
      set_error(pmsg[c]);
      return pst[c];
    }
  return plan_build(b, n_circuits, circ, prep, out);
}

ohm_status ohm_ac_plan_solve(ohm_ac_plan *p, int64_t n_freqs, const double *freqs,
                             double *const *x, ohm_batch_result *results, double *device_ms) {
  if (!p || (n_freqs > 0 && !freqs) || !results) return fail(OHM_ERR_INVALID, "null argument");
  if (n_freqs < 0) return fail(OHM_ERR_INVALID, "negative size");
  if (!backend_alive(p->b)) return fail(OHM_ERR_INVALID, "the plan's backend was destroyed");
  return plan_solve(p, n_freqs, freqs, x, results, device_ms);
}

ohm_status ohm_ac_plan_get_x(ohm_ac_plan *p, int64_t circuit, int64_t f_first, int64_t f_count,
                             double *x) {
  if (!p || (f_count > 0 && !x)) return fail(OHM_ERR_INVALID, "null argument");
  if (!backend_alive(p->b)) return fail(OHM_ERR_INVALID, "the plan's backend was destroyed");
  if (circuit < 0 || circuit >= p->nc) return fail(OHM_ERR_INVALID, "circuit %lld out of range", (long long)circuit);
  if (f_first < 0 || f_count < 0 || f_first + f_count > p->nf)
    return fail(OHM_ERR_INVALID, "frequencies [%lld, %lld) out of range (the last solve had %lld)",
                (long long)f_first, (long long)(f_first + f_count), (long long)p->nf);
  ohm_backend *b = p->b;
  if (!p->resident || p->ws_gen != b->ac_ws_gen)
    return fail(OHM_ERR_INVALID,
                "the last solve's solutions are not resident (it ran in several chunks, or "
                "another sweep has used the workspace since): pass x to ohm_ac_plan_solve");
  const int64_t n = p->n[circuit];
  if (n == 0 || f_count == 0) return OHM_OK;
  OHM_HIP(hipSetDevice(b->device));
  hipStream_t st = b->stream;
  cz *d_out = nullptr;
  OHM_HIP(hipMalloc((void **)&d_out, sizeof(cz) * n * f_count));
  const cz *dY = (const cz *)((const char *)b->ac_ws + p->y_off);
  ohm_status s = OHM_OK;
  for (const BandWave &w : p->waves) {
    const int64_t c = w.sys / p->nf, f0 = w.sys % p->nf;
    if (c != circuit) continue;
    const int64_t lo = std::max<int64_t>(f0, f_first), hi = std::min<int64_t>(f0 + w.nf, f_first + f_count);
    if (lo >= hi) continue;
    const int64_t cnt = n * (hi - lo);
    hipLaunchKernelGGL(k_band_gather_x, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st, w,
                       (int)(lo - f0), (int)(hi - f0), dY, d_out + (lo - f_first) * n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      s = hip_fail(e, "k_band_gather_x", __FILE__, __LINE__);
      break;
    }
  }
  if (s == OHM_OK) {
    hipError_t e = hipMemcpyAsync(x, d_out, sizeof(cz) * n * f_count, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) s = hip_fail(e, "get_x copy", __FILE__, __LINE__);
  }
  (void)hipFree(d_out);
  return s;
}

ohm_status ohm_ac_plan_stats(const ohm_ac_plan *p, double *phase_ms, int64_t *rows_per_part,
                             int64_t *fallback_systems, int32_t *chunks) {
  if (!p) return fail(OHM_ERR_INVALID, "null plan");
  if (phase_ms)
    for (int k = 0; k < 6; k++) phase_ms[k] = p->phase_ms[k];
  if (rows_per_part) *rows_per_part = p->part_rows;
  if (fallback_systems) *fallback_systems = p->fallback_systems;
  if (chunks) *chunks = p->chunks;
  return OHM_OK;
}

void ohm_ac_plan_free(ohm_ac_plan *p) { plan_release(p); }

ohm_status ohm_ac_sweep_circuits(ohm_backend *b, int64_t n_circuits, const ohm_ac_circuit *circ,
                                 int64_t n_freqs, const double *freqs, int32_t method,
                                 const ohm_bicgstab_params *params, int32_t precond,
                                 ohm_batch_result *results, double *device_ms) {
  if (!b || (n_circuits > 0 && !circ) || (n_freqs > 0 && !freqs) || !results)
    return fail(OHM_ERR_INVALID, "null argument");
  if (n_circuits < 0 || n_freqs < 0) return fail(OHM_ERR_INVALID, "negative size");
  if (method < OHM_AC_AUTO || method > OHM_AC_DIRECT) return fail(OHM_ERR_INVALID, "unknown AC method %d", method);
  OHM_TRY(check_circuits(n_circuits, circ));
  OHM_HIP(hipSetDevice(b->device));
  if (device_ms) *device_ms = 0.0;
  // symbolic preparation per circuit on host threads; inputs are checked
  // first so errors are deterministic
  std::vector<Prepared> prep(n_circuits);
  std::vector<ohm_status> pst(n_circuits, OHM_OK);
  std::vector<std::string> pmsg(n_circuits);
  if (method != OHM_AC_ITERATIVE)
    parallel_for(n_circuits, [&](int64_t c) { pst[c] = prepare_circuit(circ[c], prep[c], pmsg[c]); });
  std::vector<char> direct(n_circuits, 0);
  for (int64_t c = 0; c < n_circuits; c++) {
    if (pst[c] != OHM_OK) {
      set_error(pmsg[c]);
      return pst[c];
    }
    if (circ[c].n == 0) continue;
    direct[c] = method == OHM_AC_ITERATIVE ? 0 : (prep[c].band <= kMaxDirectBand);
    if (method == OHM_AC_DIRECT && !direct[c])
      return fail(OHM_ERR_INVALID,
                  "circuit %lld: half bandwidth %d after RCM exceeds the direct solver's %d",
                  (long long)c, prep[c].band, kMaxDirectBand);
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
  // direct circuits: one plan over them, solved once
  std::vector<int64_t> dl;
  for (int64_t c = 0; c < n_circuits; c++)
    if (direct[c]) dl.push_back(c);
  if (!dl.empty()) {
    std::vector<ohm_ac_circuit> dc(dl.size());
    std::vector<Prepared> dp(dl.size());
    std::vector<double *> dx(dl.size());
    for (size_t k = 0; k < dl.size(); k++) {
      dc[k] = circ[dl[k]];
      dp[k] = std::move(prep[dl[k]]);
      dx[k] = circ[dl[k]].x;
    }
    ohm_ac_plan *p = nullptr;
    OHM_TRY(plan_build(b, (int64_t)dl.size(), dc.data(), dp, &p));
    p->tuned_rows = kDefaultPartRows;  // a one-off sweep is not worth tuning
    std::vector<ohm_batch_result> dr(dl.size() * std::max<int64_t>(n_freqs, 1));
    double ms = 0.0;
    ohm_status s = plan_solve(p, n_freqs, freqs, dx.data(), dr.data(), &ms);
    std::string msg = s != OHM_OK ? ohm_last_error() : std::string();
    plan_release(p);
    if (s != OHM_OK && s != OHM_ERR_SOLVE) return fail(s, "%s", msg.c_str());
    if (s != OHM_OK) {
      set_error(msg);
      note(s);
    }
    for (size_t k = 0; k < dl.size(); k++)
      for (int64_t f = 0; f < n_freqs; f++) results[dl[k] * n_freqs + f] = dr[k * n_freqs + f];
    if (device_ms) *device_ms = ms;
  }
  if (first != OHM_OK) return fail(first, "%s", first_msg.c_str());
  return OHM_OK;
}

}  // extern "C"

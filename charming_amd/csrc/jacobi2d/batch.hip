// Batched and served forms of the Jacobi2D solver: many independent
// problems in one persistent launch of a task engine (the model is described
// in engine.h), plus the plain per-request form they are compared with.
//
//   chm_jacobi2d_batch_run       many uneven problems, one launch
//   chm_jacobi2d_batch_plain     the same problems one after another, one
//                                launch per sweep (or a hipGraph per problem)
//   chm_jacobi2d_batch_autotune  workgroup size, sweep shape, tile per problem
//   chm_jacobi2d_server_*        a request stream into the running engine
//                                (continuous batching) or into the plain form
//
// Arithmetic is the shared sweep of stencil.h, so every problem's grids are
// bit-identical whatever form, tiling or mix it ran in.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include <algorithm>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "engine_dev.h"
#include "jacobi2d_internal.h"

namespace chm {
namespace eng {

using jac::EdgeOut;
using jac::kColOff;

// ---------------------------------------------------------------------------
// task bodies (workgroup-collective)
// ---------------------------------------------------------------------------

// u_0 value of array cell (i, j): seeded interior, ring, 0 outside the grid.
__device__ inline double u0_cell(const Request& r, int64_t i, int64_t j) {
  if (i < 0 || i > r.H + 1 || j < 0 || j > r.W + 1) return 0.0;
  const bool interior = i >= 1 && i <= r.H && j >= 1 && j <= r.W;
  if (!interior) return (r.init_mode == CHM_INIT_REFERENCE) ? 10.0 : r.boundary;
  return (r.init_mode == CHM_INIT_REFERENCE) ? 0.0 : jac::cell_u01(r.seed, i, j);
}

// Seed a tile's cells, and at the grid's edge the `depth` halo rings beside
// it, in every array.
__device__ void do_init(const Request& r, const TileGeom& g) {
  const int d = r.depth;
  const int i0 = g.r0 - (g.r0 == 1 ? d : 0), i1 = g.r1 + (g.r1 == r.H + 1 ? d : 0);
  const int j0 = g.c0 - (g.c0 == 1 ? d : 0), j1 = g.c1 + (g.c1 == r.W + 1 ? d : 0);
  const int64_t cols = j1 - j0, n = (int64_t)(i1 - i0) * cols;
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x) {
    const int64_t i = i0 + e / cols, j = j0 + e % cols;
    const double v = u0_cell(r, i, j);
    const int64_t o = i * r.pitch + kColOff + j;
    r.a[o] = v;
    r.b[o] = v;
    if (r.c) r.c[o] = v;
  }
}

// Waves of the workgroup split a tile into panels of 128*NC columns x bands.
struct TileSplit {
  int S, nb, br, pw;
};
template <int NC>
__device__ inline TileSplit split_tile(const TileGeom& g) {
  TileSplit t;
  const int nW = blockDim.x / 64;
  t.pw = 128 * NC;
  t.S = (g.c1 - g.c0 + t.pw - 1) / t.pw;
  t.nb = nW / t.S;
  if (t.nb < 1) t.nb = 1;
  if (t.nb > g.r1 - g.r0) t.nb = g.r1 - g.r0;
  t.br = (g.r1 - g.r0 + t.nb - 1) / t.nb;
  return t;
}

// Workgroup-collective: the exact sum of every lane's u128 partials (waves
// reduce across lanes, then thread 0 across waves in LDS) goes to the
// problem's accumulator with one 128-bit atomic add per task, not one per
// wave: thousands of tiles add into the same two words every round.
__device__ void add_residual(uint64_t* slot, jac::u128acc acc) {
  __shared__ uint64_t part[16][2];
  acc = jac::wave_reduce_u128(acc);
  const int wave = threadIdx.x / 64, n_waves = blockDim.x / 64;
  if (__lane_id() == 0) {
    part[wave][0] = acc.lo;
    part[wave][1] = acc.hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    jac::u128acc t = {0, 0};
    for (int w = 0; w < n_waves; w++) jac::acc_add(t, jac::u128acc{part[w][0], part[w][1]});
    if (t.lo | t.hi) jac::atomic_add_u128(slot, t);
  }
  __syncthreads();
}

// Workgroup-collective: the FP64 sum of every lane's v added to *slot
// with one atomic per task.
__device__ void add_residual_d(double* slot, double v) {
  __shared__ double part_d[16];
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  const int wave = threadIdx.x / 64, n_waves = blockDim.x / 64;
  if (__lane_id() == 0) part_d[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < n_waves; w++) t += part_d[w];
    if (t != 0.0)
      __hip_atomic_fetch_add(slot, t, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
}

// One sweep of a tile from array `src` into array `dst`.
template <int NC, int U, bool RES>
__device__ void do_sweep(const Request& r, const TileGeom& g, int src, int dst,
                         uint64_t* res_slot) {
  const double* in = arr(r, src);
  double* out = arr(r, dst);
  const TileSplit ts = split_tile<NC>(g);
  const int nW = blockDim.x / 64, wave = threadIdx.x / 64;
  jac::u128acc acc = {0, 0};
  const EdgeOut eo = {nullptr, nullptr, nullptr, nullptr};
  for (int t = wave; t < ts.S * ts.nb; t += nW) {
    const int panel = t % ts.S, b = t / ts.S;
    const int rr0 = g.r0 + b * ts.br, rr1 = min(rr0 + ts.br, g.r1);
    jac::sweep_panel<NC, U, RES, false>(in, out, r.pitch, r.W, r.H,
                                        g.c0 + panel * ts.pw, rr0, rr1, acc,
                                        eo, g.c1 - 1);
  }
  if (RES) add_residual(res_slot, acc);
}

// Two fused sweeps of a tile from array `src` into array `dst`; RESA /
// RESB: residuals of the first / second sweep into res_slot[0..1] / [2..3].
template <int NC, int U, bool RESA, bool RESB>
__device__ void do_round2(const Request& r, const TileGeom& g, int src,
                          int dst, uint64_t* res_slot) {
  const double* in = arr(r, src);
  double* out = arr(r, dst);
  const TileSplit ts = split_tile<NC>(g);
  const int nW = blockDim.x / 64, wave = threadIdx.x / 64;
  jac::u128acc acc_a = {0, 0}, acc_b = {0, 0};
  const jac::EdgeOut2 eo = {nullptr, nullptr, nullptr, nullptr};
  const jac::RingSides rs = {true, true, true, true};  // the grid's ring
  for (int t = wave; t < ts.S * ts.nb; t += nW) {
    const int panel = t % ts.S, b = t / ts.S;
    const int rr0 = g.r0 + b * ts.br, rr1 = min(rr0 + ts.br, g.r1);
    jac::sweep_panel2<NC, U, RESA, RESB, false>(
        in, out, r.pitch, r.W, r.H, g.c0 + panel * ts.pw, rr0, rr1, acc_a,
        acc_b, eo, rs, g.c1 - 1);
  }
  if (RESA) add_residual(res_slot, acc_a);
  if (RESB) add_residual(res_slot + 2, acc_b);
}

// D = 3 .. kDeep (or kDeepWide) sweeps of a tile from array `src` into
// array `dst` in one pass (the deep pass of the chare program: overlapping
// 128-column windows, each wave one contiguous share of the tile's windows
// x rows).  FULL: all D sweeps (no stage copies, a compile-time fact);
// otherwise D - n_sw leading stages copy.  RES: the last sweep's residual
// into res_slot[2..3].  T (the engine's workgroup size) gives each engine
// its own copy, compiled for that workgroup's register budget: sixteen
// sweeps need the 256 registers of workgroups of 512 threads or fewer.
// Kept out of line: its register rings would otherwise weigh on the
// engine's task loop.
constexpr int kDeep = 8;       // fixed sweep counts, 1024-thread workgroups
constexpr int kDeepWide = 16;  // fixed sweep counts, up to 512 threads
constexpr int kWideRing = 6;   // input-ring slots of the sixteen-deep pass

__host__ __device__ constexpr int deep_for_threads(int threads) {
  return threads <= 512 ? kDeepWide : kDeep;
}

template <int T, int D, bool RES, bool FULL>
__device__ __noinline__ void do_round_deep(const double* in, double* out,
                                           int64_t pitch, int W, int H,
                                           int r0, int r1, int c0, int c1,
                                           int n_sw, uint64_t* res_slot) {
  in = uni(in);
  out = uni(out);
  pitch = uni(pitch);
  W = uni(W);
  H = uni(H);
  r0 = uni(r0);
  r1 = uni(r1);
  c0 = uni(c0);
  c1 = uni(c1);
  n_sw = uni(n_sw);
  constexpr int OW = jac::window_out_cols(D);
  const int rows = r1 - r0, S = (c1 - c0 + OW - 1) / OW;
  const int WV = blockDim.x / 64, wave = threadIdx.x / 64;
  const int64_t total = (int64_t)S * rows;
  const int64_t per = (total + WV - 1) / WV;
  int64_t pos = (int64_t)wave * per;
  const int64_t end = min(total, pos + per);
  jac::u128acc acc = {0, 0};
  const jac::EdgeOut2 none = {nullptr, nullptr, nullptr, nullptr};
  const jac::RingSides rs = {true, true, true, true};  // the grid's ring
  while (pos < end) {
    const int panel = (int)(pos / rows), rr = (int)(pos % rows);
    const int len = (int)min((int64_t)(rows - rr), end - pos);
    jac::sweep_panelD<D, RES, false, 0, D == kDeepWide ? kWideRing : 3>(
        in, out, pitch, W, H, c0 + panel * OW, r0 + rr, r0 + rr + len, acc,
        none, rs, c1 - 1, FULL ? 0 : D - n_sw);
    pos += len;
  }
  if (RES) add_residual(res_slot + 2, acc);
}

// Rounds of tolerance problems: 3 .. D sweeps of a tile in one pass (D -
// n_sw leading stages copy; D = kDeepTol, or kDeepTolWide in engines of up
// to 512-thread workgroups, each engine with its own copy as for
// do_round_deep).  MODE 1: the exact residual of every sweep j = 1 .. n_sw
// into res_slot[2(j-1) .. 2(j-1)+1] (the stop rule needs each one); MODE 2:
// the FP64 sum of its squared changes into resd_slot[j-1] instead (the stop
// filter); MODE 0: none (the repair of a stop inside a round).
constexpr int kDeepTol = 6;
constexpr int kDeepTolWide = 12;

__host__ __device__ constexpr int deep_tol_for_threads(int threads) {
  return threads <= 512 ? kDeepTolWide : kDeepTol;
}

// FULL: n_sw == D, known at compile time, so no stage carries a copy test.
template <int T, int D, int MODE, bool FULL>
__device__ __noinline__ void do_round_tol(const double* in, double* out,
                                          int64_t pitch, int W, int H, int r0,
                                          int r1, int c0, int c1, int n_sw,
                                          uint64_t* res_slot,
                                          double* resd_slot) {
  in = uni(in);
  out = uni(out);
  pitch = uni(pitch);
  W = uni(W);
  H = uni(H);
  r0 = uni(r0);
  r1 = uni(r1);
  c0 = uni(c0);
  c1 = uni(c1);
  n_sw = uni(n_sw);
  constexpr int OW = jac::window_out_cols(D);
  const int rows = r1 - r0, S = (c1 - c0 + OW - 1) / OW;
  const int WV = blockDim.x / 64, wave = threadIdx.x / 64;
  const int64_t total = (int64_t)S * rows;
  const int64_t per = (total + WV - 1) / WV;
  int64_t pos = (int64_t)wave * per;
  const int64_t end = min(total, pos + per);
  jac::u128acc accs[D];
  double accd[D];
#pragma unroll
  for (int s = 0; s < D; s++) {
    accs[s] = {0, 0};
    accd[s] = 0.0;
  }
  const jac::EdgeOut2 none = {nullptr, nullptr, nullptr, nullptr};
  const jac::RingSides rs = {true, true, true, true};  // the grid's ring
  const int n_copy = FULL ? 0 : D - n_sw;
  while (pos < end) {
    const int panel = (int)(pos / rows), rr = (int)(pos % rows);
    const int len = (int)min((int64_t)(rows - rr), end - pos);
    // (the twelve-deep filtered rounds, most of a request's, keep four rows
    // of loads in flight: 141 against 125 solves/s at saturation)
    jac::sweep_panelD<D, MODE == 1, false, MODE,
                      (D == kDeepTolWide && MODE == 2) ? 6 : 3>(
        in, out, pitch, W, H, c0 + panel * OW, r0 + rr, r0 + rr + len,
        accs[D - 1], none, rs, c1 - 1, n_copy, accs, accd);
    pos += len;
  }
  if (MODE == 1) {
#pragma unroll
    for (int s = 0; s < D; s++)
      if (s >= n_copy) add_residual(res_slot + 2 * (s - n_copy), accs[s]);
  } else if (MODE == 2) {
#pragma unroll
    for (int s = 0; s < D; s++)
      if (s >= n_copy) add_residual_d(resd_slot + (s - n_copy), accd[s]);
  }
}

// do_round_tol at the problem's round depth (kDeepTol, or kDeepTolWide in
// engines of up to 512-thread workgroups).
template <int T, int MODE>
__device__ inline void round_tol(int depth, const double* in, double* out,
                                 int64_t pitch, int W, int H, int r0, int r1,
                                 int c0, int c1, int n_sw, uint64_t* res_slot,
                                 double* resd_slot) {
  if constexpr (deep_tol_for_threads(T) == kDeepTolWide) {
    if (depth == kDeepTolWide) {
      if (n_sw == kDeepTolWide)
        do_round_tol<T, kDeepTolWide, MODE, true>(in, out, pitch, W, H, r0, r1,
                                                  c0, c1, n_sw, res_slot,
                                                  resd_slot);
      else
        do_round_tol<T, kDeepTolWide, MODE, false>(in, out, pitch, W, H, r0,
                                                   r1, c0, c1, n_sw, res_slot,
                                                   resd_slot);
      return;
    }
  }
  if (n_sw == kDeepTol)
    do_round_tol<T, kDeepTol, MODE, true>(in, out, pitch, W, H, r0, r1, c0, c1,
                                          n_sw, res_slot, resd_slot);
  else
    do_round_tol<T, kDeepTol, MODE, false>(in, out, pitch, W, H, r0, r1, c0,
                                           c1, n_sw, res_slot, resd_slot);
}

// Hash of the tile's interior cells of array `src`.
__device__ uint64_t do_hash(const Request& r, const TileGeom& g, int src) {
  const double* u = arr(r, src);
  const int64_t cols = g.c1 - g.c0, n = (int64_t)(g.r1 - g.r0) * cols;
  uint64_t h = 0;
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x) {
    const int64_t i = g.r0 + e / cols, j = g.c0 + e % cols;
    h += jac::hash_cell(i, j, u[i * r.pitch + kColOff + j]);
  }
  for (int off = 32; off >= 1; off >>= 1) h += __shfl_down(h, off, 64);
  return h;
}

// ---------------------------------------------------------------------------
// completion logic
// ---------------------------------------------------------------------------
enum : int {
  F_GATE = 1,      // round rr+2 of every tile may go
  F_FINALIZE = 2,
  F_PUBLISH = 4,
  F_FIX = 8,
  F_RESID = 16,    // recompute round rr with exact residuals (every tile);
                   // this tile's release of round rr+1 waits for it
  F_HELD = 32      // the held tile's release of round rr+1 may go
};

__host__ __device__ inline double u128d(uint64_t lo, uint64_t hi) {
  return ldexp((double)hi, 64) + (double)lo;
}

// Wave 0: decrement the counters (word `word`) of `tile` and its up to
// eight neighbours, one lane each so the atomics overlap; a lane that takes
// a counter to zero queues that round task (counted in flight first).
__device__ void release_deps(const EngineDesc& E, Prob& p, const Request& r,
                             uint32_t slot, uint32_t tile, int word) {
  const int lane = __lane_id();
  const int ti = (int)tile / r.tc, tj = (int)tile % r.tc;
  int target = -1;
  if (lane < 9) {
    const int i = ti + lane / 3 - 1, j = tj + lane % 3 - 1;  // lane 4: the tile
    if (i >= 0 && i < r.tr && j >= 0 && j < r.tc) target = i * r.tc + j;
  }
  if (target >= 0 && asub(&r.pend[3 * target + word], 1u) == 1) {
    aadd(&p.inflight, 1u);
    push_task(E, task_word(K_ITER, slot, (uint32_t)target),
              (int)((blockIdx.x + (uint32_t)lane) % (uint32_t)E.n_wg));
  }
}

// Rounds of a tolerance problem whose residuals are FP64 sums (a filter):
// every round but the first (it sets S_0) and the last (the cap's residual
// is the result's) of a problem with deep rounds.
__device__ inline bool filtered_round(const Request& r, int rr) {
  return r.tol2 >= 0 && r.depth >= 3 && rr > 0 && rr + 1 < r.rounds;
}

// Thread 0: whether a filtered round may hold the stop.  The exact residual
// of sweep j is S_j = sum floor(fl(d^2) 2^56) over the cells (saturating at
// 2^64 - 1), so S_j >= 2^56 T_j - N with T_j = sum fl(d^2) when no term
// saturates, and the FP64 sum A_j of the same fl(d^2) (N <= 2^30 terms, any
// order) is within a factor 1 + 2^-23 of T_j.  A_j 2^56 (1 - 2^-20) - N >
// thr (1 + 2^-20) therefore proves S_j > thr; a saturated term makes S_j >=
// 2^64 - 1, more than any thr below 2^62.
__device__ bool filter_may_stop(Prob& p, const Request& r, int sk, int n_sw) {
  const double thr = r.tol2 * u128d(ald(&p.s0[0]), ald(&p.s0[1]));
  if (!(thr < 0x1.0p62)) return true;
  const double n = (double)r.W * (double)r.H;
  for (int j = 0; j < n_sw; j++) {
    const double a = __longlong_as_double((long long)ald((uint64_t*)&p.resd[sk][j]));
    if (!(a * 0x1.0p56 * (1.0 - 0x1.0p-20) - n > thr * (1.0 + 0x1.0p-20)))
      return true;
  }
  return false;
}

// Thread 0: the decision on round rr from exact residuals (state 0, or 5
// when they were recomputed).  Stop after the first sweep k with S_k <=
// tol^2 * S_0 (the rule and the doubles of the monolithic tolerance solve):
// if k is the round's last sweep its output is the result; else the round
// stored a later grid, and every tile redoes the round's first sweeps up to
// k from the intact input (K_FIX tasks).  Returns F_GATE when round rr+2
// may go.
__device__ int decide_exact(Prob& p, const Request& r, int rr, int sk) {
  const bool tol = r.tol2 >= 0;
  const bool last = rr + 1 == r.rounds;
  const int n_sw = min(r.depth, r.max_iters - r.depth * rr);
  int stop = 0;  // sweeps in the result when stopping
  uint64_t slo = 0, shi = 0;
  if (!tol) {
    // a pass of 2+ sweeps adds its last sweep's residual to the second slot
    const int j = n_sw >= 2 ? 1 : 0;
    stop = r.max_iters;
    slo = ald(&p.res[sk][j][0]);
    shi = ald(&p.res[sk][j][1]);
  } else {
    // the residual of every sweep of the round, in order
    uint64_t s0lo, s0hi;
    if (rr == 0) {
      s0lo = ald(&p.res[sk][0][0]);
      s0hi = ald(&p.res[sk][0][1]);
      ast(&p.s0[0], s0lo);
      ast(&p.s0[1], s0hi);
    } else {
      s0lo = ald(&p.s0[0]);
      s0hi = ald(&p.s0[1]);
    }
    const double thr = r.tol2 * u128d(s0lo, s0hi);
    for (int j = 0; j < n_sw; j++) {
      const uint64_t lo = ald(&p.res[sk][j][0]), hi = ald(&p.res[sk][j][1]);
      if (u128d(lo, hi) <= thr || (last && j == n_sw - 1)) {
        stop = r.depth * rr + j + 1;
        slo = lo;
        shi = hi;
        if (j + 1 < n_sw) {
          // the round stored the grid after n_sw sweeps: the repair redoes
          // the first j+1 from the round's intact input
          ast(&p.fix_sweeps, (uint32_t)(j + 1));
          ast(&p.fix_round, (uint32_t)rr);
        }
        break;
      }
    }
  }
  if (stop) {
    ast(&p.s_last[0], slo);
    ast(&p.s_last[1], shi);
    ast(&p.stop_iters, (uint32_t)stop);
    ast(&p.state, 1u);
  } else {
    ast(&p.round_tiles[sk], 0u);
    for (int i = 0; i < n_sw; i++) {
      ast(&p.res[sk][i][0], 0ull);
      ast(&p.res[sk][i][1], 0ull);
      ast((uint64_t*)&p.resd[sk][i], 0ull);
    }
    ast(&p.state, 0u);
  }
  release_agent();
  return (!stop && rr + 2 < r.rounds) ? F_GATE : 0;
}

// Thread 0: count the tile in round rr; the last tile of a round whose
// residual matters decides whether the problem stops.  Returns F_GATE when
// the decision lets round rr+2 of every tile go (its gate is then dropped
// by the whole workgroup), F_RESID when a filtered round may hold the stop:
// every tile then recomputes it with exact residuals (its input is intact:
// round rr+2 waits for this decision) and the decision is made on those,
// while this tile's release of round rr+1 waits (so no later round can be
// decided first).
//
// With a tolerance, round rr+2 waits for the decision on round rr, and the
// decision comes before the release of round rr+1 (engine_kernel), so
// decisions are made in order and round rr's input (array rr % 3) is intact
// when it is decided even if round rr+1 is running.
__device__ int decide_round(Prob& p, const Request& r, int rr) {
  const bool tol = r.tol2 >= 0;
  const bool last = rr + 1 == r.rounds;
  if (!tol && !last) return 0;
  const uint32_t n_tiles = (uint32_t)(r.tr * r.tc);
  const int sk = tol ? (rr & 1) : 0;
  if (aadd(&p.round_tiles[sk], 1u) != n_tiles - 1) return 0;
  // every tile has finished round rr: its residuals are complete
  if (ald(&p.state) != 0) return 0;  // stopped on the previous round
  if (filtered_round(r, rr)) {
    if (filter_may_stop(p, r, sk, r.depth)) {
      ast(&p.resid_round, (uint32_t)rr);
      ast(&p.resid_tiles, 0u);
      ast(&p.state, 5u);
      aadd(&p.inflight, n_tiles);  // the recompute tasks
      release_agent();
      return F_RESID;
    }
    ast(&p.round_tiles[sk], 0u);
    for (int i = 0; i < r.depth; i++) ast((uint64_t*)&p.resd[sk][i], 0ull);
    release_agent();
    return rr + 2 < r.rounds ? F_GATE : 0;
  }
  return decide_exact(p, r, rr, sk);
}

// Thread 0: a task of the problem is finished.  The last task of a stopped
// problem starts the repair sweep (F_FIX) or the hash tasks (F_FINALIZE);
// the last repair task starts the hash tasks; the last hash task publishes
// the completion (F_PUBLISH).
__device__ int finish_task(Prob& p, uint32_t n_tiles) {
  if (asub(&p.inflight, 1u) != 1) return 0;
  acquire_agent();
  const uint32_t st = ald(&p.state);
  if (st == 1 || st == 4) {
    const bool fix = st == 1 && ald(&p.fix_round) != 0xffffffffu;
    ast(&p.state, fix ? 4u : 2u);
    aadd(&p.inflight, n_tiles);
    release_agent();
    return fix ? F_FIX : F_FINALIZE;
  }
  return st == 2 ? F_PUBLISH : 0;
}

// Thread 0: write the completion record; `valid` goes last.  After this the
// slot belongs to the host again.
__device__ void publish(const EngineDesc& E, Prob& p, uint32_t slot) {
  const uint64_t t_done = wall_clock64();
  const uint64_t idx = aadd(&E.ctl[CW_DONE], 1ull);
  Completion* c = &E.dones[idx & (E.req_cap - 1)];
  const uint64_t ticket = p.r.ticket;
  const uint64_t v[8] = {ticket,             ald(&p.hash),  ald(&p.s_last[0]),
                         ald(&p.s_last[1]),  ald(&p.s0[0]), ald(&p.s0[1]),
                         ald(&p.t_admit),    t_done};
  uint64_t* dst = &c->ticket;  // the first eight words, in field order
  for (int i = 0; i < 8; i++)
    __hip_atomic_store(dst + i, v[i], __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&c->iters, ald(&p.stop_iters), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&c->slot, slot, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_SYSTEM);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __hip_atomic_store(&c->valid, ticket + 1, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_fetch_sub(&E.ctl[CW_ACTIVE], 1ull, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// admission of a request (workgroup-collective)
// ---------------------------------------------------------------------------
__device__ void admit(const EngineDesc& E, uint64_t idx, Request* s_req) {
  if (threadIdx.x == 0) {
    const volatile Request* vr = &E.reqs[idx & (E.req_cap - 1)];
    Request q;
    q.a = vr->a;
    q.b = vr->b;
    q.c = vr->c;
    q.pend = vr->pend;
    q.pitch = vr->pitch;
    q.W = vr->W;
    q.H = vr->H;
    q.bs = vr->bs;
    q.tr = vr->tr;
    q.tc = vr->tc;
    q.max_iters = vr->max_iters;
    q.init_on_device = vr->init_on_device;
    q.init_mode = vr->init_mode;
    q.boundary = vr->boundary;
    q.tol2 = vr->tol2;
    q.seed = vr->seed;
    q.ticket = vr->ticket;
    q.slot = vr->slot;
    q.rounds = vr->rounds;
    q.depth = vr->depth;
    *s_req = q;
    Prob& p = E.probs[q.slot & (kMaxSlots - 1)];
    p.r = q;
    ast(&p.inflight, (uint32_t)(q.tr * q.tc));
    ast(&p.state, 0u);
    ast(&p.stop_iters, 0u);
    ast(&p.fix_round, 0xffffffffu);
    ast(&p.fix_sweeps, 0u);
    for (int i = 0; i < 2; i++) {
      ast(&p.round_tiles[i], 0u);
      for (int j = 0; j < 8; j++) {
        ast(&p.res[i][j][0], 0ull);
        ast(&p.res[i][j][1], 0ull);
        ast((uint64_t*)&p.resd[i][j], 0ull);
      }
    }
    ast(&p.resid_round, 0xffffffffu);
    ast(&p.resid_tiles, 0u);
    ast(&p.held_tile, 0u);
    ast(&p.s0[0], 0ull);
    ast(&p.s0[1], 0ull);
    ast(&p.s_last[0], 0ull);
    ast(&p.s_last[1], 0ull);
    ast(&p.hash, 0ull);
    ast(&p.t_admit, (uint64_t)wall_clock64());
    aadd(&E.ctl[CW_ACTIVE], 1ull);
  }
  __syncthreads();
  const Request q = uni(*s_req);
  const uint32_t n_tiles = (uint32_t)(q.tr * q.tc);
  for (uint32_t u = threadIdx.x; u < n_tiles; u += blockDim.x) {
    uint32_t nb[8];
    const uint32_t deg = (uint32_t)tile_neighbors(q, u, nb);
    q.pend[3 * u + 0] = q.init_on_device ? 1 + deg : 0;  // round 0
    q.pend[3 * u + 1] = 1 + deg;                         // round 1
    q.pend[3 * u + 2] = 0;                               // rounds done
  }
  drain();
  __syncthreads();
  if (threadIdx.x == 0) release_agent();
  __syncthreads();
  const uint32_t kind = q.init_on_device ? K_INIT : K_ITER;
  for (uint32_t u = threadIdx.x; u < n_tiles; u += blockDim.x)
    push_task(E, task_word(kind, q.slot, u),
              (int)((blockIdx.x + u) % (uint32_t)E.n_wg));
}

// ---------------------------------------------------------------------------
// the engine kernel
// ---------------------------------------------------------------------------
enum : int { W_EXIT = 0, W_TASK = 1, W_ADMIT = 2 };

// Thread 0: find the next thing to do - a task from the own queue, a task
// stolen from a random victim, or a request to admit - or learn to exit.
__device__ int find_work(const EngineDesc& E, uint32_t* task, uint64_t* req,
                         uint32_t* rng) {
  const int me = (int)blockIdx.x;
  uint64_t* ctl = E.ctl;
  for (uint32_t spins = 0;; spins++) {
    if (q_pop(E, me, task)) {
      acquire_agent();
      return W_TASK;
    }
    for (int v = 0; v < 4; v++) {
      *rng = *rng * 1664525u + 1013904223u;
      const int victim = (int)((*rng >> 8) % (uint32_t)E.n_wg);
      if (victim != me && q_pop(E, victim, task)) {
        aadd(&ctl[CW_STATS + 1], 1ull);
        acquire_agent();
        return W_TASK;
      }
    }
    if (me < E.n_pollers && (spins & 3) == 0) {
      // mirror the host-written words into device memory
      const uint32_t st = *E.host_stop;
      if (E.host_tail) {
        const uint64_t t = *E.host_tail;
        __hip_atomic_fetch_max(&ctl[CW_REQ_TAIL], t, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
      }
      if (st) {
        drain();
        ast(&ctl[CW_STOP], (uint64_t)st);
      }
    }
    const uint64_t h = ald(&ctl[CW_REQ_HEAD]);
    const uint64_t t = ald(&ctl[CW_REQ_TAIL]);
    if (h < t && acas(&ctl[CW_REQ_HEAD], h, h + 1)) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
      *req = h;
      return W_ADMIT;
    }
    if (ald(&ctl[CW_STOP]) != 0) return W_EXIT;
    if (E.expected && ald(&ctl[CW_DONE]) >= E.expected) return W_EXIT;
    if ((spins & 63) == 63) aadd(&ctl[CW_STATS + 2], 1ull);
    __builtin_amdgcn_s_sleep(2);
  }
}

// Round rr of a tile: `depth` fused sweeps (or what is left in the last
// round), with the residuals the problem needs.
template <int T, int NC, int U>
__device__ void run_round(const Request& r, const TileGeom& g, int rr,
                          Prob& p) {
  const bool tol = r.tol2 >= 0;
  const bool last = rr + 1 == r.rounds;
  const int n_sw = min(r.depth, r.max_iters - r.depth * rr);
  uint64_t* slot = &p.res[tol ? (rr & 1) : 0][0][0];
  const int src = rr % 3, dst = (rr + 1) % 3;
  if (n_sw >= 3 && tol && filtered_round(r, rr)) {
    round_tol<T, 2>(r.depth, arr(r, src), arr(r, dst), r.pitch, r.W, r.H,
                    g.r0, g.r1,
                    g.c0, g.c1, n_sw, nullptr, &p.resd[rr & 1][0]);
  } else if (n_sw >= 3 && tol) {
    round_tol<T, 1>(r.depth, arr(r, src), arr(r, dst), r.pitch, r.W, r.H,
                    g.r0, g.r1,
                    g.c0, g.c1, n_sw, slot, nullptr);
  } else if (n_sw >= 3) {
    // every round but the last sweeps `depth` times; a shorter last round
    // runs the shallowest pass that holds it (eight stages, or sixteen),
    // its leading stages copying
    const double* in = arr(r, src);
    double* out = arr(r, dst);
    if constexpr (deep_for_threads(T) == kDeepWide) {
      if (n_sw > kDeep) {
        if (!last)
          do_round_deep<T, kDeepWide, false, true>(in, out, r.pitch, r.W,
                                                   r.H, g.r0, g.r1, g.c0,
                                                   g.c1, n_sw, slot);
        else if (n_sw == kDeepWide)
          do_round_deep<T, kDeepWide, true, true>(in, out, r.pitch, r.W, r.H,
                                                  g.r0, g.r1, g.c0, g.c1,
                                                  n_sw, slot);
        else
          do_round_deep<T, kDeepWide, true, false>(in, out, r.pitch, r.W,
                                                   r.H, g.r0, g.r1, g.c0,
                                                   g.c1, n_sw, slot);
        return;
      }
    }
    if (!last)
      do_round_deep<T, kDeep, false, true>(in, out, r.pitch, r.W, r.H, g.r0,
                                           g.r1, g.c0, g.c1, n_sw, slot);
    else if (n_sw == kDeep)
      do_round_deep<T, kDeep, true, true>(in, out, r.pitch, r.W, r.H, g.r0,
                                          g.r1, g.c0, g.c1, n_sw, slot);
    else
      do_round_deep<T, kDeep, true, false>(in, out, r.pitch, r.W, r.H, g.r0,
                                           g.r1, g.c0, g.c1, n_sw, slot);
  } else if (n_sw == 1) {
    if (tol || last)
      do_sweep<NC, U, true>(r, g, src, dst, slot);
    else
      do_sweep<NC, U, false>(r, g, src, dst, slot);
  } else if (tol) {
    do_round2<NC, U, true, true>(r, g, src, dst, slot);
  } else if (last) {
    do_round2<NC, U, false, true>(r, g, src, dst, slot);
  } else {
    do_round2<NC, U, false, false>(r, g, src, dst, slot);
  }
}

template <int T, int NC, int U>
__global__ void __launch_bounds__(T) engine_kernel(EngineDesc E) {
  __shared__ uint32_t s_task;
  __shared__ uint64_t s_req;
  __shared__ int s_cmd, s_k, s_n, s_flags;
  __shared__ Request s_r;
  uint32_t rng = blockIdx.x * 2654435761u + 12345u;
  for (;;) {
    if (threadIdx.x == 0) {
      uint32_t task = 0;
      uint64_t req = 0;
      s_cmd = find_work(E, &task, &req, &rng);
      s_task = task;
      s_req = req;
    }
    __syncthreads();
    const int cmd = uni(s_cmd);
    if (cmd == W_EXIT) break;
    if (cmd == W_ADMIT) {
      admit(E, uni(s_req), &s_r);
      __syncthreads();
      continue;
    }
    const uint32_t w = uni(s_task);
    const uint32_t kind = task_kind(w), slot = task_slot(w), tile = task_tile(w);
    Prob& p = E.probs[slot];
    if (threadIdx.x == 0) {
      s_r = p.r;
      int k = 0;
      if (kind == K_ITER) {
        k = (int)ald(&s_r.pend[3 * tile + 2]);
        // re-arm this parity's counter for round k+2: the tile's and its
        // neighbours' round k+1, and the stop decision on round k
        uint32_t nb[8];
        const uint32_t deg = (uint32_t)tile_neighbors(s_r, tile, nb);
        const uint32_t gate = s_r.tol2 >= 0 ? 1u : 0u;
        ast(&s_r.pend[3 * tile + (k & 1)], 1u + deg + gate);
      } else if (kind == K_HASH) {
        k = (int)ald(&p.stop_iters);
      } else if (kind == K_FIX) {
        // a repair (state 4) or, in state 5, the exact recompute of a
        // filtered round's residuals (n = -1)
        if (ald(&p.state) == 5) {
          k = (int)ald(&p.resid_round);
          s_n = -1;
        } else {
          k = (int)ald(&p.fix_round);
          s_n = (int)ald(&p.fix_sweeps);
        }
      }
      s_k = k;
    }
    __syncthreads();
    const Request r = uni(s_r);
    const int k = uni(s_k);
    const TileGeom g = tile_geom(r, tile);
    if (kind == K_INIT) {
      do_init(r, g);
    } else if (kind == K_ITER) {
      run_round<T, NC, U>(r, g, k, p);
    } else if (kind == K_FIX && uni(s_n) < 0) {
      // round k again from its intact input, now with exact residuals (the
      // grid it writes is bit-identical to what the round wrote)
      round_tol<T, 1>(r.depth, arr(r, k % 3), arr(r, (k + 1) % 3), r.pitch,
                      r.W, r.H,
                      g.r0, g.r1, g.c0, g.c1,
                      min(r.depth, r.max_iters - r.depth * k),
                      &p.res[k & 1][0][0], nullptr);
    } else if (kind == K_FIX) {
      // redo the first n sweeps of round k from its intact input
      const int n = uni(s_n);
      if (n == 1)
        do_sweep<NC, U, false>(r, g, k % 3, (k + 1) % 3, nullptr);
      else if (n == 2)
        do_round2<NC, U, false, false>(r, g, k % 3, (k + 1) % 3, nullptr);
      else
        round_tol<T, 0>(r.depth, arr(r, k % 3), arr(r, (k + 1) % 3),
                        r.pitch, r.W, r.H, g.r0, g.r1, g.c0, g.c1, n, nullptr,
                        nullptr);
    } else {
      // the result of `k` sweeps is in array ceil(k / depth) % 3
      const uint64_t h = do_hash(r, g, ((k + r.depth - 1) / r.depth) % 3);
      if (__lane_id() == 0 && h) aadd(&p.hash, h);
    }
    drain();
    __syncthreads();
    const uint32_t n_tiles = (uint32_t)(r.tr * r.tc);
    __shared__ uint32_t s_held;
    if (threadIdx.x == 0) {
      int f = 0;
      if (kind == K_ITER) ast(&r.pend[3 * tile + 2], (uint32_t)(k + 1));
      release_agent();
      if (kind == K_ITER) {
        f = decide_round(p, r, k);
        if (f & F_RESID) {
          ast(&p.held_tile, tile);
          release_agent();
        }
        const uint64_t sweeps =
            (uint64_t)min(r.depth, r.max_iters - r.depth * k);
        aadd(&E.ctl[CW_STATS + 3],
             sweeps * (uint64_t)(g.r1 - g.r0) * (uint64_t)(g.c1 - g.c0));
      } else if (kind == K_FIX && s_n < 0) {
        // the last tile of an exact recompute decides the round; without a
        // stop the held tile's release of round k+1 goes too
        if (aadd(&p.resid_tiles, 1u) == n_tiles - 1) {
          acquire_agent();
          f = decide_exact(p, r, k, k & 1);
          if (ald(&p.state) == 0) {
            f |= F_HELD;
            s_held = ald(&p.held_tile);
          }
        }
      }
      aadd(&E.ctl[CW_STATS + 0], 1ull);
      s_flags = f;
    }
    __syncthreads();
    const int fl = uni(s_flags);
    // after the decision: release the tile's and its neighbours' next round
    // (held back while its round is recomputed)
    if (threadIdx.x < 64) {
      const int word = kind == K_INIT ? 0
                       : (kind == K_ITER && k + 1 < r.rounds) ? ((k + 1) & 1)
                                                              : -1;
      if (word >= 0 && !(fl & F_RESID)) release_deps(E, p, r, slot, tile, word);
      if (fl & F_HELD)
        release_deps(E, p, r, slot, uni(s_held), (k + 1) & 1);
    }
    if (fl & F_RESID) {  // (counted in flight by the decision)
      for (uint32_t u = threadIdx.x; u < n_tiles; u += blockDim.x)
        push_task(E, task_word(K_FIX, slot, u),
                  (int)((blockIdx.x + u) % (uint32_t)E.n_wg));
    }
    if (fl & F_GATE) {
      // the decision on round k lets round k+2 (parity k & 1) go
      if (k + 2 < r.rounds) {
        for (uint32_t u = threadIdx.x; u < n_tiles; u += blockDim.x)
          if (asub(&r.pend[3 * u + (k & 1)], 1u) == 1) {
            aadd(&p.inflight, 1u);
            push_task(E, task_word(K_ITER, slot, u),
                      (int)((blockIdx.x + u) % (uint32_t)E.n_wg));
          }
      }
    }
    drain();  // the counts of queued tasks land before this task's own
    __syncthreads();
    if (threadIdx.x == 0) s_flags = finish_task(p, n_tiles);
    __syncthreads();
    const int f = uni(s_flags);
    if (f & (F_FINALIZE | F_FIX)) {
      const uint32_t kk = (f & F_FIX) ? K_FIX : K_HASH;
      for (uint32_t u = threadIdx.x; u < n_tiles; u += blockDim.x)
        push_task(E, task_word(kk, slot, u),
                  (int)((blockIdx.x + u) % (uint32_t)E.n_wg));
    } else if ((f & F_PUBLISH) && threadIdx.x == 0) {
      publish(E, p, slot);
    }
    __syncthreads();
  }
}

// Queue cells start with sequence number = index within the queue.
__global__ void init_queues(uint64_t* cells, uint64_t n_cells, uint32_t qcap) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
       i < n_cells; i += (uint64_t)gridDim.x * blockDim.x)
    cells[i] = (uint64_t)(i & (qcap - 1)) << 32;
}

// Host-launched seeding of a whole problem: u_0 in every array (rows
// 1-depth .. H+depth), padding 0.
__global__ void init_problem(Request r) {
  const int d = r.depth;
  const int64_t n = (int64_t)(r.H + 2 * d) * r.pitch;
  const int64_t o = (int64_t)(d - 1) * r.pitch;  // row 0 from the first row
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / r.pitch - (d - 1), j = e % r.pitch - kColOff;
    const double v = (j >= 1 - d && j <= r.W + d) ? u0_cell(r, i, j) : 0.0;
    r.a[e - o] = v;
    r.b[e - o] = v;
    if (r.c) r.c[e - o] = v;
  }
}

// ---------------------------------------------------------------------------
// plain per-request form: one launch per sweep over the whole problem
// ---------------------------------------------------------------------------
template <int NC, int U, bool RES>
__global__ void __launch_bounds__(256)
    req_step_kernel(const double* __restrict__ in, double* __restrict__ out,
                    int W, int H, int64_t P, int band, int n_panels,
                    int n_units, uint64_t* res) {
  const int unit = blockIdx.x * 4 + (int)(threadIdx.x / 64);
  if (unit >= n_units) return;
  const int panel = unit % n_panels, b = unit / n_panels;
  const int r0 = 1 + b * band, r1 = min(r0 + band, H + 1);
  jac::u128acc acc = {0, 0};
  const EdgeOut eo = {nullptr, nullptr, nullptr, nullptr};
  jac::sweep_panel<NC, U, RES, false>(in, out, P, W, H, 1 + panel * 128 * NC,
                                      r0, r1, acc, eo);
  if (RES) {
    acc = jac::wave_reduce_u128(acc);
    if (__lane_id() == 0 && (acc.lo | acc.hi)) jac::atomic_add_u128(res, acc);
  }
}

// Interior hash of one array added to out[0].
__global__ void hash_req_kernel(const double* u, int W, int H, int64_t P,
                                uint64_t* out) {
  const int64_t n = (int64_t)W * H;
  uint64_t h = 0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = 1 + e / W, j = 1 + e % W;
    h += jac::hash_cell(i, j, u[i * P + kColOff + j]);
  }
  for (int off = 32; off >= 1; off >>= 1) h += __shfl_down(h, off, 64);
  if (__lane_id() == 0 && h)
    __hip_atomic_fetch_add(out, h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A fill kernel rather than a memset, so it may sit in a captured graph.
__global__ void zero_words(uint64_t* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// ---------------------------------------------------------------------------
// host side: shared pieces
// ---------------------------------------------------------------------------
static size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }
static uint32_t next_pow2(uint64_t v) {
  uint64_t p = 1;
  while (p < v) p <<= 1;
  return (uint32_t)p;
}
static double now_s() {
  timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + ts.tv_nsec * 1e-9;
}

// Tile side for a problem: `block` (> 0), or a rule that keeps tiles large
// enough to amortise a task's fences and small enough to spread the problem,
// in multiples of the deep pass's window: 112 columns at eight sweeps
// (measured on the 512-problem batch: 116-118 ms, against 158 ms with sides
// 64/128/256/512), 96 at sixteen.
static int pick_block(int W, int H, int block, int deep) {
  if (block > 0) return block;
  const int side = std::max(W, H);
  const int ow = jac::window_out_cols(deep);
  if (side <= 256) return ow;
  if (side <= 1024) return 2 * ow;
  return 4 * ow;
}

struct Built {
  Request r;
  size_t grid_bytes, pend_bytes;
};

// deep: the rounds of fixed sweep counts (kDeep, or kDeepWide for engines
// of up to 512-thread workgroups).
static Built build_request(const chm_batch_problem& pb, int block, int deep) {
  Built b;
  memset(&b.r, 0, sizeof(b.r));
  Request& r = b.r;
  r.W = pb.width;
  r.H = pb.height;
  r.bs = pick_block(pb.width, pb.height, pb.block > 0 ? pb.block : block,
                    deep);
  r.tr = (r.H + r.bs - 1) / r.bs;
  r.tc = (r.W + r.bs - 1) / r.bs;
  r.max_iters = pb.iters;
  r.init_mode = pb.init_mode;
  r.boundary = pb.boundary;
  r.tol2 = pb.tol > 0 ? pb.tol * pb.tol : -1.0;
  r.seed = pb.seed;
  // rounds of `deep` (sixteen or eight) fused sweeps for a fixed sweep
  // count, twelve or six with a tolerance (the pass then sums every sweep's
  // residual, in registers); tiles and grid at least that many cells on a
  // side, else eight (six), else two
  const int side = std::min(std::min(r.W, r.H), r.bs);
  const int deep_tol = deep == kDeepWide ? kDeepTolWide : kDeepTol;
  if (r.tol2 >= 0)
    r.depth = side >= deep_tol ? deep_tol : (side >= kDeepTol ? kDeepTol : 2);
  else if (side >= deep)
    r.depth = deep;
  else
    r.depth = side >= kDeep ? kDeep : 2;
  r.rounds = (pb.iters + r.depth - 1) / r.depth;
  r.pitch = jac::pitch_for_depth(r.W, r.depth);  // halos `depth` deep
  b.grid_bytes = round_up(
      jac::array_elems_depth(r.W, r.H, r.depth) * sizeof(double), 256);
  b.pend_bytes = round_up((size_t)r.tr * r.tc * 3 * sizeof(uint32_t), 256);
  return b;
}

static int check_problem(const chm_batch_problem& p, int block, int deep) {
  if (p.width < 1 || p.height < 1 || p.iters < 1 || p.width > (1 << 20) ||
      p.height > (1 << 20))
    return fail(CHM_E_INVALID, "bad problem %dx%d iters %d", p.width,
                p.height, p.iters);
  if (p.init_mode != CHM_INIT_SEEDED && p.init_mode != CHM_INIT_REFERENCE)
    return fail(CHM_E_INVALID, "bad init mode %d", p.init_mode);
  const int bs =
      pick_block(p.width, p.height, p.block > 0 ? p.block : block, deep);
  if (bs < 16 || bs > 8192)
    return fail(CHM_E_INVALID, "tile side %d (16..8192)", bs);
  const int64_t tiles =
      (int64_t)((p.width + bs - 1) / bs) * ((p.height + bs - 1) / bs);
  if (tiles >= (int64_t)kMaxTiles)
    return fail(CHM_E_INVALID, "%dx%d with tiles of %d exceeds %u tiles",
                p.width, p.height, bs, kMaxTiles);
  return CHM_OK;
}

// ---------------------------------------------------------------------------
// host side: the engine
// ---------------------------------------------------------------------------
struct Engine {
  EngineDesc d;
  int threads = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<void*> dev;
  Request* h_reqs = nullptr;     // pinned request ring (serving)
  Completion* h_dones = nullptr; // pinned completion ring
  uint64_t* h_tail = nullptr;    // pinned (serving)
  uint32_t* h_stop = nullptr;    // pinned
  bool running = false;
  double wall_hz = 1e8;
};

template <class T>
static int dmalloc(Engine& e, T** p, size_t bytes) {
  void* q = nullptr;
  if (hipMalloc(&q, bytes) != hipSuccess)
    return fail(CHM_E_NOMEM, "engine allocation of %zu bytes", bytes);
  e.dev.push_back(q);
  *p = (T*)q;
  return CHM_OK;
}

static void engine_free(Engine& e) {
  for (void* p : e.dev) (void)hipFree(p);
  e.dev.clear();
  if (e.h_reqs) (void)hipHostFree(e.h_reqs);
  if (e.h_dones) (void)hipHostFree(e.h_dones);
  if (e.h_tail) (void)hipHostFree(e.h_tail);
  if (e.h_stop) (void)hipHostFree(e.h_stop);
  e.h_reqs = nullptr;
  e.h_dones = nullptr;
  e.h_tail = nullptr;
  e.h_stop = nullptr;
  if (e.ev0) (void)hipEventDestroy(e.ev0);
  if (e.ev1) (void)hipEventDestroy(e.ev1);
  if (e.stream) (void)hipStreamDestroy(e.stream);
  e.ev0 = e.ev1 = nullptr;
  e.stream = nullptr;
}

static int host_alloc(void** p, size_t bytes) {
  if (hipHostMalloc(p, bytes, hipHostMallocMapped | hipHostMallocCoherent) !=
      hipSuccess)
    return fail(CHM_E_NOMEM, "pinned allocation of %zu bytes", bytes);
  memset(*p, 0, bytes);
  return CHM_OK;
}

static int check_shape(int nc, int u) {
  if ((nc == 1 && (u == 2 || u == 4)) || (nc == 2 && u == 2)) return CHM_OK;
  return fail(CHM_E_INVALID,
              "sweep shape panel_chunks=%d rows_in_flight=%d (1x2, 1x4 or 2x2)",
              nc, u);
}

typedef void (*engine_fn_t)(EngineDesc);
static engine_fn_t pick_engine(int threads, int nc, int u) {
#define CHM_PICK(T_, NC_, U_)                         \
  if (threads == T_ && nc == NC_ && u == U_)          \
    return &engine_kernel<T_, NC_, U_>;
#define CHM_PICK3(T_) CHM_PICK(T_, 1, 2) CHM_PICK(T_, 1, 4) CHM_PICK(T_, 2, 2)
  CHM_PICK3(256) CHM_PICK3(512) CHM_PICK3(1024)
#undef CHM_PICK3
#undef CHM_PICK
  return nullptr;
}

// Queues, rings and control words.  qtasks bounds the tasks that can be
// queued at once (every tile of every problem in flight).  serving: the
// request ring and its tail live in pinned host memory.
static int engine_setup(Engine& e, int threads, int nc, int u, uint32_t req_cap,
                        uint64_t qtasks, bool serving) {
  if (threads != 256 && threads != 512 && threads != 1024)
    return fail(CHM_E_INVALID, "workgroup size %d (256, 512 or 1024)", threads);
  CHM_TRY(check_shape(nc, u));
  int dev = 0;
  CHM_TRY_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  CHM_TRY_HIP(hipGetDeviceProperties(&prop, dev));
  int khz = 100000;
  (void)hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
  e.wall_hz = khz * 1e3;
  e.threads = threads;
  EngineDesc& d = e.d;
  memset(&d, 0, sizeof(d));
  // as many workgroups as stay resident together (registers permitting)
  const engine_fn_t fn = pick_engine(threads, nc, u);
  int per_cu = 0;
  CHM_TRY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void*>(fn), threads, 0));
  if (per_cu < 1) return fail(CHM_E_HIP, "engine kernel does not fit a CU");
  d.n_wg = prop.multiProcessorCount * std::min(per_cu, 1024 / threads);
  d.n_pollers = std::min(d.n_wg, 16);
  d.nc = nc;
  d.u = u;
  const uint64_t per = (qtasks + d.n_wg - 1) / d.n_wg * 4 + 64;
  d.qcap = next_pow2(std::max<uint64_t>(per, 256));
  if (d.qcap > (1u << 20))
    return fail(CHM_E_INVALID, "%llu queued tasks exceed the queues",
                (unsigned long long)qtasks);
  d.req_cap = next_pow2(std::max<uint32_t>(req_cap, 16));
  CHM_TRY(dmalloc(e, &d.probs, sizeof(Prob) * kMaxSlots));
  CHM_TRY(dmalloc(e, &d.qcells, (size_t)d.n_wg * d.qcap * 8));
  CHM_TRY(dmalloc(e, &d.qhead, (size_t)d.n_wg * 128));
  CHM_TRY(dmalloc(e, &d.qtail, (size_t)d.n_wg * 128));
  CHM_TRY(dmalloc(e, &d.ctl, CW_WORDS * 8));
  CHM_TRY_HIP(hipMemset(d.probs, 0, sizeof(Prob) * kMaxSlots));
  CHM_TRY_HIP(hipMemset(d.qhead, 0, (size_t)d.n_wg * 128));
  CHM_TRY_HIP(hipMemset(d.qtail, 0, (size_t)d.n_wg * 128));
  CHM_TRY_HIP(hipMemset(d.ctl, 0, CW_WORDS * 8));
  init_queues<<<1024, 256>>>(d.qcells, (uint64_t)d.n_wg * d.qcap, d.qcap);
  CHM_TRY_HIP(hipGetLastError());
  CHM_TRY(host_alloc((void**)&e.h_dones, sizeof(Completion) * d.req_cap));
  CHM_TRY(host_alloc((void**)&e.h_stop, 64));
  void* p = nullptr;
  CHM_TRY_HIP(hipHostGetDevicePointer(&p, e.h_dones, 0));
  d.dones = (Completion*)p;
  CHM_TRY_HIP(hipHostGetDevicePointer(&p, e.h_stop, 0));
  d.host_stop = (volatile uint32_t*)p;
  if (serving) {
    CHM_TRY(host_alloc((void**)&e.h_reqs, sizeof(Request) * d.req_cap));
    CHM_TRY(host_alloc((void**)&e.h_tail, 64));
    CHM_TRY_HIP(hipHostGetDevicePointer(&p, e.h_reqs, 0));
    d.reqs = (const Request*)p;
    CHM_TRY_HIP(hipHostGetDevicePointer(&p, e.h_tail, 0));
    d.host_tail = (volatile uint64_t*)p;
  } else {
    Request* q = nullptr;
    CHM_TRY(dmalloc(e, &q, sizeof(Request) * d.req_cap));
    d.reqs = q;
  }
  CHM_TRY_HIP(hipStreamCreateWithFlags(&e.stream, hipStreamNonBlocking));
  CHM_TRY_HIP(hipEventCreate(&e.ev0));
  CHM_TRY_HIP(hipEventCreate(&e.ev1));
  CHM_TRY_HIP(hipDeviceSynchronize());
  return CHM_OK;
}

static int engine_launch(Engine& e) {
  const engine_fn_t fn = pick_engine(e.threads, e.d.nc, e.d.u);
  if (!fn) return fail(CHM_E_INVALID, "no engine for this shape");
  CHM_TRY_HIP(hipEventRecord(e.ev0, e.stream));
  hipLaunchKernelGGL(fn, dim3(e.d.n_wg), dim3(e.threads), 0, e.stream, e.d);
  CHM_TRY_HIP(hipGetLastError());
  CHM_TRY_HIP(hipEventRecord(e.ev1, e.stream));
  e.running = true;
  return CHM_OK;
}

// Blocks until the launch ends; tells it to stop after timeout_s.
static int engine_wait(Engine& e, double timeout_s, float* ms) {
  const double t0 = now_s();
  bool aborted = false;
  for (;;) {
    const hipError_t q = hipEventQuery(e.ev1);
    if (q == hipSuccess) break;
    if (q != hipErrorNotReady) {
      e.running = false;
      return fail(CHM_E_HIP, "engine kernel failed: %s", hipGetErrorString(q));
    }
    if (!aborted && timeout_s > 0 && now_s() - t0 > timeout_s) {
      __atomic_store_n(e.h_stop, 2u, __ATOMIC_SEQ_CST);
      aborted = true;
    }
    usleep(50);
  }
  e.running = false;
  CHM_TRY_HIP(hipStreamSynchronize(e.stream));
  if (ms) CHM_TRY_HIP(hipEventElapsedTime(ms, e.ev0, e.ev1));
  if (aborted) return fail(CHM_E_TIMEOUT, "engine stopped by the watchdog");
  return CHM_OK;
}

struct BatchOut {
  chm_batch_result* res;
  uint64_t* hashes;
  uint64_t* residuals;
  int* iters_done;
};

static void store_out(const BatchOut& out, int i, const Completion& c) {
  if (out.hashes) out.hashes[i] = c.hash;
  if (out.residuals) {
    out.residuals[2 * i] = c.s_last_lo;
    out.residuals[2 * i + 1] = c.s_last_hi;
  }
  if (out.iters_done) out.iters_done[i] = (int)c.iters;
}

// Lays out the grids of n problems in one arena.
// n_arr: 3 for the engine (rounds rotate over three arrays), 2 for the
// plain form.
static int layout_batch(const chm_batch_problem* probs, int n, int block,
                        int iters_cap, const int* blocks, int n_arr, int deep,
                        std::vector<Built>& bl, size_t* total, uint64_t* tiles) {
  bl.resize(n);
  *total = 0;
  *tiles = 0;
  for (int i = 0; i < n; i++) {
    chm_batch_problem pb = probs[i];
    if (iters_cap > 0) pb.iters = std::min(pb.iters, iters_cap);
    if (blocks && blocks[i] > 0) pb.block = blocks[i];
    CHM_TRY(check_problem(pb, block, deep));
    bl[i] = build_request(pb, block, deep);
    *total += n_arr * bl[i].grid_bytes + bl[i].pend_bytes;
    *tiles += (uint64_t)bl[i].r.tr * bl[i].r.tc;
  }
  return CHM_OK;
}

static void place_batch(std::vector<Built>& bl, char* arena, int n_arr) {
  size_t off = 0;
  for (size_t i = 0; i < bl.size(); i++) {
    Request& r = bl[i].r;
    r.a = (double*)(arena + off) + (r.depth - 1) * r.pitch;  // row 0
    off += bl[i].grid_bytes;
    r.b = (double*)(arena + off) + (r.depth - 1) * r.pitch;
    off += bl[i].grid_bytes;
    r.c = nullptr;
    if (n_arr == 3) {
      r.c = (double*)(arena + off) + (r.depth - 1) * r.pitch;
      off += bl[i].grid_bytes;
    }
    r.pend = (uint32_t*)(arena + off);
    off += bl[i].pend_bytes;
    r.init_on_device = 0;
    r.ticket = (uint64_t)i;
    r.slot = (uint32_t)i;
  }
}

// One engine launch over n problems (grids seeded before the launch).
// iters_cap > 0 caps every problem's sweeps (autotuner samples); blocks
// (optional) overrides the tile side per problem.
static int batch_once(const chm_batch_problem* probs, int n,
                      const chm_batch_config* cfg, int iters_cap,
                      const int* blocks, BatchOut out) {
  if (n < 1 || n > (int)kMaxSlots)
    return fail(CHM_E_INVALID, "%d problems (1..%u)", n, kMaxSlots);
  // (default 512 threads: sixteen sweeps per round, 105-108 ms for the
  // 512-problem batch at the size rule's tiles, against 123-124 ms at 1024
  // threads and eight sweeps per round, and 136 ms at 256 threads)
  const int threads = cfg && cfg->wg_threads ? cfg->wg_threads : 512;
  const int block = cfg ? cfg->block : 0;
  const int nc = cfg && cfg->panel_chunks ? cfg->panel_chunks : 1;
  const int u = cfg && cfg->rows_in_flight ? cfg->rows_in_flight : 2;
  std::vector<Built> bl;
  size_t total = 0;
  uint64_t tiles = 0;
  CHM_TRY(layout_batch(probs, n, block, iters_cap, blocks, 3,
                       deep_for_threads(threads), bl, &total, &tiles));
  Engine e;
  int rc = engine_setup(e, threads, nc, u, (uint32_t)n, tiles, false);
  char* arena = nullptr;
  if (rc == CHM_OK && hipMalloc(&arena, total) != hipSuccess)
    rc = fail(CHM_E_NOMEM, "batch grids: %zu bytes", total);
  if (rc == CHM_OK) {
    place_batch(bl, arena, 3);
    std::vector<Request> reqs(n);
    for (int i = 0; i < n; i++) {
      init_problem<<<512, 256, 0, e.stream>>>(bl[i].r);
      reqs[i] = bl[i].r;
    }
    const uint64_t tail = (uint64_t)n;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync((void*)e.d.reqs, reqs.data(), sizeof(Request) * n,
                       hipMemcpyHostToDevice, e.stream) != hipSuccess ||
        hipMemcpyAsync(e.d.ctl + CW_REQ_TAIL, &tail, 8, hipMemcpyHostToDevice,
                       e.stream) != hipSuccess ||
        hipStreamSynchronize(e.stream) != hipSuccess)
      rc = fail(CHM_E_HIP, "seeding the batch failed");
  }
  float ms = 0;
  if (rc == CHM_OK) {
    e.d.expected = (uint64_t)n;
    rc = engine_launch(e);
    if (rc == CHM_OK)
      rc = engine_wait(e, cfg && cfg->timeout_s > 0 ? cfg->timeout_s : 120.0,
                       &ms);
  }
  if (rc == CHM_OK) {
    uint64_t ctl[CW_WORDS];
    if (hipMemcpy(ctl, e.d.ctl, sizeof(ctl), hipMemcpyDeviceToHost) !=
        hipSuccess)
      rc = fail(CHM_E_HIP, "reading the engine counters failed");
    std::vector<int> seen(n, 0);
    for (uint32_t k = 0; rc == CHM_OK && k < (uint32_t)n; k++) {
      const Completion& c = e.h_dones[k & (e.d.req_cap - 1)];
      if (c.valid == 0 || c.valid - 1 >= (uint64_t)n) continue;
      const int i = (int)(c.valid - 1);
      seen[i]++;
      store_out(out, i, c);
    }
    for (int i = 0; rc == CHM_OK && i < n; i++)
      if (seen[i] != 1)
        rc = fail(CHM_E_DEVICE, "problem %d completed %d times", i, seen[i]);
    if (rc == CHM_OK && out.res) {
      memset(out.res, 0, sizeof(*out.res));
      out.res->device_ms = ms;
      out.res->tasks = (long long)ctl[CW_STATS + 0];
      out.res->steals = (long long)ctl[CW_STATS + 1];
      out.res->idle_polls = (long long)ctl[CW_STATS + 2];
      out.res->cell_updates = (long long)ctl[CW_STATS + 3];
      out.res->n_wg = e.d.n_wg;
      out.res->wg_threads = threads;
    }
  }
  if (arena) (void)hipFree(arena);
  engine_free(e);
  return rc;
}

// ---------------------------------------------------------------------------
// host side: the plain per-request form
// ---------------------------------------------------------------------------
struct PlainCtx {
  hipStream_t s = nullptr;
  uint64_t* d_words = nullptr;  // [0..1] running residual sum, [2] hash
  uint64_t* h_words = nullptr;  // pinned: 4 readback slots of 2, [8] hash
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
};

static int plain_ctx_init(PlainCtx& c) {
  CHM_TRY_HIP(hipStreamCreateWithFlags(&c.s, hipStreamNonBlocking));
  CHM_TRY_HIP(hipMalloc(&c.d_words, 64));
  CHM_TRY(host_alloc((void**)&c.h_words, 128));
  for (auto& e : c.ev)
    CHM_TRY_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return CHM_OK;
}
static void plain_ctx_free(PlainCtx& c) {
  for (auto& e : c.ev)
    if (e) (void)hipEventDestroy(e);
  if (c.d_words) (void)hipFree(c.d_words);
  if (c.h_words) (void)hipHostFree(c.h_words);
  if (c.s) (void)hipStreamDestroy(c.s);
  c = PlainCtx();
}

struct StepPlan {
  int nc, band, n_panels, n_units, blocks;
};
// One wave per panel x band; bands shrink (to 8 rows) until small problems
// still get about 2048 waves.
static StepPlan plan_req(const Request& r) {
  StepPlan p;
  p.nc = r.W >= 2048 ? 2 : 1;
  p.n_panels = (r.W + 128 * p.nc - 1) / (128 * p.nc);
  int band = (int)(((int64_t)r.H * p.n_panels) / 2048);
  band = std::max(8, std::min(128, (band + 7) / 8 * 8));
  p.band = band;
  p.n_units = p.n_panels * ((r.H + band - 1) / band);
  p.blocks = (p.n_units + 3) / 4;
  return p;
}

static void launch_req_step(const Request& r, const StepPlan& sp, int k,
                            bool res, uint64_t* acc, hipStream_t s) {
  const double* in = (k & 1) ? r.b : r.a;
  double* out = (k & 1) ? r.a : r.b;
  const dim3 g(sp.blocks), b(256);
#define CHM_REQ(NC_, U_, R_)                                              \
  req_step_kernel<NC_, U_, R_><<<g, b, 0, s>>>(in, out, r.W, r.H, r.pitch, \
                                               sp.band, sp.n_panels,      \
                                               sp.n_units, acc)
  if (sp.nc == 2) {
    if (res)
      CHM_REQ(2, 4, true);
    else
      CHM_REQ(2, 4, false);
  } else {
    if (res)
      CHM_REQ(1, 2, true);
    else
      CHM_REQ(1, 2, false);
  }
#undef CHM_REQ
}

static inline void u128_sub(uint64_t alo, uint64_t ahi, uint64_t blo,
                            uint64_t bhi, uint64_t* lo, uint64_t* hi) {
  *lo = alo - blo;
  *hi = ahi - bhi - (alo < blo ? 1 : 0);
}

// Solve one request (grids seeded earlier on c.s), one launch per sweep.
// Sweeps add their residual to one running 128-bit sum; S_k is the
// difference of two readbacks (exact modulo 2^128).  With a tolerance the
// sum after sweep k is read back (pinned, async) while sweep k+1 runs; a
// stop after sweep k leaves u_{k+1} intact since sweep k+1 writes the other
// array.  Fills hash, residuals and iters of *out.
static int plain_request(PlainCtx& c, const Request& r, Completion* out) {
  const StepPlan sp = plan_req(r);
  const bool tol = r.tol2 >= 0;
  zero_words<<<1, 64, 0, c.s>>>(c.d_words, 8);
  uint64_t s0lo = 0, s0hi = 0, slo = 0, shi = 0, prev_lo = 0, prev_hi = 0;
  int done = 0;
  for (int k = 0; k < r.max_iters && done == 0; k++) {
    const bool res = tol || k == r.max_iters - 1;
    launch_req_step(r, sp, k, res, c.d_words, c.s);
    if (res) {
      CHM_TRY_HIP(hipMemcpyAsync(c.h_words + 2 * (k & 3), c.d_words, 16,
                                 hipMemcpyDeviceToHost, c.s));
      CHM_TRY_HIP(hipEventRecord(c.ev[k & 3], c.s));
    }
    // decide on sweep j = k-1 while sweep k runs (on the last sweep too)
    for (int j = (tol ? k - 1 : k); j <= k && done == 0; j++) {
      if (j < 0 || (j == k && k != r.max_iters - 1)) continue;
      CHM_TRY_HIP(hipEventSynchronize(c.ev[j & 3]));
      const uint64_t clo = c.h_words[2 * (j & 3)];
      const uint64_t chi = c.h_words[2 * (j & 3) + 1];
      u128_sub(clo, chi, prev_lo, prev_hi, &slo, &shi);
      prev_lo = clo;
      prev_hi = chi;
      if (j == 0) {
        s0lo = slo;
        s0hi = shi;
      }
      if (j == r.max_iters - 1 ||
          (tol && u128d(slo, shi) <= r.tol2 * u128d(s0lo, s0hi)))
        done = j + 1;
    }
  }
  hash_req_kernel<<<512, 256, 0, c.s>>>((done & 1) ? r.b : r.a, r.W, r.H,
                                         r.pitch, c.d_words + 2);
  CHM_TRY_HIP(hipGetLastError());
  CHM_TRY_HIP(hipMemcpyAsync(c.h_words + 8, c.d_words + 2, 8,
                             hipMemcpyDeviceToHost, c.s));
  CHM_TRY_HIP(hipStreamSynchronize(c.s));
  out->hash = c.h_words[8];
  out->s_last_lo = slo;
  out->s_last_hi = shi;
  out->s0_lo = (tol || r.max_iters == 1) ? s0lo : 0;
  out->s0_hi = (tol || r.max_iters == 1) ? s0hi : 0;
  out->iters = (uint32_t)done;
  return CHM_OK;
}

// The hipGraph of one fixed-iteration problem: its sweeps, the residual of
// the last one and the hash, captured once.
static int capture_plain(PlainCtx& c, const Request& r, hipGraphExec_t* ge) {
  const StepPlan sp = plan_req(r);
  hipGraph_t g;
  CHM_TRY_HIP(hipStreamBeginCapture(c.s, hipStreamCaptureModeThreadLocal));
  zero_words<<<1, 64, 0, c.s>>>(c.d_words, 8);
  for (int k = 0; k < r.max_iters; k++)
    launch_req_step(r, sp, k, k == r.max_iters - 1, c.d_words, c.s);
  hash_req_kernel<<<512, 256, 0, c.s>>>((r.max_iters & 1) ? r.b : r.a, r.W,
                                         r.H, r.pitch, c.d_words + 2);
  CHM_TRY_HIP(hipStreamEndCapture(c.s, &g));
  const hipError_t e = hipGraphInstantiate(ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess)
    return fail(CHM_E_HIP, "graph instantiate: %s", hipGetErrorString(e));
  return CHM_OK;
}

// ---------------------------------------------------------------------------
// host side: servers
// ---------------------------------------------------------------------------
struct Server {
  int mode = CHM_SERVE_ENGINE;
  Engine e;
  char* pool = nullptr;
  size_t pool_size = 0;
  std::map<size_t, size_t> free_blocks;  // offset -> size
  std::vector<uint32_t> free_slots;
  struct Live {
    size_t off, bytes;
    uint32_t slot;
    double t_submit;
  };
  std::map<uint64_t, Live> live;  // ticket -> allocation
  uint64_t next_ticket = 0;
  uint64_t seen = 0;  // completion records consumed (engine ring)
  long long completed = 0;
  size_t used = 0;
  int block = 0;
  int max_active = 0;
  // plain mode: one worker thread serving a FIFO
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Request> fifo;
  std::deque<Completion> finished;  // t_admit/t_done in host ns
  bool quit = false;
  int worker_rc = CHM_OK;
  PlainCtx pc;
};

// Tile side of a served request when none is given: a lone request is cut
// into about as many tiles as there are workgroups, so one round of it
// occupies the whole GPU; with k requests resident each gets about 1/(k+1)
// of them, and tiles grow (fewer fences per cell) as the load rises.
static int serve_block(int W, int H, int n_wg, int resident, int ow) {
  const int64_t target = std::max<int64_t>(1, n_wg / (1 + resident));
  // multiples of the tolerance pass's window (ow output columns)
  const int kSides[] = {4 * ow, 2 * ow, ow};
  for (int bs : kSides) {
    const int64_t tiles = (int64_t)((W + bs - 1) / bs) * ((H + bs - 1) / bs);
    if (tiles >= target) return bs;
  }
  return 64;
}

static bool pool_alloc(Server& s, size_t bytes, size_t* off) {
  for (auto it = s.free_blocks.begin(); it != s.free_blocks.end(); ++it) {
    if (it->second >= bytes) {
      *off = it->first;
      const size_t rest = it->second - bytes;
      const size_t noff = it->first + bytes;
      s.free_blocks.erase(it);
      if (rest) s.free_blocks[noff] = rest;
      s.used += bytes;
      return true;
    }
  }
  return false;
}

static void pool_free(Server& s, size_t off, size_t bytes) {
  s.used -= bytes;
  auto it = s.free_blocks.emplace(off, bytes).first;
  auto nx = std::next(it);
  if (nx != s.free_blocks.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    s.free_blocks.erase(nx);
  }
  if (it != s.free_blocks.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      s.free_blocks.erase(it);
    }
  }
}

static void plain_worker(Server* s, int dev) {
  (void)hipSetDevice(dev);
  for (;;) {
    Request r;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait(lk, [s] { return s->quit || !s->fifo.empty(); });
      if (s->fifo.empty()) return;
      r = s->fifo.front();
      s->fifo.pop_front();
    }
    Completion c;
    memset(&c, 0, sizeof(c));
    c.ticket = r.ticket;
    c.slot = r.slot;
    c.t_admit = (uint64_t)(now_s() * 1e9);
    init_problem<<<512, 256, 0, s->pc.s>>>(r);
    const int rc = plain_request(s->pc, r, &c);
    c.t_done = (uint64_t)(now_s() * 1e9);
    c.valid = r.ticket + 1;
    std::lock_guard<std::mutex> lk(s->mu);
    if (rc != CHM_OK && s->worker_rc == CHM_OK) s->worker_rc = rc;
    s->finished.push_back(c);
  }
}

static void server_free(Server* s) {
  if (s->worker.joinable()) {
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->quit = true;
    }
    s->cv.notify_all();
    s->worker.join();
  }
  plain_ctx_free(s->pc);
  if (s->pool) (void)hipFree(s->pool);
  s->pool = nullptr;
  engine_free(s->e);
  (void)hipDeviceSynchronize();
  delete s;
}

}  // namespace eng
}  // namespace chm

using namespace chm;
using namespace chm::eng;

extern "C" {

int chm_jacobi2d_batch_run(const chm_batch_problem* probs, int n,
                           const chm_batch_config* cfg, chm_batch_result* res,
                           uint64_t* hashes, uint64_t* residuals,
                           int* iters_done) {
  if (!probs) return fail(CHM_E_INVALID, "null problems");
  return batch_once(probs, n, cfg, 0, nullptr,
                    BatchOut{res, hashes, residuals, iters_done});
}

int chm_jacobi2d_batch_plain(const chm_batch_problem* probs, int n,
                             int use_graph, chm_batch_result* res,
                             uint64_t* hashes, uint64_t* residuals,
                             int* iters_done) {
  if (!probs || n < 1) return fail(CHM_E_INVALID, "no problems");
  for (int i = 0; i < n; i++)
    if (probs[i].tol > 0 && use_graph)
      return fail(CHM_E_INVALID, "graph form runs fixed iteration counts");
  std::vector<Built> bl;
  size_t total = 0;
  uint64_t tiles = 0;
  CHM_TRY(layout_batch(probs, n, 0, 0, nullptr, 2, kDeep, bl, &total,
                       &tiles));
  PlainCtx c;
  int rc = plain_ctx_init(c);
  char* arena = nullptr;
  std::vector<hipGraphExec_t> graphs;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (rc == CHM_OK && hipMalloc(&arena, total) != hipSuccess)
    rc = fail(CHM_E_NOMEM, "batch grids: %zu bytes", total);
  if (rc == CHM_OK) {
    place_batch(bl, arena, 2);
    for (int i = 0; i < n; i++) init_problem<<<512, 256, 0, c.s>>>(bl[i].r);
    if (hipStreamSynchronize(c.s) != hipSuccess)
      rc = fail(CHM_E_HIP, "seeding the batch failed");
  }
  for (int i = 0; rc == CHM_OK && use_graph && i < n; i++) {
    hipGraphExec_t ge = nullptr;
    rc = capture_plain(c, bl[i].r, &ge);
    if (rc == CHM_OK) graphs.push_back(ge);
  }
  if (rc == CHM_OK && (hipEventCreate(&e0) != hipSuccess ||
                       hipEventCreate(&e1) != hipSuccess))
    rc = fail(CHM_E_HIP, "event create failed");
  std::vector<Completion> done(n);
  if (rc == CHM_OK) {
    (void)hipEventRecord(e0, c.s);
    for (int i = 0; rc == CHM_OK && i < n; i++) {
      memset(&done[i], 0, sizeof(Completion));
      if (use_graph) {
        if (hipGraphLaunch(graphs[i], c.s) != hipSuccess) {
          rc = fail(CHM_E_HIP, "graph launch failed");
          break;
        }
        // the last sweep's residual and the hash, read once it is done
        if (hipMemcpyAsync(c.h_words, c.d_words, 24, hipMemcpyDeviceToHost,
                           c.s) != hipSuccess ||
            hipStreamSynchronize(c.s) != hipSuccess) {
          rc = fail(CHM_E_HIP, "graph readback failed");
          break;
        }
        done[i].s_last_lo = c.h_words[0];
        done[i].s_last_hi = c.h_words[1];
        done[i].hash = c.h_words[2];
        done[i].iters = (uint32_t)bl[i].r.max_iters;
      } else {
        rc = plain_request(c, bl[i].r, &done[i]);
      }
    }
    (void)hipEventRecord(e1, c.s);
    if (rc == CHM_OK && hipEventSynchronize(e1) != hipSuccess)
      rc = fail(CHM_E_HIP, "plain batch failed");
  }
  if (rc == CHM_OK) {
    const BatchOut out{res, hashes, residuals, iters_done};
    for (int i = 0; i < n; i++) store_out(out, i, done[i]);
    if (res) {
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      memset(res, 0, sizeof(*res));
      res->device_ms = ms;
      long long cells = 0;
      for (int i = 0; i < n; i++)
        cells += (long long)bl[i].r.W * bl[i].r.H * (long long)done[i].iters;
      res->cell_updates = cells;
    }
  }
  for (auto ge : graphs) (void)hipGraphExecDestroy(ge);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (arena) (void)hipFree(arena);
  plain_ctx_free(c);
  return rc;
}

int chm_jacobi2d_batch_autotune(const chm_batch_problem* probs, int n,
                                int sample_iters, chm_batch_config* best,
                                int* blocks_out, double* best_rate) {
  if (!probs || !best || n < 1) return fail(CHM_E_INVALID, "null argument");
  // whole rounds for every candidate: a sample that ends inside a round of
  // sixteen pays for the round's copying stages and misranks the engines.
  // 0: no cap.  (Each timed run's setup, seeding and readback cost more than
  // its sweeps on the 512-problem batch, so full-length runs tune in less
  // time than 32-sweep samples, 16.4 against 21.8 s, and rank better: their
  // pick ran 102.6-103.2 ms, the 32-sweep samples' pick 114-122 ms.)
  const int cap = sample_iters > 0 ? (sample_iters + 15) / 16 * 16 : 0;
  // tile sides tried per size class: the powers of two and the multiples
  // of the deep pass's window (112 columns at eight sweeps, 96 at sixteen)
  static const int kBlocks8[] = {64, 112, 128, 224, 256, 448, 512};
  static const int kBlocks16[] = {64, 96, 128, 192, 256, 384, 512};
  // size classes: one per octave of the longer side
  std::vector<int> cls(n);
  int n_cls = 0;
  for (int i = 0; i < n; i++) {
    const int side = std::max(probs[i].width, probs[i].height);
    int c = 0;
    while ((1 << (c + 7)) < side) c++;  // <=128, <=256, ...
    cls[i] = c;
    n_cls = std::max(n_cls, c + 1);
  }
  // the size rule's tiles for an engine of `threads`
  auto rule_blocks = [&](int threads) {
    std::vector<int> b(n);
    for (int i = 0; i < n; i++)
      b[i] = pick_block(probs[i].width, probs[i].height, 0,
                        deep_for_threads(threads));
    return b;
  };
  auto rate_of = [&](const chm_batch_config& c, const std::vector<int>& b,
                     double* rate) {
    chm_batch_result r;
    const int rc = batch_once(probs, n, &c, cap, b.data(),
                              BatchOut{&r, nullptr, nullptr, nullptr});
    if (rc != CHM_OK) return rc;
    *rate = (double)r.cell_updates / (r.device_ms * 1e-3);
    return CHM_OK;
  };
  chm_batch_config cfg;
  memset(&cfg, 0, sizeof(cfg));
  cfg.wg_threads = 1024;
  cfg.panel_chunks = 1;
  cfg.rows_in_flight = 2;
  cfg.timeout_s = 120;
  double top = -1;
  // 1. workgroup size x sweep shape, heuristic tiles
  static const int kThreads[] = {256, 512, 1024};
  static const int kShapes[][2] = {{1, 2}, {1, 4}, {2, 2}};
  chm_batch_config top_cfg = cfg;
  for (int t : kThreads)
    for (const auto& sh : kShapes) {
      chm_batch_config c = cfg;
      c.wg_threads = t;
      c.panel_chunks = sh[0];
      c.rows_in_flight = sh[1];
      double rate = 0;
      if (rate_of(c, rule_blocks(t), &rate) == CHM_OK && rate > top) {
        top = rate;
        top_cfg = c;
      }
    }
  if (top < 0) return fail(CHM_E_DEVICE, "no autotune candidate ran");
  std::vector<int> blocks = rule_blocks(top_cfg.wg_threads);
  const bool wide = deep_for_threads(top_cfg.wg_threads) == kDeepWide;
  const int* kBlocks = wide ? kBlocks16 : kBlocks8;
  const int n_blocks = wide ? (int)(sizeof(kBlocks16) / sizeof(int))
                            : (int)(sizeof(kBlocks8) / sizeof(int));
  // 2. tile side per size class, one class at a time (largest first)
  for (int c = n_cls - 1; c >= 0; c--) {
    std::vector<int> trial = blocks;
    int best_b = -1;
    for (int bi = 0; bi < n_blocks; bi++) {
      const int b = kBlocks[bi];
      bool any = false;
      for (int i = 0; i < n; i++)
        if (cls[i] == c) {
          trial[i] = b;
          any = true;
        }
      if (!any) break;
      if (trial == blocks) continue;  // the current choice is already timed
      double rate = 0;
      if (rate_of(top_cfg, trial, &rate) == CHM_OK && rate > top) {
        top = rate;
        best_b = b;
      }
    }
    if (best_b > 0)
      for (int i = 0; i < n; i++)
        if (cls[i] == c) blocks[i] = best_b;
  }
  *best = top_cfg;
  best->timeout_s = 0;
  if (blocks_out)
    for (int i = 0; i < n; i++) blocks_out[i] = blocks[i];
  if (best_rate) *best_rate = top;
  return CHM_OK;
}

int chm_jacobi2d_server_start(const chm_server_config* cfg, void** handle) {
  if (!cfg || !handle) return fail(CHM_E_INVALID, "null argument");
  if (cfg->mode != CHM_SERVE_ENGINE && cfg->mode != CHM_SERVE_PLAIN)
    return fail(CHM_E_INVALID, "server mode %d", cfg->mode);
  const size_t budget = cfg->budget_bytes ? cfg->budget_bytes : (size_t)16 << 30;
  Server* s = new Server();
  s->mode = cfg->mode;
  s->block = cfg->block;
  s->max_active = cfg->max_active > 0 ? cfg->max_active : (int)kMaxSlots;
  if (s->max_active > (int)kMaxSlots) s->max_active = (int)kMaxSlots;
  int rc = CHM_OK;
  if (s->mode == CHM_SERVE_ENGINE) {
    // (default 512 threads: tolerance rounds of twelve sweeps; on a
    // saturating stream 124.8 solves/s against 119.4 at 1024 threads and
    // rounds of six, before the filtered rounds' six-slot input ring)
    const int threads = cfg->wg_threads ? cfg->wg_threads : 512;
    const int nc = cfg->panel_chunks ? cfg->panel_chunks : 1;
    const int u = cfg->rows_in_flight ? cfg->rows_in_flight : 2;
    // queued tasks are bounded by the tiles of the resident problems: at
    // least one 64x64 tile's grids (2 x 33 KiB) per task
    const uint64_t qtasks = std::min<uint64_t>(budget / (2 * 66 * 66 * 8), 1u << 22);
    rc = engine_setup(s->e, threads, nc, u, 2 * kMaxSlots, qtasks, true);
  } else {
    rc = plain_ctx_init(s->pc);
  }
  // the budget is reserved up front: admission never allocates
  if (rc == CHM_OK && hipMalloc(&s->pool, budget) != hipSuccess) {
    (void)hipGetLastError();
    size_t fr = 0, tot = 0;
    (void)hipMemGetInfo(&fr, &tot);
    rc = fail(CHM_E_NOMEM, "serving pool of %zu bytes (free HBM %zu of %zu)",
              budget, fr, tot);
  }
  if (rc == CHM_OK) {
    s->pool_size = budget;
    s->free_blocks[0] = budget;
    for (int i = (int)kMaxSlots - 1; i >= 0; i--) s->free_slots.push_back(i);
    if (s->mode == CHM_SERVE_ENGINE) {
      rc = engine_launch(s->e);
    } else {
      int dev = 0;
      (void)hipGetDevice(&dev);
      s->worker = std::thread(plain_worker, s, dev);
    }
  }
  if (rc != CHM_OK) {
    const int keep = rc;
    if (s->e.running) {
      __atomic_store_n(s->e.h_stop, 2u, __ATOMIC_SEQ_CST);
      (void)engine_wait(s->e, 60, nullptr);
    }
    server_free(s);
    return keep;
  }
  *handle = s;
  return CHM_OK;
}

int chm_jacobi2d_server_submit(void* handle, const chm_batch_problem* pb,
                               uint64_t* ticket) {
  Server* s = (Server*)handle;
  if (!s || !pb || !ticket) return fail(CHM_E_INVALID, "null argument");
  const bool engine = s->mode == CHM_SERVE_ENGINE;
  int block = s->block;
  if (engine && pb->block <= 0 && block <= 0)
    block = serve_block(pb->width, pb->height, s->e.d.n_wg,
                        (int)s->live.size(),
                        jac::window_out_cols(deep_tol_for_threads(s->e.threads)));
  const int deep = engine ? deep_for_threads(s->e.threads) : kDeep;
  CHM_TRY(check_problem(*pb, block, deep));
  Built b = build_request(*pb, block, deep);
  const size_t bytes = (engine ? 3 : 2) * b.grid_bytes + (engine ? b.pend_bytes : 0);
  if (bytes > s->pool_size)
    return fail(CHM_E_INVALID, "request needs %zu bytes, budget %zu", bytes,
                s->pool_size);
  if ((int)s->live.size() >= s->max_active || s->free_slots.empty())
    return fail(CHM_E_BUSY, "%d requests resident", (int)s->live.size());
  size_t off = 0;
  if (!pool_alloc(*s, bytes, &off))
    return fail(CHM_E_BUSY, "budget: %zu of %zu bytes in use", s->used,
                s->pool_size);
  Request& r = b.r;
  const int64_t row0 = (int64_t)(r.depth - 1) * r.pitch;  // rows 1-depth ..
  r.a = (double*)(s->pool + off) + row0;
  r.b = (double*)(s->pool + off + b.grid_bytes) + row0;
  r.c = engine ? (double*)(s->pool + off + 2 * b.grid_bytes) + row0 : nullptr;
  r.pend = engine ? (uint32_t*)(s->pool + off + 3 * b.grid_bytes) : nullptr;
  r.init_on_device = 1;
  r.ticket = s->next_ticket++;
  r.slot = s->free_slots.back();
  s->free_slots.pop_back();
  s->live[r.ticket] = Server::Live{off, bytes, r.slot, now_s()};
  if (engine) {
    s->e.h_reqs[r.ticket & (s->e.d.req_cap - 1)] = r;
    __atomic_store_n(s->e.h_tail, r.ticket + 1, __ATOMIC_SEQ_CST);
  } else {
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->fifo.push_back(r);
    }
    s->cv.notify_one();
  }
  *ticket = r.ticket;
  return CHM_OK;
}

int chm_jacobi2d_server_poll(void* handle, chm_server_completion* out,
                             int max_out) {
  Server* s = (Server*)handle;
  if (!s || (!out && max_out > 0)) return fail(CHM_E_INVALID, "null argument");
  int n = 0;
  const bool engine = s->mode == CHM_SERVE_ENGINE;
  if (!engine) {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->worker_rc != CHM_OK) return s->worker_rc;
  }
  while (n < max_out) {
    Completion c;
    if (engine) {
      Completion* rec = &s->e.h_dones[s->seen & (s->e.d.req_cap - 1)];
      const uint64_t v = __atomic_load_n(&rec->valid, __ATOMIC_ACQUIRE);
      if (v == 0) break;
      c = *rec;
      c.valid = v;
      rec->valid = 0;
      s->seen++;
    } else {
      std::lock_guard<std::mutex> lk(s->mu);
      if (s->finished.empty()) break;
      c = s->finished.front();
      s->finished.pop_front();
    }
    const double now = now_s();
    const uint64_t t = c.valid - 1;
    auto it = s->live.find(t);
    chm_server_completion& o = out[n++];
    o.ticket = t;
    o.iters = (int)c.iters;
    o.hash = c.hash;
    o.residual_lo = c.s_last_lo;
    o.residual_hi = c.s_last_hi;
    o.residual0_lo = c.s0_lo;
    o.residual0_hi = c.s0_hi;
    const double hz = engine ? s->e.wall_hz : 1e9;
    o.device_ms = (double)(c.t_done - c.t_admit) / hz * 1e3;
    o.latency_ms = it != s->live.end() ? (now - it->second.t_submit) * 1e3 : -1;
    if (it != s->live.end()) {
      pool_free(*s, it->second.off, it->second.bytes);
      s->free_slots.push_back(it->second.slot);
      s->live.erase(it);
    }
    s->completed++;
  }
  return n;
}

int chm_jacobi2d_server_stats(void* handle, chm_server_stats* out) {
  Server* s = (Server*)handle;
  if (!s || !out) return fail(CHM_E_INVALID, "null argument");
  memset(out, 0, sizeof(*out));
  out->active = (int)s->live.size();
  out->bytes_in_use = s->used;
  out->budget_bytes = s->pool_size;
  out->submitted = (long long)s->next_ticket;
  out->completed = s->completed;
  return CHM_OK;
}

int chm_jacobi2d_server_stop(void* handle, int abort_now, double timeout_s) {
  Server* s = (Server*)handle;
  if (!s) return CHM_OK;
  int rc = CHM_OK;
  if (!abort_now) {
    // drain: collect what is still in flight (the caller has seen the rest)
    const double t0 = now_s();
    std::vector<chm_server_completion> tmp(64);
    while (!s->live.empty()) {
      const int got = chm_jacobi2d_server_poll(s, tmp.data(), 64);
      if (got < 0) {
        rc = got;
        break;
      }
      if (timeout_s > 0 && now_s() - t0 > timeout_s) {
        rc = fail(CHM_E_TIMEOUT, "%d requests still running",
                  (int)s->live.size());
        break;
      }
      if (got == 0) usleep(100);
    }
  }
  if (s->mode == CHM_SERVE_ENGINE && s->e.running) {
    __atomic_store_n(s->e.h_stop, (rc == CHM_OK && !abort_now) ? 1u : 2u,
                     __ATOMIC_SEQ_CST);
    const int rc2 = engine_wait(s->e, 60.0, nullptr);
    if (rc == CHM_OK) rc = rc2;
  }
  server_free(s);
  return rc;
}

}  // extern "C"

// Message-driven Jacobi2D on the persistent runtime: the chare program of
// examples/jacobi2d/jacobi2d.cu:26-335 written against the current charm::
// API, plus the host side of the persistent form of the Jacobi2D C ABI.
//
// Differences from the reference program, each deliberate:
//   - halos carry refnum = iteration and the chare's refnum advances with
//     its iteration, so a neighbour one iteration ahead is deferred instead
//     of being miscounted (the reference counts halos without refnums,
//     jacobi2d.cu:222);
//   - the stencil, pack and unpack run inside the PE's workgroup (HIP has no
//     device-side launch, which the reference uses at jacobi2d.cu:145-235);
//     the sweep writes the four outgoing edges straight into the staging
//     buffers the next sends copy from (pack fused into the sweep);
//   - grids are allocated once by the host in HBM and seeded by a kernel
//     (no device `new`, jacobi2d.cu:132-141);
//   - the chare-column count divides the width (the reference divides the
//     width by the chare-row count, jacobi2d.cu:64-65, a transposition that
//     is harmless only for square chare grids);
//   - the last row/column of chares takes the remainder of an uneven split;
//   - a chare with no neighbours iterates without waiting;
//   - temporal blocking (halo depth D = 2 .. 16): chares
//     exchange halos D cells deep (four edges of D rows/columns plus four
//     (D-1) x (D-1) corner blocks from the diagonal neighbours) once per D
//     sweeps, and sweep D times in one pass over their block with the
//     intermediate grids held in registers; the arithmetic, and so every
//     grid, is unchanged;
//   - the sweep rounds run in the runtime's work hook (requested by the
//     halo entry method once every halo is in), which the scheduler calls
//     directly: compiled for the 512-thread PE (256 VGPRs per lane) rather
//     than an entry method's 128, so sixteen fused sweeps fit in registers
//     and no callee-saved registers are spilled around an indirect call.
#ifndef CHARMING_PE_THREADS
#define CHARMING_PE_THREADS 512
#endif
__device__ void jacobi_work(void* token);
#define CHARMING_WORK_HOOK jacobi_work
#include <charming.h>

#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "jacobi2d_internal.h"

#define LEFT 0
#define RIGHT 1
#define TOP 2
#define BOTTOM 3
#define TOP_LEFT 4
#define TOP_RIGHT 5
#define BOTTOM_LEFT 6
#define BOTTOM_RIGHT 7
#define N_NEIGHBORS 8

namespace chm {
namespace jac {

// Problem description shared with the device (one per solve).
struct JacobiDesc {
  double* arena;          // chare c: arrays at arena + (2c + k) * chare_elems
  int64_t chare_elems;
  double* staging;        // chare c: 4 edges at staging + (4c + d) * stage_elems
  int64_t stage_elems;
  uint64_t* res_acc;      // exact residual of the last sweep {lo, hi}
  uint64_t* phase;        // per chare: wall-clock ticks in sweep, send, wait
  int W, H, cr, cc, iters;
  int tune_nc, tune_u;
  int tune_band;          // deep fused sweeps: rows per band, < 0 = one
                          // contiguous share of the block per wave, 0 = by
                          // chare height
  int overlap;            // sweep edges first, send, then the interior
  uint64_t* recv_ticks;   // per chare: ticks inside recv_ghosts (unpack)
  int depth;              // halo depth = sweeps per exchange (1 .. 16)
  int halo_copy;          // 1: halos staged and copied by invoke (else the
                          // sweep writes them into the messages in place)
};

// Chare arrays always hold halos of the deepest supported depth, sixteen:
// rows -15 .. bh+16 of pitch_for_depth(bw, 16) doubles; grids[] point at
// row 0.
constexpr int kHaloRows = 16;
__host__ __device__ inline int64_t chare_pitch(int bw) {
  return pitch_for_depth(bw, kHaloRows);
}

// The side a message arrives on at the receiver, by the direction it was
// sent in.
__host__ __device__ inline int opposite_side(int k) {
  const int opp[8] = {RIGHT, LEFT, BOTTOM, TOP, BOTTOM_RIGHT, BOTTOM_LEFT,
                      TOP_RIGHT, TOP_LEFT};
  return opp[k];
}

__device__ JacobiDesc g_jdesc;

__host__ __device__ inline ChareGeom chare_geom(int W, int H, int cr, int cc,
                                               int idx) {
  ChareGeom g;
  split_extent(H, cr, idx / cc, &g.row0, &g.bh);
  split_extent(W, cc, idx % cc, &g.col0, &g.bw);
  return g;
}

}  // namespace jac
}  // namespace chm

using namespace chm::jac;

// ---------------------------------------------------------------------------
// the chare program
// ---------------------------------------------------------------------------
struct Block : charm::chare {
  int row, col;
  int neighbor_index[N_NEIGHBORS];
  int neighbor_count;
  int recv_count;
  int end_count;
  int block_width, block_height;
  int64_t pitch;
  int n_iters, iter;
  int depth;             // sweeps per halo exchange
  int cur;               // u_iter lives in grids[cur]; each round flips it
  double* grids[2];      // (pointers to row 0 of each array)
  double* boundaries[N_NEIGHBORS];  // [dir slot | depth x edge] staging
  // per-chare phase accounting (jacobi2d.h:31-42 keeps send/recv/update
  // times the same way): device wall-clock ticks
  uint64_t t_sweep, t_send, t_wait, t_mark;

  __device__ Block() {}
  __device__ void init(void* arg);
  __device__ size_t halo_bytes(int k) const;
  __device__ void fill_corners(double* const* bufs);
  __device__ void send_boundaries();
  __device__ void put_edges(const double* g, int D, int parity);
  __device__ void recv_ghosts(void* arg);
  __device__ void update();
  __device__ void end(void* arg);
};

__device__ void entry_init(Block& b, void* arg) { b.init(arg); }
__device__ void entry_recv_ghosts(Block& b, void* arg) { b.recv_ghosts(arg); }
__device__ void entry_end(Block& b, void* arg) { b.end(arg); }

// The work hook: the sweep rounds of a chare whose halos are all in.  Only
// reached from the scheduler's direct call (never through an entry method),
// so everything below it is compiled for the PE's 512 threads.
__device__ void jacobi_work(void* token) { ((Block*)token)->update(); }

__shared__ charm::chare_proxy<Block>* block_proxy;

__device__ void charm::main(int argc, char** argv, size_t* argvs, int pe) {
  const JacobiDesc& d = g_jdesc;
  if (threadIdx.x == 0) {
    void* mem = charm::rt::heap_alloc(sizeof(charm::chare_proxy<Block>), 16);
    block_proxy = new (mem) charm::chare_proxy<Block>();
    block_proxy->add_entry_method<&entry_init>();
    block_proxy->add_entry_method<&entry_recv_ghosts>();
    block_proxy->add_entry_method<&entry_end>();
    block_proxy->create(d.cr * d.cc);
  }
  __syncthreads();
  if (pe == 0) {
    __shared__ int params[5];
    if (threadIdx.x == 0) {
      params[0] = d.W;
      params[1] = d.H;
      params[2] = d.iters;
      params[3] = d.cr;
      params[4] = d.cc;
    }
    __syncthreads();
    block_proxy->invoke_all(0, params, sizeof(params));
  }
}

// The sweep of one chare by its PE's workgroup, in two phases so the halo
// sends overlap the bulk of the work: phase 0 sweeps the edge panels (first
// and last 128*NC columns, full height) and thin top/bottom bands of the
// middle panels, which produce every outgoing edge; phase 1 sweeps the rest.
// Waves take tiles round-robin.  RES accumulates the exact residual.
template <int NC, int U, bool RES>
__device__ __noinline__ void chare_sweep(const double* in, double* out,
                                         int64_t P, int bw, int bh,
                                         EdgeOut eo, int phase) {
  const int wave = threadIdx.x / 64, WV = blockDim.x / 64;
  const int pw = 128 * NC;
  const int S = (bw + pw - 1) / pw;  // panels
  const int e = bh > 16 ? 4 : bh;     // rows of the thin bands
  const int n_mid = S > 2 ? S - 2 : 0;
  u128acc acc = {0, 0};
  if (phase == 2) {  // the whole block at once: panels x bands
    int nb = WV / S;
    if (nb < 1) nb = 1;
    if (nb > bh) nb = bh;
    const int br = (bh + nb - 1) / nb;
    for (int t = wave; t < S * nb; t += WV) {
      const int panel = t % S, b = t / S;
      const int r0 = 1 + b * br, r1 = min(r0 + br, bh + 1);
      sweep_panel<NC, U, RES, true>(in, out, P, bw, bh, 1 + panel * pw, r0,
                                    r1, acc, eo);
    }
  } else if (phase == 0) {
    const int n_edge = S == 1 ? 1 : 2;
    int nb = WV / n_edge;
    if (nb < 1) nb = 1;
    if (nb > bh) nb = bh;
    const int br = (bh + nb - 1) / nb;
    const int tiles_edge = n_edge * nb;
    for (int t = wave; t < tiles_edge + 2 * n_mid; t += WV) {
      int panel, r0, r1;
      if (t < tiles_edge) {
        panel = (t / nb == 0) ? 0 : S - 1;
        r0 = 1 + (t % nb) * br;
        r1 = min(r0 + br, bh + 1);
      } else {
        const int q = t - tiles_edge;
        panel = 1 + q / 2;
        if (q % 2 == 0) {
          r0 = 1;
          r1 = min(1 + e, bh + 1);
        } else {
          r0 = max(1 + e, bh + 1 - e);
          r1 = bh + 1;
        }
      }
      sweep_panel<NC, U, RES, true>(in, out, P, bw, bh, 1 + panel * pw, r0,
                                    r1, acc, eo);
    }
  } else if (n_mid > 0) {
    const int rlo = 1 + e, rhi = max(1 + e, bh + 1 - e);
    if (rhi > rlo) {
      int nb = (3 * WV + n_mid - 1) / n_mid;  // about three tiles per wave
      if (nb > rhi - rlo) nb = rhi - rlo;
      const int br = (rhi - rlo + nb - 1) / nb;
      for (int t = wave; t < n_mid * nb; t += WV) {
        // band-major: neighbouring panels are swept at the same time
        const int panel = 1 + t % n_mid, b = t / n_mid;
        const int r0 = rlo + b * br, r1 = min(r0 + br, rhi);
        sweep_panel<NC, U, RES, true>(in, out, P, bw, bh, 1 + panel * pw, r0,
                                      r1, acc, eo);
      }
    }
  }
  if (RES) {
    acc = wave_reduce_u128(acc);
    if (__lane_id() == 0) atomic_add_u128(g_jdesc.res_acc, acc);
  }
}

__device__ void sweep_dispatch(const double* in, double* out, int64_t P,
                               int bw, int bh, bool res, const EdgeOut& eo,
                               int phase) {
  int nc = g_jdesc.tune_nc, u = g_jdesc.tune_u;
  // The sweep runs from the work hook (256 VGPRs at 512-thread PEs).  Default
  // 2x2: two 128-column chunks, two rows in flight per wave, 8 waves x 8 KiB
  // of reads in flight per CU (16384^2 x 1000 at one sweep per exchange:
  // 1x1 1157 ms, 1x2 1028, 1x4 959, 2x1 947, 2x2 914, 4x1 911-920, 4x2 923,
  // 2x4 930, 4x4 933, 8x1 966, 8x2 979; the wider shapes spill).
  if (nc == 0) {
    nc = 2;
    u = 2;
  }
  if (u == 0) u = (4 / nc) > 0 ? 4 / nc : 1;
#define CHM_CASE(NC_, U_)                                               \
  if (nc == NC_ && u == U_) {                                           \
    if (res)                                                            \
      chare_sweep<NC_, U_, true>(in, out, P, bw, bh, eo, phase);        \
    else                                                                \
      chare_sweep<NC_, U_, false>(in, out, P, bw, bh, eo, phase);       \
    return;                                                             \
  }
  CHM_CASE(1, 2) CHM_CASE(1, 4) CHM_CASE(2, 2) CHM_CASE(1, 1) CHM_CASE(2, 1)
  CHM_CASE(4, 1) CHM_CASE(4, 2) CHM_CASE(2, 4) CHM_CASE(8, 1) CHM_CASE(4, 4)
  CHM_CASE(8, 2)
#undef CHM_CASE
  if (res)
    chare_sweep<1, 2, true>(in, out, P, bw, bh, eo, phase);
  else
    chare_sweep<1, 2, false>(in, out, P, bw, bh, eo, phase);
}

// Two fused sweeps of the whole block (halos two deep), panels x bands.
template <int NC, int U, bool RES, bool PACK>
__device__ __noinline__ void chare_sweep2(const double* in, double* out,
                                          int64_t P, int bw, int bh,
                                          EdgeOut2 eo, RingSides rs) {
  const int wave = threadIdx.x / 64, WV = blockDim.x / 64;
  const int pw = 128 * NC;
  const int S = (bw + pw - 1) / pw;
  int nb = WV / S;
  if (nb < 1) nb = 1;
  if (nb > bh) nb = bh;
  const int br = (bh + nb - 1) / nb;
  u128acc acc = {0, 0}, unused = {0, 0};
  for (int t = wave; t < S * nb; t += WV) {
    const int panel = t % S, b = t / S;
    const int r0 = 1 + b * br, r1 = min(r0 + br, bh + 1);
    sweep_panel2<NC, U, false, RES, PACK>(in, out, P, bw, bh, 1 + panel * pw,
                                          r0, r1, unused, acc, eo, rs);
  }
  if (RES) {
    acc = wave_reduce_u128(acc);
    if (__lane_id() == 0) atomic_add_u128(g_jdesc.res_acc, acc);
  }
}

// Values the whole wave shares, moved into scalar registers: an entry
// method's callees receive their arguments in vector registers, which the
// deeper fused sweeps cannot spare.
__device__ inline int wuni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline int64_t wuni(int64_t v) {
  const uint64_t u = (uint64_t)v;
  return (int64_t)(((uint64_t)(uint32_t)wuni((int)(u >> 32)) << 32) |
                   (uint32_t)wuni((int)(uint32_t)u));
}
template <class T>
__device__ inline T* wuni(T* p) {
  return (T*)wuni((int64_t)p);
}

// Input rows each wave of the deep pass keeps in its register ring: 6,
// four rows in flight ahead of the step that first needs them (with two
// waves per SIMD, one row in flight left the loads exposed: 16384^2 x 1000
// at sixteen fused sweeps 123.6 ms with 6 slots against 133.8 ms with 3;
// 9 and 12 slots, 130-131 ms, raise register pressure more than they hide)
#ifndef CHM_JACOBI_IN_RING
#define CHM_JACOBI_IN_RING 6
#endif
constexpr int kInRing = CHM_JACOBI_IN_RING;

// D >= 3 fused sweeps of the whole block (halos D deep), overlapping
// windows x bands.  The outgoing edges are packed afterwards (pack_edges):
// kept out of the sweep, they would cost its inner loop registers it needs
// for the intermediate grids.
// Which cells of the window whose output starts at column 1 + panel * OW
// are ring columns of the grid (column 0 with rs.left, bw + 1 with
// rs.right), by the lanes' own test in PanelW::setup over the window's 128
// columns (lane l holds columns x0 + 2l and x0 + 2l + 1, x0 odd): 0 none,
// 1 only lanes' first cells, 2 only their second cells, 3 both.
template <int D>
__host__ __device__ inline int ring_window(int panel, int bw, RingSides rs) {
  const int x0 = 1 + panel * window_out_cols(D) - window_margin(D);
  int k = 0;
  if (rs.left && x0 <= 0 && x0 + 127 >= 0) k |= (0 - x0) % 2 ? 2 : 1;
  if (rs.right && x0 <= bw + 1 && x0 + 127 >= bw + 1)
    k |= (bw + 1 - x0) % 2 ? 2 : 1;
  return k;
}

// This wave's tiles (window, band) of one kind: KIND = 1 or 2, the windows
// whose ring columns fall on that cell of the lanes (their steady steps
// keep them by selects, in a steady loop of their own, with a three-slot
// input ring: with six the selects raise the register pressure of the
// sixteen-deep pass enough to spill); KIND = 0, the others (a window with
// ring columns on both cells runs every step guarded).  Waves take the
// tiles round-robin in one order over all kinds, so neighbouring windows
// (which read overlapping columns) run at the same rows at about the same
// time.
template <int D, bool RES, int KIND>
__device__ __forceinline__ void band_tiles(const double* in, double* out,
                                           int64_t P, int bw, int bh,
                                           RingSides rs, int band,
                                           u128acc& acc) {
  const int wave = threadIdx.x / 64, WV = blockDim.x / 64;
  constexpr int OW = window_out_cols(D);
  const int S = (bw + OW - 1) / OW;
  const int nb = (bh + band - 1) / band;
  const EdgeOut2 none = {nullptr, nullptr, nullptr, nullptr};
  for (int t = wave; t < S * nb; t += WV) {
    const int panel = t % S, b = t / S;
    const int k = ring_window<D>(panel, bw, rs);
    if ((KIND == 0) ? (k == 1 || k == 2) : (k != KIND)) continue;
    const int r0 = 1 + b * band, r1 = min(r0 + band, bh + 1);
    sweep_panelD<D, RES, false, 0, KIND ? 3 : kInRing, KIND>(
        in, out, P, bw, bh, 1 + panel * OW, r0, r1, acc, none, rs);
  }
}

template <int D, bool RES>
__device__ __noinline__ void chare_sweepD(const double* in, double* out,
                                          int64_t P, int bw, int bh,
                                          RingSides rs) {
  in = wuni(in);
  out = wuni(out);
  P = wuni(P);
  bw = wuni(bw);
  bh = wuni(bh);
  rs.top = wuni((int)rs.top);
  rs.bottom = wuni((int)rs.bottom);
  rs.left = wuni((int)rs.left);
  rs.right = wuni((int)rs.right);
  // (the wave index stays a vector value: moved to a scalar register, the
  // row bookkeeping derived from it raises scalar-register pressure enough
  // to spill in the deeper sweeps, measured slower)
  const int wave = threadIdx.x / 64, WV = blockDim.x / 64;
  constexpr int OW = window_out_cols(D);
  const int S = (bw + OW - 1) / OW;
  u128acc acc = {0, 0};
  const EdgeOut2 none = {nullptr, nullptr, nullptr, nullptr};
  // bands of 256 rows for tall chares (measured best on 1024-row chares:
  // at sixteen fused sweeps 132.6 ms per 16384^2 solve against 138.9 with
  // 384-row bands, 143.1 with 192 and 185.4 with contiguous shares);
  // shorter chares are cut into contiguous shares
  int band = wuni(g_jdesc.tune_band);
  if (band == 0) band = bh >= 1024 ? 256 : -1;
  if (band > 0) {
    // bands of `band` rows, windows of a band side by side
    if (rs.left || rs.right) {
      band_tiles<D, RES, 1>(in, out, P, bw, bh, rs, band, acc);
      band_tiles<D, RES, 2>(in, out, P, bw, bh, rs, band, acc);
    }
    band_tiles<D, RES, 0>(in, out, P, bw, bh, rs, band, acc);
  } else {
    // the S windows x bh rows, window after window, cut into one contiguous
    // range per wave: equal work whatever S is, at most a few bands each
    const int64_t total = (int64_t)S * bh;
    const int64_t per = (total + WV - 1) / WV;
    int64_t pos = (int64_t)wave * per;
    const int64_t end = min(total, pos + per);
    while (pos < end) {
      const int panel = (int)(pos / bh), r = (int)(pos % bh);
      const int len = (int)min((int64_t)(bh - r), end - pos);
      sweep_panelD<D, RES, false, 0, kInRing>(in, out, P, bw, bh,
                                              1 + panel * OW, 1 + r,
                                              1 + r + len, acc, none, rs);
      pos += len;
    }
  }
  if (RES) {
    acc = wave_reduce_u128(acc);
    if (__lane_id() == 0) atomic_add_u128(g_jdesc.res_acc, acc);
  }
}

// `steps` (3 .. kHaloRows) fused sweeps in one pass: 3 .. 8 on 112-column
// windows, 9 .. 16 on 96-column windows.  Every depth has its own
// instantiation, so no stage carries a run-time copy test (a uniform branch
// per stage would split the step loop into a basic block per stage and
// leave the stages' latencies exposed).
__device__ void sweepD_dispatch(int steps, const double* in, double* out,
                                int64_t P, int bw, int bh, bool res,
                                RingSides rs) {
#define CHM_CASED(D_)                                            \
  case D_:                                                       \
    if (res)                                                     \
      chare_sweepD<D_, true>(in, out, P, bw, bh, rs);            \
    else                                                         \
      chare_sweepD<D_, false>(in, out, P, bw, bh, rs);           \
    break;
  switch (steps) {
    CHM_CASED(3) CHM_CASED(4) CHM_CASED(5) CHM_CASED(6) CHM_CASED(7)
    CHM_CASED(8) CHM_CASED(9) CHM_CASED(10) CHM_CASED(11) CHM_CASED(12)
    CHM_CASED(13) CHM_CASED(14) CHM_CASED(15) CHM_CASED(16)
    default: break;
  }
#undef CHM_CASED
}

// The whole workgroup: the outgoing edges of the array g, D deep, into the
// messages' payloads (layout of EdgeOut2: top = rows 1..D, bottom = rows
// bh-D+1..bh, w each; left = columns 1..D, right = columns bw-D+1..bw, h
// each).  Other waves of the workgroup wrote g just before, partly with
// streaming stores, so the loads go to L2 (agent-scope atomic loads) rather
// than trust the CU's L1.
__device__ __noinline__ void pack_edges(const double* g, int64_t P, int bw,
                                        int bh, int D, EdgeOut2 eo) {
  auto ld = [](const double* p) {
    const uint64_t v = __hip_atomic_load((const uint64_t*)p, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
    return __longlong_as_double((long long)v);
  };
  const int n_row = D * bw, n_col = D * bh;
  for (int k = threadIdx.x; k < n_row; k += blockDim.x) {
    const int i = k / bw, j = 1 + k % bw;
    if (eo.top) eo.top[k] = ld(g + (int64_t)(1 + i) * P + kColOff + j);
    if (eo.bottom)
      eo.bottom[k] = ld(g + (int64_t)(bh - D + 1 + i) * P + kColOff + j);
  }
  for (int k = threadIdx.x; k < n_col; k += blockDim.x) {
    const int j = k / bh, i = 1 + k % bh;
    if (eo.left) eo.left[k] = ld(g + (int64_t)i * P + kColOff + 1 + j);
    if (eo.right)
      eo.right[k] = ld(g + (int64_t)i * P + kColOff + bw - D + 1 + j);
  }
}

__device__ void sweep2_dispatch(const double* in, double* out, int64_t P,
                                int bw, int bh, bool res, bool pack,
                                const EdgeOut2& eo, RingSides rs) {
  // default 1 chunk x 1 row in flight: the fused sweep keeps two rings of
  // rows in registers, and with the 128-VGPR cap of entry methods deeper
  // shapes spill (measured on 16384^2: 1x1 517, 1x2 513, 1x4 473 Gcell/s)
  int nc = g_jdesc.tune_nc, u = g_jdesc.tune_u;
  if (nc == 0) {
    nc = 1;
    u = 1;
  }
  if (u == 0) u = 1;
#define CHM_CASE2(NC_, U_)                                                   \
  if (nc == NC_ && u == U_) {                                                \
    if (res)                                                                 \
      chare_sweep2<NC_, U_, true, false>(in, out, P, bw, bh, eo, rs);        \
    else if (pack)                                                           \
      chare_sweep2<NC_, U_, false, true>(in, out, P, bw, bh, eo, rs);        \
    else                                                                     \
      chare_sweep2<NC_, U_, false, false>(in, out, P, bw, bh, eo, rs);       \
    return;                                                                  \
  }
  CHM_CASE2(1, 1) CHM_CASE2(1, 2) CHM_CASE2(1, 4) CHM_CASE2(2, 1)
  CHM_CASE2(2, 2)
#undef CHM_CASE2
  if (res)
    chare_sweep2<1, 1, true, false>(in, out, P, bw, bh, eo, rs);
  else if (pack)
    chare_sweep2<1, 1, false, true>(in, out, P, bw, bh, eo, rs);
  else
    chare_sweep2<1, 1, false, false>(in, out, P, bw, bh, eo, rs);
}

__device__ void Block::init(void* arg) {
  const JacobiDesc& d = g_jdesc;
  if (threadIdx.x == 0) {
    const int* params = (const int*)arg;
    n_iters = params[2];
    const int cr = params[3], cc = params[4];
    const int index = charm::chare::i;
    const ChareGeom g = chare_geom(params[0], params[1], cr, cc, index);
    block_width = g.bw;
    block_height = g.bh;
    pitch = chare_pitch(g.bw);
    iter = 0;
    cur = 0;
    depth = d.depth;
    row = index / cc;
    col = index % cc;
    neighbor_index[LEFT] = (col == 0) ? -1 : index - 1;
    neighbor_index[RIGHT] = (col == cc - 1) ? -1 : index + 1;
    neighbor_index[TOP] = (row == 0) ? -1 : index - cc;
    neighbor_index[BOTTOM] = (row == cr - 1) ? -1 : index + cc;
    // diagonal neighbours supply the corner blocks of deeper halos
    const bool dg = depth >= 2;
    neighbor_index[TOP_LEFT] = (dg && row > 0 && col > 0) ? index - cc - 1 : -1;
    neighbor_index[TOP_RIGHT] =
        (dg && row > 0 && col < cc - 1) ? index - cc + 1 : -1;
    neighbor_index[BOTTOM_LEFT] =
        (dg && row < cr - 1 && col > 0) ? index + cc - 1 : -1;
    neighbor_index[BOTTOM_RIGHT] =
        (dg && row < cr - 1 && col < cc - 1) ? index + cc + 1 : -1;
    neighbor_count = 0;
    for (int k = 0; k < N_NEIGHBORS; k++)
      neighbor_count += (neighbor_index[k] >= 0);
    recv_count = 0;
    end_count = 0;
    t_sweep = t_send = t_wait = 0;
    grids[0] = d.arena + (int64_t)(2 * index) * d.chare_elems +
               (kHaloRows - 1) * pitch;
    grids[1] = grids[0] + d.chare_elems;
    for (int k = 0; k < N_NEIGHBORS; k++) {
      boundaries[k] = d.staging + (int64_t)(N_NEIGHBORS * index + k) * d.stage_elems;
      // dir tag = the receiver's side (jacobi2d.cu:139-140)
      ((int*)boundaries[k])[0] = opposite_side(k);
      ((int*)boundaries[k])[1] = 0;
    }
    block_proxy->set_refnum(index, 0);
  }
  __syncthreads();
  // pack the edges of u_0, `depth` deep (later rounds pack inside the sweep)
  const double* t = grids[0];
  const int bw = block_width, bh = block_height, D = depth;
  const int64_t P = pitch;
  for (int dd = 0; dd < D; dd++) {
    for (int k = threadIdx.x; k < bw; k += blockDim.x) {
      boundaries[TOP][1 + dd * bw + k] = t[P * (1 + dd) + kColOff + 1 + k];
      boundaries[BOTTOM][1 + dd * bw + k] =
          t[P * (bh - D + 1 + dd) + kColOff + 1 + k];
    }
    for (int k = threadIdx.x; k < bh; k += blockDim.x) {
      boundaries[LEFT][1 + dd * bh + k] = t[P * (1 + k) + kColOff + 1 + dd];
      boundaries[RIGHT][1 + dd * bh + k] =
          t[P * (1 + k) + kColOff + bw - D + 1 + dd];
    }
  }
  __syncthreads();
  send_boundaries();
  if (threadIdx.x == 0) t_mark = wall_clock64();
  if (neighbor_count == 0) charm::request_work(this);
}

// Bytes of the halo message toward neighbour k: the dir slot, then `depth`
// rows/columns of the edge, or a (depth-1) x (depth-1) corner block.
__device__ size_t Block::halo_bytes(int k) const {
  size_t len = (size_t)(depth - 1) * (depth - 1);
  if (k == LEFT || k == RIGHT) len = (size_t)depth * block_height;
  if (k == TOP || k == BOTTOM) len = (size_t)depth * block_width;
  return sizeof(double) + sizeof(double) * len;
}

// Thread 0: with halos D >= 2 deep the four (D-1) x (D-1) corner blocks go
// to the diagonal neighbours, taken from the top and bottom edges already
// written (a diagonal neighbour implies both): the block nearest each
// corner of the chare, rows in order.  bufs[k] = [dir slot | payload].
// (A receiver needs cells within D-1 steps of its corner, which the
// (D-1) x (D-1) block covers; the farthest one is unused.)
__device__ void Block::fill_corners(double* const* bufs) {
  const int D = depth;
  if (D < 2) return;
  const int bw = block_width, n = D - 1;
  const double* top = bufs[TOP] + 1;                 // rows 1 .. D
  const double* bot = bufs[BOTTOM] + 1 + (int64_t)bw;  // rows bh-D+2 .. bh
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      if (neighbor_index[TOP_LEFT] >= 0)
        bufs[TOP_LEFT][1 + i * n + j] = top[(int64_t)i * bw + j];
      if (neighbor_index[TOP_RIGHT] >= 0)
        bufs[TOP_RIGHT][1 + i * n + j] = top[(int64_t)i * bw + bw - n + j];
      if (neighbor_index[BOTTOM_LEFT] >= 0)
        bufs[BOTTOM_LEFT][1 + i * n + j] = bot[(int64_t)i * bw + j];
      if (neighbor_index[BOTTOM_RIGHT] >= 0)
        bufs[BOTTOM_RIGHT][1 + i * n + j] = bot[(int64_t)i * bw + bw - n + j];
    }
}

__device__ void Block::send_boundaries() {
  // (jacobi2d.cu:152-194) sends tagged with the iteration: four edges
  // `depth` deep, and with deeper halos the four corner blocks, copied
  // from the staging buffers
  if (threadIdx.x == 0) fill_corners(boundaries);
  __syncthreads();
  for (int k = 0; k < N_NEIGHBORS; k++) {
    if (neighbor_index[k] < 0) continue;
    block_proxy->invoke(neighbor_index[k], 1, boundaries[k], halo_bytes(k),
                        iter);
  }
}

// Deep rounds with zero-copy halos: the workgroup writes the new grid's
// edges, D deep, and its (D-1) x (D-1) corner blocks straight into the halo
// of each neighbour's array of the same parity (the one its next round
// reads), exactly where the neighbour's unpack would put them.  No
// conflict: the neighbour's current round reads its other array and writes
// only the interior of this one, and its previous round (which read this
// halo) has ended, or this chare could not have run its own.  The messages
// that follow only notify (slot word 1 = 1: placed).  g was just written by
// other waves, partly with streaming stores: loads at agent scope.
__device__ void Block::put_edges(const double* g, int D, int parity) {
  const JacobiDesc& d = g_jdesc;
  const int bw = block_width, bh = block_height;
  const int64_t P = pitch;
  for (int k = 0; k < N_NEIGHBORS; k++) {
    const int n = neighbor_index[k];
    if (n < 0) continue;
    const ChareGeom gn = chare_geom(d.W, d.H, d.cr, d.cc, n);
    const int64_t PN = chare_pitch(gn.bw);
    double* t = d.arena + (int64_t)(2 * n + parity) * d.chare_elems +
                (kHaloRows - 1) * PN;
    // rows x columns of the source block in g (from si0, sj0) and of its
    // place in the neighbour's halo (from di0, dj0)
    int si0, sj0, di0, dj0, ni, nj;
    switch (k) {
      case LEFT:  // its right halo
        si0 = 1; sj0 = 1; di0 = 1; dj0 = gn.bw + 1; ni = bh; nj = D; break;
      case RIGHT:  // its left halo
        si0 = 1; sj0 = bw - D + 1; di0 = 1; dj0 = 1 - D; ni = bh; nj = D;
        break;
      case TOP:  // its bottom halo
        si0 = 1; sj0 = 1; di0 = gn.bh + 1; dj0 = 1; ni = D; nj = bw; break;
      case BOTTOM:  // its top halo
        si0 = bh - D + 1; sj0 = 1; di0 = 1 - D; dj0 = 1; ni = D; nj = bw;
        break;
      case TOP_LEFT:  // its bottom-right corner
        si0 = 1; sj0 = 1; di0 = gn.bh + 1; dj0 = gn.bw + 1; ni = nj = D - 1;
        break;
      case TOP_RIGHT:  // its bottom-left corner
        si0 = 1; sj0 = bw - D + 2; di0 = gn.bh + 1; dj0 = 2 - D;
        ni = nj = D - 1;
        break;
      case BOTTOM_LEFT:  // its top-right corner
        si0 = bh - D + 2; sj0 = 1; di0 = 2 - D; dj0 = gn.bw + 1;
        ni = nj = D - 1;
        break;
      default:  // BOTTOM_RIGHT: its top-left corner
        si0 = bh - D + 2; sj0 = bw - D + 2; di0 = 2 - D; dj0 = 2 - D;
        ni = nj = D - 1;
        break;
    }
    for (int e = threadIdx.x; e < ni * nj; e += blockDim.x) {
      const int i = e / nj, j = e % nj;
      const uint64_t v = __hip_atomic_load(
          (const uint64_t*)(g + (int64_t)(si0 + i) * P + kColOff + sj0 + j),
          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      t[(int64_t)(di0 + i) * PN + kColOff + dj0 + j] =
          __longlong_as_double((long long)v);
    }
  }
}

__device__ void Block::recv_ghosts(void* arg) {
  // (jacobi2d.cu:196-229) unpack into the halo of u_iter (unless the
  // sender placed the cells itself: slot word 1)
  const uint64_t tr0 = wall_clock64();
  const int dir = *(const int*)arg;
  const bool placed = ((const int*)arg)[1] == 1;
  const double* ghost = (const double*)arg + 1;
  double* t = grids[cur];
  const int bw = block_width, bh = block_height;
  const int64_t P = pitch;
  const int D = depth;
  // edge messages hold `depth` rows/columns, nearest the sender's far side
  // first: they land in halo rows/columns 1-D .. 0 or n+1 .. n+D
  if (dir < 0 || dir >= N_NEIGHBORS) {
    if (threadIdx.x == 0) charm::rt::raise_error(charm::rt::ERR_BAD_MESSAGE, dir);
  } else if (!placed) switch (dir) {
    case LEFT:
      for (int k = threadIdx.x; k < D * bh; k += blockDim.x)
        t[P * (1 + k % bh) + kColOff + 1 - D + k / bh] = ghost[k];
      break;
    case RIGHT:
      for (int k = threadIdx.x; k < D * bh; k += blockDim.x)
        t[P * (1 + k % bh) + kColOff + bw + 1 + k / bh] = ghost[k];
      break;
    case TOP:
      for (int k = threadIdx.x; k < D * bw; k += blockDim.x)
        t[P * (1 - D + k / bw) + kColOff + 1 + k % bw] = ghost[k];
      break;
    case BOTTOM:
      for (int k = threadIdx.x; k < D * bw; k += blockDim.x)
        t[P * (bh + 1 + k / bw) + kColOff + 1 + k % bw] = ghost[k];
      break;
    // corner blocks of (D-1) x (D-1): rows 2-D .. 0 or bh+1 .. bh+D-1,
    // columns 2-D .. 0 or bw+1 .. bw+D-1
    case TOP_LEFT:
    case TOP_RIGHT:
    case BOTTOM_LEFT:
    case BOTTOM_RIGHT: {
      const int n = D - 1;
      const int i0 = (dir == TOP_LEFT || dir == TOP_RIGHT) ? 2 - D : bh + 1;
      const int j0 = (dir == TOP_LEFT || dir == BOTTOM_LEFT) ? 2 - D : bw + 1;
      for (int k = threadIdx.x; k < n * n; k += blockDim.x)
        t[P * (i0 + k / n) + kColOff + j0 + k % n] = ghost[k];
      break;
    }
    default:
      if (threadIdx.x == 0) charm::rt::raise_error(charm::rt::ERR_BAD_MESSAGE, dir);
      break;
  }
  __shared__ int go;
  __syncthreads();
  if (threadIdx.x == 0) {
    go = (++recv_count == neighbor_count);
    if (go) recv_count = 0;
    if (g_jdesc.recv_ticks)
      g_jdesc.recv_ticks[charm::chare::i] += wall_clock64() - tr0;
  }
  __syncthreads();
  if (go) charm::request_work(this);
}

__device__ void Block::update() {
  const int bw = block_width, bh = block_height;
  const int64_t P = pitch;
  for (;;) {
    uint64_t t0 = 0;
    if (threadIdx.x == 0) {
      t0 = wall_clock64();
      t_wait += t0 - t_mark;
    }
    // `depth` fused sweeps per exchange (fewer in a final round that has
    // fewer left)
    const int steps = min(depth, n_iters - iter);
    const bool last = (iter + steps == n_iters);
    double* t = grids[cur];
    double* nt = grids[cur ^ 1];
    // overlap: phase 0 = everything that feeds the outgoing halos, sent
    // before phase 1 sweeps the interior; otherwise one whole-block sweep
    const bool ov = g_jdesc.overlap != 0 && depth == 1;
    // zero-copy halos: the outgoing messages are allocated in the
    // neighbours' inboxes first and the sweep writes the edges straight
    // into them (no staging copy); sent after the sweep
    const bool zc = !ov && !last && g_jdesc.halo_copy == 0;
    // deep rounds put their edges straight into the neighbours' halos
    // (put_edges): their messages only notify
    const bool put = zc && steps >= 3;
    // (kept in LDS: the values are the workgroup's, and the stack of an
    // entry method is small)
    __shared__ charm::envelope* s_env[N_NEIGHBORS];
    __shared__ size_t s_off[N_NEIGHBORS];
    __shared__ double* s_dst[N_NEIGHBORS];
    for (int k = 0; k < N_NEIGHBORS; k++) {
      double* d = (neighbor_index[k] >= 0) ? boundaries[k] : nullptr;
      if (zc && neighbor_index[k] >= 0) {
        charm::message m;
        d = (double*)block_proxy->alloc(m, neighbor_index[k], 1,
                                        put ? sizeof(double) : halo_bytes(k),
                                        iter + steps);
        if (threadIdx.x == 0) {
          s_env[k] = m.env;
          s_off[k] = m.offset;
          if (d) {
            ((int*)d)[0] = opposite_side(k);
            ((int*)d)[1] = put ? 1 : 0;
          }
        }
      }
      if (threadIdx.x == 0) s_dst[k] = d;
    }
    __syncthreads();
    EdgeOut eo;
    eo.top = s_dst[TOP] ? s_dst[TOP] + 1 : nullptr;
    eo.bottom = s_dst[BOTTOM] ? s_dst[BOTTOM] + 1 : nullptr;
    eo.left = s_dst[LEFT] ? s_dst[LEFT] + 1 : nullptr;
    eo.right = s_dst[RIGHT] ? s_dst[RIGHT] + 1 : nullptr;
    if (steps >= 2) {
      EdgeOut2 eo2 = {eo.top, eo.bottom, eo.left, eo.right};
      const RingSides rs = {neighbor_index[TOP] < 0, neighbor_index[BOTTOM] < 0,
                            neighbor_index[LEFT] < 0,
                            neighbor_index[RIGHT] < 0};
      if (steps >= 3) {
        sweepD_dispatch(steps, t, nt, P, bw, bh, last, rs);
        if (!last) {
          __syncthreads();
          if (put)
            put_edges(nt, steps, cur ^ 1);
          else
            pack_edges(nt, P, bw, bh, steps, eo2);
        }
      } else
        sweep2_dispatch(t, nt, P, bw, bh, last, !last, eo2, rs);
    } else if (ov) {
      sweep_dispatch(t, nt, P, bw, bh, last, eo, 0);
    } else {
      sweep_dispatch(t, nt, P, bw, bh, last, eo, 2);
    }
    __syncthreads();
    __shared__ int done;
    uint64_t t1 = 0, t2 = 0;
    if (threadIdx.x == 0) {
      t1 = wall_clock64();
      t_sweep += t1 - t0;
      iter += steps;
      cur ^= 1;
      done = (iter == n_iters);
      if (!done) block_proxy->set_refnum(charm::chare::i, iter);
    }
    __syncthreads();
    // the halos of u_{iter} leave before the interior is swept
    if (!done && zc) {
      if (threadIdx.x == 0 && !put) fill_corners(s_dst);
      for (int k = 0; k < N_NEIGHBORS; k++) {
        if (neighbor_index[k] < 0) continue;
        charm::message m;
        m.env = s_env[k];
        m.offset = s_off[k];
        m.dst_pe = block_proxy->pe_of(neighbor_index[k]);
        block_proxy->send(m);
      }
    } else if (!done) {
      send_boundaries();
      if (ov) charm::flush_sends();
    }
    if (threadIdx.x == 0) {
      t2 = wall_clock64();
      t_send += t2 - t1;
    }
    if (ov) sweep_dispatch(t, nt, P, bw, bh, last, eo, 1);
    __syncthreads();
    if (threadIdx.x == 0) {
      t_mark = wall_clock64();
      t_sweep += t_mark - t2;
    }
    if (done) {
      if (threadIdx.x == 0 && g_jdesc.phase) {
        uint64_t* ph = g_jdesc.phase + 4 * (int64_t)charm::chare::i;
        ph[0] = t_sweep;
        ph[1] = t_send;
        ph[2] = t_wait;
        ph[3] = (uint64_t)n_iters;
      }
      block_proxy->invoke(0, 2);
      return;
    }
    if (neighbor_count > 0) return;
  }
}

__device__ void Block::end(void* arg) {
  __shared__ int fin;
  if (threadIdx.x == 0) fin = (++end_count == charm::chare::n);
  __syncthreads();
  if (fin) charm::end();
}

// ---------------------------------------------------------------------------
// host side of the persistent form
// ---------------------------------------------------------------------------
namespace chm {
namespace jac {

// u_0 for every chare: interior from the counter-based generator at global
// coordinates (decomposition invariant), halo ring = ring value, padding 0.
// Halo cells get u_0 of the global cell they mirror (ring value on the
// ring), so the values there are right before the first exchange; cells
// outside the global grid and padding get 0.
__global__ void init_chares_kernel(JacobiDesc d, uint64_t seed, int mode,
                                   double boundary) {
  const int c = blockIdx.y;
  const ChareGeom g = chare_geom(d.W, d.H, d.cr, d.cc, c);
  const int64_t P = chare_pitch(g.bw);
  const int64_t n = (int64_t)(g.bh + 2 * kHaloRows) * P;
  double* a = d.arena + (int64_t)(2 * c) * d.chare_elems;
  double* b = a + d.chare_elems;
  const double ring = (mode == CHM_INIT_REFERENCE) ? 10.0 : boundary;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / P - (kHaloRows - 1), j = e % P - kColOff;
    const int64_t gi = g.row0 + i, gj = g.col0 + j;
    double v = 0.0;
    if (j >= 1 - kHaloRows && j <= g.bw + kHaloRows && gi >= 0 &&
        gi <= d.H + 1 && gj >= 0 && gj <= d.W + 1) {
      const bool interior = gi >= 1 && gi <= d.H && gj >= 1 && gj <= d.W;
      if (!interior)
        v = ring;
      else
        v = (mode == CHM_INIT_REFERENCE) ? 0.0 : cell_u01(seed, gi, gj);
    }
    a[e] = v;
    b[e] = v;
  }
}

// Hash of the cells of global rows [r0,r1) x cols [c0,c1) over all chares.
__global__ void hash_chares_kernel(JacobiDesc d, int parity, int r0, int r1,
                                   int c0, int c1, uint64_t* out) {
  const int c = blockIdx.y;
  const ChareGeom g = chare_geom(d.W, d.H, d.cr, d.cc, c);
  const int64_t P = chare_pitch(g.bw);
  const double* a =
      d.arena + (int64_t)(2 * c + parity) * d.chare_elems + (kHaloRows - 1) * P;
  const int lr0 = max(r0, g.row0 + 1), lr1 = min(r1, g.row0 + 1 + g.bh);
  const int lc0 = max(c0, g.col0 + 1), lc1 = min(c1, g.col0 + 1 + g.bw);
  uint64_t h = 0;
  if (lr0 < lr1 && lc0 < lc1) {
    const int64_t cols = lc1 - lc0, n = (int64_t)(lr1 - lr0) * cols;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
      const int64_t gi = lr0 + e / cols, gj = lc0 + e % cols;
      const int64_t li = gi - g.row0, lj = gj - g.col0;
      h += hash_cell(gi, gj, a[li * P + kColOff + lj]);
    }
  }
  for (int off = 32; off >= 1; off >>= 1) h += __shfl_down(h, off, 64);
  if (__lane_id() == 0 && h)
    __hip_atomic_fetch_add(out, h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct PersistentState {
  JacobiDesc desc;
  ChareGeom max_geom;
  int n_chares;
  int min_side;  // shortest chare side
  bool own_runtime;
};

static int g_runtime_users = 0;

static void surface_split(int n, int W, int H, int* cr, int* cc) {
  // jacobi2d.cu:37-65: minimise the halo surface; rows divide the height
  double best = 2.0 * ((double)H + (double)W);
  *cr = 1;
  *cc = n;
  for (int ipx = 1; ipx <= n; ipx++) {
    if (n % ipx) continue;
    const int ipy = n / ipx;
    const double surf = 2.0 * ((double)H / ipx + (double)W / ipy);
    if (surf < best) {
      best = surf;
      *cr = ipx;
      *cc = ipy;
    }
  }
}

static int upload_desc(const JacobiDesc& d) {
  CHM_TRY_HIP(hipMemcpyToSymbol(HIP_SYMBOL(g_jdesc), &d, sizeof(d)));
  return CHM_OK;
}

int persistent_create(Handle* h) {
  PersistentState* ps = new PersistentState();
  memset(&ps->desc, 0, sizeof(ps->desc));
  h->ps = ps;
  if (g_runtime_users == 0) {
    CHM_TRY(chm_select_program("persistent.hip") == CHM_OK
                ? CHM_OK
                : fail(CHM_E_STATE, "Jacobi2D chare program not registered"));
    chm_config cfg;
    memset(&cfg, 0, sizeof(cfg));
    // watchdog: generous for the largest grids this form is used on
    // (16384^2 x 1000 sweeps takes about a second)
    cfg.timeout_s = 120.0;
    const int rc = chm_init(0, nullptr, &cfg);
    if (rc != CHM_OK) return fail(rc, "chm_init: %s", chm_last_error());
  }
  g_runtime_users++;
  ps->own_runtime = true;
  chm_stats st;
  CHM_TRY(chm_get_stats(&st) == CHM_OK
              ? CHM_OK
              : fail(CHM_E_STATE, "runtime stats: %s", chm_last_error()));
  const int n_pes = st.n_pes;
  int cr = h->prob.chares_r, cc = h->prob.chares_c;
  if (cr <= 0 || cc <= 0) {
    const int n = h->prob.n_chares > 0 ? h->prob.n_chares : n_pes;
    if (n < 1) return fail(CHM_E_INVALID, "no chares to place (n_pes %d)", n_pes);
    surface_split(n, h->W, h->H, &cr, &cc);
  }
  if (cr < 1 || cc < 1 || cr > h->H || cc > h->W)
    return fail(CHM_E_INVALID, "%dx%d chares do not fit a %dx%d grid", cr, cc,
                h->W, h->H);
  JacobiDesc& d = ps->desc;
  d.W = h->W;
  d.H = h->H;
  d.cr = cr;
  d.cc = cc;
  d.iters = h->prob.iters;
  ps->n_chares = cr * cc;
  // the last chare row/column is the largest
  const ChareGeom gmax = chare_geom(h->W, h->H, cr, cc, cr * cc - 1);
  ps->max_geom = gmax;
  d.chare_elems = (int64_t)array_elems_depth(gmax.bw, gmax.bh, kHaloRows);
  d.chare_elems = (d.chare_elems + 31) / 32 * 32;  // 256-byte aligned slots
  d.stage_elems =
      ((int64_t)kHaloRows * std::max(gmax.bw, gmax.bh) + 2 + 15) / 16 * 16;
  // the smallest chare decides whether two-deep halos are possible
  const ChareGeom gmin = chare_geom(h->W, h->H, cr, cc, 0);
  ps->min_side = std::min(gmin.bw, gmin.bh);
  const size_t arena_bytes =
      (size_t)ps->n_chares * 2 * (size_t)d.chare_elems * sizeof(double);
  const size_t stage_bytes = (size_t)ps->n_chares * N_NEIGHBORS *
                             (size_t)d.stage_elems * sizeof(double);
  if (hipMalloc(&d.arena, arena_bytes) != hipSuccess ||
      hipMalloc(&d.staging, stage_bytes) != hipSuccess ||
      hipMalloc(&d.res_acc, 4 * sizeof(uint64_t)) != hipSuccess ||
      hipMalloc(&d.phase, (size_t)ps->n_chares * 4 * sizeof(uint64_t)) !=
          hipSuccess ||
      hipMalloc(&d.recv_ticks, (size_t)ps->n_chares * sizeof(uint64_t)) !=
          hipSuccess)
    return fail(CHM_E_NOMEM, "persistent form allocation (%zu + %zu bytes)",
                arena_bytes, stage_bytes);
  return persistent_reset(h);
}

int persistent_reset(Handle* h) {
  PersistentState* ps = h->ps;
  init_chares_kernel<<<dim3(64, ps->n_chares), 256, 0, h->stream>>>(
      ps->desc, h->prob.seed, h->prob.init_mode, h->prob.boundary);
  CHM_TRY_HIP(hipGetLastError());
  CHM_TRY_HIP(hipStreamSynchronize(h->stream));
  h->cur = 0;
  return CHM_OK;
}

int persistent_solve(Handle* h, chm_jacobi_result* res) {
  PersistentState* ps = h->ps;
  JacobiDesc& d = ps->desc;
  if (h->cur != 0)
    return fail(CHM_E_STATE, "persistent form restarts from u_0: call reset");
  d.iters = h->prob.iters;
  // zero = choose per chare width on the device
  d.tune_nc = h->tune.panel_chunks;
  d.tune_u = h->tune.rows_in_flight;
  d.tune_band = h->tune.band_rows;
  d.overlap = h->tune.overlap;
  d.halo_copy = h->tune.halo_copy;
  // halos D deep need every chare at least D x D (all of a halo then comes
  // from the adjacent chares)
  // (default: as deep as the chares allow, up to kHaloRows: fewer exchange
  // rounds and fewer HBM passes, faster on every decomposition measured,
  // from 128^2 to 1024^2 chares.  The deep pass has one sweep shape, so a
  // request for another shape keeps the two-sweep pass.)
  if (h->tune.depth) {
    d.depth = h->tune.depth;
  } else if (d.tune_nc <= 1 && d.tune_u <= 1) {
    d.depth = std::min(ps->min_side, kHaloRows);
  } else {
    d.depth = ps->min_side >= 2 ? 2 : 1;
  }
  if (d.depth < 1 || d.depth > kHaloRows)
    return fail(CHM_E_INVALID, "halo depth %d (1 .. %d)", d.depth, kHaloRows);
  if (ps->min_side < d.depth)
    return fail(CHM_E_INVALID, "halo depth %d needs chares of at least %dx%d",
                d.depth, d.depth, d.depth);
  if (d.depth >= 3 && (d.tune_nc > 1 || d.tune_u > 1))
    // the deep fused pass (sweepD_dispatch) has one shape
    return fail(CHM_E_INVALID,
                "halo depth %d sweeps with panel_chunks x rows_in_flight of "
                "1x1 only (asked %dx%d)", d.depth, d.tune_nc, d.tune_u);
  if (d.depth == 2 && d.tune_nc) {
    // shapes of the fused two-sweep pass (sweep2_dispatch)
    const int nc = d.tune_nc, u = d.tune_u ? d.tune_u : 1;
    const bool ok = (nc == 1 && (u == 1 || u == 2 || u == 4)) ||
                    (nc == 2 && (u == 1 || u == 2));
    if (!ok)
      return fail(CHM_E_INVALID,
                  "halo depth 2 sweeps with panel_chunks x rows_in_flight of "
                  "1x1, 1x2, 1x4, 2x1 or 2x2 (asked %dx%d)", nc, u);
  }
  CHM_TRY(upload_desc(d));
  CHM_TRY_HIP(hipMemset(d.res_acc, 0, 4 * sizeof(uint64_t)));
  CHM_TRY_HIP(hipMemset(d.recv_ticks, 0, (size_t)ps->n_chares * 8));
  CHM_TRY_HIP(hipDeviceSynchronize());
  if (d.iters == 0) {
    h->cur = 0;
  } else {
    const int rc = chm_run();
    if (rc != CHM_OK) return fail(rc, "persistent run: %s", chm_last_error());
    // one array flip per exchange round
    const int rounds = (d.iters + d.depth - 1) / d.depth;
    h->cur = rounds & 1;
  }
  if (res) {
    memset(res, 0, sizeof(*res));
    chm_stats st;
    chm_get_stats(&st);
    res->device_ms = d.iters ? st.kernel_ms : 0.0;
    res->iters_done = d.iters;
    res->launches = d.iters ? 1 : 0;
    res->messages = (long long)st.delivered;
    res->deferred = (long long)st.deferred;
    res->sched_overhead_ms = st.overhead_ms_per_pe;
    uint64_t acc[2];
    CHM_TRY_HIP(hipMemcpy(acc, d.res_acc, sizeof(acc), hipMemcpyDeviceToHost));
    res->residual_lo = acc[0];
    res->residual_hi = acc[1];
    CHM_TRY(persistent_hash_region(h, 1, h->H + 1, 1, h->W + 1, &res->hash));
  }
  return CHM_OK;
}

int persistent_hash_region(Handle* h, int r0, int r1, int c0, int c1,
                           uint64_t* hash) {
  PersistentState* ps = h->ps;
  CHM_TRY_HIP(hipMemsetAsync(h->d_acc, 0, sizeof(uint64_t), h->stream));
  hash_chares_kernel<<<dim3(32, ps->n_chares), 256, 0, h->stream>>>(
      ps->desc, h->cur, r0, r1, c0, c1, h->d_acc);
  CHM_TRY_HIP(hipGetLastError());
  CHM_TRY_HIP(hipMemcpyAsync(hash, h->d_acc, sizeof(uint64_t),
                             hipMemcpyDeviceToHost, h->stream));
  CHM_TRY_HIP(hipStreamSynchronize(h->stream));
  return CHM_OK;
}

int persistent_read_grid(Handle* h, double* out) {
  PersistentState* ps = h->ps;
  const JacobiDesc& d = ps->desc;
  const int64_t GP = (int64_t)h->W + 2;
  const double ring =
      (h->prob.init_mode == CHM_INIT_REFERENCE) ? 10.0 : h->prob.boundary;
  for (int64_t e = 0; e < GP * (h->H + 2); e++) out[e] = ring;
  for (int c = 0; c < ps->n_chares; c++) {
    const ChareGeom g = chare_geom(d.W, d.H, d.cr, d.cc, c);
    const int64_t P = chare_pitch(g.bw);
    const double* a = d.arena + (int64_t)(2 * c + h->cur) * d.chare_elems +
                      (kHaloRows - 1) * P;
    CHM_TRY_HIP(hipMemcpy2DAsync(
        out + (int64_t)(g.row0 + 1) * GP + g.col0 + 1, GP * sizeof(double),
        a + P + kColOff + 1, P * sizeof(double), (size_t)g.bw * sizeof(double),
        (size_t)g.bh, hipMemcpyDeviceToHost, h->stream));
  }
  CHM_TRY_HIP(hipStreamSynchronize(h->stream));
  return CHM_OK;
}

int persistent_chare_count(Handle* h, int* n) {
  *n = h->ps->n_chares;
  return CHM_OK;
}

int persistent_chare_phases(Handle* h, double* out) {
  PersistentState* ps = h->ps;
  std::vector<uint64_t> ph((size_t)ps->n_chares * 4);
  CHM_TRY_HIP(hipMemcpy(ph.data(), ps->desc.phase, ph.size() * 8,
                        hipMemcpyDeviceToHost));
  int khz = 100000;
  (void)hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, 0);
  const double ms_per_tick = 1.0 / (double)khz;
  std::vector<uint64_t> rt((size_t)ps->n_chares);
  CHM_TRY_HIP(hipMemcpy(rt.data(), ps->desc.recv_ticks, rt.size() * 8,
                        hipMemcpyDeviceToHost));
  for (int c = 0; c < ps->n_chares; c++) ph[4 * c + 3] = rt[c];
  for (size_t e = 0; e < ph.size(); e++) out[e] = (double)ph[e] * ms_per_tick;
  return CHM_OK;
}

int persistent_phase_stats(Handle* h, double* out) {
  PersistentState* ps = h->ps;
  std::vector<double> per((size_t)ps->n_chares * 4);
  CHM_TRY(persistent_chare_phases(h, per.data()));
  for (int k = 0; k < 4; k++) {
    double sum = 0, mx = 0;
    for (int c = 0; c < ps->n_chares; c++) {
      const double v = per[4 * c + k];
      sum += v;
      mx = std::max(mx, v);
    }
    out[k] = sum / ps->n_chares;
    out[4 + k] = mx;
  }
  return CHM_OK;
}

void persistent_destroy(Handle* h) {
  PersistentState* ps = h->ps;
  if (!ps) return;
  if (ps->desc.arena) (void)hipFree(ps->desc.arena);
  if (ps->desc.staging) (void)hipFree(ps->desc.staging);
  if (ps->desc.res_acc) (void)hipFree(ps->desc.res_acc);
  if (ps->desc.phase) (void)hipFree(ps->desc.phase);
  if (ps->desc.recv_ticks) (void)hipFree(ps->desc.recv_ticks);
  if (ps->own_runtime && --g_runtime_users == 0) chm_finalize();
  delete ps;
  h->ps = nullptr;
}

}  // namespace jac
}  // namespace chm

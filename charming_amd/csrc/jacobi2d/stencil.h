// FP64 5-point Jacobi sweep for gfx950, shared by every form of the solver
// (plain per-launch, hipGraph, message-driven persistent, batched, served).
//
// Arithmetic follows examples/jacobi2d/jacobi2d.cu:327-335 exactly:
//   u'[i][j] = (u[i][j] + u[i][j-1] + u[i][j+1] + u[i-1][j] + u[i+1][j]) * 0.2
// evaluated left to right, compiled with -ffp-contract=off.
//
// Array layout (one array = one chare block or one whole grid):
//   (h + 2) rows of `pitch` doubles; row i holds the halo column 0, interior
//   columns 1..w and the halo column w+1 at element offsets 15 .. 16 + w, so
//   interior column 1 starts on a 128-byte boundary and every wave load of
//   two doubles per lane (dwordx4) is aligned and coalesced.  Rows 0 and h+1
//   are the halo rows.  pitch = roundup(w + 18, 16) leaves a spare element
//   after the right halo so a pair load that starts on column w+1 stays in
//   the row.
//
// Work decomposition: one wavefront sweeps a panel of NC chunks (a chunk is
// 64 lanes x 2 doubles = 128 columns) down a band of rows.  Each lane keeps
// the rows above/at/below in registers (a rolling window), so every input
// element is loaded once per panel; the west/east neighbours come from the
// adjacent lane (cross-lane moves), from the adjacent chunk (readlane) or,
// at the panel's two ends, from one extra element load by lane 0 / lane 63.
// Loads for the next U rows are issued before the current U rows are
// computed, so each wave keeps U * NC KiB of reads in flight.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace chm {
namespace jac {

constexpr int kColOff = 15;  // element offset of halo column 0 within a row

__host__ __device__ inline int64_t pitch_for(int w) {
  return ((int64_t)w + 18 + 15) / 16 * 16;
}
__host__ __device__ inline size_t array_elems(int w, int h) {
  return (size_t)(h + 2) * (size_t)pitch_for(w);
}
// Arrays with halos `d` cells deep (temporal blocking over d sweeps): rows
// 1-d .. h+d, columns 1-d .. w+d at element offset 15 + col; the pitch keeps
// two spare elements after column w+d for the pair loads.
__host__ __device__ inline int64_t pitch_for_depth(int w, int d) {
  return ((int64_t)w + 16 + 2 * d + 15) / 16 * 16;
}
__host__ __device__ inline size_t array_elems_depth(int w, int h, int d) {
  return (size_t)(h + 2 * d) * (size_t)pitch_for_depth(w, d);
}

// ----------------------------------------------------------------------------
// exact fixed-point residual: sum floor((a-b)^2 * 2^56) in 128 bits
// ----------------------------------------------------------------------------
struct u128acc {
  uint64_t lo, hi;
};
__device__ inline void acc_add(u128acc& a, uint64_t v) {
  uint64_t lo = a.lo + v;
  a.hi += (lo < a.lo) ? 1u : 0u;
  a.lo = lo;
}
__device__ inline void acc_add(u128acc& a, u128acc b) {
  uint64_t lo = a.lo + b.lo;
  a.hi += b.hi + ((lo < a.lo) ? 1u : 0u);
  a.lo = lo;
}
__device__ inline uint64_t res_term(double a, double b) {
  double d = a - b;
  double q = d * d * 0x1.0p56;
  if (!(q < 0x1.0p64)) return ~0ull;
  return (uint64_t)q;
}

// Exact 128-bit add into a global accumulator {lo, hi} with two 64-bit
// atomics: the carry out of the low word is added to the high word, so the
// final value is the exact sum whatever the interleaving.
__device__ inline void atomic_add_u128(uint64_t* dst, u128acc v) {
  uint64_t old = __hip_atomic_fetch_add(dst, v.lo, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
  uint64_t carry = (old + v.lo < old) ? 1u : 0u;
  uint64_t hi = v.hi + carry;
  if (hi)
    __hip_atomic_fetch_add(dst + 1, hi, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// ----------------------------------------------------------------------------
// splitmix64 counter-based input and order-independent grid hash
// ----------------------------------------------------------------------------
__host__ __device__ inline uint64_t sm64_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t cell_key(int64_t row, int64_t col) {
  return ((uint64_t)row << 32) | (uint64_t)(uint32_t)col;
}
__host__ __device__ inline double cell_u01(uint64_t seed, int64_t row,
                                           int64_t col) {
  uint64_t z = seed + (cell_key(row, col) + 1) * 0x9E3779B97F4A7C15ull;
  return (double)(sm64_mix(z) >> 11) * 0x1.0p-53;
}
__device__ inline uint64_t hash_cell(int64_t row, int64_t col, double v) {
  return sm64_mix((uint64_t)__double_as_longlong(v) +
                  cell_key(row, col) * 0x9E3779B97F4A7C15ull);
}

// ----------------------------------------------------------------------------
// cross-lane helpers (wave64)
// ----------------------------------------------------------------------------
__device__ inline double readlane_d(double v, int lane) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_readlane(lo, lane);
  hi = __builtin_amdgcn_readlane(hi, lane);
  return __hiloint2double(hi, lo);
}
// value of lane-1 (lane 0 gets garbage: callers overwrite it)
__device__ inline double from_left_lane(double v) { return __shfl_up(v, 1, 64); }
// value of lane+1 (lane 63 gets garbage: callers overwrite it)
__device__ inline double from_right_lane(double v) {
  return __shfl_down(v, 1, 64);
}

// Where a sweep writes the outgoing halo edges of its result (chare form).
// Each pointer is the first double of an edge (nullptr = not wanted).
struct EdgeOut {
  double* top;     // row 1, w values
  double* bottom;  // row h, w values
  double* left;    // column 1, h values
  double* right;   // column w, h values
};

// Pointers into global memory (address space 1): loads and stores through
// them are global_* instructions even when the pointer came from a generic
// one (chare state), not flat_* ones that must also assume LDS.
typedef __attribute__((address_space(1))) const double gcdouble;
typedef __attribute__((address_space(1))) double gdouble;
typedef double v2d __attribute__((ext_vector_type(2)));
#ifndef CHM_NT_LOADS
#define CHM_NT_LOADS 0
#endif
#ifndef CHM_NT_STORES
#define CHM_NT_STORES 1
#endif
// Streaming accesses: every input row is read once per sweep and every
// output row written once, so stores carry the non-temporal hint: they
// leave few dirty lines in L2, which keeps the L2 writeback of a message
// release short.
__device__ inline double2 gload2(const gcdouble* p) {
  const __attribute__((address_space(1))) v2d* q =
      (const __attribute__((address_space(1))) v2d*)p;
#if CHM_NT_LOADS
  const v2d v = __builtin_nontemporal_load(q);
#else
  const v2d v = *q;
#endif
  return make_double2(v.x, v.y);
}
__device__ inline void gstore2(gdouble* p, double a, double b) {
  v2d v;
  v.x = a;
  v.y = b;
  __attribute__((address_space(1))) v2d* q =
      (__attribute__((address_space(1))) v2d*)p;
#if CHM_NT_STORES
  __builtin_nontemporal_store(v, q);
#else
  *q = v;
#endif
}

template <int NC>
struct RowRegs {
  double2 v[NC];
  double e;  // lane 0: column c0-1; lane 63: column c0+128*NC (east edge)
};

template <int NC>
struct Panel {
  const gcdouble* in;
  gdouble* out;
  int64_t pitch;
  int w, h;    // interior of the array
  int c0;      // first interior column of the panel (1-based)
  int cend;    // last column this panel may write (<= w)
  int lane;
  int ecol;    // column this lane loads into RowRegs::e (-1: none)

  __device__ inline void setup(const double* in_, double* out_, int64_t p,
                               int w_, int h_, int c0_, int cend_) {
    in = (const gcdouble*)in_;
    out = (gdouble*)out_;
    pitch = p;
    w = w_;
    h = h_;
    c0 = c0_;
    cend = cend_;
    lane = __lane_id();
    ecol = -1;
    if (lane == 0) ecol = c0 - 1;
    if (lane == 63 && c0 + 128 * NC <= w + 1) ecol = c0 + 128 * NC;
  }
  __device__ inline int col(int k) const { return c0 + 128 * k + 2 * lane; }

  __device__ inline void load(RowRegs<NC>& r, int row) const {
    const gcdouble* rp = in + (int64_t)row * pitch + kColOff;
#pragma unroll
    for (int k = 0; k < NC; k++) {
      const int c = col(k);
      if (c <= w + 1)
        r.v[k] = gload2(rp + c);
      else
        r.v[k] = make_double2(0.0, 0.0);
    }
    r.e = (ecol >= 0) ? rp[ecol] : 0.0;
  }

  template <bool RES, bool PACK>
  __device__ inline void compute(const RowRegs<NC>& up, const RowRegs<NC>& cu,
                                 const RowRegs<NC>& dn, int row,
                                 u128acc& acc, const EdgeOut& eo) const {
    gdouble* orow = out + (int64_t)row * pitch + kColOff;
#pragma unroll
    for (int k = 0; k < NC; k++) {
      const double a = cu.v[k].x, b = cu.v[k].y;
      double bl = from_left_lane(b);
      double ar = from_right_lane(a);
      if (lane == 0) bl = (k == 0) ? cu.e : readlane_d(cu.v[k - 1].y, 63);
      if (lane == 63)
        ar = (k == NC - 1) ? cu.e : readlane_d(cu.v[(k + 1) % NC].x, 0);
      // c + W + E + N + S, then * 0.2 (jacobi2d.cu:332-333)
      const double na = (a + bl + b + up.v[k].x + dn.v[k].x) * 0.2;
      const double nb = (b + a + ar + up.v[k].y + dn.v[k].y) * 0.2;
      const int c = col(k);
      const bool two = c + 1 <= cend, one = c == cend;
      if (two) {
        gstore2(orow + c, na, nb);
        if (RES) {
          acc_add(acc, res_term(na, a));
          acc_add(acc, res_term(nb, b));
        }
      } else if (one) {
        orow[c] = na;
        if (RES) acc_add(acc, res_term(na, a));
      }
      if (PACK && (two || one)) {
        if (row == 1 && eo.top) {
          eo.top[c - 1] = na;
          if (two) eo.top[c] = nb;
        }
        if (row == h && eo.bottom) {
          eo.bottom[c - 1] = na;
          if (two) eo.bottom[c] = nb;
        }
        if (c == 1 && eo.left) eo.left[row - 1] = na;
        if (eo.right) {
          if (c == w) eo.right[row - 1] = na;
          if (two && c + 1 == w) eo.right[row - 1] = nb;
        }
      }
    }
  }
};

// Sweep rows [r0, r1) (1-based, r1 <= h+1) of the panel starting at column
// c0.  Reads rows r0-1 .. r1 of `in`, writes rows r0 .. r1-1 of `out`,
// columns c0 .. min(c0 + 128*NC - 1, cend) (cend <= w; default w).
// Registers hold a ring of U+2 rows: computing row r frees row r-1's slot,
// which immediately receives the load of row r+U+1, so U rows stay in
// flight.  The row loop is unrolled by the ring length so every slot index
// is static.
template <int NC, int U, bool RES, bool PACK>
__device__ inline void sweep_panel(const double* __restrict__ in,
                                   double* __restrict__ out, int64_t pitch,
                                   int w, int h, int c0, int r0, int r1,
                                   u128acc& acc, const EdgeOut& eo,
                                   int cend = -1) {
  if (r0 >= r1) return;
  constexpr int K = U + 2;
  Panel<NC> p;
  p.setup(in, out, pitch, w, h, c0, cend < 0 ? w : cend);
  RowRegs<NC> R[K];
#pragma unroll
  for (int s = 0; s < K; s++)
    if (r0 - 1 + s <= r1) p.load(R[s], r0 - 1 + s);
  for (int r = r0; r < r1; r += K) {
#pragma unroll
    for (int s = 0; s < K; s++) {
      if (r + s < r1) {
        p.template compute<RES, PACK>(R[s], R[(s + 1) % K], R[(s + 2) % K],
                                      r + s, acc, eo);
        if (r + s + U + 1 <= r1) p.load(R[s], r + s + U + 1);
      }
    }
  }
}

// ----------------------------------------------------------------------------
// Two sweeps fused (temporal blocking, depth-2 halos)
// ----------------------------------------------------------------------------
// u_{k+1} is computed in registers on one extra row and column on every side
// of the output region and never stored; u_{k+2} is computed from it with
// the same expression, so both are bit-identical to two separate sweeps.
// HBM traffic per two sweeps: one read of u_k and one write of u_{k+2}.

// Outgoing halos two deep (chare form): top = rows 1,2 (w each), bottom =
// rows h-1,h, left = columns 1,2 (h each), right = columns w-1,w.
struct EdgeOut2 {
  double* top;
  double* bottom;
  double* left;
  double* right;
};

// Which sides of the array are the fixed Dirichlet ring (not a neighbour's
// halo): there u_{k+1} is the ring value, not a stencil result.
struct RingSides {
  bool top, bottom, left, right;
};

template <int NC>
struct Row2 {      // a row of u_k
  double2 v[NC];
  double e1, e2;   // lane 0: columns c0-1, c0-2; lane 63: c0+128NC, c0+128NC+1
};
template <int NC>
struct MidRow {    // a row of u_{k+1}
  double2 v[NC];
  double e;        // lane 0: column c0-1; lane 63: column c0+128NC
};

template <int NC>
struct Panel2 {
  const gcdouble* in;
  gdouble* out;
  int64_t pitch;
  int w, h, c0, cend, lane;
  int col1, col2;  // columns this lane loads into e1 / e2 (-1000: none)
  RingSides ring;

  __device__ inline void setup(const double* in_, double* out_, int64_t p,
                               int w_, int h_, int c0_, int cend_,
                               RingSides rs) {
    in = (const gcdouble*)in_;
    out = (gdouble*)out_;
    pitch = p;
    w = w_;
    h = h_;
    c0 = c0_;
    cend = cend_;
    ring = rs;
    lane = __lane_id();
    col1 = col2 = -1000;
    if (lane == 0) {
      col1 = c0 - 1;
      col2 = c0 - 2;
    }
    if (lane == 63) {
      if (c0 + 128 * NC <= w + 2) col1 = c0 + 128 * NC;
      if (c0 + 128 * NC + 1 <= w + 2) col2 = c0 + 128 * NC + 1;
    }
  }
  __device__ inline int col(int k) const { return c0 + 128 * k + 2 * lane; }

  __device__ inline void load(Row2<NC>& r, int row) const {
    const gcdouble* rp = in + (int64_t)row * pitch + kColOff;
#pragma unroll
    for (int k = 0; k < NC; k++) {
      const int c = col(k);
      if (c <= w + 2)
        r.v[k] = gload2(rp + c);
      else
        r.v[k] = make_double2(0.0, 0.0);
    }
    r.e1 = (col1 > -1000) ? rp[col1] : 0.0;
    r.e2 = (col2 > -1000) ? rp[col2] : 0.0;
  }

  // u_{k+1} of row q over the panel's columns and one more on each side;
  // RESA: add its residual against u_k over the output region (row q in
  // [r0, r1) when `inside`, columns c0 .. cend)
  template <bool RESA>
  __device__ inline void mid(const Row2<NC>& up, const Row2<NC>& cu,
                             const Row2<NC>& dn, int q, MidRow<NC>& m,
                             bool inside, u128acc& acc_a) const {
    if ((q == 0 && ring.top) || (q == h + 1 && ring.bottom)) {
#pragma unroll
      for (int k = 0; k < NC; k++) m.v[k] = cu.v[k];
      m.e = cu.e1;
      return;
    }
#pragma unroll
    for (int k = 0; k < NC; k++) {
      const double a = cu.v[k].x, b = cu.v[k].y;
      double bl = from_left_lane(b);
      double ar = from_right_lane(a);
      if (lane == 0) bl = (k == 0) ? cu.e1 : readlane_d(cu.v[k - 1].y, 63);
      if (lane == 63)
        ar = (k == NC - 1) ? cu.e1 : readlane_d(cu.v[(k + 1) % NC].x, 0);
      // c + W + E + N + S, then * 0.2 (jacobi2d.cu:332-333)
      double na = (a + bl + b + up.v[k].x + dn.v[k].x) * 0.2;
      double nb = (b + a + ar + up.v[k].y + dn.v[k].y) * 0.2;
      const int c = col(k);
      if (ring.right) {
        if (c == w + 1) na = a;
        if (c + 1 == w + 1) nb = b;
      }
      m.v[k] = make_double2(na, nb);
      if (RESA && inside) {
        if (c <= cend) acc_add(acc_a, res_term(na, a));
        if (c + 1 <= cend) acc_add(acc_a, res_term(nb, b));
      }
    }
    // the edge columns: lane 0 holds c0-1 (west c0-2, east c0), lane 63
    // holds c0+128NC (west c0+128NC-1, east c0+128NC+1)
    const double ew = (lane == 0) ? cu.e2 : cu.v[NC - 1].y;
    const double ee = (lane == 0) ? cu.v[0].x : cu.e2;
    double me = (cu.e1 + ew + ee + up.e1 + dn.e1) * 0.2;
    if (lane == 0 && ring.left && c0 == 1) me = cu.e1;
    if (lane == 63 && ring.right && c0 + 128 * NC == w + 1) me = cu.e1;
    m.e = me;
  }

  // u_{k+2} of row `row` from three rows of u_{k+1}; stores, residual
  // against u_{k+1}, and the two-deep outgoing edges
  template <bool RES, bool PACK>
  __device__ inline void outrow(const MidRow<NC>& up, const MidRow<NC>& cu,
                                const MidRow<NC>& dn, int row, u128acc& acc,
                                const EdgeOut2& eo) const {
    gdouble* orow = out + (int64_t)row * pitch + kColOff;
#pragma unroll
    for (int k = 0; k < NC; k++) {
      const double a = cu.v[k].x, b = cu.v[k].y;
      double bl = from_left_lane(b);
      double ar = from_right_lane(a);
      if (lane == 0) bl = (k == 0) ? cu.e : readlane_d(cu.v[k - 1].y, 63);
      if (lane == 63)
        ar = (k == NC - 1) ? cu.e : readlane_d(cu.v[(k + 1) % NC].x, 0);
      const double na = (a + bl + b + up.v[k].x + dn.v[k].x) * 0.2;
      const double nb = (b + a + ar + up.v[k].y + dn.v[k].y) * 0.2;
      const int c = col(k);
      const bool two = c + 1 <= cend, one = c == cend;
      if (two) {
        gstore2(orow + c, na, nb);
        if (RES) {
          acc_add(acc, res_term(na, a));
          acc_add(acc, res_term(nb, b));
        }
      } else if (one) {
        orow[c] = na;
        if (RES) acc_add(acc, res_term(na, a));
      }
      if (PACK && (two || one)) {
        if (row <= 2 && eo.top) {
          eo.top[(int64_t)(row - 1) * w + c - 1] = na;
          if (two) eo.top[(int64_t)(row - 1) * w + c] = nb;
        }
        if (row >= h - 1 && eo.bottom) {
          eo.bottom[(int64_t)(row - (h - 1)) * w + c - 1] = na;
          if (two) eo.bottom[(int64_t)(row - (h - 1)) * w + c] = nb;
        }
        if (eo.left) {
          if (c <= 2) eo.left[(int64_t)(c - 1) * h + row - 1] = na;
          if (two && c + 1 <= 2) eo.left[(int64_t)c * h + row - 1] = nb;
        }
        if (eo.right) {
          if (c >= w - 1) eo.right[(int64_t)(c - (w - 1)) * h + row - 1] = na;
          if (two && c + 1 >= w - 1)
            eo.right[(int64_t)(c + 1 - (w - 1)) * h + row - 1] = nb;
        }
      }
    }
  }
};

// Two fused sweeps of rows [r0, r1) of the panel at column c0: reads u_k
// rows r0-2 .. r1+1 (columns c0-2 .. c0+128NC+1), writes u_{k+2} rows
// r0 .. r1-1, columns c0 .. min(c0+128NC-1, cend).  u_k rows sit in a ring
// of U+2 register slots (U rows in flight, static indices by unrolling);
// the three live rows of u_{k+1} shift through registers.  RESA / RES add
// the exact residuals of the first / second sweep over the output region
// to acc_a / acc.
template <int NC, int U, bool RESA, bool RES, bool PACK>
__device__ inline void sweep_panel2(const double* __restrict__ in,
                                    double* __restrict__ out, int64_t pitch,
                                    int w, int h, int c0, int r0, int r1,
                                    u128acc& acc_a, u128acc& acc,
                                    const EdgeOut2& eo, RingSides rs,
                                    int cend = -1) {
  if (r0 >= r1) return;
  constexpr int K = U + 2;
  Panel2<NC> p;
  p.setup(in, out, pitch, w, h, c0, cend < 0 ? w : cend, rs);
  Row2<NC> R[K];
  MidRow<NC> m0, m1, m2;
#pragma unroll
  for (int s = 0; s < K; s++)
    if (r0 - 2 + s <= r1 + 1) p.load(R[s], r0 - 2 + s);
  // step t computes u_{k+1} row q = r0-1+t from u_k rows q-1 (slot t % K),
  // q, q+1; then emits u_{k+2} row q-1 once three u_{k+1} rows exist
  const int n_steps = r1 - r0 + 2;
  for (int t0 = 0; t0 < n_steps; t0 += K) {
#pragma unroll
    for (int s = 0; s < K; s++) {
      const int t = t0 + s;
      if (t < n_steps) {
        const int q = r0 - 1 + t;
        m0 = m1;
        m1 = m2;
        p.template mid<RESA>(R[s], R[(s + 1) % K], R[(s + 2) % K], q, m2,
                             q >= r0 && q < r1, acc_a);
        if (q - 1 + K <= r1 + 1) p.load(R[s], q - 1 + K);
        if (t >= 2) p.template outrow<RES, PACK>(m0, m1, m2, q - 1, acc, eo);
      }
    }
  }
}

// ----------------------------------------------------------------------------
// D sweeps fused (temporal blocking, halos D deep)
// ----------------------------------------------------------------------------
// The same scheme D-1 stages deep: the grid after s sweeps is computed on
// D-s extra rows and columns around the output region, the grid after D
// sweeps is stored.

// Stores u' of the pair at columns (c, c+1) of an output row, its residual
// and its share of outgoing edges `D` deep (chare form; see EdgeOut2 for the
// layout with D = 2: rows 1..D, rows h-D+1..h, columns 1..D, columns
// w-D+1..w).
template <int D, bool RES, bool PACK>
__device__ inline void store_pair(gdouble* orow, int c, int cend, int w, int h,
                                  int row, double na, double nb, double a,
                                  double b, u128acc& acc, const EdgeOut2& eo) {
  const bool two = c + 1 <= cend, one = c == cend;
  if (two) {
    gstore2(orow + c, na, nb);
    if (RES) {
      acc_add(acc, res_term(na, a));
      acc_add(acc, res_term(nb, b));
    }
  } else if (one) {
    orow[c] = na;
    if (RES) acc_add(acc, res_term(na, a));
  }
  if (PACK && (two || one)) {
    if (row <= D && eo.top) {
      eo.top[(int64_t)(row - 1) * w + c - 1] = na;
      if (two) eo.top[(int64_t)(row - 1) * w + c] = nb;
    }
    if (row >= h - D + 1 && eo.bottom) {
      eo.bottom[(int64_t)(row - (h - D + 1)) * w + c - 1] = na;
      if (two) eo.bottom[(int64_t)(row - (h - D + 1)) * w + c] = nb;
    }
    if (eo.left) {
      if (c <= D) eo.left[(int64_t)(c - 1) * h + row - 1] = na;
      if (two && c + 1 <= D) eo.left[(int64_t)c * h + row - 1] = nb;
    }
    if (eo.right) {
      if (c >= w - D + 1) eo.right[(int64_t)(c - (w - D + 1)) * h + row - 1] = na;
      if (two && c + 1 >= w - D + 1)
        eo.right[(int64_t)(c + 1 - (w - D + 1)) * h + row - 1] = nb;
    }
  }
}

// Deep fusion works on overlapping windows: a wave holds 128 columns of a
// row (lane l: the pair at w0+2l, w0+2l+1), each sweep leaves one more
// column at either end without a valid neighbour, and after D <= 8 sweeps
// the middle 112 columns are stored (margins of H = 8 columns whatever D
// is), after 9 <= D <= 16 sweeps the middle 96 (H = 16): a window's output
// is seven or six whole 128-byte lines (column 1 starts a line) and every
// store is line-aligned.  Neighbouring windows overlap by 2H columns; the
// lane shifts need no boundary fix-ups.
struct RowW {
  double2 v;
};

// Lane l gets lane l-1's v (l+1's for lane_shl1z); the end lane gets 0.
__device__ inline double lane_shr1z(double v) {
  const long long x = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)x, 0x138, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(x >> 32), 0x138, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ inline double lane_shl1z(double v) {
  const long long x = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)x, 0x130, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(x >> 32), 0x130, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Columns of a window's margin, and its output columns, for D <= 16 fused
// sweeps.
__host__ __device__ constexpr int window_margin(int D) {
  return D <= 8 ? 8 : (D <= 16 ? 16 : -1);
}
__host__ __device__ constexpr int window_out_cols(int D) {
  return 128 - 2 * window_margin(D);
}

struct PanelW {
  const gcdouble* in;
  gdouble* out;
  int64_t pitch;
  int w, h, c, cend;
  bool vload;  // this lane's pair lies within the columns read
  bool vstore; // this lane's pair lies within the window's output columns
  // cells of this lane on the grid's ring (they keep their value), and
  // whether the wave has any (a uniform branch skips the selects if not)
  bool ring_a, ring_b, ringcols;
  RingSides ring;

  // the window whose output starts at column c0 (odd)
  __device__ inline void setup(const double* in_, double* out_, int64_t p,
                               int w_, int h_, int c0, int cend_,
                               RingSides rs, int D) {
    in = (const gcdouble*)in_;
    out = (gdouble*)out_;
    pitch = p;
    w = w_;
    h = h_;
    cend = cend_;
    ring = rs;
    const int H = window_margin(D);
    const int lane = __lane_id();
    c = c0 - H + 2 * lane;
    vload = c <= w + D;
    vstore = lane >= H / 2 && lane < 64 - H / 2;
    ring_a = (rs.left && c == 0) || (rs.right && c == w + 1);
    ring_b = (rs.left && c + 1 == 0) || (rs.right && c + 1 == w + 1);
    ringcols = __ballot(ring_a || ring_b) != 0;
  }

  // row `row` of the input; lanes past column w+D keep what they held
  __device__ inline void load(RowW& r, int row) const {
    if (vload) r.v = gload2(in + (int64_t)row * pitch + kColOff + c);
  }

  // One sweep: row q of the next grid from three rows of the current one
  // (COPY: the row itself, a stage that does no sweep).  CHECK = false is
  // for rows that meet no ring row; with it, COLS keeps the ring columns by
  // selects without a branch (waves whose window holds one): bit 0 those
  // among the lanes' first cells, bit 1 among their second cells.
  // RS = 1: add the sweep's exact residual over the cells this window
  // stores on rows [q0, q1) to *acc; RS = 2: add the plain FP64 sum of
  // (u' - u)^2 over them to *accd (a bound for the stop filter, not a
  // result).
  template <bool CHECK, int RS = 0, int COLS = 0>
  __device__ inline void stage(const RowW& up, const RowW& cu, const RowW& dn,
                               int q, RowW& o, bool copy = false,
                               u128acc* acc = nullptr, double* accd = nullptr,
                               int q0 = 0, int q1 = 0) const {
    if (copy ||
        (CHECK && ((q == 0 && ring.top) || (q == h + 1 && ring.bottom)))) {
      o = cu;
      return;
    }
    const double a = cu.v.x, b = cu.v.y;
    const double bl = lane_shr1z(b);  // west of a
    const double ar = lane_shl1z(a);  // east of b
    // c + W + E + N + S, then * 0.2 (jacobi2d.cu:332-333)
    double na = (a + bl + b + up.v.x + dn.v.x) * 0.2;
    double nb = (b + a + ar + up.v.y + dn.v.y) * 0.2;
    if (CHECK && ringcols) {
      if (ring_a) na = a;
      if (ring_b) nb = b;
    } else if (!CHECK) {
      if (COLS & 1) na = ring_a ? a : na;
      if (COLS & 2) nb = ring_b ? b : nb;
    }
    o.v = make_double2(na, nb);
    if (RS != 0 && q >= q0 && q < q1 && vstore) {
      if (RS == 1) {
        if (c <= cend) acc_add(*acc, res_term(na, a));
        if (c + 1 <= cend) acc_add(*acc, res_term(nb, b));
      } else {
        add_sq(*accd, na, a, nb, b);
      }
    }
  }

  // The FP64 sum of squared changes of the stored cells of a pair.
  __device__ inline void add_sq(double& accd, double na, double a, double nb,
                                double b) const {
    const double da = na - a, db = nb - b;
    if (c <= cend) accd += da * da;
    if (c + 1 <= cend) accd += db * db;
  }

  // The stored sweep: row `row` of the output, its residual against the
  // input of the sweep, and the outgoing edges D deep.
  // (APX: the FP64 sum of squared changes into *accd as well)
  template <int D, bool RES, bool PACK, bool APX = false>
  __device__ inline void outrow(const RowW& up, const RowW& cu, const RowW& dn,
                                int row, u128acc& acc, const EdgeOut2& eo,
                                double* accd = nullptr) const {
    const double a = cu.v.x, b = cu.v.y;
    const double bl = lane_shr1z(b);
    const double ar = lane_shl1z(a);
    const double na = (a + bl + b + up.v.x + dn.v.x) * 0.2;
    const double nb = (b + a + ar + up.v.y + dn.v.y) * 0.2;
    if (vstore) {
      store_pair<D, RES, PACK>(out + (int64_t)row * pitch + kColOff, c, cend,
                               w, h, row, na, nb, a, b, acc, eo);
      if (APX) add_sq(*accd, na, a, nb, b);
    }
  }
};

// D fused sweeps (D >= 3) of rows [r0, r1) of the window whose output
// starts at column c0 (odd): reads the input rows r0-D .. r1+D-1 (columns
// c0-H .. c0-H+127, those up to w+D), writes the output rows r0 .. r1-1,
// columns c0 .. min(c0+window_out_cols(D)-1, cend).  The input rows sit in
// a ring of RIN register slots (RIN = 3, 6, 9 or 12: RIN - 2 rows in
// flight ahead of the step that first needs them), and each intermediate
// grid keeps its three live rows in a ring of its own; ring slots are
// static indices because the step loop is unrolled by RIN.  RES adds the exact residual of the
// last sweep.  n_copy (0 .. D-1) leading stages only copy their input, so
// one instantiation also does D - n_copy sweeps (the halos read stay D
// deep).  RALL = 1 (with RES): stage s = 1 .. D-1 also adds its sweep's
// exact residual over the stored cells to accs[s-1] (acc is the last
// stage's); RALL = 2: every stage s = 1 .. D adds the FP64 sum of its
// squared changes over the stored cells to accd[s-1].  RINGC != 0: the
// window holds ring columns of the grid only among its lanes' first cells
// (1) or only among their second cells (2), and its steady steps keep them
// by selects; otherwise a window with a ring column runs every step
// guarded.
template <int D, bool RES, bool PACK, int RALL = 0, int RIN = 3,
          int RINGC = 0>
__device__ inline void sweep_panelD(const double* __restrict__ in,
                                    double* __restrict__ out, int64_t pitch,
                                    int w, int h, int c0, int r0, int r1,
                                    u128acc& acc, const EdgeOut2& eo,
                                    RingSides rs, int cend = -1,
                                    int n_copy = 0, u128acc* accs = nullptr,
                                    double* accd = nullptr) {
  static_assert(RALL != 1 || RES, "per-sweep residuals include the last one");
  static_assert(RALL != 2 || !RES, "the FP64 sums replace the exact ones");
  static_assert(D >= 3 && D <= 16, "deep fused sweep: 3 <= D <= 16");
  static_assert(RIN >= 3 && RIN <= 12 && RIN % 3 == 0,
                "input ring of 3, 6, 9 or 12 rows");
  constexpr int U = RIN;  // lcm(3, RIN)
  if (r0 >= r1) return;
  PanelW p;
  p.setup(in, out, pitch, w, h, c0, cend < 0 ? w : cend, rs, D);
  RowW R[RIN];        // input rows: row r0-D+i in slot i % RIN
  RowW S[D - 1][3];   // S[s-1]: rows of the grid after s sweeps
  const int last_in = r1 + D - 1;  // the last input row of the band
#pragma unroll
  for (int k = 0; k < RIN; k++) {
    R[k].v = make_double2(0.0, 0.0);
    if (r0 - D + k <= last_in) p.load(R[k], r0 - D + k);
  }
  // step t: the first sweep gives row q1 = r0-D+1+t from input rows q1-1
  // (slot t % RIN), q1, q1+1; sweep s >= 2 gives row q1-(s-1) once
  // t >= 2(s-1) from rows made at steps t-2, t-1, t; the last one is
  // stored.  After step t, input row q1-1 is dead and its slot takes row
  // q1-1+RIN.
  const int n_steps = r1 - r0 + 2 * D - 2;
  // Steps [tA, tB) form the steady state: every sweep is under way, no row
  // of the grid's ring is met (row 0 only occurs among the first 2D-2
  // steps, row h+1 only among the last D-1 and only if the band ends within
  // D-2 rows of it: other bands run their drain in the steady state too,
  // 109.5 -> 105.7 ms of sweeping per chare of a 16384^2 solve) and no load
  // runs past the band, so those steps run without guards or ring checks;
  // a wave that holds a ring column runs them with that column's selects
  // (RINGC, its own instantiation: a select test inside one shared steady
  // loop raised the register pressure of the pass enough to spill in every
  // step, 571 ms instead of 124 per 16384^2 solve) or else guarded.  Both
  // bounds are multiples of U, as the ring slots require.  (Running the
  // pipeline fill in the steady loop as well, the sweeps not yet under way
  // computing rows nothing reads, measured slower: 107.9 against 105.7 ms;
  // it does every stage's work in every fill step.)
  const int tA = (2 * D - 2 + U - 1) / U * U;
  const bool ring_below = rs.bottom && h + 1 <= r1 + D - 2;
  const int t_hi = ring_below ? min(n_steps - (D - 1), n_steps + 2 - RIN)
                              : n_steps + 2 - RIN;
  int tB = tA + (t_hi - tA) / U * U;
  if ((!RINGC && p.ringcols) || tB < tA) tB = tA;
  auto guarded = [&](int t0) {
#pragma unroll
    for (int k = 0; k < U; k++) {
      const int t = t0 + k;
      const int ks = k % 3;
      if (t < n_steps) {
        const int q1 = r0 - D + 1 + t;
        p.template stage<true, RALL>(R[k % RIN], R[(k + 1) % RIN],
                                     R[(k + 2) % RIN], q1, S[0][ks],
                                     n_copy >= 1, accs, accd, r0, r1);
        if (q1 - 1 + RIN <= last_in) p.load(R[k % RIN], q1 - 1 + RIN);
#pragma unroll
        for (int s = 2; s < D; s++)
          if (t >= 2 * (s - 1))
            p.template stage<true, RALL>(S[s - 2][(ks + 1) % 3],
                                         S[s - 2][(ks + 2) % 3], S[s - 2][ks],
                                         q1 - (s - 1), S[s - 1][ks],
                                         n_copy >= s,
                                         RALL == 1 ? accs + (s - 1) : nullptr,
                                         RALL == 2 ? accd + (s - 1) : nullptr,
                                         r0, r1);
        if (t >= 2 * (D - 1))
          p.template outrow<D, RES, PACK, RALL == 2>(
              S[D - 2][(ks + 1) % 3], S[D - 2][(ks + 2) % 3], S[D - 2][ks],
              q1 - (D - 1), acc, eo, RALL == 2 ? accd + (D - 1) : nullptr);
      }
    }
  };
  // the first tA steps, then the steady state, then the rest (one loop for
  // the guarded steps, so their code exists once)
  int t0 = 0;
  for (;;) {
    for (; t0 < n_steps && !(t0 >= tA && t0 < tB); t0 += U) guarded(t0);
    if (t0 >= n_steps) break;
    for (; t0 < tB; t0 += U) {
#pragma unroll
      for (int k = 0; k < U; k++) {
        const int ks = k % 3;
        const int q1 = r0 - D + 1 + t0 + k;
        p.template stage<false, RALL, RINGC>(R[k % RIN], R[(k + 1) % RIN],
                                             R[(k + 2) % RIN], q1, S[0][ks],
                                             n_copy >= 1, accs, accd, r0, r1);
        p.load(R[k % RIN], q1 - 1 + RIN);
#pragma unroll
        for (int s = 2; s < D; s++)
          p.template stage<false, RALL, RINGC>(
              S[s - 2][(ks + 1) % 3], S[s - 2][(ks + 2) % 3], S[s - 2][ks],
              q1 - (s - 1), S[s - 1][ks], n_copy >= s,
              RALL == 1 ? accs + (s - 1) : nullptr,
              RALL == 2 ? accd + (s - 1) : nullptr, r0, r1);
        p.template outrow<D, RES, PACK, RALL == 2>(
            S[D - 2][(ks + 1) % 3], S[D - 2][(ks + 2) % 3], S[D - 2][ks],
            q1 - (D - 1), acc, eo, RALL == 2 ? accd + (D - 1) : nullptr);
      }
    }
  }
}

// Wave-level reduction of a u128 accumulator to lane 0 (exact).
__device__ inline u128acc wave_reduce_u128(u128acc a) {
  for (int off = 32; off >= 1; off >>= 1) {
    u128acc o;
    o.lo = __shfl_down(a.lo, off, 64);
    o.hi = __shfl_down(a.hi, off, 64);
    acc_add(a, o);
  }
  return a;
}

}  // namespace jac
}  // namespace chm

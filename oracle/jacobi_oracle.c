/*
 * CPU oracle for the message-driven FP64 Jacobi2D path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C, the arithmetic
 * and the message semantics that the HIP runtime must reproduce bit for bit.
 * Only the test suite, the smoke check and the CPU-baseline leg of the bench
 * load it, and only as the checker; the product path never calls it.
 *
 * Built with -O3 -ffp-contract=off (no FMA contraction, no reassociation), the
 * same floating-point contract the device code is compiled under.
 *
 * Reference restated (paths relative to the reference tree):
 *   stencil            examples/jacobi2d/jacobi2d.cu:327-335 (operand order
 *                      c + W + E + N + S, then * 0.2), identical order in
 *                      examples/jacobi2d-mpi/jacobi2d.cu:296-302 and
 *                      examples/jacobi-kernel/jacobi_kernel.cu:26-34
 *   array layout       examples/jacobi2d/jacobi2d.cu:14 IDX(i,j)=(bw+2)*i+j
 *   reference init     examples/jacobi2d/jacobi2d.cu:282-293 (interior 0,
 *                      halo ring 10)
 *   decomposition      examples/jacobi2d/jacobi2d.cu:37-65
 *   neighbours         examples/jacobi2d/jacobi2d.cu:104-118
 *   halo pack/unpack   examples/jacobi2d/jacobi2d.cu:152-229, 295-325
 *                      (payload = [int dir in an 8-byte slot][edge doubles],
 *                      dir tag = receiver's side, :139-140)
 *   iteration driver   examples/jacobi2d/jacobi2d.cu:231-279
 *   block map          include/chare.h:125-177
 *   refnum rule        include/chare.h:181-195, 243-253;
 *                      src/comm/get/get.cu:422-444, 1316-1345
 *   wire sizes         include/message.h:19-82
 *
 * Project-defined pieces (no counterpart in the reference, pinned by the
 * tests with published vectors and closed forms):
 *   - counter-based input generator: splitmix64 (Steele, Lea, Flood,
 *     "Fast splittable pseudorandom number generators", OOPSLA 2014; the
 *     public-domain splitmix64.c by S. Vigna) evaluated at state
 *     seed + key * 0x9E3779B97F4A7C15 with key = (row << 32) | col;
 *     the double is (z >> 11) * 2^-53, uniform in [0, 1);
 *   - residual as an exact fixed-point sum: S = sum floor((a-b)^2 * 2^56)
 *     accumulated in 128-bit integers, so every summation order gives the
 *     same bits;
 *   - grid hash: sum over interior cells of mix(bits(v) + key * golden)
 *     modulo 2^64 (order independent).
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define GOLDEN 0x9E3779B97F4A7C15ULL
#define RES_SCALE 0x1.0p56

/* ------------------------------------------------------------------------ */
/* splitmix64                                                                 */
/* ------------------------------------------------------------------------ */

static inline uint64_t sm64_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

/* One step of the sequential generator: state += golden; return mix(state). */
uint64_t orc_splitmix64_next(uint64_t* state) {
  *state += GOLDEN;
  return sm64_mix(*state);
}

static inline uint64_t cell_key(int64_t row, int64_t col) {
  return ((uint64_t)row << 32) | (uint64_t)(uint32_t)col;
}

/* Interior value at global (row, col), both 1-based interior coordinates. */
double orc_cell_u01(uint64_t seed, int64_t row, int64_t col) {
  uint64_t z = seed + (cell_key(row, col) + 1) * GOLDEN;
  return (double)(sm64_mix(z) >> 11) * 0x1.0p-53;
}

/* ------------------------------------------------------------------------ */
/* Grid init, step, residual, hash (monolithic (H+2) x (W+2) row-major)       */
/* ------------------------------------------------------------------------ */

/* mode 0: seeded uniform [0,1) interior, boundary ring = boundary.
 * mode 1: reference init_kernel (jacobi2d.cu:282-293): interior 0, ring 10. */
void orc_init_grid(int W, int H, uint64_t seed, int mode, double boundary,
                   double* g) {
  const int64_t P = (int64_t)W + 2;
  double ring = (mode == 1) ? 10.0 : boundary;
  for (int64_t i = 0; i < H + 2; i++) {
    for (int64_t j = 0; j < W + 2; j++) {
      int interior = (i > 0 && i < H + 1 && j > 0 && j < W + 1);
      double v;
      if (!interior) v = ring;
      else v = (mode == 1) ? 0.0 : orc_cell_u01(seed, i, j);
      g[i * P + j] = v;
    }
  }
}

/* The 5-point update, reference operand order (jacobi2d.cu:332-333). */
static inline double stencil5(double c, double w, double e, double n, double s) {
  return (c + w + e + n + s) * 0.2;
}

/* One sweep over the interior of a (h+2) x pitch array; the ring of `out`
 * is left untouched. */
static void sweep(const double* in, double* out, int w, int h, int64_t pitch) {
  for (int64_t i = 1; i <= h; i++) {
    const double* up = in + (i - 1) * pitch;
    const double* ce = in + i * pitch;
    const double* dn = in + (i + 1) * pitch;
    double* o = out + i * pitch;
    for (int64_t j = 1; j <= w; j++) {
      o[j] = stencil5(ce[j], ce[j - 1], ce[j + 1], up[j], dn[j]);
    }
  }
}

void orc_jacobi_step(int W, int H, const double* in, double* out) {
  sweep(in, out, W, H, (int64_t)W + 2);
}

typedef struct { uint64_t lo, hi; } u128_t;

static inline void u128_add64(u128_t* a, uint64_t v) {
  uint64_t lo = a->lo + v;
  a->hi += (lo < a->lo);
  a->lo = lo;
}
static inline void u128_add(u128_t* a, u128_t b) {
  uint64_t lo = a->lo + b.lo;
  a->hi += b.hi + (lo < a->lo);
  a->lo = lo;
}

/* floor((a-b)^2 * 2^56) as u64; (a-b)^2 < 2^7 for all grids used here. */
static inline uint64_t res_term(double a, double b) {
  double d = a - b;
  double q = d * d * RES_SCALE;
  if (!(q < 0x1.0p64)) return UINT64_MAX; /* saturate (never reached) */
  return (uint64_t)q;
}

static u128_t residual_region(const double* a, const double* b, int w, int h,
                              int64_t pitch) {
  u128_t s = {0, 0};
  for (int64_t i = 1; i <= h; i++)
    for (int64_t j = 1; j <= w; j++)
      u128_add64(&s, res_term(a[i * pitch + j], b[i * pitch + j]));
  return s;
}

void orc_residual_fixed(int W, int H, const double* a, const double* b,
                        uint64_t* lo, uint64_t* hi) {
  u128_t s = residual_region(a, b, W, H, (int64_t)W + 2);
  *lo = s.lo;
  *hi = s.hi;
}

/* u128 -> double, each limb conversion correctly rounded, one final add. */
double orc_u128_to_double(uint64_t lo, uint64_t hi) {
  return ldexp((double)hi, 64) + (double)lo;
}

static inline uint64_t dbl_bits(double v) {
  uint64_t u;
  memcpy(&u, &v, 8);
  return u;
}

static inline uint64_t hash_cell(int64_t row, int64_t col, double v) {
  return sm64_mix(dbl_bits(v) + cell_key(row, col) * GOLDEN);
}

/* Order-independent hash of the interior rows [r0, r1) x cols [c0, c1)
 * (global 1-based coordinates) of a monolithic grid. */
uint64_t orc_grid_hash_region(int W, int H, const double* g, int r0, int r1,
                              int c0, int c1) {
  (void)H;
  const int64_t P = (int64_t)W + 2;
  uint64_t h = 0;
  for (int64_t i = r0; i < r1; i++)
    for (int64_t j = c0; j < c1; j++) h += hash_cell(i, j, g[i * P + j]);
  return h;
}

/* Monolithic reference run: u_0 from orc_init_grid, `iters` sweeps.  Writes
 * u_iters into out (full (H+2)x(W+2) grid) and the exact residual of the
 * last sweep, sum floor((u_n - u_{n-1})^2 * 2^56). */
int orc_jacobi_run(int W, int H, int iters, uint64_t seed, int mode,
                   double boundary, double* out, uint64_t* res_lo,
                   uint64_t* res_hi) {
  size_t n = (size_t)(W + 2) * (size_t)(H + 2);
  double* a = (double*)malloc(n * sizeof(double));
  double* b = (double*)malloc(n * sizeof(double));
  if (!a || !b) { free(a); free(b); return -1; }
  orc_init_grid(W, H, seed, mode, boundary, a);
  memcpy(b, a, n * sizeof(double));
  u128_t r = {0, 0};
  for (int it = 0; it < iters; it++) {
    orc_jacobi_step(W, H, a, b);
    if (it == iters - 1) r = residual_region(b, a, W, H, (int64_t)W + 2);
    double* t = a; a = b; b = t;
  }
  memcpy(out, a, n * sizeof(double));
  if (res_lo) *res_lo = r.lo;
  if (res_hi) *res_hi = r.hi;
  free(a);
  free(b);
  return 0;
}

/* Solve to a relative residual: after each sweep k (producing u_{k+1}),
 * S_k = sum floor((u_{k+1}-u_k)^2 2^56); stop when
 * double(S_k) <= tol*tol * double(S_0) or after max_iters sweeps.  Returns
 * the number of sweeps done; out = final grid; S_last and S_0 reported. */
int orc_jacobi_solve_tol(int W, int H, int max_iters, double tol,
                         uint64_t seed, int mode, double boundary, double* out,
                         uint64_t* s_last_lo, uint64_t* s_last_hi,
                         uint64_t* s0_lo, uint64_t* s0_hi) {
  size_t n = (size_t)(W + 2) * (size_t)(H + 2);
  double* a = (double*)malloc(n * sizeof(double));
  double* b = (double*)malloc(n * sizeof(double));
  if (!a || !b) { free(a); free(b); return -1; }
  orc_init_grid(W, H, seed, mode, boundary, a);
  memcpy(b, a, n * sizeof(double));
  const double tol2 = tol * tol;
  u128_t s0 = {0, 0}, sk = {0, 0};
  int done = 0;
  for (int it = 0; it < max_iters; it++) {
    orc_jacobi_step(W, H, a, b);
    sk = residual_region(b, a, W, H, (int64_t)W + 2);
    if (it == 0) s0 = sk;
    double* t = a; a = b; b = t;
    done = it + 1;
    if (orc_u128_to_double(sk.lo, sk.hi) <= tol2 * orc_u128_to_double(s0.lo, s0.hi))
      break;
  }
  memcpy(out, a, n * sizeof(double));
  *s_last_lo = sk.lo; *s_last_hi = sk.hi;
  *s0_lo = s0.lo; *s0_hi = s0.hi;
  free(a);
  free(b);
  return done;
}

/* ------------------------------------------------------------------------ */
/* Decomposition, block map, wire sizes                                       */
/* ------------------------------------------------------------------------ */

/* Surface-minimising factorisation (jacobi2d.cu:37-65): chare rows `*cr`
 * divide the height, chare columns `*cc` the width.  Returns 0 on success. */
int orc_decompose(int n_chares, int W, int H, int* cr, int* cc) {
  double area0 = H, area1 = W;
  double best = 2.0 * (area0 + area1);
  int bx = -1, by = -1;
  for (int ipx = 1; ipx <= n_chares; ipx++) {
    if (n_chares % ipx == 0) {
      int ipy = n_chares / ipx;
      double surf = 2.0 * (area0 / ipx + area1 / ipy);
      if (surf < best) { best = surf; bx = ipx; by = ipy; }
    }
  }
  if (bx < 0 || bx * by != n_chares) return -1;
  *cr = bx;
  *cc = by;
  return 0;
}

/* Default block map (chare.h:136-176): each PE gets n/P chares, the first
 * n%P PEs one extra; or counts from block_map.  loc_map[i] = PE of chare i. */
int orc_block_map(int n_chares, int n_pes, const int* block_map, int* loc_map) {
  int per = n_chares / n_pes, rem = n_chares % n_pes, start = 0;
  for (int pe = 0; pe < n_pes; pe++) {
    int cnt = block_map ? block_map[pe] : per + (pe < rem ? 1 : 0);
    for (int i = start; i < start + cnt && i < n_chares; i++) loc_map[i] = pe;
    start += cnt;
  }
  return start == n_chares ? 0 : -1;
}

/* envelope::alloc_size (message.h:61-81): round16(16 + header + payload);
 * header 16 B for regular/user, 48 for request, 32 for forward, 0 for the
 * termination types.  type: 0 regular 1 request 2 forward 3 user 4 begin_term
 * 5 do_term. */
uint64_t orc_alloc_size(int type, uint64_t payload) {
  uint64_t hdr = 0;
  if (type == 0 || type == 3) hdr = 16;
  else if (type == 1) hdr = 48;
  else if (type == 2) hdr = 32;
  uint64_t s = 16 + hdr + payload;
  return (s + 15) / 16 * 16;
}

/* Extents of chare-row r of `n` over `len` cells: equal blocks of len/n, the
 * last one taking the remainder (a ragged edge). */
static void split(int len, int n, int r, int* start, int* size) {
  int b = len / n;
  *start = r * b;
  *size = (r == n - 1) ? len - b * (n - 1) : b;
}

/* ------------------------------------------------------------------------ */
/* Message-driven, chare-decomposed runner on host threads                    */
/* ------------------------------------------------------------------------ */
/* One std-thread-like PE per worker thread.  Each PE owns an inbox (FIFO),
 * a deferred list for refnum-mismatched messages and the chares the block map
 * gives it.  Entry methods follow jacobi2d.cu:93-279 with two deliberate
 * fixes: halos carry refnum = iteration (the reference omits it, so a
 * neighbour one iteration ahead can be miscounted), and a chare with no
 * neighbours updates without waiting. */

enum { DIR_LEFT = 0, DIR_RIGHT = 1, DIR_TOP = 2, DIR_BOTTOM = 3 };
enum { EP_RECV = 1, EP_END = 2 };

typedef struct msg_s {
  struct msg_s* next;
  int chare;
  int ep;
  int refnum;
  size_t size;
  unsigned char payload[]; /* [int dir | 4 pad][doubles] */
} msg_t;

typedef struct {
  int r, c, idx;
  int gr0, gc0;   /* global 1-based coordinates of interior cell (1,1) */
  int bw, bh;
  int64_t pitch;
  double* t;      /* current */
  double* nt;     /* next */
  int nbr[4];
  int nbr_count, recv_count;
  int iter, n_iters;
  int refnum;
  u128_t res;     /* residual of the last sweep */
  double* stage;  /* outgoing edge staging */
} chare_t;

struct runner;

typedef struct {
  struct runner* R;
  int pe;
  pthread_mutex_t mu;
  pthread_cond_t cv;
  msg_t* head;
  msg_t* tail;
  msg_t* deferred;   /* refnum-mismatched, in arrival order */
  msg_t* revived;    /* ready to be re-dispatched, FIFO */
  msg_t* revived_tail;
  long delivered, deferred_count;
} pe_t;

typedef struct runner {
  int W, H, cr, cc, iters, n_pes;
  int n_chares;
  chare_t* ch;
  int* loc;
  pe_t* pes;
  volatile int term;
  int end_count;
  pthread_mutex_t end_mu;
} runner_t;

static void pe_push(pe_t* p, msg_t* m) {
  m->next = NULL;
  pthread_mutex_lock(&p->mu);
  if (p->tail) p->tail->next = m; else p->head = m;
  p->tail = m;
  pthread_cond_signal(&p->cv);
  pthread_mutex_unlock(&p->mu);
}

static void send_msg(runner_t* R, int chare, int ep, int refnum,
                     const void* buf, size_t size) {
  msg_t* m = (msg_t*)malloc(sizeof(msg_t) + size + 16);
  m->chare = chare; m->ep = ep; m->refnum = refnum; m->size = size;
  if (size) memcpy(m->payload, buf, size);
  pe_push(&R->pes[R->loc[chare]], m);
}

static void send_boundaries(runner_t* R, chare_t* c) {
  const int64_t P = c->pitch;
  double* st = c->stage;
  /* payload: 8-byte dir slot then the edge (jacobi2d.cu:139-140, 171-190) */
  for (int d = 0; d < 4; d++) {
    if (c->nbr[d] < 0) continue;
    int len = (d < 2) ? c->bh : c->bw;
    double* e = st + 1;
    *(int*)st = (d % 2 == 0) ? d + 1 : d - 1;
    ((int*)st)[1] = 0;
    if (d == DIR_LEFT) for (int i = 0; i < len; i++) e[i] = c->t[(i + 1) * P + 1];
    if (d == DIR_RIGHT) for (int i = 0; i < len; i++) e[i] = c->t[(i + 1) * P + c->bw];
    if (d == DIR_TOP) memcpy(e, c->t + P + 1, len * sizeof(double));
    if (d == DIR_BOTTOM) memcpy(e, c->t + P * c->bh + 1, len * sizeof(double));
    send_msg(R, c->nbr[d], EP_RECV, c->iter, st, 8 + 8 * (size_t)len);
  }
}

static void set_refnum(runner_t* R, pe_t* p, chare_t* c, int v);

static void update(runner_t* R, pe_t* p, chare_t* c) {
  for (;;) {
    sweep(c->t, c->nt, c->bw, c->bh, c->pitch);
    if (c->iter == c->n_iters - 1)
      c->res = residual_region(c->nt, c->t, c->bw, c->bh, c->pitch);
    if (++c->iter == c->n_iters) {
      double* tmp = c->t; c->t = c->nt; c->nt = tmp; /* result in c->t */
      send_msg(R, 0, EP_END, -1, NULL, 0);
      return;
    }
    double* tmp = c->t; c->t = c->nt; c->nt = tmp;
    set_refnum(R, p, c, c->iter);
    send_boundaries(R, c);
    if (c->nbr_count > 0) return;
  }
}

static void recv_ghosts(runner_t* R, pe_t* p, chare_t* c, const unsigned char* pl) {
  int dir = *(const int*)pl;
  const double* g = (const double*)(pl + 8);
  const int64_t P = c->pitch;
  switch (dir) {
    case DIR_LEFT: for (int i = 0; i < c->bh; i++) c->t[(i + 1) * P] = g[i]; break;
    case DIR_RIGHT: for (int i = 0; i < c->bh; i++) c->t[(i + 1) * P + c->bw + 1] = g[i]; break;
    case DIR_TOP: memcpy(c->t + 1, g, c->bw * sizeof(double)); break;
    case DIR_BOTTOM: memcpy(c->t + P * (c->bh + 1) + 1, g, c->bw * sizeof(double)); break;
  }
  if (++c->recv_count == c->nbr_count) {
    c->recv_count = 0;
    update(R, p, c);
  }
}

/* Dispatch with the refnum gate (chare.h:185-189). Returns 0 if deferred. */
static int dispatch(runner_t* R, pe_t* p, msg_t* m) {
  chare_t* c = &R->ch[m->chare];
  if (m->ep == EP_END) {
    pthread_mutex_lock(&R->end_mu);
    int done = (++R->end_count == R->n_chares);
    pthread_mutex_unlock(&R->end_mu);
    if (done) {
      R->term = 1;
      for (int q = 0; q < R->n_pes; q++) {
        pthread_mutex_lock(&R->pes[q].mu);
        pthread_cond_broadcast(&R->pes[q].cv);
        pthread_mutex_unlock(&R->pes[q].mu);
      }
    }
    return 1;
  }
  if (m->refnum != -1 && m->refnum != c->refnum) return 0;
  recv_ghosts(R, p, c, m->payload);
  return 1;
}

/* set_refnum + revive_mismatches (chare.h:243-253, get.cu:1316-1345): matching
 * deferred messages move, in arrival order, to the revived queue. */
static void set_refnum(runner_t* R, pe_t* p, chare_t* c, int v) {
  (void)R;
  c->refnum = v;
  msg_t** pp = &p->deferred;
  while (*pp) {
    msg_t* m = *pp;
    if (m->chare == c->idx && m->refnum == v) {
      *pp = m->next;
      m->next = NULL;
      if (p->revived_tail) p->revived_tail->next = m; else p->revived = m;
      p->revived_tail = m;
    } else {
      pp = &m->next;
    }
  }
}

static void defer(pe_t* p, msg_t* m) {
  m->next = NULL;
  msg_t** pp = &p->deferred;
  while (*pp) pp = &(*pp)->next;
  *pp = m;
  p->deferred_count++;
}

static void* pe_main(void* arg) {
  pe_t* p = (pe_t*)arg;
  runner_t* R = p->R;
  for (;;) {
    msg_t* m = NULL;
    if (p->revived) {
      m = p->revived;
      p->revived = m->next;
      if (!p->revived) p->revived_tail = NULL;
    } else {
      pthread_mutex_lock(&p->mu);
      while (!p->head && !R->term) pthread_cond_wait(&p->cv, &p->mu);
      if (p->head) {
        m = p->head;
        p->head = m->next;
        if (!p->head) p->tail = NULL;
      }
      pthread_mutex_unlock(&p->mu);
      if (!m) break; /* terminated */
    }
    if (dispatch(R, p, m)) { p->delivered++; free(m); }
    else defer(p, m);
  }
  return NULL;
}

/* Chare init (jacobi2d.cu:93-150 with the seeded/reference fill). */
static void chare_init(runner_t* R, chare_t* c, uint64_t seed, int mode,
                       double boundary) {
  int idx = c->idx;
  c->r = idx / R->cc;
  c->c = idx % R->cc;
  split(R->H, R->cr, c->r, &c->gr0, &c->bh);
  split(R->W, R->cc, c->c, &c->gc0, &c->bw);
  c->gr0 += 1;
  c->gc0 += 1;
  c->pitch = c->bw + 2;
  c->nbr[DIR_LEFT] = (c->c == 0) ? -1 : idx - 1;
  c->nbr[DIR_RIGHT] = (c->c == R->cc - 1) ? -1 : idx + 1;
  c->nbr[DIR_TOP] = (c->r == 0) ? -1 : idx - R->cc;
  c->nbr[DIR_BOTTOM] = (c->r == R->cr - 1) ? -1 : idx + R->cc;
  c->nbr_count = 0;
  for (int d = 0; d < 4; d++) c->nbr_count += (c->nbr[d] >= 0);
  c->recv_count = 0;
  c->iter = 0;
  c->n_iters = R->iters;
  c->refnum = 0;
  c->res.lo = c->res.hi = 0;
  size_t n = (size_t)(c->bh + 2) * (size_t)c->pitch;
  c->t = (double*)malloc(n * sizeof(double));
  c->nt = (double*)malloc(n * sizeof(double));
  c->stage = (double*)malloc((size_t)(1 + (c->bw > c->bh ? c->bw : c->bh)) * 8);
  double ring = (mode == 1) ? 10.0 : boundary;
  for (int64_t i = 0; i < c->bh + 2; i++)
    for (int64_t j = 0; j < c->bw + 2; j++) {
      int interior = (i > 0 && i <= c->bh && j > 0 && j <= c->bw);
      double v = ring;
      if (interior)
        v = (mode == 1) ? 0.0 : orc_cell_u01(seed, c->gr0 + i - 1, c->gc0 + j - 1);
      c->t[i * c->pitch + j] = v;
      c->nt[i * c->pitch + j] = v;
    }
}

static double g_last_run_seconds = 0.0;

/* Wall time of the threaded (message-driven) phase of the last
 * orc_jacobi_chares_run: iterations only, without init or gather. */
double orc_last_run_seconds(void) { return g_last_run_seconds; }

/* Run the message-driven decomposition: chare grid cr x cc over a W x H
 * interior, n_pes host threads, `iters` iterations.  out (if non-NULL) gets
 * the full (H+2)x(W+2) grid; res_* the exact residual of the last sweep.
 * Returns the number of deferred (refnum-mismatched) messages, or <0. */
long orc_jacobi_chares_run(int W, int H, int cr, int cc, int iters, int n_pes,
                           uint64_t seed, int mode, double boundary,
                           double* out, uint64_t* res_lo, uint64_t* res_hi) {
  if (cr <= 0 || cc <= 0 || cr > H || cc > W || iters <= 0 || n_pes <= 0)
    return -1;
  runner_t R;
  memset(&R, 0, sizeof(R));
  R.W = W; R.H = H; R.cr = cr; R.cc = cc; R.iters = iters;
  R.n_chares = cr * cc;
  R.n_pes = n_pes < R.n_chares ? n_pes : R.n_chares;
  R.ch = (chare_t*)calloc(R.n_chares, sizeof(chare_t));
  R.loc = (int*)malloc(sizeof(int) * R.n_chares);
  R.pes = (pe_t*)calloc(R.n_pes, sizeof(pe_t));
  pthread_mutex_init(&R.end_mu, NULL);
  orc_block_map(R.n_chares, R.n_pes, NULL, R.loc);
  for (int i = 0; i < R.n_chares; i++) {
    R.ch[i].idx = i;
    chare_init(&R, &R.ch[i], seed, mode, boundary);
  }
  for (int q = 0; q < R.n_pes; q++) {
    R.pes[q].R = &R;
    R.pes[q].pe = q;
    pthread_mutex_init(&R.pes[q].mu, NULL);
    pthread_cond_init(&R.pes[q].cv, NULL);
  }
  /* iteration 0 halos, sent before the threads start (init entry method) */
  for (int i = 0; i < R.n_chares; i++) send_boundaries(&R, &R.ch[i]);
  for (int i = 0; i < R.n_chares; i++)
    if (R.ch[i].nbr_count == 0) update(&R, &R.pes[R.loc[i]], &R.ch[i]);
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * R.n_pes);
  for (int q = 0; q < R.n_pes; q++) pthread_create(&th[q], NULL, pe_main, &R.pes[q]);
  for (int q = 0; q < R.n_pes; q++) pthread_join(th[q], NULL);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  g_last_run_seconds = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
  long deferred = 0;
  for (int q = 0; q < R.n_pes; q++) {
    deferred += R.pes[q].deferred_count;
    msg_t* m = R.pes[q].head;
    while (m) { msg_t* n = m->next; free(m); m = n; }
    m = R.pes[q].deferred;
    while (m) { msg_t* n = m->next; free(m); m = n; }
  }
  u128_t res = {0, 0};
  const int64_t GP = (int64_t)W + 2;
  if (out) {
    double ring = (mode == 1) ? 10.0 : boundary;
    for (int64_t i = 0; i < H + 2; i++)
      for (int64_t j = 0; j < W + 2; j++) out[i * GP + j] = ring;
  }
  for (int i = 0; i < R.n_chares; i++) {
    chare_t* c = &R.ch[i];
    u128_add(&res, c->res);
    if (out)
      for (int64_t a = 1; a <= c->bh; a++)
        memcpy(out + (c->gr0 + a - 1) * GP + c->gc0, c->t + a * c->pitch + 1,
               sizeof(double) * c->bw);
    free(c->t); free(c->nt); free(c->stage);
  }
  if (res_lo) *res_lo = res.lo;
  if (res_hi) *res_hi = res.hi;
  free(th); free(R.ch); free(R.loc); free(R.pes);
  return deferred;
}

/* Per-chare hashes of the result of a chare-decomposed run without keeping
 * the full grid: used to make full-size fixtures.  hashes[i] for chare i. */
long orc_jacobi_chares_hashes(int W, int H, int cr, int cc, int iters,
                              int n_pes, uint64_t seed, int mode,
                              double boundary, uint64_t* hashes,
                              uint64_t* res_lo, uint64_t* res_hi) {
  size_t n = (size_t)(W + 2) * (size_t)(H + 2);
  double* g = (double*)malloc(n * sizeof(double));
  if (!g) return -1;
  long d = orc_jacobi_chares_run(W, H, cr, cc, iters, n_pes, seed, mode,
                                 boundary, g, res_lo, res_hi);
  if (d >= 0) {
    for (int i = 0; i < cr * cc; i++) {
      int r0, bh, c0, bw;
      split(H, cr, i / cc, &r0, &bh);
      split(W, cc, i % cc, &c0, &bw);
      hashes[i] = orc_grid_hash_region(W, H, g, r0 + 1, r0 + 1 + bh, c0 + 1,
                                       c0 + 1 + bw);
    }
  }
  free(g);
  return d;
}

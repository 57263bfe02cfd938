/*
 * C ABI of the FP64 Jacobi2D hot path on one MI355X (gfx950).
 *
 * The reference has no FFI: its Jacobi2D is a device-side chare program
 * (examples/jacobi2d/jacobi2d.cu:26-335) started by the library-owned `main`
 * (src/charming.cu:50-263), plus two host-driven per-launch programs
 * (examples/jacobi2d-mpi/jacobi2d.cu:71-76, examples/jacobi-kernel/
 * jacobi_kernel.cu:104-148).  These entry points are what a host binding
 * (ctypes, cgo, JNI) would call to run the same computation with device-
 * resident inputs and outputs.  Plain pointers and sizes only.
 *
 * Every function returns 0 on success or a negative CHM_E* code; the message
 * of the last failure on the calling thread is chm_jacobi2d_last_error().
 */
#ifndef CHM_JACOBI2D_H_
#define CHM_JACOBI2D_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum chm_jacobi_init {
  CHM_INIT_SEEDED = 0,    /* interior splitmix64 uniform [0,1), ring = boundary */
  CHM_INIT_REFERENCE = 1  /* init_kernel, jacobi2d.cu:282-293: interior 0, ring 10 */
};

enum chm_jacobi_form {
  CHM_FORM_PLAIN = 0,      /* one stencil launch per iteration (host loop) */
  CHM_FORM_GRAPH = 1,      /* the same loop captured once as a hipGraph */
  CHM_FORM_PERSISTENT = 2  /* message-driven chares in one persistent launch */
};

typedef struct chm_jacobi_problem {
  int width;      /* interior columns W */
  int height;     /* interior rows H */
  int chares_r;   /* chare rows; 0 with chares_c 0 = surface-minimising split */
  int chares_c;   /* chare columns */
  int n_chares;   /* chare count used by the automatic split (0 = n_pes) */
  int iters;      /* Jacobi sweeps */
  uint64_t seed;  /* CHM_INIT_SEEDED */
  int init_mode;  /* enum chm_jacobi_init */
  double boundary;/* Dirichlet value of the ring for CHM_INIT_SEEDED */
} chm_jacobi_problem;

typedef struct chm_jacobi_result {
  double device_ms;      /* device time of the solve (HIP events on its stream) */
  uint64_t residual_lo;  /* exact sum floor((u_n - u_{n-1})^2 * 2^56), low word */
  uint64_t residual_hi;  /* ... high word */
  uint64_t hash;         /* order-independent hash of the interior of u_n */
  int iters_done;
  int launches;          /* kernel launches issued by the solve */
  long long messages;    /* messages delivered (persistent form) */
  long long deferred;    /* refnum-mismatched messages deferred (persistent) */
  double sched_overhead_ms; /* persistent: mean per PE in the scheduler around
                               entry methods (decode, fences, release) */
} chm_jacobi_result;

/* Tuning of the sweep: chunks of 128 columns per wave panel and rows kept in
 * flight per wave.  Zero fields take the built-in defaults. */
typedef struct chm_jacobi_tuning {
  int panel_chunks;  /* 1, 2, 4 or 8 */
  int rows_in_flight;/* 1, 2, 4 or 8 */
  int band_rows;     /* rows per wave band (plain form; persistent form with
                        halos 3+ deep: < 0 = one contiguous share of the
                        chare per wave, 0 = 256 for chares of 1024+ rows,
                        else contiguous shares) */
  int wg_waves;      /* waves per workgroup (plain form) */
  int overlap;       /* persistent form: 1 = sweep the edges first and send the
                        halos before the interior (default 0; depth 1 only) */
  int depth;         /* persistent form: halo depth = sweeps per exchange,
                        1 .. 16 (every chare at least depth x depth; 0 = as
                        deep as the smallest chare allows, up to 16, with
                        the default or 1x1 sweep shape, else 2) */
  int halo_copy;     /* persistent form: 1 = pack halos into staging and send
                        them by copy (invoke); 0 = zero-copy: the sweep writes
                        them into messages allocated in the neighbours'
                        inboxes */
} chm_jacobi_tuning;

/* Allocate the grids of one problem in HBM for the given form and fill u_0.
 * Replaces the allocations and init_kernel launch of Block::init
 * (jacobi2d.cu:93-150, jacobi2d-mpi/jacobi2d.cu:91-131). */
int chm_jacobi2d_create(const chm_jacobi_problem* prob, int form, void** handle);

/* Set sweep tuning for later solves (NULL restores defaults). */
int chm_jacobi2d_set_tuning(void* handle, const chm_jacobi_tuning* t);

/* Refill u_0 on the device (so repeated solves start from the same state). */
int chm_jacobi2d_reset(void* handle);

/* Run prob->iters sweeps in the handle's form, blocking until done; inputs
 * and outputs stay in HBM.  Fills *res (may be NULL).  Plain form: the host
 * loop of jacobi2d-mpi/jacobi2d.cu:71-76 / jacobi_kernel.cu:108-138;
 * persistent form: one launch of the message-driven chare program. */
int chm_jacobi2d_solve(void* handle, chm_jacobi_result* res);

/* Copy u (the current grid, (H+2) x (W+2) row-major doubles, ring included)
 * to host memory. */
int chm_jacobi2d_read_grid(void* handle, double* out, size_t out_elems);

/* Order-independent hash of rows [r0, r1) x cols [c0, c1) (1-based interior
 * coordinates) of the current grid, computed on the device. */
int chm_jacobi2d_hash_region(void* handle, int r0, int r1, int c0, int c1,
                             uint64_t* hash);

/* Per-chare time of the last persistent solve, in ms: mean over chares of
 * {sweep, send, wait-for-halos, inside halo-receive entry methods} in
 * out8[0..3], maximum in out8[4..7].  The reference keeps the same
 * per-chare phase timers and prints them (examples/jacobi2d/jacobi2d.cu:
 * 148-166, 224-262). */
int chm_jacobi2d_phase_stats(void* handle, double* out8);

/* The same times per chare: out[4c .. 4c+3] = {sweep, send, wait-for-halos,
 * inside halo-receive entry methods} of chare c, in ms (out_elems >= 4 x
 * chares).  *n_chares receives the chare count; with out NULL only the count
 * is written.  The reference prints these per chare (examples/jacobi2d/
 * jacobi2d.cu:148-166). */
int chm_jacobi2d_chare_phases(void* handle, double* out, size_t out_elems,
                              int* n_chares);

int chm_jacobi2d_destroy(void* handle);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, serving loops and harnesses that
 * forbid hipMalloc.  The plain per-launch solve of chm_jacobi2d_solve
 * (CHM_FORM_PLAIN: one sweep launch per iteration, the ring held fixed) on a
 * grid the caller owns:
 *   grid   device memory, 128-byte aligned, in the library's layout:
 *          height+2 rows of pitch = ((width + 18 + 15) / 16) * 16 doubles;
 *          row i holds ring column 0, interior columns 1..width and ring
 *          column width+1 at element offsets 15 .. 16+width.  u_0 on entry
 *          (the ring is the fixed boundary), u_iters on return.
 *   out3   device memory, 3 words: the order-independent hash of the
 *          interior of u_iters, then the exact residual of the last sweep
 *          (lo, hi), as in chm_jacobi_result.
 *   stream a hipStream_t (NULL = the null stream).  The work is enqueued
 *          and the call returns without waiting; it may be captured.
 * The second grid of the ping-pong and the reduction accumulators live in
 * one static device buffer of CHM_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process exits; a grid whose (height+2) x pitch doubles exceed it less
 * 256 bytes is refused with CHM_E_INVALID (up to about 2870 x 2870).  Every
 * call on one GPU shares that buffer, so calls that overlap on one GPU (two
 * streams, two threads) race.  Calls no hipMalloc, hipFree or hipMemset. */
#define CHM_NOALLOC_SCRATCH_BYTES (64ull << 20)
int chm_jacobi2d_solve_noalloc(double* grid, int width, int height, int iters,
                               void* stream, uint64_t* out3);

const char* chm_jacobi2d_last_error(void);

/* Measured HBM ceiling: device time of one streaming copy of `bytes` bytes
 * (read + write), averaged over `reps` launches.  No reference counterpart:
 * it calibrates the roofline and traffic figures of the sweep. */
int chm_hbm_copy_probe(size_t bytes, int reps, double* ms_per_copy);

/* Free and total HBM of the current device, in bytes (what an admission
 * budget can be sized from).  No reference counterpart. */
int chm_device_memory(size_t* free_bytes, size_t* total_bytes);

/* Sanity hook with no GPU work: returns the ABI version. */
int chm_jacobi2d_abi_version(void);

/* ------------------------------------------------------------------------
 * Batched and served forms.  No reference counterpart: the reference runs
 * one Jacobi2D problem per job (examples/jacobi2d/jacobi2d.cu:26-77).  Here
 * many independent problems share one persistent launch of a task engine:
 * each problem is cut into square chare tiles, one task is one round (two
 * fused sweeps) of one tile, a tile's completion is the (zero-copy) halo
 * message that releases its neighbours, and idle workgroups steal ready
 * tasks from the others.
 * Every problem's result is bit-identical to the same problem solved alone
 * (and to the monolithic CPU sweep).
 * ------------------------------------------------------------------------ */

typedef struct chm_batch_problem {
  int width;       /* interior columns W */
  int height;      /* interior rows H */
  int iters;       /* sweeps (the cap when tol > 0) */
  int block;       /* chare tile side; 0 = the configuration's choice */
  double tol;      /* > 0: stop after the first sweep k whose residual
                      S_k <= tol^2 * S_0 (relative L2 change <= tol) */
  uint64_t seed;   /* CHM_INIT_SEEDED */
  int init_mode;   /* enum chm_jacobi_init */
  double boundary; /* ring value for CHM_INIT_SEEDED */
} chm_batch_problem;

typedef struct chm_batch_config {
  int wg_threads;     /* 256, 512 or 1024 (0 = 512): up to 1024 / wg_threads
                         workgroups per CU, as many as stay resident; fixed
                         sweep counts go in rounds of sixteen fused sweeps
                         at 512 threads or fewer, of eight at 1024 */
  int block;          /* tile side for problems with block 0 (0 = by size) */
  int panel_chunks;   /* sweep shape: 1 or 2 (0 = 1) */
  int rows_in_flight; /* sweep shape: 2 or 4 (0 = 2) */
  double timeout_s;   /* watchdog (0 = 120 s) */
} chm_batch_config;

typedef struct chm_batch_result {
  double device_ms;       /* the engine launch (or the plain sequence) */
  long long tasks;        /* tile tasks executed */
  long long steals;       /* tasks taken from another workgroup's queue */
  long long idle_polls;   /* idle poll rounds (x64) */
  long long cell_updates; /* interior cells written by sweeps */
  int n_wg, wg_threads;
} chm_batch_result;

/* Solve n problems in one persistent launch; grids are seeded in HBM before
 * the timed launch.  Per problem i: hashes[i] = hash of the interior of the
 * result, residuals[2i..2i+1] = exact residual of the last sweep (lo, hi),
 * iters_done[i].  Output arrays may be NULL. */
int chm_jacobi2d_batch_run(const chm_batch_problem* probs, int n,
                           const chm_batch_config* cfg, chm_batch_result* res,
                           uint64_t* hashes, uint64_t* residuals,
                           int* iters_done);

/* The same problems solved one after another in the plain per-iteration
 * form (use_graph = 0: one launch per sweep; 1: each problem's sweeps
 * captured beforehand as a hipGraph and replayed).  Same outputs. */
int chm_jacobi2d_batch_plain(const chm_batch_problem* probs, int n,
                             int use_graph, chm_batch_result* res,
                             uint64_t* hashes, uint64_t* residuals,
                             int* iters_done);

/* Autotuner: picks the workgroup size and sweep shape, and a tile side per
 * problem (blocks_out[i]) from {64, 96, 128, 192, 256, 384, 512} for
 * workgroups of 512 threads or fewer, {64, 112, 128, 224, 256, 448, 512}
 * for 1024, by coordinate descent over size classes, timing short engine
 * runs of the whole batch with every problem capped at sample_iters sweeps
 * rounded up to a multiple of 16 (0 = no cap: full-length runs).  best_rate
 * = cell updates per second of the best. */
int chm_jacobi2d_batch_autotune(const chm_batch_problem* probs, int n,
                                int sample_iters, chm_batch_config* best,
                                int* blocks_out, double* best_rate);

enum chm_server_mode {
  CHM_SERVE_ENGINE = 0, /* continuous batching in the persistent engine */
  CHM_SERVE_PLAIN = 1   /* one request at a time, one launch per sweep */
};

typedef struct chm_server_config {
  int mode;            /* enum chm_server_mode */
  size_t budget_bytes; /* HBM for resident grids (0 = 16 GiB) */
  int max_active;      /* requests resident at once (0 = 4096) */
  int wg_threads, block, panel_chunks, rows_in_flight; /* as chm_batch_config
                          (wg_threads 0 = 512: tolerance rounds of twelve);
                          block 0 = by load: a request gets about
                          workgroups / (1 + resident requests) tiles */
} chm_server_config;

typedef struct chm_server_completion {
  uint64_t ticket;
  int iters;             /* sweeps in the result */
  uint64_t hash;         /* hash of the interior of the result */
  uint64_t residual_lo, residual_hi;   /* S of the last sweep */
  uint64_t residual0_lo, residual0_hi; /* S_0 (tol > 0) */
  double device_ms;      /* admission to completion on the device clock */
  double latency_ms;     /* submit to poll on the host clock */
} chm_server_completion;

typedef struct chm_server_stats {
  int active;
  size_t bytes_in_use, budget_bytes;
  long long submitted, completed;
} chm_server_stats;

/* Start a server: allocates the grid pool (budget) once, launches the engine
 * (mode ENGINE) or a host worker thread (mode PLAIN). */
int chm_jacobi2d_server_start(const chm_server_config* cfg, void** server);

/* Admit one request: CHM_OK and *ticket, or CHM_E_BUSY when its grids do not
 * fit the free budget or max_active requests are resident; the caller keeps
 * it and submits again after completions free memory.  Grids are seeded on
 * the device. */
int chm_jacobi2d_server_submit(void* server, const chm_batch_problem* p,
                               uint64_t* ticket);

/* Collect up to max_out completions (non-blocking); returns the count. */
int chm_jacobi2d_server_poll(void* server, chm_server_completion* out,
                             int max_out);

int chm_jacobi2d_server_stats(void* server, chm_server_stats* out);

/* Wait for every admitted request to complete (up to timeout_s), then stop
 * the engine and free the pool.  abort_now = 1 stops at once. */
int chm_jacobi2d_server_stop(void* server, int abort_now, double timeout_s);

#ifdef __cplusplus
}
#endif

#endif /* CHM_JACOBI2D_H_ */

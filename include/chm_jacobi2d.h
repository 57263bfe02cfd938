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
  int band_rows;     /* rows per wave band (plain form) */
  int wg_waves;      /* waves per workgroup (plain form) */
  int overlap;       /* persistent form: 1 = sweep the edges first and send the
                        halos before the interior (default 0) */
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

int chm_jacobi2d_destroy(void* handle);

const char* chm_jacobi2d_last_error(void);

/* Measured HBM ceiling: device time of one streaming copy of `bytes` bytes
 * (read + write), averaged over `reps` launches.  No reference counterpart:
 * it calibrates the roofline and traffic figures of the sweep. */
int chm_hbm_copy_probe(size_t bytes, int reps, double* ms_per_copy);

/* Sanity hook with no GPU work: returns the ABI version. */
int chm_jacobi2d_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* CHM_JACOBI2D_H_ */

/*
 * C ABI of the charming_amd runtime host library (libcharming_amd.so).
 *
 * The reference owns `int main` and does everything inside it
 * (src/charming.cu:50-263): bootstrap, device checks, proxy tables, argv in
 * managed memory, device constants, comm buffers (src/comm/get/get.cu:69-149),
 * one cooperative launch of the scheduler and a blocking sync, then teardown.
 * Here those steps are split into entry points a host binding can call, and
 * a library-provided `main` (libcharming_main.a) calls them in the
 * reference's order: charm::main_host, chm_init, chm_run, chm_finalize.
 *
 * A chare program is one translation unit that includes <charming.h>; the
 * header instantiates the persistent scheduler kernel there and registers it
 * with chm_register_program() when the program is loaded.
 */
#ifndef CHARMING_AMD_H_
#define CHARMING_AMD_H_

#include <stddef.h>
#include <stdint.h>

#include "chm_status.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct chm_config {
  int device;               /* HIP device ordinal (0) */
  int n_pes;                /* PEs = workgroups of the persistent kernel; 0 = one per CU */
  size_t ring_bytes;        /* per-PE inbox ring, power of two (0 = 8 MiB; get.cu:21) */
  size_t max_msg_bytes;     /* largest message incl. headers (0 = 1 MiB + 64) */
  size_t deferred_bytes;    /* per-PE store for refnum-deferred messages (0 = 2 MiB) */
  size_t heap_bytes;        /* per-PE heap for chare objects (0 = 1 MiB) */
  double timeout_s;         /* watchdog: stop the kernel after this long (0 = none) */
  unsigned long long spin_limit; /* polls before a blocked send fails (0 = none) */
  int verbose;              /* print the launch configuration (charming.cu:214-227) */
  size_t stack_bytes;       /* per-lane stack of the scheduler kernel (0 = 4096).
                               Entry methods are called through pointers, so the
                               compiler cannot size the stack of the chain below
                               them; the runtime's default of 1 KiB is too small
                               for some entry methods */
  int workgroups;           /* GPU-level programs: workgroups of the one PE
                               (0 = one per CU); ignored otherwise */
  size_t user_bytes;        /* GPU-level programs: per-PE arena of malloc_user
                               (0 = 64 MiB) */
} chm_config;

typedef struct chm_stats {
  unsigned long long delivered;  /* entry-method calls, revived ones included */
  unsigned long long sent;       /* messages published */
  unsigned long long deferred;   /* refnum mismatches stored */
  unsigned long long revived;    /* deferred messages re-dispatched */
  unsigned long long send_spins; /* polls spent waiting for ring space */
  double kernel_ms;              /* device time of the last chm_run */
  double overhead_ms_per_pe;     /* mean time per PE in the scheduler around
                                    entry methods (decode, fences, release) */
  int n_pes;
  int threads_per_pe;
  unsigned int error_code;       /* device error word of the last run (0 = none) */
  int error_pe;
} chm_stats;

/* Registration record passed by a chare program (filled in by <charming.h>). */
typedef struct chm_program {
  const void* kernel;            /* the persistent scheduler kernel */
  int threads_per_pe;            /* workgroup size it was compiled for */
  void (*main_host)(int, char**);/* charm::main_host, or NULL if not defined */
  int (*set_ctx)(const void* ctx, size_t bytes); /* copies Ctx to the TU's symbol */
  const char* name;
  int gpu_level;                 /* 1: one PE spans every workgroup (built
                                    with CHARMING_GPU_LEVEL, the reference's
                                    build without SM_LEVEL) */
} chm_program;

/* Called at load time by every program that includes <charming.h>;
 * the last registration wins.  No reference counterpart (its scheduler is
 * linked into libcharming.a with relocatable device code, Makefile:38). */
int chm_register_program(const chm_program* prog);

/* Choose among several registered programs by the source file name they
 * were registered with (default: the last one loaded): the latest program
 * whose name ends with name_part, else the latest that contains it.
 * GPU-level programs are registered as "<source> (gpu-level)". */
int chm_select_program(const char* name_part);

/* charming.cu:76-235: pick the device, check cooperative launch, size and
 * allocate rings, control words and per-PE state, copy argv to the device.
 * cfg may be NULL (defaults). */
int chm_init(int argc, char** argv, const chm_config* cfg);

/* charming.cu:239-243: one cooperative launch of the scheduler, blocking
 * until charm::end() (or an error / the watchdog).  May be called again
 * after it returns; each run starts from empty inboxes. */
int chm_run(void);

/* charming.cu:250-260: free everything chm_init allocated. */
int chm_finalize(void);

/* The body of the library-provided main (charming.cu:50-263). */
int chm_main(int argc, char** argv);

/* Counters and error word of the last run. */
int chm_get_stats(chm_stats* out);

/* Opaque application pointer handed to the device runtime
 * (readable on the device with charm::rt::user_ptr()). */
int chm_set_user_ptr(void* device_ptr);

/* Device-side HIP stream used by chm_run (for event timing by callers). */
void* chm_stream(void);

const char* chm_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* CHARMING_AMD_H_ */

// Host side of the charming_amd runtime (libcharming_amd.so).
//
// Replaces the library-owned main of src/charming.cu:50-263 and the host
// half of the comm module (src/comm/get/get.cu:69-149) with C entry points:
// chm_init sizes and allocates everything in HBM, chm_run makes the single
// cooperative launch of the registered scheduler kernel and watches it,
// chm_finalize frees.  No NVSHMEM/MPI bootstrap: one process, one GPU.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include <vector>

#include "charming_amd.h"
#include "charming_detail/ctx.h"
#include "common/status.h"

using charm::rt::Ctx;
using charm::rt::PeState;
using charm::rt::Stats;

namespace {

struct Runtime {
  std::vector<chm_program> programs;
  int selected = -1;
  bool inited = false;
  int device = 0;
  chm_config cfg;
  chm_program prog;
  Ctx h;  // host copy of the device context
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  uint32_t* host_abort = nullptr;  // host-coherent
  std::vector<void*> dev_allocs;
  Stats* stats_buf = nullptr;  // n_pes packed statistics (chm_get_stats)
  chm_stats last;
  int grid_wgs = 0;  // workgroups launched (n_pes, or the GPU-level PE's)
};

// Packs the statistics of every PE (one field of each PeState) into one
// contiguous buffer, so reading them is a single small copy rather than a
// strided copy with one row per PE (milliseconds at 256 PEs).
__global__ void gather_stats_kernel(const PeState* pes, Stats* out, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x)
    out[i] = pes[i].stats;
}

Runtime& R() {
  static Runtime r;
  return r;
}

const char* err_name(uint32_t c) {
  switch (c) {
    case charm::rt::ERR_NONE: return "none";
    case charm::rt::ERR_MSG_TOO_LARGE: return "message larger than max_msg_bytes";
    case charm::rt::ERR_SELF_RING_FULL: return "inbox full on a send to self";
    case charm::rt::ERR_DEFERRED_FULL: return "refnum deferral store full";
    case charm::rt::ERR_BAD_MESSAGE: return "malformed message";
    case charm::rt::ERR_HEAP_FULL: return "per-PE heap exhausted";
    case charm::rt::ERR_HOST_ABORT: return "stopped by the host watchdog";
    case charm::rt::ERR_USER_ABORT: return "charm::abort() called";
    case charm::rt::ERR_SPIN_LIMIT: return "send spin limit reached (inbox full)";
    case charm::rt::ERR_BAD_CHARE: return "bad chare index, entry method or table";
    case charm::rt::ERR_UNSUPPORTED: return "unsupported feature";
    case charm::rt::ERR_WAITS_FULL: return "too many pending async_waits on a PE";
  }
  return "unknown";
}

double now_s() {
  timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + ts.tv_nsec * 1e-9;
}

size_t round_pow2(size_t v) {
  size_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

template <class T>
int dev_alloc(T** p, size_t bytes, bool zero) {
  void* q = nullptr;
  if (hipMalloc(&q, bytes) != hipSuccess)
    return chm::fail(CHM_E_NOMEM, "hipMalloc of %zu bytes failed", bytes);
  R().dev_allocs.push_back(q);
  if (zero && hipMemset(q, 0, bytes) != hipSuccess)
    return chm::fail(CHM_E_HIP, "hipMemset of %zu bytes failed", bytes);
  *p = (T*)q;
  return CHM_OK;
}

void free_all() {
  Runtime& r = R();
  for (void* p : r.dev_allocs) (void)hipFree(p);
  r.dev_allocs.clear();
  if (r.host_abort) (void)hipHostFree(r.host_abort);
  r.host_abort = nullptr;
  if (r.ev0) (void)hipEventDestroy(r.ev0);
  if (r.ev1) (void)hipEventDestroy(r.ev1);
  if (r.stream) (void)hipStreamDestroy(r.stream);
  r.ev0 = r.ev1 = nullptr;
  r.stream = nullptr;
  r.stats_buf = nullptr;
  r.inited = false;
}

// Process exit with the runtime still initialised (a host program that never
// called chm_finalize): release the stream, events, pinned word and device
// memory while the HIP runtime is still up.  Registered at the first
// chm_init, so it runs before the HIP runtime's own exit handlers, which
// were registered earlier (at the process's first HIP call).
void runtime_atexit() {
  Runtime& r = R();
  if (!r.inited) return;
  if (r.stream) (void)hipStreamSynchronize(r.stream);
  free_all();
}

void default_config(chm_config* c) {
  memset(c, 0, sizeof(*c));
  c->ring_bytes = 8ull << 20;
  c->max_msg_bytes = (1ull << 20) + 64;
  c->deferred_bytes = 2ull << 20;
  c->heap_bytes = 1ull << 20;
  c->stack_bytes = 4096;
  c->user_bytes = 64ull << 20;
}

// Environment overrides for programs started through the library main.
void env_config(chm_config* c) {
  if (const char* s = getenv("CHM_N_PES")) c->n_pes = atoi(s);
  if (const char* s = getenv("CHM_TIMEOUT_S")) c->timeout_s = atof(s);
  if (const char* s = getenv("CHM_RING_BYTES")) c->ring_bytes = strtoull(s, 0, 0);
  if (const char* s = getenv("CHM_VERBOSE")) c->verbose = atoi(s);
  if (const char* s = getenv("CHM_STACK_BYTES")) c->stack_bytes = strtoull(s, 0, 0);
  if (const char* s = getenv("CHM_MAX_MSG_BYTES"))
    c->max_msg_bytes = strtoull(s, 0, 0);
  if (const char* s = getenv("CHM_DEFERRED_BYTES"))
    c->deferred_bytes = strtoull(s, 0, 0);
  if (const char* s = getenv("CHM_SPIN_LIMIT")) c->spin_limit = strtoull(s, 0, 0);
  if (const char* s = getenv("CHM_GPU_WGS")) c->workgroups = atoi(s);
  if (const char* s = getenv("CHM_USER_BYTES")) c->user_bytes = strtoull(s, 0, 0);
}

}  // namespace

extern "C" {

const char* chm_last_error(void) { return chm::last_error_buf(); }

int chm_register_program(const chm_program* prog) {
  if (!prog || !prog->kernel || !prog->set_ctx || prog->threads_per_pe < 64)
    return chm::fail(CHM_E_INVALID, "bad program registration");
  R().programs.push_back(*prog);
  R().selected = (int)R().programs.size() - 1;
  return CHM_OK;
}

int chm_select_program(const char* name_part) {
  Runtime& r = R();
  const size_t np = name_part ? strlen(name_part) : 0;
  for (int pass = 0; pass < 2 && name_part; pass++)
    for (int i = (int)r.programs.size() - 1; i >= 0; i--) {
      const char* nm = r.programs[i].name;
      const size_t nl = strlen(nm);
      const bool hit = pass == 0 ? (nl >= np && strcmp(nm + nl - np, name_part) == 0)
                                 : strstr(nm, name_part) != nullptr;
      if (hit) {
        r.selected = i;
        return CHM_OK;
      }
    }
  return chm::fail(CHM_E_INVALID, "no registered program matches '%s'",
                   name_part ? name_part : "(null)");
}

int chm_init(int argc, char** argv, const chm_config* cfg_in) {
  Runtime& r = R();
  if (r.inited) return chm::fail(CHM_E_STATE, "chm_init called twice");
  if (r.selected < 0)
    return chm::fail(CHM_E_STATE,
                     "no chare program registered (include <charming.h> in "
                     "the program's translation unit)");
  r.prog = r.programs[r.selected];
  chm_config cfg;
  default_config(&cfg);
  if (cfg_in) {
    cfg.device = cfg_in->device;
    cfg.n_pes = cfg_in->n_pes;
    if (cfg_in->ring_bytes) cfg.ring_bytes = cfg_in->ring_bytes;
    if (cfg_in->max_msg_bytes) cfg.max_msg_bytes = cfg_in->max_msg_bytes;
    if (cfg_in->deferred_bytes) cfg.deferred_bytes = cfg_in->deferred_bytes;
    if (cfg_in->heap_bytes) cfg.heap_bytes = cfg_in->heap_bytes;
    cfg.timeout_s = cfg_in->timeout_s;
    cfg.spin_limit = cfg_in->spin_limit;
    cfg.verbose = cfg_in->verbose;
    if (cfg_in->stack_bytes) cfg.stack_bytes = cfg_in->stack_bytes;
    cfg.workgroups = cfg_in->workgroups;
    if (cfg_in->user_bytes) cfg.user_bytes = cfg_in->user_bytes;
  } else {
    env_config(&cfg);
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return chm::fail(CHM_E_NOGPU, "no HIP device visible");
  if (cfg.device < 0 || cfg.device >= ndev)
    return chm::fail(CHM_E_INVALID, "device %d out of range", cfg.device);
  CHM_TRY_HIP(hipSetDevice(cfg.device));
  static bool exit_hook = false;
  if (!exit_hook) {
    atexit(runtime_atexit);
    exit_hook = true;
  }
  // the scheduler kernel's stack is dynamic (entry methods are indirect
  // calls): its per-lane size comes from this limit, not from the compiler
  if (cfg.stack_bytes < 1024) cfg.stack_bytes = 1024;
  CHM_TRY_HIP(hipDeviceSetLimit(hipLimitStackSize, cfg.stack_bytes));
  hipDeviceProp_t prop;
  CHM_TRY_HIP(hipGetDeviceProperties(&prop, cfg.device));
  if (!prop.cooperativeLaunch)
    return chm::fail(CHM_E_UNSUPPORTED, "device lacks cooperative launch");
  const int cus = prop.multiProcessorCount;
  const int threads = r.prog.threads_per_pe;
  int per_cu = 0;
  CHM_TRY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, r.prog.kernel, threads, 0));
  // workgroup-level programs: one workgroup per PE; GPU-level programs: one
  // PE (this GPU, as the reference's c_n_pes on one device) spanning
  // `workgroups` workgroups
  const bool gpu_level = r.prog.gpu_level != 0;
  const int n_pes = gpu_level ? 1 : (cfg.n_pes > 0 ? cfg.n_pes : cus);
  const int grid_wgs =
      gpu_level ? (cfg.workgroups > 0 ? cfg.workgroups : cus) : n_pes;
  if (per_cu < 1 || grid_wgs > per_cu * cus)
    return chm::fail(CHM_E_INVALID,
                     "%d workgroups of %d threads exceed the co-resident "
                     "capacity (%d per CU x %d CUs)",
                     grid_wgs, threads, per_cu, cus);
  r.grid_wgs = grid_wgs;
  cfg.user_bytes = (cfg.user_bytes + 255) / 256 * 256;
  if (cfg.ring_bytes < (64u << 10)) cfg.ring_bytes = 64u << 10;
  cfg.ring_bytes = round_pow2(cfg.ring_bytes);
  cfg.max_msg_bytes =
      (cfg.max_msg_bytes + charm::rt::kGranule - 1) / charm::rt::kGranule *
      charm::rt::kGranule;
  if (cfg.max_msg_bytes > cfg.ring_bytes)
    return chm::fail(CHM_E_INVALID, "max_msg_bytes %zu exceeds ring_bytes %zu",
                     cfg.max_msg_bytes, cfg.ring_bytes);
  cfg.deferred_bytes = (cfg.deferred_bytes + 255) / 256 * 256;
  if (cfg.deferred_bytes > 0xFFFF0000ull) cfg.deferred_bytes = 0xFFFF0000ull;
  cfg.heap_bytes = (cfg.heap_bytes + 255) / 256 * 256;
  r.cfg = cfg;
  r.device = cfg.device;

  Ctx& h = r.h;
  memset(&h, 0, sizeof(h));
  h.n_pes = n_pes;
  h.threads_per_pe = threads;
  h.ring_bytes = cfg.ring_bytes;
  h.max_msg_bytes = cfg.max_msg_bytes;
  h.ring_stride = cfg.ring_bytes + cfg.max_msg_bytes;
  h.deferred_bytes = cfg.deferred_bytes;
  h.heap_bytes = cfg.heap_bytes;
  h.spin_limit = cfg.spin_limit;
  h.user_bytes = gpu_level ? cfg.user_bytes : 0;
  int rc = CHM_OK;
  const size_t ng = cfg.ring_bytes / charm::rt::kGranule;
#define CHM_STEP(x)       \
  if ((rc = (x)) < 0) { \
    free_all();         \
    return rc;          \
  }
  CHM_STEP(dev_alloc(&h.rings, (size_t)n_pes * h.ring_stride, false));
  CHM_STEP(dev_alloc(&h.ready, (size_t)n_pes * ng * 8, true));
  CHM_STEP(dev_alloc(&h.tails, (size_t)n_pes * 128, true));
  CHM_STEP(dev_alloc(&h.heads, (size_t)n_pes * 128, true));
  CHM_STEP(dev_alloc(&h.pes, (size_t)n_pes * sizeof(PeState), true));
  CHM_STEP(dev_alloc(&h.deferred, (size_t)n_pes * cfg.deferred_bytes, false));
  CHM_STEP(dev_alloc(&h.heap, (size_t)n_pes * cfg.heap_bytes, false));
  CHM_STEP(dev_alloc(&h.ctl, charm::rt::CTL_WORDS * sizeof(uint32_t), true));
  CHM_STEP(dev_alloc(&r.stats_buf, (size_t)n_pes * sizeof(Stats), true));
  CHM_STEP(dev_alloc(&h.bcast, charm::rt::kBcastWords * sizeof(uint64_t), true));
  if (gpu_level)
    CHM_STEP(dev_alloc(&h.user_arena, (size_t)n_pes * cfg.user_bytes, false));
  // argv in device memory (charming.cu:145-169 used managed memory)
  size_t total = 0;
  for (int i = 0; i < argc; i++) total += strlen(argv[i]) + 1;
  std::vector<char> hargs(total + 1, 0);
  std::vector<size_t> hsizes(argc > 0 ? argc : 1, 0);
  std::vector<size_t> hoffs(argc > 0 ? argc : 1, 0);
  size_t off = 0;
  for (int i = 0; i < argc; i++) {
    hsizes[i] = strlen(argv[i]);
    hoffs[i] = off;
    memcpy(&hargs[off], argv[i], hsizes[i] + 1);
    off += hsizes[i] + 1;
  }
  char* d_args = nullptr;
  char** d_argv = nullptr;
  size_t* d_argvs = nullptr;
  CHM_STEP(dev_alloc(&d_args, total + 1, false));
  CHM_STEP(dev_alloc(&d_argv, sizeof(char*) * hsizes.size(), false));
  CHM_STEP(dev_alloc(&d_argvs, sizeof(size_t) * hsizes.size(), false));
  std::vector<char*> hptrs(hsizes.size(), nullptr);
  for (int i = 0; i < argc; i++) hptrs[i] = d_args + hoffs[i];
  if (hipMemcpy(d_args, hargs.data(), total + 1, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(d_argv, hptrs.data(), sizeof(char*) * hptrs.size(),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_argvs, hsizes.data(), sizeof(size_t) * hsizes.size(),
                hipMemcpyHostToDevice) != hipSuccess) {
    free_all();
    return chm::fail(CHM_E_HIP, "argv copy failed");
  }
  h.argc = argc;
  h.argv = d_argv;
  h.argvs = d_argvs;
  if (hipHostMalloc((void**)&r.host_abort, 64,
                    hipHostMallocMapped | hipHostMallocCoherent) !=
      hipSuccess) {
    free_all();
    return chm::fail(CHM_E_NOMEM, "pinned host word allocation failed");
  }
  memset(r.host_abort, 0, 64);
  void* dabort = nullptr;
  if (hipHostGetDevicePointer(&dabort, r.host_abort, 0) != hipSuccess) {
    free_all();
    return chm::fail(CHM_E_HIP, "hipHostGetDevicePointer failed");
  }
  h.host_abort = (volatile uint32_t*)dabort;
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate,
                            cfg.device) != hipSuccess ||
      khz <= 0)
    khz = 100000;
  h.wall_hz = (double)khz * 1e3;
  if (hipStreamCreateWithFlags(&r.stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&r.ev0) != hipSuccess ||
      hipEventCreate(&r.ev1) != hipSuccess) {
    free_all();
    return chm::fail(CHM_E_HIP, "stream/event creation failed");
  }
  if (r.prog.set_ctx(&h, sizeof(Ctx)) != 0) {
    free_all();
    return chm::fail(CHM_E_HIP, "copying the runtime context failed");
  }
#undef CHM_STEP
  if (cfg.verbose) {
    printf("[INFO] charming_amd: %d PEs x %d threads on %s (%d CUs), ring %zu B"
           " per PE, max message %zu B\n",
           n_pes, threads, prop.gcnArchName, cus, (size_t)cfg.ring_bytes,
           (size_t)cfg.max_msg_bytes);
    fflush(stdout);
  }
  memset(&r.last, 0, sizeof(r.last));
  r.last.n_pes = n_pes;
  r.last.threads_per_pe = threads;
  r.inited = true;
  return CHM_OK;
}

int chm_set_user_ptr(void* device_ptr) {
  Runtime& r = R();
  if (!r.inited) return chm::fail(CHM_E_STATE, "chm_init first");
  r.h.user = device_ptr;
  if (r.prog.set_ctx(&r.h, sizeof(Ctx)) != 0)
    return chm::fail(CHM_E_HIP, "copying the runtime context failed");
  return CHM_OK;
}

void* chm_stream(void) { return R().stream; }

int chm_run(void) {
  Runtime& r = R();
  if (!r.inited) return chm::fail(CHM_E_STATE, "chm_run before chm_init");
  CHM_TRY_HIP(hipSetDevice(r.device));
  // fresh control words (not a graph capture: plain memset is fine here)
  CHM_TRY_HIP(hipMemsetAsync(r.h.ctl, 0,
                             charm::rt::CTL_WORDS * sizeof(uint32_t), r.stream));
  __atomic_store_n(r.host_abort, 0u, __ATOMIC_SEQ_CST);
  int dummy = 0;
  void* args[] = {&dummy};
  CHM_TRY_HIP(hipEventRecord(r.ev0, r.stream));
  CHM_TRY_HIP(hipLaunchCooperativeKernel(
      r.prog.kernel, dim3(r.grid_wgs), dim3(r.h.threads_per_pe), args, 0,
      r.stream));
  CHM_TRY_HIP(hipEventRecord(r.ev1, r.stream));
  // watchdog (no counterpart in the reference, which blocks forever)
  const double t0 = now_s();
  bool aborted = false;
  double t_abort = 0;
  for (;;) {
    hipError_t q = hipEventQuery(r.ev1);
    if (q == hipSuccess) break;
    if (q != hipErrorNotReady)
      return chm::fail(CHM_E_HIP, "scheduler kernel failed: %s",
                       hipGetErrorString(q));
    const double t = now_s();
    if (!aborted && r.cfg.timeout_s > 0 && t - t0 > r.cfg.timeout_s) {
      __atomic_store_n(r.host_abort, 1u, __ATOMIC_SEQ_CST);
      aborted = true;
      t_abort = t;
    }
    if (aborted && t - t_abort > 30.0)
      return chm::fail(CHM_E_TIMEOUT,
                       "scheduler kernel did not stop 30 s after the abort");
    usleep(aborted ? 1000 : 20);
  }
  CHM_TRY_HIP(hipStreamSynchronize(r.stream));
  float ms = 0;
  CHM_TRY_HIP(hipEventElapsedTime(&ms, r.ev0, r.ev1));
  uint32_t ctl[charm::rt::CTL_WORDS];
  CHM_TRY_HIP(hipMemcpy(ctl, r.h.ctl, sizeof(ctl), hipMemcpyDeviceToHost));
  r.last.kernel_ms = ms;
  r.last.error_code = ctl[charm::rt::CTL_ERROR];
  r.last.error_pe = (int)ctl[charm::rt::CTL_ERROR_PE];
  r.last.n_pes = r.h.n_pes;
  r.last.threads_per_pe = r.h.threads_per_pe;
  if (aborted)
    return chm::fail(CHM_E_TIMEOUT, "watchdog stopped the run after %.1f s",
                     r.cfg.timeout_s);
  if (r.last.error_code)
    return chm::fail(CHM_E_DEVICE, "device error %u (%s) on PE %d, detail %u",
                     ctl[charm::rt::CTL_ERROR],
                     err_name(ctl[charm::rt::CTL_ERROR]),
                     (int)ctl[charm::rt::CTL_ERROR_PE],
                     ctl[charm::rt::CTL_ERROR_AUX]);
  return CHM_OK;
}

int chm_get_stats(chm_stats* out) {
  Runtime& r = R();
  if (!out) return chm::fail(CHM_E_INVALID, "null stats pointer");
  if (!r.inited) return chm::fail(CHM_E_STATE, "chm_init first");
  std::vector<Stats> s(r.h.n_pes);
  gather_stats_kernel<<<(r.h.n_pes + 255) / 256, 256, 0, r.stream>>>(
      r.h.pes, r.stats_buf, r.h.n_pes);
  CHM_TRY_HIP(hipGetLastError());
  CHM_TRY_HIP(hipMemcpyAsync(s.data(), r.stats_buf, sizeof(Stats) * s.size(),
                             hipMemcpyDeviceToHost, r.stream));
  CHM_TRY_HIP(hipStreamSynchronize(r.stream));
  chm_stats o = r.last;
  o.delivered = o.sent = o.deferred = o.revived = o.send_spins = 0;
  double ticks = 0;
  for (const Stats& x : s) {
    ticks += (double)x.overhead_ticks;
    o.delivered += x.delivered;
    o.sent += x.sent;
    o.deferred += x.deferred;
    o.revived += x.revived;
    o.send_spins += x.send_spins;
  }
  o.overhead_ms_per_pe = ticks / r.h.n_pes / r.h.wall_hz * 1e3;
  *out = o;
  return CHM_OK;
}

int chm_finalize(void) {
  Runtime& r = R();
  if (!r.inited) return CHM_OK;
  (void)hipStreamSynchronize(r.stream);
  free_all();
  return CHM_OK;
}

int chm_main(int argc, char** argv) {
  Runtime& r = R();
  if (r.selected >= 0 && r.programs[r.selected].main_host)
    r.programs[r.selected].main_host(argc, argv);
  int rc = chm_init(argc, argv, nullptr);
  if (rc == CHM_OK) {
    rc = chm_run();
    if (rc != CHM_OK)
      fprintf(stderr, "[ERROR] charming_amd: %s\n", chm_last_error());
    const char* st = getenv("CHM_STATS");
    chm_stats s;
    if (st && atoi(st) && chm_get_stats(&s) == CHM_OK) {
      fflush(stdout);
      printf("[STATS] pes=%d wgs=%d threads=%d delivered=%llu sent=%llu "
             "deferred=%llu revived=%llu send_spins=%llu kernel_ms=%.3f "
             "overhead_ms=%.3f error=%u\n",
             s.n_pes, r.grid_wgs, s.threads_per_pe, s.delivered, s.sent,
             s.deferred,
             s.revived, s.send_spins, s.kernel_ms, s.overhead_ms_per_pe,
             s.error_code);
    }
    chm_finalize();
  } else {
    fprintf(stderr, "[ERROR] charming_amd: %s\n", chm_last_error());
  }
  fflush(stdout);
  return rc == CHM_OK ? 0 : 1;
}

}  // extern "C"

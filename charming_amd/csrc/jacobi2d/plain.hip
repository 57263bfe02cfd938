// Jacobi2D C ABI and the plain per-launch forms.
//
// Plain form = the host loop of examples/jacobi-kernel/jacobi_kernel.cu:
// 104-148 (one stencil launch per iteration over the whole grid, ring held
// fixed) which is also what examples/jacobi2d-mpi/jacobi2d.cu:71-76 computes
// when its halos are exchanged in place.  Graph form = the same launches
// captured once into a hipGraph and replayed.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>

#include "jacobi2d_internal.h"

namespace chm {
namespace jac {

// ----------------------------------------------------------------------------
// kernels over one (H+2) x P array
// ----------------------------------------------------------------------------

// u_0: seeded interior (or 0 in the reference init), ring = boundary (or 10),
// padding = 0.  Fills both arrays so the ring is fixed in either.
__global__ void init_global_kernel(double* a, double* b, int W, int H,
                                   int64_t P, uint64_t seed, int mode,
                                   double boundary) {
  const int64_t n = (int64_t)(H + 2) * P;
  const double ring = (mode == CHM_INIT_REFERENCE) ? 10.0 : boundary;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / P;
    const int64_t j = e % P - kColOff;
    double v = 0.0;
    if (j >= 0 && j <= W + 1) {
      const bool interior = i >= 1 && i <= H && j >= 1 && j <= W;
      if (!interior)
        v = ring;
      else
        v = (mode == CHM_INIT_REFERENCE) ? 0.0 : cell_u01(seed, i, j);
    }
    a[e] = v;
    if (b) b[e] = v;
  }
}

// Maps a hardware workgroup id to a logical tile so that consecutive logical
// tiles run on the same XCD (workgroups are dealt to the 8 XCDs round-robin);
// neighbouring tiles then share halo lines in one L2.  The grid is launched
// rounded up to a multiple of 8; ids past `n` do nothing.
__device__ inline int xcd_logical_block(int n) {
  const int per = (n + 7) / 8;
  const int i = blockIdx.x;
  return (i % 8) * per + i / 8;
}

template <int NC, int U>
__global__ void __launch_bounds__(512)
    plain_step_kernel(const double* __restrict__ in, double* __restrict__ out,
                      int W, int H, int64_t P, int band_rows, int n_panels,
                      int n_tiles, int n_blocks) {
  const int lb = xcd_logical_block(n_blocks);
  if (lb >= n_blocks) return;
  const int waves = blockDim.x / 64;
  const int tile = lb * waves + (int)(threadIdx.x / 64);
  if (tile >= n_tiles) return;
  const int panel = tile % n_panels;
  const int band = tile / n_panels;
  const int c0 = 1 + panel * 128 * NC;
  const int r0 = 1 + band * band_rows;
  const int r1 = min(r0 + band_rows, H + 1);
  u128acc acc = {0, 0};
  EdgeOut eo = {nullptr, nullptr, nullptr, nullptr};
  sweep_panel<NC, U, false, false>(in, out, P, W, H, c0, r0, r1, acc, eo);
}

// Order-independent hash over interior rows [r0,r1) x cols [c0,c1).
__global__ void hash_global_kernel(const double* g, int64_t P, int r0, int r1,
                                   int c0, int c1, uint64_t* out) {
  const int64_t rows = r1 - r0, cols = c1 - c0;
  const int64_t n = rows * cols;
  uint64_t h = 0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = r0 + e / cols, j = c0 + e % cols;
    h += hash_cell(i, j, g[i * P + kColOff + j]);
  }
  for (int off = 32; off >= 1; off >>= 1) h += __shfl_down(h, off, 64);
  if (__lane_id() == 0)
    __hip_atomic_fetch_add(out, h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exact fixed-point residual between two arrays over the interior.
__global__ void residual_global_kernel(const double* a, const double* b,
                                       int W, int H, int64_t P,
                                       uint64_t* out) {
  const int64_t n = (int64_t)W * H;
  u128acc acc = {0, 0};
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = 1 + e / W, j = 1 + e % W;
    const int64_t o = i * P + kColOff + j;
    acc_add(acc, res_term(a[o], b[o]));
  }
  acc = wave_reduce_u128(acc);
  if (__lane_id() == 0) atomic_add_u128(out, acc);
}

// Streaming copy of n16 16-byte words: the measured on-device copy rate the
// sweep is compared with, and the calibration kernel for traffic counters.
// Each thread issues kCopyU independent 16-byte loads before its first store
// (several lines in flight per lane), stores are non-temporal (NT) or plain,
// and the grid is sized by occupancy with a grid-stride loop over chunks.
constexpr int kCopyU = 8;
constexpr int kCopyThreads = 256;
typedef double cv2d __attribute__((ext_vector_type(2)));
template <bool NT>
__global__ void __launch_bounds__(kCopyThreads)
    hbm_copy_kernel(const cv2d* __restrict__ src, cv2d* __restrict__ dst,
                    int64_t n16) {
  const int64_t chunk = (int64_t)kCopyThreads * kCopyU;
  for (int64_t base = (int64_t)blockIdx.x * chunk + threadIdx.x; base < n16;
       base += (int64_t)gridDim.x * chunk) {
    cv2d v[kCopyU];
#pragma unroll
    for (int u = 0; u < kCopyU; u++) {
      const int64_t i = base + (int64_t)u * kCopyThreads;
      if (i < n16) v[u] = src[i];
    }
#pragma unroll
    for (int u = 0; u < kCopyU; u++) {
      const int64_t i = base + (int64_t)u * kCopyThreads;
      if (i >= n16) continue;
      if (NT)
        __builtin_nontemporal_store(v[u], &dst[i]);
      else
        dst[i] = v[u];
    }
  }
}

// ----------------------------------------------------------------------------
// launch helpers
// ----------------------------------------------------------------------------

typedef void (*step_fn_t)(const double*, double*, int, int, int64_t, int, int,
                          int, int);

static step_fn_t pick_step(int nc, int u) {
#define CHM_PICK(NC_, U_) \
  if (nc == NC_ && u == U_) return &plain_step_kernel<NC_, U_>;
  CHM_PICK(1, 1) CHM_PICK(1, 2) CHM_PICK(1, 4) CHM_PICK(1, 8)
  CHM_PICK(2, 1) CHM_PICK(2, 2) CHM_PICK(2, 4) CHM_PICK(2, 8)
  CHM_PICK(4, 1) CHM_PICK(4, 2) CHM_PICK(4, 4)
  CHM_PICK(8, 1) CHM_PICK(8, 2)
#undef CHM_PICK
  return nullptr;
}

static void default_tuning(chm_jacobi_tuning* t) {
  t->panel_chunks = 2;
  t->rows_in_flight = 4;
  t->band_rows = 128;
  t->wg_waves = 4;
}

struct StepLaunch {
  step_fn_t fn;
  dim3 grid, block;
  int band_rows, n_panels, n_tiles, n_blocks;
};

static int plan_step(const Handle* h, StepLaunch* L) {
  chm_jacobi_tuning t;
  default_tuning(&t);
  if (h->tune.panel_chunks) t.panel_chunks = h->tune.panel_chunks;
  if (h->tune.rows_in_flight) t.rows_in_flight = h->tune.rows_in_flight;
  if (h->tune.band_rows) t.band_rows = h->tune.band_rows;
  if (h->tune.wg_waves) t.wg_waves = h->tune.wg_waves;
  L->fn = pick_step(t.panel_chunks, t.rows_in_flight);
  if (!L->fn)
    return fail(CHM_E_INVALID, "unsupported tuning panel_chunks=%d rows=%d",
                t.panel_chunks, t.rows_in_flight);
  if (t.wg_waves < 1 || t.wg_waves > 8 || t.band_rows < 1)
    return fail(CHM_E_INVALID, "bad wg_waves=%d band_rows=%d", t.wg_waves,
                t.band_rows);
  const int pw = 128 * t.panel_chunks;
  L->band_rows = t.band_rows;
  L->n_panels = (h->W + pw - 1) / pw;
  const int n_bands = (h->H + t.band_rows - 1) / t.band_rows;
  L->n_tiles = L->n_panels * n_bands;
  L->n_blocks = (L->n_tiles + t.wg_waves - 1) / t.wg_waves;
  const int launched = (L->n_blocks + 7) / 8 * 8;
  L->grid = dim3(launched);
  L->block = dim3(64 * t.wg_waves);
  return CHM_OK;
}

static int launch_init(Handle* h) {
  init_global_kernel<<<2048, 256, 0, h->stream>>>(
      h->grid[0], h->grid[1], h->W, h->H, h->P, h->prob.seed,
      h->prob.init_mode, h->prob.boundary);
  CHM_TRY_HIP(hipGetLastError());
  h->cur = 0;
  return CHM_OK;
}

static int plain_solve(Handle* h, chm_jacobi_result* res) {
  StepLaunch L;
  CHM_TRY(plan_step(h, &L));
  const int iters = h->prob.iters;
  int launches = 0;
  if (h->form == CHM_FORM_GRAPH) {
    if (!h->gexec || h->graph_iters != iters) {
      if (h->gexec) {
        (void)hipGraphExecDestroy(h->gexec);
        h->gexec = nullptr;
      }
      // Capture a full solve starting from grid[0]; replays always start
      // from the state in grid[cur], so capture both parities once.
      hipGraph_t g;
      CHM_TRY_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeGlobal));
      for (int it = 0; it < iters; it++) {
        const double* in = h->grid[it & 1];
        double* out = h->grid[(it + 1) & 1];
        hipLaunchKernelGGL(L.fn, L.grid, L.block, 0, h->stream, in, out, h->W,
                           h->H, h->P, L.band_rows, L.n_panels, L.n_tiles,
                           L.n_blocks);
      }
      CHM_TRY_HIP(hipStreamEndCapture(h->stream, &g));
      hipError_t e = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      CHM_TRY_HIP(e);
      h->graph_iters = iters;
    }
    if (h->cur != 0)
      return fail(CHM_E_STATE, "graph form needs u_k in grid[0]; call reset");
    CHM_TRY_HIP(hipEventRecord(h->ev0, h->stream));
    CHM_TRY_HIP(hipGraphLaunch(h->gexec, h->stream));
    CHM_TRY_HIP(hipEventRecord(h->ev1, h->stream));
    launches = iters;
  } else {
    CHM_TRY_HIP(hipEventRecord(h->ev0, h->stream));
    for (int it = 0; it < iters; it++) {
      const double* in = h->grid[(h->cur + it) & 1];
      double* out = h->grid[(h->cur + it + 1) & 1];
      hipLaunchKernelGGL(L.fn, L.grid, L.block, 0, h->stream, in, out, h->W,
                         h->H, h->P, L.band_rows, L.n_panels, L.n_tiles,
                         L.n_blocks);
    }
    CHM_TRY_HIP(hipGetLastError());
    CHM_TRY_HIP(hipEventRecord(h->ev1, h->stream));
    launches = iters;
  }
  CHM_TRY_HIP(hipStreamSynchronize(h->stream));
  float ms = 0;
  CHM_TRY_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->cur = (h->cur + iters) & 1;
  if (res) {
    memset(res, 0, sizeof(*res));
    res->device_ms = ms;
    res->iters_done = iters;
    res->launches = launches;
    CHM_TRY_HIP(hipMemsetAsync(h->d_acc, 0, 4 * sizeof(uint64_t), h->stream));
    if (iters > 0)
      residual_global_kernel<<<1024, 256, 0, h->stream>>>(
          h->grid[h->cur], h->grid[h->cur ^ 1], h->W, h->H, h->P,
          h->d_acc + 1);
    hash_global_kernel<<<1024, 256, 0, h->stream>>>(
        h->grid[h->cur], h->P, 1, h->H + 1, 1, h->W + 1, h->d_acc);
    CHM_TRY_HIP(hipGetLastError());
    uint64_t acc[4];
    CHM_TRY_HIP(hipMemcpyAsync(acc, h->d_acc, sizeof(acc),
                               hipMemcpyDeviceToHost, h->stream));
    CHM_TRY_HIP(hipStreamSynchronize(h->stream));
    res->hash = acc[0];
    res->residual_lo = acc[1];
    res->residual_hi = acc[2];
  }
  return CHM_OK;
}

static int check_problem(const chm_jacobi_problem* p) {
  if (!p) return fail(CHM_E_INVALID, "null problem");
  if (p->width < 1 || p->height < 1 || p->iters < 0)
    return fail(CHM_E_INVALID, "bad size %dx%d iters %d", p->width, p->height,
                p->iters);
  if ((int64_t)p->width > (1 << 30) || (int64_t)p->height > (1 << 30))
    return fail(CHM_E_INVALID, "grid too large");
  if (p->init_mode != CHM_INIT_SEEDED && p->init_mode != CHM_INIT_REFERENCE)
    return fail(CHM_E_INVALID, "bad init_mode %d", p->init_mode);
  return CHM_OK;
}

// ----------------------------------------------------------------------------
// the no-alloc plain solve: its second grid and accumulators live in one
// static device buffer
// ----------------------------------------------------------------------------

// Bytes of the static buffer; its first kNoAllocAccBytes hold the hash and
// residual accumulators, the rest the second grid of the ping-pong.
constexpr size_t kNoAllocScratchBytes = CHM_NOALLOC_SCRATCH_BYTES;
constexpr size_t kNoAllocAccBytes = 256;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / 8];

// scratch's device address on each device (0 until looked up there)
constexpr int kMaxDevices = 64;
static void* g_scratch_addr[kMaxDevices];
static std::mutex g_scratch_mu;

static int scratch_on_device(int dev, char** out) {
  if (dev < 0 || dev >= kMaxDevices)
    return fail(CHM_E_UNSUPPORTED, "device %d beyond the %d this build tracks",
                dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(CHM_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s",
                  dev, hipGetErrorString(e));
    g_scratch_addr[dev] = ptr;
  }
  *out = (char*)g_scratch_addr[dev];
  return CHM_OK;
}

// Clears n words (a kernel rather than hipMemsetAsync, so the call can be
// captured into a graph).
__global__ void zero_words_kernel(uint64_t* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// Copies the ring (rows 0 and H+1, columns 0 and W+1) of src to dst.
__global__ void copy_ring_kernel(const double* src, double* dst, int W, int H,
                                 int64_t P) {
  const int64_t n = 2 * (int64_t)(W + 2) + 2 * (int64_t)H;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t i, j;
    if (e < 2 * (int64_t)(W + 2)) {
      i = e < W + 2 ? 0 : H + 1;
      j = e % (W + 2);
    } else {
      const int64_t k = e - 2 * (int64_t)(W + 2);
      i = 1 + k / 2;
      j = (k & 1) ? W + 1 : 0;
    }
    const int64_t o = i * P + kColOff + j;
    dst[o] = src[o];
  }
}

// Copies the interior (rows 1..H, columns 1..W) of src to dst.
__global__ void copy_interior_kernel(const double* src, double* dst, int W,
                                     int H, int64_t P) {
  const int64_t n = (int64_t)W * H;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = (1 + e / W) * P + kColOff + 1 + e % W;
    dst[o] = src[o];
  }
}

__global__ void copy_words_kernel(const uint64_t* src, uint64_t* dst, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

static int solve_noalloc(double* grid, int W, int H, int iters,
                         hipStream_t stream, uint64_t* out3) {
  if (!grid || !out3) return fail(CHM_E_INVALID, "null grid or output");
  if (W < 1 || H < 1 || iters < 0)
    return fail(CHM_E_INVALID, "bad size %dx%d iters %d", W, H, iters);
  if (((uintptr_t)grid & 127) != 0)
    return fail(CHM_E_INVALID, "grid must be 128-byte aligned");
  const int64_t P = pitch_for(W);
  const size_t bytes = array_elems(W, H) * sizeof(double);
  if ((int64_t)W > (1 << 30) || (int64_t)H > (1 << 30) ||
      bytes > kNoAllocScratchBytes - kNoAllocAccBytes)
    return fail(CHM_E_INVALID,
                "a %dx%d grid needs %zu bytes of scratch; the no-alloc solve "
                "holds at most %zu",
                W, H, bytes, kNoAllocScratchBytes - kNoAllocAccBytes);
  int dev = 0;
  CHM_TRY_HIP(hipGetDevice(&dev));
  char* base = nullptr;
  CHM_TRY(scratch_on_device(dev, &base));
  uint64_t* acc = (uint64_t*)base;
  double* other = (double*)(base + kNoAllocAccBytes);
  Handle tmp;
  memset(&tmp, 0, sizeof(tmp));
  tmp.W = W;
  tmp.H = H;
  StepLaunch L;
  CHM_TRY(plan_step(&tmp, &L));
  copy_ring_kernel<<<64, 256, 0, stream>>>(grid, other, W, H, P);
  double* g[2] = {grid, other};
  for (int it = 0; it < iters; it++)
    hipLaunchKernelGGL(L.fn, L.grid, L.block, 0, stream, g[it & 1],
                       g[(it + 1) & 1], W, H, P, L.band_rows, L.n_panels,
                       L.n_tiles, L.n_blocks);
  const int cur = iters & 1;
  zero_words_kernel<<<1, 64, 0, stream>>>(acc, 4);
  if (iters > 0)
    residual_global_kernel<<<1024, 256, 0, stream>>>(g[cur], g[cur ^ 1], W, H,
                                                     P, acc + 1);
  hash_global_kernel<<<1024, 256, 0, stream>>>(g[cur], P, 1, H + 1, 1, W + 1,
                                               acc);
  if (cur) copy_interior_kernel<<<1024, 256, 0, stream>>>(other, grid, W, H, P);
  copy_words_kernel<<<1, 64, 0, stream>>>(acc, out3, 3);
  CHM_TRY_HIP(hipGetLastError());
  return CHM_OK;
}

}  // namespace jac
}  // namespace chm

using namespace chm;
using namespace chm::jac;

extern "C" {

int chm_jacobi2d_abi_version(void) { return 1; }

const char* chm_jacobi2d_last_error(void) { return last_error_buf(); }

int chm_jacobi2d_create(const chm_jacobi_problem* prob, int form,
                        void** handle) {
  CHM_TRY(check_problem(prob));
  if (!handle) return fail(CHM_E_INVALID, "null handle pointer");
  if (form != CHM_FORM_PLAIN && form != CHM_FORM_GRAPH &&
      form != CHM_FORM_PERSISTENT)
    return fail(CHM_E_INVALID, "bad form %d", form);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(CHM_E_NOGPU, "no HIP device visible");
  Handle* h = (Handle*)calloc(1, sizeof(Handle));
  if (!h) return fail(CHM_E_NOMEM, "host alloc");
  h->prob = *prob;
  h->form = form;
  h->W = prob->width;
  h->H = prob->height;
  h->P = pitch_for(h->W);
  memset(&h->tune, 0, sizeof(h->tune));  // zero fields = per-form defaults
  int rc = CHM_OK;
  do {
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) !=
            hipSuccess ||
        hipEventCreate(&h->ev0) != hipSuccess ||
        hipEventCreate(&h->ev1) != hipSuccess) {
      rc = fail(CHM_E_HIP, "stream/event creation failed");
      break;
    }
    if (hipMalloc(&h->d_acc, 4 * sizeof(uint64_t)) != hipSuccess) {
      rc = fail(CHM_E_NOMEM, "scratch alloc");
      break;
    }
    if (form == CHM_FORM_PERSISTENT) {
      rc = persistent_create(h);
      break;
    }
    const size_t bytes = array_elems(h->W, h->H) * sizeof(double);
    if (hipMalloc(&h->grid[0], bytes) != hipSuccess ||
        hipMalloc(&h->grid[1], bytes) != hipSuccess) {
      rc = fail(CHM_E_NOMEM, "grid alloc of 2 x %zu bytes", bytes);
      break;
    }
    rc = launch_init(h);
    if (rc == CHM_OK && hipStreamSynchronize(h->stream) != hipSuccess)
      rc = fail(CHM_E_HIP, "init failed");
  } while (0);
  if (rc != CHM_OK) {
    char saved[1024];
    strncpy(saved, last_error_buf(), sizeof(saved) - 1);
    saved[sizeof(saved) - 1] = 0;
    chm_jacobi2d_destroy(h);
    strcpy(last_error_buf(), saved);
    return rc;
  }
  *handle = h;
  return CHM_OK;
}

int chm_jacobi2d_set_tuning(void* handle, const chm_jacobi_tuning* t) {
  Handle* h = (Handle*)handle;
  if (!h) return fail(CHM_E_INVALID, "null handle");
  chm_jacobi_tuning d;
  memset(&d, 0, sizeof(d));
  if (t) d = *t;
  chm_jacobi_tuning chk;
  default_tuning(&chk);
  if (d.panel_chunks) chk.panel_chunks = d.panel_chunks;
  if (d.rows_in_flight) chk.rows_in_flight = d.rows_in_flight;
  if (!pick_step(chk.panel_chunks, chk.rows_in_flight))
    return fail(CHM_E_INVALID, "unsupported tuning %d/%d", chk.panel_chunks,
                chk.rows_in_flight);
  h->tune = d;
  if (h->gexec) {
    (void)hipGraphExecDestroy(h->gexec);
    h->gexec = nullptr;
  }
  return CHM_OK;
}

int chm_jacobi2d_reset(void* handle) {
  Handle* h = (Handle*)handle;
  if (!h) return fail(CHM_E_INVALID, "null handle");
  if (h->form == CHM_FORM_PERSISTENT) return persistent_reset(h);
  CHM_TRY(launch_init(h));
  CHM_TRY_HIP(hipStreamSynchronize(h->stream));
  return CHM_OK;
}

int chm_jacobi2d_solve(void* handle, chm_jacobi_result* res) {
  Handle* h = (Handle*)handle;
  if (!h) return fail(CHM_E_INVALID, "null handle");
  if (h->form == CHM_FORM_PERSISTENT) return persistent_solve(h, res);
  return plain_solve(h, res);
}

int chm_jacobi2d_read_grid(void* handle, double* out, size_t out_elems) {
  Handle* h = (Handle*)handle;
  if (!h || !out) return fail(CHM_E_INVALID, "null argument");
  const size_t need = (size_t)(h->W + 2) * (size_t)(h->H + 2);
  if (out_elems < need)
    return fail(CHM_E_INVALID, "output holds %zu elements, need %zu",
                out_elems, need);
  if (h->form == CHM_FORM_PERSISTENT) return persistent_read_grid(h, out);
  CHM_TRY_HIP(hipMemcpy2DAsync(
      out, (size_t)(h->W + 2) * sizeof(double), h->grid[h->cur] + kColOff,
      (size_t)h->P * sizeof(double), (size_t)(h->W + 2) * sizeof(double),
      (size_t)(h->H + 2), hipMemcpyDeviceToHost, h->stream));
  CHM_TRY_HIP(hipStreamSynchronize(h->stream));
  return CHM_OK;
}

int chm_jacobi2d_hash_region(void* handle, int r0, int r1, int c0, int c1,
                             uint64_t* hash) {
  Handle* h = (Handle*)handle;
  if (!h || !hash) return fail(CHM_E_INVALID, "null argument");
  if (r0 < 1 || r1 > h->H + 1 || c0 < 1 || c1 > h->W + 1 || r0 > r1 ||
      c0 > c1)
    return fail(CHM_E_INVALID, "bad region [%d,%d)x[%d,%d)", r0, r1, c0, c1);
  if (h->form == CHM_FORM_PERSISTENT)
    return persistent_hash_region(h, r0, r1, c0, c1, hash);
  CHM_TRY_HIP(hipMemsetAsync(h->d_acc, 0, sizeof(uint64_t), h->stream));
  hash_global_kernel<<<1024, 256, 0, h->stream>>>(h->grid[h->cur], h->P, r0,
                                                  r1, c0, c1, h->d_acc);
  CHM_TRY_HIP(hipGetLastError());
  CHM_TRY_HIP(hipMemcpyAsync(hash, h->d_acc, sizeof(uint64_t),
                             hipMemcpyDeviceToHost, h->stream));
  CHM_TRY_HIP(hipStreamSynchronize(h->stream));
  return CHM_OK;
}

int chm_jacobi2d_phase_stats(void* handle, double* out8) {
  Handle* h = (Handle*)handle;
  if (!h || !out8) return fail(CHM_E_INVALID, "null argument");
  if (h->form != CHM_FORM_PERSISTENT || !h->ps)
    return fail(CHM_E_UNSUPPORTED, "phase times exist for the persistent form");
  return persistent_phase_stats(h, out8);
}

int chm_jacobi2d_chare_phases(void* handle, double* out, size_t out_elems,
                              int* n_chares) {
  Handle* h = (Handle*)handle;
  if (!h || !n_chares) return fail(CHM_E_INVALID, "null argument");
  if (h->form != CHM_FORM_PERSISTENT || !h->ps)
    return fail(CHM_E_UNSUPPORTED, "phase times exist for the persistent form");
  int n = 0;
  CHM_TRY(persistent_chare_count(h, &n));
  *n_chares = n;
  if (!out) return CHM_OK;
  if (out_elems < (size_t)n * 4)
    return fail(CHM_E_INVALID, "output holds %zu elements, need %d", out_elems,
                n * 4);
  return persistent_chare_phases(h, out);
}

int chm_device_memory(size_t* free_bytes, size_t* total_bytes) {
  if (!free_bytes || !total_bytes) return fail(CHM_E_INVALID, "null argument");
  CHM_TRY_HIP(hipMemGetInfo(free_bytes, total_bytes));
  return CHM_OK;
}

int chm_hbm_copy_probe(size_t bytes, int reps, double* ms_per_copy) {
  if (bytes < 16 || reps < 1 || !ms_per_copy)
    return fail(CHM_E_INVALID, "bad copy probe arguments");
  const int64_t n16 = (int64_t)(bytes / 16);
  cv2d *a = nullptr, *b = nullptr;
  hipStream_t s;
  hipEvent_t e0, e1;
  CHM_TRY_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  CHM_TRY_HIP(hipEventCreate(&e0));
  CHM_TRY_HIP(hipEventCreate(&e1));
  int rc = CHM_OK;
  if (hipMalloc(&a, n16 * 16) != hipSuccess ||
      hipMalloc(&b, n16 * 16) != hipSuccess) {
    rc = fail(CHM_E_NOMEM, "copy probe alloc");
  } else {
    (void)hipMemsetAsync(a, 0, n16 * 16, s);
    // as many workgroups as stay resident, each looping over its chunks;
    // both store kinds are timed and the faster one is reported
    int dev = 0, cus = 0, per_cu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, (const void*)hbm_copy_kernel<true>, kCopyThreads, 0) !=
            hipSuccess ||
        per_cu < 1)
      per_cu = 8;
    const int64_t chunks = (n16 + (int64_t)kCopyThreads * kCopyU - 1) /
                           ((int64_t)kCopyThreads * kCopyU);
    const int blocks = (int)std::min<int64_t>(chunks, (int64_t)cus * per_cu);
    double best = 0;
    for (int nt = 0; nt < 2 && rc == CHM_OK; nt++) {
      auto k = nt ? hbm_copy_kernel<true> : hbm_copy_kernel<false>;
      k<<<blocks, kCopyThreads, 0, s>>>(a, b, n16);  // warm-up
      (void)hipEventRecord(e0, s);
      for (int r = 0; r < reps; r++) k<<<blocks, kCopyThreads, 0, s>>>(a, b, n16);
      (void)hipEventRecord(e1, s);
      if (hipStreamSynchronize(s) != hipSuccess) {
        rc = fail(CHM_E_HIP, "copy probe failed");
      } else {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (nt == 0 || ms / reps < best) best = ms / reps;
      }
    }
    if (rc == CHM_OK) *ms_per_copy = best;
  }
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  return rc;
}

int chm_jacobi2d_solve_noalloc(double* grid, int width, int height,
                               int iters, void* stream, uint64_t* out3) {
  return solve_noalloc(grid, width, height, iters, (hipStream_t)stream, out3);
}

int chm_jacobi2d_destroy(void* handle) {
  Handle* h = (Handle*)handle;
  if (!h) return CHM_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->ps) persistent_destroy(h);
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  if (h->grid[0]) (void)hipFree(h->grid[0]);
  if (h->grid[1]) (void)hipFree(h->grid[1]);
  if (h->d_acc) (void)hipFree(h->d_acc);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  free(h);
  return CHM_OK;
}

}  // extern "C"

// Internal (host + device) declarations shared by the Jacobi2D translation
// units: the problem handle and the device kernels every form reuses.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../common/status.h"
#include "chm_jacobi2d.h"
#include "stencil.h"

namespace chm {
namespace jac {

// Chare geometry of the decomposed forms: chare (r, c) covers interior rows
// [row0, row0 + bh) and columns [col0, col0 + bw) (1-based); equal blocks
// with the last row/column of chares taking the remainder.
struct ChareGeom {
  int row0, col0, bh, bw;
};

__host__ __device__ inline void split_extent(int len, int n, int r, int* start,
                                             int* size) {
  int b = len / n;
  *start = r * b;
  *size = (r == n - 1) ? len - b * (n - 1) : b;
}

struct PersistentState;  // defined by the message-driven form

struct Handle {
  chm_jacobi_problem prob;
  int form;
  int W, H;
  int64_t P;         // pitch of the global array (plain forms)
  double* grid[2];   // plain forms: u_k lives in grid[cur]
  int cur;
  hipStream_t stream;
  hipEvent_t ev0, ev1;
  uint64_t* d_acc;   // 4 x u64 scratch: hash, residual lo/hi
  chm_jacobi_tuning tune;
  // graph form
  hipGraphExec_t gexec;
  int graph_iters;
  // persistent form
  PersistentState* ps;
};

// Forms implemented in other translation units.
int persistent_create(Handle* h);
int persistent_reset(Handle* h);
int persistent_solve(Handle* h, chm_jacobi_result* res);
int persistent_read_grid(Handle* h, double* out);
int persistent_hash_region(Handle* h, int r0, int r1, int c0, int c1,
                           uint64_t* hash);
int persistent_phase_stats(Handle* h, double* out);
int persistent_chare_phases(Handle* h, double* out);
int persistent_chare_count(Handle* h, int* n);
void persistent_destroy(Handle* h);

}  // namespace jac
}  // namespace chm

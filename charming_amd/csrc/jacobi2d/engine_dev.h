// Device side of the Jacobi2D task engine (batched and served forms); see
// engine.h for the model.  Included by batch.hip only.
#pragma once
#include <hip/hip_runtime.h>

#include "engine.h"
#include "stencil.h"

namespace chm {
namespace eng {

using jac::u128acc;

__device__ inline uint64_t ald(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint32_t ald(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void ast(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void ast(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint32_t aadd(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint32_t asub(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_sub(p, v, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint64_t aadd(uint64_t* p, uint64_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline bool acas(uint64_t* p, uint64_t e, uint64_t d) {
  return __hip_atomic_compare_exchange_strong(
      p, &e, d, __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Release at agent scope: L2 writeback then an explicit wait (the compiler's
// own sequence can lose the wait, see the runtime's release_fence).
__device__ inline void release_agent() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
__device__ inline void acquire_agent() {
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}
__device__ inline void drain() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Wave-uniform copies (scalar registers) of values every lane holds alike,
// e.g. values read from LDS, which arrive in vector registers.
__device__ inline uint32_t uni(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline uint64_t uni(uint64_t v) {
  return ((uint64_t)uni((uint32_t)(v >> 32)) << 32) | uni((uint32_t)v);
}
__device__ inline int64_t uni(int64_t v) { return (int64_t)uni((uint64_t)v); }
__device__ inline double uni(double v) {
  return __longlong_as_double((long long)uni((uint64_t)__double_as_longlong(v)));
}
template <class T>
__device__ inline T* uni(T* p) {
  return (T*)uni((uint64_t)p);
}
__device__ inline Request uni(const Request& s) {
  Request r;
  r.a = uni(s.a);
  r.b = uni(s.b);
  r.c = uni(s.c);
  r.pend = uni(s.pend);
  r.pitch = uni(s.pitch);
  r.W = uni(s.W);
  r.H = uni(s.H);
  r.bs = uni(s.bs);
  r.tr = uni(s.tr);
  r.tc = uni(s.tc);
  r.max_iters = uni(s.max_iters);
  r.init_on_device = uni(s.init_on_device);
  r.init_mode = uni(s.init_mode);
  r.boundary = uni(s.boundary);
  r.tol2 = uni(s.tol2);
  r.seed = uni(s.seed);
  r.ticket = uni(s.ticket);
  r.slot = uni(s.slot);
  r.rounds = uni(s.rounds);
  r.depth = uni(s.depth);
  return r;
}

// Array i (0, 1, 2) of a request.
__device__ inline double* arr(const Request& r, int i) {
  return i == 0 ? r.a : (i == 1 ? r.b : r.c);
}

// ---------------------------------------------------------------------------
// queues: one bounded MPMC ring per workgroup (cells pack a 32-bit sequence
// number with the task word so one 64-bit atomic moves both)
// ---------------------------------------------------------------------------
__device__ inline bool q_push(const EngineDesc& E, int q, uint32_t task) {
  uint64_t* cells = E.qcells + (uint64_t)q * E.qcap;
  uint64_t* tail = E.qtail + (uint64_t)q * 16;
  for (int tries = 0; tries < 64; tries++) {
    const uint64_t pos = ald(tail);
    const uint64_t c = ald(&cells[pos & (E.qcap - 1)]);
    const int32_t dif = (int32_t)((uint32_t)(c >> 32) - (uint32_t)pos);
    if (dif == 0) {
      if (acas(tail, pos, pos + 1)) {
        ast(&cells[pos & (E.qcap - 1)], ((pos + 1) << 32) | task);
        return true;
      }
    } else if (dif < 0) {
      return false;  // full
    }
  }
  return false;
}

__device__ inline bool q_pop(const EngineDesc& E, int q, uint32_t* task) {
  uint64_t* cells = E.qcells + (uint64_t)q * E.qcap;
  uint64_t* head = E.qhead + (uint64_t)q * 16;
  for (int tries = 0; tries < 64; tries++) {
    const uint64_t pos = ald(head);
    const uint64_t c = ald(&cells[pos & (E.qcap - 1)]);
    const int32_t dif = (int32_t)((uint32_t)(c >> 32) - (uint32_t)(pos + 1));
    if (dif == 0) {
      if (acas(head, pos, pos + 1)) {
        ast(&cells[pos & (E.qcap - 1)], (pos + E.qcap) << 32);
        *task = (uint32_t)c;
        return true;
      }
    } else if (dif < 0) {
      return false;  // empty
    }
  }
  return false;
}

// Push to queue `q0`, else to the next ones (a full queue).  Queues hold four
// times the tasks that can be ready at once, so some queue always has room.
__device__ inline void push_task(const EngineDesc& E, uint32_t task, int q0) {
  for (;;) {
    for (int k = 0; k < E.n_wg; k++)
      if (q_push(E, (q0 + k) % E.n_wg, task)) return;
    if (*E.host_stop >= 2) return;  // the host is aborting the launch
    __builtin_amdgcn_s_sleep(8);
  }
}
__device__ inline void push_task(const EngineDesc& E, uint32_t task) {
  push_task(E, task, (int)blockIdx.x);
}

// ---------------------------------------------------------------------------
// tile geometry
// ---------------------------------------------------------------------------
struct TileGeom {
  int r0, r1, c0, c1;  // interior rows [r0, r1), cols [c0, c1), 1-based
};
__device__ inline TileGeom tile_geom(const Request& r, uint32_t t) {
  const int ti = (int)t / r.tc, tj = (int)t % r.tc;
  TileGeom g;
  g.r0 = 1 + ti * r.bs;
  g.r1 = min(g.r0 + r.bs, r.H + 1);
  g.c0 = 1 + tj * r.bs;
  g.c1 = min(g.c0 + r.bs, r.W + 1);
  return g;
}
// The up to eight tiles a fused round of tile t reads from (sides and
// corners); returns the count.
__device__ inline int tile_neighbors(const Request& r, uint32_t t,
                                     uint32_t* nb) {
  const int ti = (int)t / r.tc, tj = (int)t % r.tc;
  int n = 0;
  for (int di = -1; di <= 1; di++)
    for (int dj = -1; dj <= 1; dj++) {
      const int i = ti + di, j = tj + dj;
      if ((di | dj) == 0 || i < 0 || i >= r.tr || j < 0 || j >= r.tc) continue;
      nb[n++] = (uint32_t)(i * r.tc + j);
    }
  return n;
}

}  // namespace eng
}  // namespace chm

// Device runtime of charming_amd: the persistent scheduler, the HBM inbox
// rings, refnum deferral/revival, grid barrier and termination.  Included
// (once) by <charming.h> in the chare program's translation unit, which
// therefore holds the scheduler kernel and the user's entry methods in one
// device code object: no relocatable device code is needed (the reference
// links its scheduler with nvcc -dc, Makefile:38).
//
// One PE = one workgroup of the cooperative launch (the reference's SM_LEVEL
// mode, src/sched/scheduler.cu:183-234), with no CE workgroup reserved.
//
// Inbox protocol (replaces the per-pair slot mailbox + min-heap of
// src/comm/get/get.cu:209-781 and ringbuf.cu/heap.cu):
//   producer (thread 0): pos = fetch_add(tail[dst], slot)          (relaxed)
//                        wait until head[dst] >= pos + slot - R     (relaxed)
//   all threads:         write envelope + header + payload at pos
//   each wave:           s_waitcnt 0; workgroup barrier
//   producer (thread 0): ready[dst][pos/64 mod N] = pos/64 + 1     (release;
//                        batched: one release covers every message an
//                        entry method sent, see flush_publish)
//   consumer (thread 0): poll ready[me][read/64 mod N] == read/64+1 (relaxed)
//                        then an agent-scope acquire fence
//   all threads:         run the entry method on the message in place
//   consumer (thread 0): read += slot; head[me] = read             (relaxed,
//                        after the reads completed)
// Tags are logical granule numbers, unique for the life of the rings, so a
// stale word can never look published; messages are consumed in FIFO order
// and a message that straddles the ring end continues into the overflow tail.
#ifndef CHARMING_DETAIL_DEVICE_RT_H_
#define CHARMING_DETAIL_DEVICE_RT_H_

#include <hip/hip_runtime.h>

#include "charming_amd.h"
#include "charming_detail/ctx.h"

#ifndef CHARMING_PE_THREADS
#define CHARMING_PE_THREADS 512  // threads per PE workgroup (charming.cu:208)
#endif

namespace charm {
namespace rt {

static __constant__ Ctx g_ctx;

constexpr uint64_t kFail = ~0ull;
constexpr int kMaxPending = 16;
enum : int { CMD_EXIT = 0, CMD_INBOX = 1, CMD_REVIVED = 2, CMD_ASYNC = 3 };

// Workgroup-shared scratch used to broadcast thread 0's decisions.  The
// pending list holds publish words of messages already written but not yet
// released: one L2 writeback then releases them all (see flush_publish).
// Thread 0's arguments of a send: invoke is workgroup-collective, and as in
// the reference (get.cu:866-881) only thread 0's buffer, size, destination
// and header fields count; the other threads may pass anything (the
// reference's hello passes a null buffer on every thread but 0).
struct SendArgs {
  const void* buf;
  uint64_t payload;
  int dst, type, chare_id, chare_idx, ep, refnum;
};

struct Lds {
  char* ptr;
  uint64_t pos;
  SendArgs snd;
  uint64_t g_slot;   // GPU-level leader: ring bytes of the current message
  int g_kind;        // GPU-level leader: what poll_next found (CMD_*)
  int g_run;         // GPU-level leader: the entry method runs (1) or not
  void* work;        // token of the program's work hook (request_work)
  int has_work;
  uint64_t t_found;  // wall clock when the current message was found
  int cmd;
  int arg;
  int n;
  int n_pend;
  uint64_t* pend_addr[kMaxPending];
  uint64_t pend_tag[kMaxPending];
  uint32_t order[CHARMING_DEFERRED_MAX];
};
static __shared__ Lds s_lds;

#ifdef CHARMING_GPU_LEVEL
// GPU-level mode (the reference's build without SM_LEVEL, scheduler.cu:
// 240-333, get.cu:990-1313): one PE spans every workgroup of the launch.
// Its leader (workgroup 0, thread 0) polls the inbox, decides the refnum
// gate and reserves messages; the decisions reach the other workgroups
// through g_ctx.bcast and a grid barrier; every thread of every workgroup
// runs main, the entry methods and invoke.
__device__ inline int pe_index() { return 0; }
#else
// One PE per workgroup (the reference's SM_LEVEL mode).
__device__ inline int pe_index() { return (int)blockIdx.x; }
#endif

// ---------------------------------------------------------------------------
// memory-model helpers (agent scope = the whole GPU, all 8 XCDs)
// ---------------------------------------------------------------------------
__device__ inline uint64_t ld_relaxed(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint32_t ld_relaxed(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_relaxed(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Release store at agent scope.  Spelled out as fence + explicit wait +
// relaxed store: the compiler's own release sequence (L2 writeback, then
// wait, then store) loses its wait when the wave's counters are already
// known to be drained, since the writeback itself is not tracked; the
// store could then overtake the writeback of the data it publishes.
__device__ inline void release_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
__device__ inline void st_release(uint64_t* p, uint64_t v) {
  release_fence();
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_release(uint32_t* p, uint32_t v) {
  release_fence();
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void fence_acquire() {
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}
// Wait until every memory operation this wave issued has completed.
__device__ inline void wave_drain() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ PeState* pe_state() { return g_ctx.pes + pe_index(); }
__device__ inline Lds& lds() { return s_lds; }
__device__ inline uint64_t n_granules() { return g_ctx.ring_bytes / kGranule; }

__device__ void raise_error(uint32_t code, uint32_t aux) {
  uint32_t* ctl = g_ctx.ctl;
  uint32_t expected = 0;
  if (__hip_atomic_compare_exchange_strong(&ctl[CTL_ERROR], &expected, code,
                                           __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT)) {
    __hip_atomic_store(&ctl[CTL_ERROR_PE], (uint32_t)pe_index(),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&ctl[CTL_ERROR_AUX], aux, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  st_release(&ctl[CTL_TERM], 1u);
}

__device__ inline bool stopping() {
  return ld_relaxed(&g_ctx.ctl[CTL_TERM]) != 0 || *g_ctx.host_abort != 0;
}

// Bump allocator over the PE's heap slice (single-threaded).
__device__ void* heap_alloc(size_t bytes, size_t align) {
  PeState* ps = pe_state();
  uint64_t off = (ps->heap_used + align - 1) / align * align;
  if (off + bytes > g_ctx.heap_bytes) {
    raise_error(ERR_HEAP_FULL, (uint32_t)bytes);
    return nullptr;
  }
  ps->heap_used = off + bytes;
  return g_ctx.heap + (uint64_t)pe_index() * g_ctx.heap_bytes + off;
}

__device__ inline void* user_ptr() { return g_ctx.user; }

// ---------------------------------------------------------------------------
// workgroup copy: 16-byte words when the source allows it, 4 in flight per
// thread, then an alignment-tiered tail.  dst is 16-byte aligned.
// ---------------------------------------------------------------------------
__device__ void wg_copy(char* dst, const char* src, size_t n) {
  const size_t t = threadIdx.x, T = blockDim.x;
  const uintptr_t sa = (uintptr_t)src;
  if ((sa & 15) == 0) {
    const size_t n16 = n >> 4;
    const uint4* s = (const uint4*)src;
    uint4* d = (uint4*)dst;
    size_t i = t;
    for (; i + 3 * T < n16; i += 4 * T) {
      uint4 a = s[i], b = s[i + T], c = s[i + 2 * T], e = s[i + 3 * T];
      d[i] = a;
      d[i + T] = b;
      d[i + 2 * T] = c;
      d[i + 3 * T] = e;
    }
    for (; i < n16; i += T) d[i] = s[i];
    for (size_t k = (n16 << 4) + t; k < n; k += T) dst[k] = src[k];
  } else if ((sa & 7) == 0) {
    const size_t n8 = n >> 3;
    for (size_t i = t; i < n8; i += T)
      ((uint64_t*)dst)[i] = ((const uint64_t*)src)[i];
    for (size_t k = (n8 << 3) + t; k < n; k += T) dst[k] = src[k];
  } else if ((sa & 3) == 0) {
    const size_t n4 = n >> 2;
    for (size_t i = t; i < n4; i += T)
      ((uint32_t*)dst)[i] = ((const uint32_t*)src)[i];
    for (size_t k = (n4 << 2) + t; k < n; k += T) dst[k] = src[k];
  } else {
    for (size_t k = t; k < n; k += T) dst[k] = src[k];
  }
}

// ---------------------------------------------------------------------------
// send
// ---------------------------------------------------------------------------

// Thread 0: release every message this PE has written but not published.
// A release at agent scope writes back the XCD's whole L2, so messages sent
// by one entry method share one writeback instead of paying one each; the
// list is flushed when an entry method (or main) returns, when it is full,
// before any wait on another PE, and at grid barriers.
__device__ void flush_publish() {
  Lds& L = lds();
  if (L.n_pend == 0) return;
  release_fence();
  for (int i = 0; i < L.n_pend; i++)
    __hip_atomic_store(L.pend_addr[i], L.pend_tag[i], __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  L.n_pend = 0;
}

// Reserve `slot` bytes in dst's ring (thread 0).  Returns kFail when the
// runtime is stopping or the request is invalid.
__device__ uint64_t reserve(int dst, uint64_t slot) {
  const Ctx& c = g_ctx;
  if (slot > c.max_msg_bytes) {
    raise_error(ERR_MSG_TOO_LARGE, (uint32_t)(slot >> 4));
    return kFail;
  }
  if (dst < 0 || dst >= c.n_pes) {
    raise_error(ERR_BAD_CHARE, (uint32_t)dst);
    return kFail;
  }
  const uint64_t pos = __hip_atomic_fetch_add(&c.tails[(uint64_t)dst * 16],
                                              slot, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
  const uint64_t head0 = ld_relaxed(&c.heads[(uint64_t)dst * 16]);
  if (pos + slot - head0 > c.ring_bytes) {
    const uint64_t need = pos + slot - c.ring_bytes;
    uint64_t spins = 0;
    flush_publish();  // never wait on another PE with messages held back
    while (ld_relaxed(&c.heads[(uint64_t)dst * 16]) < need) {
      if (dst == pe_index()) {
        raise_error(ERR_SELF_RING_FULL, (uint32_t)slot);
        return kFail;
      }
      if (stopping()) return kFail;
      if (c.spin_limit && ++spins > c.spin_limit) {
        raise_error(ERR_SPIN_LIMIT, (uint32_t)dst);
        return kFail;
      }
      pe_state()->stats.send_spins++;
      __builtin_amdgcn_s_sleep(2);
    }
  }
  return pos;
}

// Thread 0: queue the publication of the message at ring position pos of
// dst (released with the others by flush_publish).
__device__ void queue_publish(int dst, uint64_t pos) {
  const Ctx& c = g_ctx;
  Lds& L = lds();
  const uint64_t g = pos / kGranule;
  if (L.n_pend == kMaxPending) flush_publish();
  L.pend_addr[L.n_pend] =
      &c.ready[(uint64_t)dst * n_granules() + (g & (n_granules() - 1))];
  L.pend_tag[L.n_pend] = g + 1;
  L.n_pend++;
  pe_state()->stats.sent++;
}

// Workgroup-collective: reserve a message in the destination's ring and
// write its headers.  Thread 0's arguments decide (they are broadcast
// through L.snd, which also keeps buf and payload for the caller's copy);
// returns the message's address (nullptr on failure) and ring position.
__device__ char* alloc_msg(int dst, msgtype type, int chare_id, int chare_idx,
                           int ep, int refnum, const void* buf, size_t payload,
                           uint64_t* pos_out) {
  const Ctx& c = g_ctx;
  Lds& L = lds();
  __syncthreads();
  if (threadIdx.x == 0) {
    L.snd.buf = buf;
    L.snd.payload = payload;
    L.snd.dst = dst;
    L.snd.type = (int)type;
    L.snd.chare_id = chare_id;
    L.snd.chare_idx = chare_idx;
    L.snd.ep = ep;
    L.snd.refnum = refnum;
    const uint64_t alloc = envelope::alloc_size(type, payload);
    L.pos = reserve(dst, (alloc + kGranule - 1) & ~(kGranule - 1));
    if (L.pos != kFail) {
      char* base = c.rings + (uint64_t)dst * c.ring_stride +
                   (L.pos & (c.ring_bytes - 1));
      new (base) envelope(type, alloc);
      new (base + sizeof(envelope))
          regular_msg(chare_id, chare_idx, ep, refnum);
    }
  }
  __syncthreads();
  const uint64_t pos = L.pos;
  *pos_out = pos;
  if (pos == kFail) return nullptr;
  return c.rings + (uint64_t)L.snd.dst * c.ring_stride +
         (pos & (c.ring_bytes - 1));
}

// Workgroup-collective send of one message (thread 0's arguments).
__device__ void send_msg(int dst, msgtype type, int chare_id, int chare_idx,
                         int ep, int refnum, const void* buf, size_t payload) {
  uint64_t pos = 0;
  char* base = alloc_msg(dst, type, chare_id, chare_idx, ep, refnum, buf,
                         payload, &pos);
  if (!base) return;
  Lds& L = lds();
  const uint64_t n = L.snd.payload;
  const int to = L.snd.dst;
  if (n) wg_copy(base + sizeof(envelope) + sizeof(regular_msg),
                 (const char*)L.snd.buf, n);
  wave_drain();
  __syncthreads();
  if (threadIdx.x == 0) queue_publish(to, pos);
}

// ---------------------------------------------------------------------------
// deferral store (refnum mismatches), owner-only
// ---------------------------------------------------------------------------
__device__ inline char* arena() {
  return g_ctx.deferred + (uint64_t)pe_index() * g_ctx.deferred_bytes;
}

// Move [src, src+bytes) down to dst < src, in workgroup-wide chunks whose
// loads all finish before any of their stores.
__device__ void move_down(char* dst, const char* src, uint32_t bytes) {
  const uint32_t n16 = bytes >> 4;
  for (uint32_t base = 0; base < n16; base += blockDim.x) {
    const uint32_t i = base + threadIdx.x;
    uint4 v;
    if (i < n16) v = ((const uint4*)src)[i];
    __syncthreads();
    if (i < n16) ((uint4*)dst)[i] = v;
    __syncthreads();
  }
}

// Collective: pack live deferred messages to the start of the arena.
__device__ void compact() {
  PeState* ps = pe_state();
  Lds& L = lds();
  if (threadIdx.x == 0) {
    int n = 0;
    for (int i = 0; i < ps->hw; i++)
      if (ps->deferred[i].state != 0) {
        int k = n++;  // insertion sort by offset
        while (k > 0 && ps->deferred[L.order[k - 1]].offset >
                            ps->deferred[i].offset) {
          L.order[k] = L.order[k - 1];
          k--;
        }
        L.order[k] = (uint32_t)i;
      }
    L.n = n;
  }
  __syncthreads();
  const int n = L.n;
  uint32_t off = 0;
  for (int k = 0; k < n; k++) {
    DeferredEntry& d = ps->deferred[L.order[k]];
    const uint32_t src = d.offset, bytes = (d.bytes + 15) & ~15u;
    if (src != off) move_down(arena() + off, arena() + src, bytes);
    __syncthreads();
    if (threadIdx.x == 0) d.offset = off;
    off += bytes;
    __syncthreads();
  }
  if (threadIdx.x == 0) ps->arena_used = off;
  __syncthreads();
}

// Collective: copy the message at msg into the deferral store.
__device__ void defer(const char* msg, uint64_t alloc, int chare_id,
                      int chare_idx, int refnum) {
  PeState* ps = pe_state();
  Lds& L = lds();
  const uint32_t need = (uint32_t)((alloc + 15) & ~15ull);
  if (threadIdx.x == 0)
    L.arg = (ps->arena_used + (uint64_t)need > g_ctx.deferred_bytes) ? 1 : 0;
  __syncthreads();
  if (L.arg) compact();
  if (threadIdx.x == 0) {
    int e = -1;
    for (int i = 0; i < CHARMING_DEFERRED_MAX; i++)
      if (ps->deferred[i].state == 0) {
        e = i;
        break;
      }
    if (e >= ps->hw) ps->hw = e + 1;
    if (e < 0 || ps->arena_used + (uint64_t)need > g_ctx.deferred_bytes) {
      raise_error(ERR_DEFERRED_FULL, (uint32_t)chare_idx);
      L.arg = -1;
    } else {
      DeferredEntry& d = ps->deferred[e];
      d.offset = ps->arena_used;
      d.bytes = (uint32_t)alloc;
      d.chare_id = chare_id;
      d.chare_idx = chare_idx;
      d.refnum = refnum;
      d.state = 1;
      d.seq = ps->seq++;
      ps->arena_used += need;
      ps->n_deferred++;
      ps->stats.deferred++;
      L.arg = e;
      L.ptr = arena() + d.offset;
    }
  }
  __syncthreads();
  if (L.arg >= 0) wg_copy(L.ptr, msg, alloc);
  __syncthreads();
}

// ---------------------------------------------------------------------------
// dispatch and the scheduler loop
// ---------------------------------------------------------------------------

// Collective.  Decodes the envelope (scheduler.cu:23-76) and runs the entry
// method through the refnum gate; false = must be deferred.
__device__ __noinline__ bool dispatch(envelope* env) {
  if (env->type == msgtype::regular || env->type == msgtype::user) {
    regular_msg* m = (regular_msg*)((char*)env + sizeof(envelope));
    ProxyTable& t = pe_state()->table;
    const int id = m->chare_id;
    if (id < 0 || id >= t.count) {
      if (threadIdx.x == 0) raise_error(ERR_BAD_MESSAGE, (uint32_t)id);
      return true;
    }
    if (threadIdx.x == 0)
      pe_state()->stats.overhead_ticks += wall_clock64() - lds().t_found;
    const bool ok = t.proxies[id]->call(m->chare_idx, m->ep_id,
                                        (char*)m + sizeof(regular_msg),
                                        m->refnum);
    if (threadIdx.x == 0) lds().t_found = wall_clock64();
    return ok;
  }
  if (threadIdx.x == 0) raise_error(ERR_BAD_MESSAGE, (uint32_t)env->type);
  return true;
}

// Thread 0: does every ivar of the wait compare true?
__device__ inline bool wait_satisfied(const AsyncWait& w) {
  for (uint64_t k = 0; k < w.nelems; k++) {
    const uint64_t v = ld_relaxed(w.ivars + k);
    bool ok;
    switch (w.cmp) {
      case charm::CMP_EQ: ok = v == w.cmp_value; break;
      case charm::CMP_NE: ok = v != w.cmp_value; break;
      case charm::CMP_GT: ok = v > w.cmp_value; break;
      case charm::CMP_LE: ok = v <= w.cmp_value; break;
      case charm::CMP_LT: ok = v < w.cmp_value; break;
      default: ok = v >= w.cmp_value; break;
    }
    if (!ok) return false;
  }
  return true;
}

// Thread 0: wait for work.  Revived messages go first (in arrival order),
// then the inbox in FIFO order, then satisfied async waits.
__device__ int poll_next() {
  const Ctx& c = g_ctx;
  PeState* ps = pe_state();
  Lds& L = lds();
  const uint64_t NG = n_granules();
  const uint64_t* rdy = c.ready + (uint64_t)pe_index() * NG;
  char* ring = c.rings + (uint64_t)pe_index() * c.ring_stride;
  for (uint32_t spins = 0;; spins++) {
    if (ps->n_revived > 0) {
      int best = -1;
      for (int i = 0; i < ps->hw; i++)
        if (ps->deferred[i].state == 2 &&
            (best < 0 || ps->deferred[i].seq < ps->deferred[best].seq))
          best = i;
      if (best >= 0) {
        L.t_found = wall_clock64();
        ps->deferred[best].state = 3;  // being processed
        ps->n_revived--;
        L.arg = best;
        L.ptr = arena() + ps->deferred[best].offset;
        return CMD_REVIVED;
      }
      ps->n_revived = 0;
    }
    const uint64_t g = ps->read_pos / kGranule;
    if (ld_relaxed(&rdy[g & (NG - 1)]) == g + 1) {
      L.t_found = wall_clock64();
      fence_acquire();
      L.ptr = ring + (ps->read_pos & (c.ring_bytes - 1));
      return CMD_INBOX;
    }
    if (ps->n_waits > 0) {
      for (int i = 0; i < ASYNC_WAIT_MAX; i++) {
        AsyncWait& w = ps->waits[i];
        if (!w.valid || !wait_satisfied(w)) continue;
        L.t_found = wall_clock64();
        fence_acquire();
        w.valid = 0;
        ps->n_waits--;
        L.arg = w.proxy;
        L.n = w.idx;
        L.pos = (uint64_t)w.ep;
        return CMD_ASYNC;
      }
    }
    if (ld_relaxed(&c.ctl[CTL_TERM]) != 0) return CMD_EXIT;
    if ((spins & 255) == 255) {
      if (*c.host_abort) {
        raise_error(ERR_HOST_ABORT, 0);
        return CMD_EXIT;
      }
      ps->stats.idle_polls++;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

__device__ __noinline__ void process_inbox() {
  const Ctx& c = g_ctx;
  PeState* ps = pe_state();
  envelope* env = (envelope*)lds().ptr;
  const uint64_t alloc = env->size;
  const uint64_t slot = (alloc + kGranule - 1) & ~(kGranule - 1);
  bool ok = true;
  if (alloc < sizeof(envelope) || slot > c.max_msg_bytes) {
    if (threadIdx.x == 0) raise_error(ERR_BAD_MESSAGE, (uint32_t)alloc);
  } else {
    ok = dispatch(env);
    if (!ok) {
      regular_msg* m = (regular_msg*)((char*)env + sizeof(envelope));
      defer((const char*)env, alloc, m->chare_id, m->chare_idx, m->refnum);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    flush_publish();
    wave_drain();
    ps->read_pos += slot;
    st_relaxed(&c.heads[(uint64_t)pe_index() * 16], ps->read_pos);
    if (ok) ps->stats.delivered++;
    ps->stats.overhead_ticks += wall_clock64() - lds().t_found;
  }
}

__device__ __noinline__ void process_revived() {
  PeState* ps = pe_state();
  const int e = lds().arg;
  envelope* env = (envelope*)lds().ptr;
  const bool ok = dispatch(env);
  __syncthreads();
  if (threadIdx.x == 0) {
    flush_publish();
    DeferredEntry& d = ps->deferred[e];
    if (ok) {
      d.state = 0;
      ps->n_deferred--;
      ps->stats.delivered++;
      ps->stats.revived++;
      if (ps->n_deferred == 0) {
        ps->arena_used = 0;
        ps->hw = 0;
      }
    } else {
      d.state = 1;  // refnum moved on again: stay deferred
    }
  }
}

// A satisfied async wait: its entry method, with a null argument.
__device__ __noinline__ void process_async() {
  Lds& L = lds();
  PeState* ps = pe_state();
  const int id = L.arg, idx = L.n, ep = (int)L.pos;
  ProxyTable& t = ps->table;
  if (id < 0 || id >= t.count || !t.proxies[id]) {
    if (threadIdx.x == 0) raise_error(ERR_BAD_CHARE, (uint32_t)id);
  } else {
    t.proxies[id]->call(idx, ep, nullptr, -1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    flush_publish();
    ps->stats.delivered++;
    ps->stats.overhead_ticks += wall_clock64() - L.t_found;
  }
}

// The program's work hook: a chare program that defines CHARMING_WORK_HOOK
// (the name of a `__device__ void f(void* token)`) before including
// <charming.h> may ask from an entry method for that function to run once
// the entry method has returned (charm::request_work(token)).  The scheduler
// calls it directly, not through a function pointer, so the compiler sizes
// it for the scheduler kernel's workgroup (up to 256 VGPRs per lane for
// 512-thread PEs, and a static stack) instead of the 128 VGPRs and dynamic
// stack of code reached through an entry-method pointer.  It runs
// workgroup-collectively like an entry method and may send; its sends are
// published when it returns.  (No counterpart in the reference.)
__device__ __noinline__ void run_work() {
#ifdef CHARMING_WORK_HOOK
  Lds& L = lds();
  for (;;) {
    __syncthreads();
    if (!L.has_work) break;
    void* tok = L.work;
    __syncthreads();
    if (threadIdx.x == 0) L.has_work = 0;
    __syncthreads();
    CHARMING_WORK_HOOK(tok);  // may request more work
    __syncthreads();
    if (threadIdx.x == 0) flush_publish();
  }
#endif
}

// One turn of the scheduler: wait for work and run it; false once the PE
// is told to stop.
__device__ __noinline__ bool pe_step() {
  Lds& L = lds();
  if (threadIdx.x == 0) L.cmd = poll_next();
  __syncthreads();
  const int cmd = L.cmd;
  if (cmd == CMD_EXIT) return false;
  if (cmd == CMD_INBOX)
    process_inbox();
  else if (cmd == CMD_REVIVED)
    process_revived();
  else
    process_async();
  return true;
}

// The scheduler loop.  The work hook is called from here, between turns,
// where nothing is held in vector registers across the call: the hook may
// use every register the workgroup size allows, and a value this loop kept
// across it would be parked in accumulation registers, which count towards
// the kernel's register total on top of the hook's.
__device__ __noinline__ void pe_loop() {
  while (pe_step()) {
    run_work();
    __syncthreads();
  }
}

// Collective barrier over every workgroup of the launch (generation
// counter): all PEs in the one-PE-per-workgroup mode, the one PE in
// GPU-level mode.  Replaces nvshmem_barrier_all + grid.sync
// (scheduler.cu:336-350).
__device__ void grid_barrier() {
  __syncthreads();
  if (threadIdx.x == 0) {
    flush_publish();
    uint32_t* cnt = &g_ctx.ctl[CTL_BAR_COUNT];
    uint32_t* gen = &g_ctx.ctl[CTL_BAR_GEN];
    const uint32_t g = ld_relaxed(gen);
    release_fence();
    const uint32_t arrived = __hip_atomic_fetch_add(
        cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (arrived == gridDim.x - 1) {
      fence_acquire();
      __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      st_release(gen, g + 1);
    } else {
      for (uint32_t spins = 0; ld_relaxed(gen) == g; spins++) {
        if ((spins & 255) == 255 && *g_ctx.host_abort) {
          raise_error(ERR_HOST_ABORT, 1);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      fence_acquire();
    }
  }
  __syncthreads();
}

// Thread 0 at kernel start: empty the inbox (skip anything a previous run
// left in flight), reset owner state.
__device__ void pe_begin() {
#ifdef CHARMING_GPU_LEVEL
  // every workgroup empties its own publication list; the leader's
  // workgroup resets the one PE's state
  if (threadIdx.x == 0) {
    lds().n_pend = 0;
    lds().has_work = 0;
  }
  if (threadIdx.x == 0 && blockIdx.x == 0) {
#else
  if (threadIdx.x == 0) {
#endif
    PeState* ps = pe_state();
    const uint64_t t = ld_relaxed(&g_ctx.tails[(uint64_t)pe_index() * 16]);
    ps->read_pos = t;
    st_relaxed(&g_ctx.heads[(uint64_t)pe_index() * 16], t);
    ps->table.count = 0;
    ps->heap_used = 0;
    ps->arena_used = 0;
    ps->seq = 0;
    ps->n_deferred = 0;
    ps->n_revived = 0;
    ps->ended = 0;
    ps->hw = 0;
    ps->n_waits = 0;
    for (int i = 0; i < ASYNC_WAIT_MAX; i++) ps->waits[i].valid = 0;
    lds().n_pend = 0;
    lds().has_work = 0;
    for (int i = 0; i < CHARMING_DEFERRED_MAX; i++) ps->deferred[i].state = 0;
    ps->stats = Stats{};
    ps->n_user_free = g_ctx.user_bytes ? 1 : 0;
    ps->user_free[0].offset = 0;
    ps->user_free[0].size = g_ctx.user_bytes;
  }
  __syncthreads();
}

#ifdef CHARMING_GPU_LEVEL
// ---------------------------------------------------------------------------
// GPU-level mode: one PE spanning every workgroup
// ---------------------------------------------------------------------------
// words of g_ctx.bcast, written by the leader before a grid barrier and read
// by every workgroup after it
enum : int { BC_CMD = 0, BC_PTR = 1, BC_PROXY = 2, BC_IDX = 3, BC_EP = 4,
             BC_POS = 5, BC_BASE = 6, BC_DST = 7, BC_BUF = 8, BC_SIZE = 9 };
enum : int { CMD_RUN = 16, CMD_SKIP = 17 };

__device__ inline bool is_leader() {
  return blockIdx.x == 0 && threadIdx.x == 0;
}
__device__ inline uint64_t bc_get(int k) { return ld_relaxed(&g_ctx.bcast[k]); }
__device__ inline void bc_put(int k, uint64_t v) { st_relaxed(&g_ctx.bcast[k], v); }

// Copy by every thread of the launch: 16-byte words when the source allows
// it, bytes otherwise.  dst is 16-byte aligned.
__device__ void grid_copy(char* dst, const char* src, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t T = (size_t)gridDim.x * blockDim.x;
  if (((uintptr_t)src & 15) == 0) {
    const size_t n16 = n >> 4;
    for (size_t i = t; i < n16; i += T)
      ((uint4*)dst)[i] = ((const uint4*)src)[i];
    for (size_t k = (n16 << 4) + t; k < n; k += T) dst[k] = src[k];
  } else {
    for (size_t k = t; k < n; k += T) dst[k] = src[k];
  }
}

// Leader: reserve a message of `payload` bytes in dst's ring and write its
// headers; returns the ring position (kFail on failure) and the address.
__device__ uint64_t leader_reserve(int dst, msgtype type, int chare_id,
                                   int chare_idx, int ep, int refnum,
                                   size_t payload, char** base_out) {
  const Ctx& c = g_ctx;
  const uint64_t alloc = envelope::alloc_size(type, payload);
  const uint64_t pos = reserve(dst, (alloc + kGranule - 1) & ~(kGranule - 1));
  *base_out = nullptr;
  if (pos == kFail) return pos;
  char* base = c.rings + (uint64_t)dst * c.ring_stride + (pos & (c.ring_bytes - 1));
  new (base) envelope(type, alloc);
  new (base + sizeof(envelope)) regular_msg(chare_id, chare_idx, ep, refnum);
  *base_out = base;
  return pos;
}

// Grid-collective send (get.cu:1250-1275): barrier, the leader reserves
// and writes the headers, barrier, every workgroup copies part of the
// payload from the leader's buffer, barrier, the leader queues the
// publication.  The buffer must be readable from every workgroup (global
// memory), as in the reference's GPU-level programs.
__device__ void gpu_send_msg(int dst, msgtype type, int chare_id, int chare_idx,
                             int ep, int refnum, const void* buf,
                             size_t payload) {
  grid_barrier();
  if (is_leader()) {
    char* base;
    const uint64_t pos = leader_reserve(dst, type, chare_id, chare_idx, ep,
                                        refnum, payload, &base);
    bc_put(BC_POS, pos);
    bc_put(BC_BASE, (uint64_t)base);
    bc_put(BC_DST, (uint64_t)dst);
    bc_put(BC_BUF, (uint64_t)buf);
    bc_put(BC_SIZE, payload);
  }
  grid_barrier();
  const uint64_t pos = bc_get(BC_POS);
  if (pos == kFail) return;
  char* base = (char*)bc_get(BC_BASE);
  const uint64_t n = bc_get(BC_SIZE);
  const int to = (int)bc_get(BC_DST);
  if (n)
    grid_copy(base + sizeof(envelope) + sizeof(regular_msg),
              (const char*)bc_get(BC_BUF), n);
  wave_drain();
  grid_barrier();
  if (is_leader()) queue_publish(to, pos);
}

// Workgroup 0, collectively: poll the inbox, decide the refnum gate, defer a
// mismatched message, and broadcast what every workgroup runs.
__device__ __noinline__ void gpu_next() {
  Lds& L = lds();
  PeState* ps = pe_state();
  if (threadIdx.x == 0) {
    L.g_kind = poll_next();
    L.g_run = 0;
    L.g_slot = 0;
  }
  __syncthreads();
  const int kind = L.g_kind;
  int proxy = -1, idx = -1, ep = -1, defer_it = 0;
  void* arg = nullptr;
  if (threadIdx.x == 0 && kind != CMD_EXIT) {
    ProxyTable& t = ps->table;
    if (kind == CMD_ASYNC) {
      proxy = L.arg;
      idx = L.n;
      ep = (int)L.pos;
      if (proxy >= 0 && proxy < t.count && t.proxies[proxy] &&
          t.proxies[proxy]->gate(idx, ep, -1) > 0)
        L.g_run = 1;
    } else {
      envelope* env = (envelope*)L.ptr;
      const uint64_t alloc = env->size;
      if (kind == CMD_INBOX) L.g_slot = (alloc + kGranule - 1) & ~(kGranule - 1);
      regular_msg* m = (regular_msg*)((char*)env + sizeof(envelope));
      if (alloc < sizeof(envelope) + sizeof(regular_msg) ||
          alloc > g_ctx.max_msg_bytes ||
          (env->type != msgtype::regular && env->type != msgtype::user) ||
          m->chare_id < 0 || m->chare_id >= t.count) {
        raise_error(ERR_BAD_MESSAGE, (uint32_t)alloc);
      } else {
        proxy = m->chare_id;
        idx = m->chare_idx;
        ep = m->ep_id;
        arg = (char*)m + sizeof(regular_msg);
        const int g = t.proxies[proxy]->gate(idx, ep, m->refnum);
        L.g_run = g > 0;
        defer_it = (g == 0 && kind == CMD_INBOX);
      }
    }
    L.n = defer_it;
    bc_put(BC_CMD, L.g_run ? CMD_RUN : CMD_SKIP);
    bc_put(BC_PROXY, (uint64_t)proxy);
    bc_put(BC_IDX, (uint64_t)idx);
    bc_put(BC_EP, (uint64_t)ep);
    bc_put(BC_PTR, (uint64_t)arg);
  }
  if (threadIdx.x == 0 && kind == CMD_EXIT) bc_put(BC_CMD, CMD_EXIT);
  __syncthreads();
  if (kind == CMD_INBOX && L.n) {
    envelope* env = (envelope*)L.ptr;
    regular_msg* m = (regular_msg*)((char*)env + sizeof(envelope));
    defer((const char*)env, env->size, m->chare_id, m->chare_idx, m->refnum);
  }
}

// Leader, after every workgroup finished the entry method: publish its
// sends, release the inbox space or settle the revived entry.
__device__ void gpu_finish() {
  Lds& L = lds();
  PeState* ps = pe_state();
  flush_publish();
  if (L.g_kind == CMD_INBOX) {
    wave_drain();
    ps->read_pos += L.g_slot;
    st_relaxed(&g_ctx.heads[(uint64_t)pe_index() * 16], ps->read_pos);
  } else if (L.g_kind == CMD_REVIVED) {
    DeferredEntry& d = ps->deferred[L.arg];
    if (L.g_run) {
      d.state = 0;
      ps->n_deferred--;
      ps->stats.revived++;
      if (ps->n_deferred == 0) {
        ps->arena_used = 0;
        ps->hw = 0;
      }
    } else {
      d.state = 1;  // refnum moved on again: stay deferred
    }
  }
  if (L.g_run) ps->stats.delivered++;
}

__device__ __noinline__ void gpu_pe_loop() {
  for (;;) {
    if (blockIdx.x == 0) gpu_next();
    grid_barrier();
    const int cmd = (int)bc_get(BC_CMD);
    if (cmd == CMD_EXIT) break;
    if (cmd == CMD_RUN) {
      ProxyTable& t = pe_state()->table;
      t.proxies[(int)bc_get(BC_PROXY)]->run((int)bc_get(BC_IDX),
                                           (int)bc_get(BC_EP),
                                           (void*)bc_get(BC_PTR));
    }
    grid_barrier();
    if (is_leader()) gpu_finish();
  }
}
#endif  // CHARMING_GPU_LEVEL

}  // namespace rt

// ---------------------------------------------------------------------------
// charm:: runtime functions (charming.h, message.h)
// ---------------------------------------------------------------------------
#ifdef CHARMING_GPU_LEVEL
// GPU-level: grid-collective (every thread of every workgroup), the
// leader's arguments count.
__device__ void send_reg_msg(int chare_id, int chare_idx, int ep_id, void* buf,
                             size_t payload_size, int dst_pe, int refnum) {
  rt::gpu_send_msg(dst_pe, msgtype::regular, chare_id, chare_idx, ep_id,
                   refnum, buf, payload_size);
}

__device__ void* alloc_user_msg(message& msg, int chare_id, int chare_idx,
                                int ep_id, size_t payload_size, int dst_pe,
                                int refnum) {
  rt::grid_barrier();
  if (rt::is_leader()) {
    char* base;
    const uint64_t pos = rt::leader_reserve(dst_pe, msgtype::user, chare_id,
                                            chare_idx, ep_id, refnum,
                                            payload_size, &base);
    rt::bc_put(rt::BC_POS, pos);
    rt::bc_put(rt::BC_BASE, (uint64_t)base);
    rt::bc_put(rt::BC_DST, (uint64_t)dst_pe);
  }
  rt::grid_barrier();
  const uint64_t pos = rt::bc_get(rt::BC_POS);
  msg.env = pos == rt::kFail ? nullptr : (envelope*)rt::bc_get(rt::BC_BASE);
  msg.offset = pos;
  msg.dst_pe = (int)rt::bc_get(rt::BC_DST);
  return msg.payload();
}

__device__ void send_user_msg(message& msg) {
  if (!msg.env) return;
  rt::wave_drain();
  rt::grid_barrier();
  if (rt::is_leader()) rt::queue_publish(msg.dst_pe, msg.offset);
  msg.env = nullptr;
}

// charming.h:31-34 (GPU-level only, as in the reference, get.cu:1347-1361).
// Single-threaded.  Memory from the PE's user arena (chm_config.user_bytes),
// first fit over a sorted list of free extents; free_user returns an extent
// and merges it with its neighbours.  nullptr when nothing fits.
__device__ void* malloc_user(size_t size, size_t& offset) {
  rt::PeState* ps = rt::pe_state();
  const uint64_t need = (size + 15) & ~15ull;
  for (int i = 0; i < ps->n_user_free; i++) {
    rt::UserExtent& e = ps->user_free[i];
    if (e.size < need) continue;
    offset = e.offset;
    e.offset += need;
    e.size -= need;
    if (e.size == 0) {
      for (int k = i; k + 1 < ps->n_user_free; k++)
        ps->user_free[k] = ps->user_free[k + 1];
      ps->n_user_free--;
    }
    return rt::g_ctx.user_arena + (uint64_t)rt::pe_index() * rt::g_ctx.user_bytes +
           offset;
  }
  printf("[ERROR] PE %d: malloc_user of %llu bytes: user arena exhausted\n",
         rt::pe_index(), (unsigned long long)size);
  return nullptr;
}

__device__ void free_user(size_t size, size_t offset) {
  rt::PeState* ps = rt::pe_state();
  const uint64_t len = (size + 15) & ~15ull;
  if (offset + len > rt::g_ctx.user_bytes) {
    rt::raise_error(rt::ERR_BAD_CHARE, (uint32_t)(offset >> 4));
    return;
  }
  int i = 0;
  while (i < ps->n_user_free && ps->user_free[i].offset < offset) i++;
  const bool left = i > 0 && ps->user_free[i - 1].offset +
                                     ps->user_free[i - 1].size == offset;
  const bool right = i < ps->n_user_free &&
                     offset + len == ps->user_free[i].offset;
  if (left && right) {
    ps->user_free[i - 1].size += len + ps->user_free[i].size;
    for (int k = i; k + 1 < ps->n_user_free; k++)
      ps->user_free[k] = ps->user_free[k + 1];
    ps->n_user_free--;
  } else if (left) {
    ps->user_free[i - 1].size += len;
  } else if (right) {
    ps->user_free[i].offset = offset;
    ps->user_free[i].size += len;
  } else if (ps->n_user_free < rt::kUserExtents) {
    for (int k = ps->n_user_free; k > i; k--)
      ps->user_free[k] = ps->user_free[k - 1];
    ps->user_free[i].offset = offset;
    ps->user_free[i].size = len;
    ps->n_user_free++;
  } else {
    rt::raise_error(rt::ERR_HEAP_FULL, (uint32_t)rt::kUserExtents);
  }
}
#else
__device__ void send_reg_msg(int chare_id, int chare_idx, int ep_id, void* buf,
                             size_t payload_size, int dst_pe, int refnum) {
  rt::send_msg(dst_pe, msgtype::regular, chare_id, chare_idx, ep_id, refnum,
               buf, payload_size);
}

__device__ void* alloc_user_msg(message& msg, int chare_id, int chare_idx,
                                int ep_id, size_t payload_size, int dst_pe,
                                int refnum) {
  uint64_t pos = 0;
  char* base = rt::alloc_msg(dst_pe, msgtype::user, chare_id, chare_idx, ep_id,
                             refnum, nullptr, payload_size, &pos);
  msg.env = (envelope*)base;
  msg.offset = pos;
  msg.dst_pe = rt::lds().snd.dst;
  return msg.payload();
}

__device__ void send_user_msg(message& msg) {
  if (!msg.env) return;
  rt::wave_drain();
  __syncthreads();
  if (threadIdx.x == 0) rt::queue_publish(msg.dst_pe, msg.offset);
  msg.env = nullptr;
}
#endif

// Single-threaded (get.cu:1316-1345): deferred messages of (chare, refnum)
// become eligible again; the scheduler re-dispatches them before new inbox
// messages, in arrival order.
__device__ void revive_mismatches(int chare_id, int chare_idx, int refnum) {
  rt::PeState* ps = rt::pe_state();
  if (ps->n_deferred == 0) return;
  for (int i = 0; i < ps->hw; i++) {
    rt::DeferredEntry& d = ps->deferred[i];
    if (d.state == 1 && d.chare_id == chare_id && d.chare_idx == chare_idx &&
        d.refnum == refnum) {
      d.state = 2;
      ps->n_revived++;
    }
  }
}

// charm::end() (charming.cu:265-267): the first call anywhere terminates
// every PE's scheduler loop; messages still in flight are dropped, as in
// the reference (no quiescence detection).  Any thread may call it.
__device__ void send_begin_term_msg() {
  rt::st_release(&rt::g_ctx.ctl[rt::CTL_TERM], 1u);
}

__device__ void end() {
  if (threadIdx.x == 0) rt::pe_state()->ended = 1;
  send_begin_term_msg();
}

// charm::abort(): stops every PE and records ERR_USER_ABORT.  (The
// reference stops only the calling PE, charming.cu:269-278, which leaves the
// others waiting at the final barrier.)
__device__ void abort() { rt::raise_error(rt::ERR_USER_ABORT, rt::pe_index()); }

__device__ int my_pe() { return rt::pe_index(); }
__device__ int n_pes() { return rt::g_ctx.n_pes; }
__device__ int my_pe_node() { return rt::pe_index(); }
__device__ int n_pes_node() { return rt::g_ctx.n_pes; }
__device__ int n_nodes() { return 1; }

__device__ void barrier() { rt::grid_barrier(); }
__device__ void barrier_local() { rt::grid_barrier(); }

// charming.cu:294-304
__device__ int device_atoi(const char* str, int strlen) {
  int tmp = 0;
  for (int i = 0; i < strlen; i++) {
    int multiplier = 1;
    for (int j = 0; j < strlen - i - 1; j++) multiplier *= 10;
    tmp += (str[i] - 48) * multiplier;
  }
  return tmp;
}

__device__ void flush_sends() {
#ifdef CHARMING_GPU_LEVEL
  rt::grid_barrier();  // the leader's pending publications are flushed there
#else
  __syncthreads();
  if (threadIdx.x == 0) rt::flush_publish();
  __syncthreads();
#endif
}

// Ask for the program's work hook to run with `token` once the current
// entry method returns (any thread; the last request before the return
// wins).  Workgroup-level mode.
__device__ void request_work(void* token) {
  if (threadIdx.x == 0) {
    rt::lds().work = token;
    rt::lds().has_work = 1;
  }
}

__device__ double wall_time() {
  return (double)wall_clock64() / rt::g_ctx.wall_hz;
}

}  // namespace charm

// ---------------------------------------------------------------------------
// the persistent scheduler kernel (scheduler.cu:183-234) and its registration
// ---------------------------------------------------------------------------
namespace charm {
namespace rt {

// The user's main and the loop run in (non-inlined) functions so the kernel
// body keeps almost nothing live across calls: every function called through
// an entry-method pointer is compiled for up to 1024-thread workgroups
// (128 VGPRs), and the kernel must not add registers of its own on top.
__device__ __noinline__ void run_user_main() {
  charm::main(g_ctx.argc, g_ctx.argv, g_ctx.argvs, pe_index());
  __syncthreads();
  if (threadIdx.x == 0) flush_publish();
#ifndef CHARMING_GPU_LEVEL
  run_work();
#endif
}

__device__ __noinline__ void pe_run() {
  pe_begin();
  grid_barrier();
  run_user_main();
  grid_barrier();
#ifdef CHARMING_GPU_LEVEL
  gpu_pe_loop();
#else
  pe_loop();
#endif
  grid_barrier();
}

template <int kThreads>
__global__ void __launch_bounds__(kThreads) scheduler_kernel(int unused) {
  (void)unused;
  pe_run();
}

static int set_ctx_host(const void* ctx, size_t bytes) {
  if (bytes != sizeof(Ctx)) return -1;
  return hipMemcpyToSymbol(HIP_SYMBOL(g_ctx), ctx, sizeof(Ctx)) == hipSuccess
             ? 0
             : -1;
}

struct ProgramRegistrar {
  ProgramRegistrar() {
    chm_program p;
    p.kernel = (const void*)&scheduler_kernel<CHARMING_PE_THREADS>;
    p.threads_per_pe = CHARMING_PE_THREADS;
    p.main_host = &charm::main_host;
    p.set_ctx = &set_ctx_host;
#ifdef CHARMING_GPU_LEVEL
    p.name = __BASE_FILE__ " (gpu-level)";
    p.gpu_level = 1;
#else
    p.name = __BASE_FILE__;  // the including program, not this header
    p.gpu_level = 0;
#endif
    chm_register_program(&p);
  }
};
static ProgramRegistrar s_program_registrar;

}  // namespace rt
}  // namespace charm

#endif  // CHARMING_DETAIL_DEVICE_RT_H_

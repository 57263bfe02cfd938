// Chare and chare-proxy object model of the charm:: interface.
//
// Same surface as the reference (include/chare.h:24-270): chare{i, n},
// chare_proxy<C>() self-registering per PE, add_entry_method<&F>(),
// create(n[, block_map]), invoke/invoke_all, set_refnum, and the refnum gate
// in call().  Differences, all internal: entry methods are dispatched through
// a per-proxy table of function pointers (no virtual functors), chare objects
// and tables come from the PE's slice of a host-sized HBM heap instead of the
// device malloc heap, and the default block map is computed arithmetically
// instead of storing loc_map on every PE.
#ifndef _CHARE_H_
#define _CHARE_H_

#include "charming_detail/ctx.h"
#include "common.h"
#include "kernel.h"
#include "message.h"

namespace charm {

__device__ int my_pe();
__device__ int n_pes();

namespace rt {
__device__ PeState* pe_state();
__device__ void* heap_alloc(size_t bytes, size_t align);
__device__ void raise_error(uint32_t code, uint32_t aux);
}  // namespace rt

struct chare {
  int i;
  int n;

  __device__ chare() : i(-1), n(0) {}
};

typedef void (*entry_fn_t)(void* obj, void* arg);

// Entry-method trampoline: obj is a C*, arg the message payload.
template <class C, void Func(C&, void*)>
struct entry_method {
  static __device__ void call(void* obj, void* arg) { Func(*(C*)obj, arg); }
};

// Comparison operators of async_wait, with the values of the NVSHMEM
// constants the reference passes through (get.cu:1181).
enum cmp_op : int { CMP_EQ = 0, CMP_NE = 1, CMP_GT = 2, CMP_LE = 3, CMP_LT = 4,
                    CMP_GE = 5 };

// Kept for source compatibility (chare.h:41-49); the pending waits live in
// the PE's state (rt::AsyncWait).
struct async_wait_t {
  bool valid;
  uint64_t* ivars;
  size_t nelems;
  int cmp;
  uint64_t cmp_value;
  int idx;
  int ep;
};

struct mismatch_t {
  int found_idx;
  uint64_t comp;
  int chare_idx;
  int refnum;

  __device__ mismatch_t() : found_idx(-1) {}
};

struct chare_proxy_base {
  int id;         // Chare array ID (registration order, same on every PE)
  int n_local;    // Number of local chares
  int n_total;    // Total number of chares
  int start_idx;  // Starting index of local chares
  int end_idx;    // Ending index of local chares
  int* loc_map;   // Chare-PE map; nullptr for the default block map
  int em_count;   // Number of registered entry methods
  async_wait_t* async_waits;
  int* refnum;    // Reference numbers of the local chares
  mismatch_t* mismatches;
  void** objs;    // local chare objects, type-erased
  int map_per, map_rem;  // default block map: first map_rem PEs get map_per+1
  entry_fn_t entries[EM_CNT_MAX];

  __device__ chare_proxy_base()
      : id(-1), n_local(0), n_total(0), start_idx(-1), end_idx(-1),
        loc_map(nullptr), em_count(0), async_waits(nullptr), refnum(nullptr),
        mismatches(nullptr), objs(nullptr), map_per(0), map_rem(0) {}

  // PE of chare idx: loc_map, or the block rule of chare.h:136-176.
  __device__ int pe_of(int idx) const {
    if (idx < 0 || idx >= n_total) {
      rt::raise_error(rt::ERR_BAD_CHARE, (uint32_t)idx);
      return 0;
    }
    if (loc_map) return loc_map[idx];
    const int big = map_rem * (map_per + 1);
    if (idx < big) return idx / (map_per + 1);
    return map_rem + (idx - big) / map_per;
  }

  // Single-threaded (chare.h:213-239): once every ivars[k], k < nelems,
  // compares true (cmp) against cmp_value, the scheduler runs entry method
  // ep of local chare idx once, with a null argument and no refnum check.
  // ivars are device memory that other PEs (or the host) write; a writer
  // that publishes data with its ivar releases before the store, and the
  // scheduler acquires after the test succeeds.
  __device__ void async_wait(uint64_t* ivars, size_t nelems, int cmp,
                             uint64_t cmp_value, int idx, int ep) {
    if (idx < start_idx || idx > end_idx || ep < 0 || ep >= em_count ||
        cmp < CMP_EQ || cmp > CMP_GE) {
      rt::raise_error(rt::ERR_BAD_CHARE, (uint32_t)idx);
      return;
    }
    rt::PeState* ps = rt::pe_state();
    for (int i = 0; i < ASYNC_WAIT_MAX; i++) {
      rt::AsyncWait& w = ps->waits[i];
      if (w.valid) continue;
      w.ivars = ivars;
      w.nelems = nelems;
      w.cmp = cmp;
      w.cmp_value = cmp_value;
      w.proxy = id;
      w.idx = idx;
      w.ep = ep;
      w.valid = 1;
      ps->n_waits++;
      return;
    }
    rt::raise_error(rt::ERR_WAITS_FULL, (uint32_t)ASYNC_WAIT_MAX);
  }

  // The refnum gate of call() on its own (GPU-level mode, where the leader
  // decides and then every workgroup runs the entry method): 1 = run now,
  // 0 = defer, -1 = bad chare or entry method (an error is raised).
  __device__ int gate(int idx, int ep, int ref) const {
    if (idx < start_idx || idx > end_idx || ep < 0 || ep >= em_count) {
      rt::raise_error(rt::ERR_BAD_CHARE, (uint32_t)idx);
      return -1;
    }
    return (ref == -1 || ref == refnum[idx - start_idx]) ? 1 : 0;
  }
  __device__ void run(int idx, int ep, void* arg) {
    entries[ep](objs[idx - start_idx], arg);
  }

  // Refnum gate then entry method (chare.h:181-195).  Workgroup-collective.
  // Returns false when the message must be deferred.
  __device__ bool call(int idx, int ep, void* arg, int ref) {
    if (idx < start_idx || idx > end_idx || ep < 0 || ep >= em_count) {
      rt::raise_error(rt::ERR_BAD_CHARE, (uint32_t)idx);
      return true;
    }
    const int local_idx = idx - start_idx;
    if (ref != -1 && ref != refnum[local_idx]) return false;
    entries[ep](objs[local_idx], arg);
    return true;
  }
};

template <typename C>
struct chare_proxy : chare_proxy_base {
  C** objects;  // Chare objects on this PE (local chares)

  // Registers this proxy in the calling PE's table (single-threaded).
  __device__ chare_proxy() : chare_proxy_base(), objects(nullptr) {
    rt::ProxyTable& t = rt::pe_state()->table;
    if (t.count >= CHARE_TYPE_CNT_MAX) {
      rt::raise_error(rt::ERR_BAD_CHARE, (uint32_t)t.count);
      return;
    }
    id = t.count++;
    t.proxies[id] = this;
  }

  // Add entry method to table (single-threaded)
  template <void Func(C&, void*)>
  __device__ void add_entry_method() {
    if (em_count >= EM_CNT_MAX) {
      rt::raise_error(rt::ERR_BAD_CHARE, (uint32_t)em_count);
      return;
    }
    entries[em_count++] = &entry_method<C, Func>::call;
  }

  // Called on all PEs before main's sends (single-threaded): block-map the
  // chares (chare.h:125-177) and construct the local ones.
  __device__ void create(int n_chares, int* block_map) {
    const int P = charm::n_pes();
    const int me = charm::my_pe();
    n_total = n_chares;
    int start = 0, cnt = 0;
    if (block_map) {
      loc_map = (int*)rt::heap_alloc(sizeof(int) * (size_t)n_chares, 4);
      if (!loc_map) return;
      int s = 0;
      for (int pe = 0; pe < P; pe++) {
        const int c = block_map[pe];
        if (pe == me) {
          start = s;
          cnt = c;
        }
        for (int i = s; i < s + c && i < n_chares; i++) loc_map[i] = pe;
        s += c;
      }
    } else {
      map_per = n_chares / P;
      map_rem = n_chares % P;
      cnt = map_per + (me < map_rem ? 1 : 0);
      start = (me < map_rem) ? me * (map_per + 1)
                             : map_rem * (map_per + 1) + (me - map_rem) * map_per;
    }
    n_local = cnt;
    start_idx = start;
    end_idx = start + cnt - 1;
    if (n_local > 0) {
      objects = (C**)rt::heap_alloc(sizeof(C*) * (size_t)n_local, 8);
      refnum = (int*)rt::heap_alloc(sizeof(int) * (size_t)n_local, 4);
      if (!objects || !refnum) return;
      objs = (void**)objects;
      for (int idx = 0; idx < n_local; idx++) {
        void* mem = rt::heap_alloc(sizeof(C), alignof(C));
        if (!mem) return;
        objects[idx] = new (mem) C;
        objects[idx]->i = start_idx + idx;
        objects[idx]->n = n_total;
        refnum[idx] = -1;
      }
    }
  }

  inline __device__ void create(int n_chares) { create(n_chares, nullptr); }

  __device__ bool call(int idx, int ep, void* arg, int ref) {
    return chare_proxy_base::call(idx, ep, arg, ref);
  }

  // invoke copies buf into the destination's inbox before returning
  // (chare.h:197-203).  Workgroup-collective.
  inline __device__ void invoke(int idx, int ep) { invoke(idx, ep, nullptr, 0); }
  inline __device__ void invoke(int idx, int ep, void* buf, size_t size) {
    send_reg_msg(id, idx, ep, buf, size, pe_of(idx), -1);
  }
  inline __device__ void invoke(int idx, int ep, void* buf, size_t size,
                                int refnum) {
    send_reg_msg(id, idx, ep, buf, size, pe_of(idx), refnum);
  }

  inline __device__ void invoke_all(int ep) { invoke_all(ep, nullptr, 0); }
  inline __device__ void invoke_all(int ep, void* buf, size_t size) {
    for (int i = 0; i < n_total; i++) invoke(i, ep, buf, size);
  }

  // Zero-copy send (workgroup-collective): allocate a message for entry
  // method ep of chare idx in its PE's inbox, fill the returned payload in
  // place, then send(msg).  No waits on other PEs in between.
  inline __device__ void* alloc(message& msg, int idx, int ep, size_t size,
                                int refnum = -1) {
    return alloc_user_msg(msg, id, idx, ep, size, pe_of(idx), refnum);
  }
  inline __device__ void send(message& msg) { send_user_msg(msg); }

  // Single-threaded (chare.h:243-253): set the refnum of a local chare and
  // revive its deferred messages carrying that refnum.
  inline __device__ void set_refnum(int idx, int val) {
    if (idx < start_idx || idx > end_idx) {
      rt::raise_error(rt::ERR_BAD_CHARE, (uint32_t)idx);
      return;
    }
    refnum[idx - start_idx] = val;
    revive_mismatches(id, idx, val);
  }
};

}  // namespace charm

#endif  // _CHARE_H_

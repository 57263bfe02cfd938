// Failure paths of the runtime, each forced from the host's configuration so
// that every kernel still exits normally (no trap, no hang):
//
//   ring     chare 1 (on PE 1) holds its PE in an entry method for a bounded
//            time while chare 0 sends `count` messages of `bytes` to it; with
//            a small inbox ring and a send spin limit (CHM_RING_BYTES,
//            CHM_MAX_MSG_BYTES, CHM_SPIN_LIMIT) the sender gives up:
//            ERR_SPIN_LIMIT, and chm_run returns CHM_E_DEVICE.
//   defer    chare 0 sends `count` messages of `bytes` whose refnum chare 1
//            never takes; with a small deferral store (CHM_DEFERRED_BYTES)
//            the store overflows: ERR_DEFERRED_FULL.
//   never    chare 0 sends one message whose refnum never matches and nobody
//            calls end(): the host watchdog (CHM_TIMEOUT_S) stops the run
//            through the host abort word: CHM_E_TIMEOUT.
//
// The reference has no such paths: a full ring or mismatch store is a
// device assert there (get.cu:436-439, 738-742) and a hang has no watchdog.
//
// argv: mode count bytes
#include <stdio.h>
#include <string.h>

#include <charming.h>

struct Node : charm::chare {
  __device__ Node() {}
  __device__ void start(void* arg);
  __device__ void hold(void* arg);
  __device__ void take(void* arg);
};

__device__ void entry_start(Node& c, void* arg) { c.start(arg); }
__device__ void entry_hold(Node& c, void* arg) { c.hold(arg); }
__device__ void entry_take(Node& c, void* arg) { c.take(arg); }

__shared__ charm::chare_proxy<Node>* node_proxy;
__device__ int g_mode, g_count, g_bytes;
__device__ char g_payload[1 << 16];

__device__ void charm::main(int argc, char** argv, size_t* argvs, int pe) {
  if (threadIdx.x == 0) {
    void* mem = charm::rt::heap_alloc(sizeof(charm::chare_proxy<Node>), 16);
    node_proxy = new (mem) charm::chare_proxy<Node>();
    node_proxy->add_entry_method<&entry_start>();
    node_proxy->add_entry_method<&entry_hold>();
    node_proxy->add_entry_method<&entry_take>();
    // chare 0 on PE 0, chare 1 on the last PE
    const int P = charm::n_pes();
    int* block_map = (int*)charm::rt::heap_alloc(sizeof(int) * P, 4);
    for (int i = 0; i < P; i++) block_map[i] = 0;
    block_map[0] += 1;
    block_map[P - 1] += 1;
    node_proxy->create(2, block_map);
    if (pe == 0) {
      const char* m = argc > 1 ? argv[1] : "never";
      g_mode = (m[0] == 'r') ? 0 : (m[0] == 'd') ? 1 : 2;
      g_count = argc > 2 ? charm::device_atoi(argv[2], argvs[2]) : 1;
      g_bytes = argc > 3 ? charm::device_atoi(argv[3], argvs[3]) : 8;
      if (g_bytes > (int)sizeof(g_payload)) g_bytes = sizeof(g_payload);
    }
  }
  __syncthreads();
  barrier();
  if (pe == 0) node_proxy->invoke(0, 0);
}

// chare 0: the flood (or the one unmatched message)
__device__ void Node::start(void* arg) {
  const int mode = g_mode;
  if (mode == 0) node_proxy->invoke(1, 1);  // keep PE 1 busy first
  charm::flush_sends();
  for (int k = 0; k < g_count; k++) {
    if (mode == 0)
      node_proxy->invoke(1, 2, g_payload, g_bytes);
    else
      node_proxy->invoke(1, 2, g_payload, g_bytes, 1000 + k);  // never taken
    charm::flush_sends();
  }
  if (threadIdx.x == 0)
    printf("failure program: mode %d sent %d messages of %d bytes\n", mode,
           g_count, g_bytes);
}

// chare 1: occupy the PE for a bounded time (the sender's spin limit is far
// shorter), or until the runtime stops
__device__ void Node::hold(void* arg) {
  if (threadIdx.x == 0) {
    const double t0 = charm::wall_time();
    while (charm::wall_time() - t0 < 0.5 && !charm::rt::stopping())
      __builtin_amdgcn_s_sleep(8);
  }
  __syncthreads();
}

__device__ void Node::take(void* arg) {}

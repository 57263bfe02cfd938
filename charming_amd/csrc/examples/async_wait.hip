// async_wait on one GPU: chare 0 waits on device ivars that the other
// chares (one per PE) set, and its entry method runs once when every ivar
// compares true (reference chare.h:213-239, get.cu:1172-1202, where the
// ivars are NVSHMEM symmetric variables of other GPUs).
//
// Round r (r = 1..rounds): chare 0 registers async_wait(ivars, n-1, EQ, r)
// and a second wait on a counter with GE, then asks every other chare to
// publish; chare i writes its slot's payload, releases, and stores
// ivars[i-1] = r, then adds 1 to the counter.  When the waits fire, chare 0
// checks every payload and starts the next round.
//
// argv: [rounds]
#include <hip/hip_runtime.h>
#include <stdio.h>

#include <charming.h>

struct Node : charm::chare {
  int round, rounds, fired_eq, fired_ge;
  unsigned long long errors;

  __device__ Node() {}
  __device__ void init(void* arg);
  __device__ void publish(void* arg);
  __device__ void all_equal(void* arg);
  __device__ void counted(void* arg);
  __device__ void start_round();
};

__device__ void entry_init(Node& c, void* a) { c.init(a); }
__device__ void entry_publish(Node& c, void* a) { c.publish(a); }
__device__ void entry_all_equal(Node& c, void* a) { c.all_equal(a); }
__device__ void entry_counted(Node& c, void* a) { c.counted(a); }

__shared__ charm::chare_proxy<Node>* node_proxy;
// set by main_host: n_pes ivars, a counter, n_pes payload words
__device__ uint64_t* g_ivars;
__device__ uint64_t* g_counter;
__device__ uint64_t* g_payload;

void charm::main_host(int argc, char** argv) {
  const size_t words = 3 * 4096;
  uint64_t* buf = nullptr;
  if (hipMalloc(&buf, words * 8) != hipSuccess) return;
  (void)hipMemset(buf, 0, words * 8);
  uint64_t* ivars = buf;
  uint64_t* counter = buf + 4096;
  uint64_t* payload = buf + 8192;
  (void)hipMemcpyToSymbol(HIP_SYMBOL(g_ivars), &ivars, 8);
  (void)hipMemcpyToSymbol(HIP_SYMBOL(g_counter), &counter, 8);
  (void)hipMemcpyToSymbol(HIP_SYMBOL(g_payload), &payload, 8);
}

__device__ void charm::main(int argc, char** argv, size_t* argvs, int pe) {
  if (threadIdx.x == 0) {
    void* mem = charm::rt::heap_alloc(sizeof(charm::chare_proxy<Node>), 16);
    node_proxy = new (mem) charm::chare_proxy<Node>();
    node_proxy->add_entry_method<&entry_init>();
    node_proxy->add_entry_method<&entry_publish>();
    node_proxy->add_entry_method<&entry_all_equal>();
    node_proxy->add_entry_method<&entry_counted>();
    node_proxy->create(charm::n_pes());
  }
  __syncthreads();
  if (pe == 0) {
    __shared__ int rounds;
    if (threadIdx.x == 0) {
      rounds = (argc > 1) ? charm::device_atoi(argv[1], argvs[1]) : 5;
      printf("Async wait: %d chares, %d rounds\n", charm::n_pes(), rounds);
    }
    __syncthreads();
    node_proxy->invoke(0, 0, &rounds, sizeof(int));
  }
}

__device__ void Node::start_round() {
  __shared__ int r;
  if (threadIdx.x == 0) {
    ++round;
    fired_eq = fired_ge = 0;
    r = round;
    if (n > 1) {
      node_proxy->async_wait(g_ivars, (size_t)(n - 1), charm::CMP_EQ,
                             (uint64_t)round, 0, 2);
      node_proxy->async_wait(g_counter, 1, charm::CMP_GE,
                             (uint64_t)round * (uint64_t)(n - 1), 0, 3);
    }
  }
  __syncthreads();
  for (int k = 1; k < n; k++) node_proxy->invoke(k, 1, &r, sizeof(int));
}

__device__ void Node::init(void* arg) {
  if (threadIdx.x == 0) {
    rounds = *(int*)arg;
    round = 0;
    errors = 0;
  }
  __syncthreads();
  if (n == 1) {
    if (threadIdx.x == 0) printf("Async wait done: nothing to wait for\n");
    charm::end();
    return;
  }
  start_round();
}

__device__ void Node::publish(void* arg) {
  const int r = *(int*)arg;
  if (threadIdx.x == 0) {
    // payload, then release, then the ivar and the counter
    g_payload[i] = (uint64_t)r * 1000003ull + (uint64_t)i;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __builtin_amdgcn_s_waitcnt(0);
    __hip_atomic_store(&g_ivars[i - 1], (uint64_t)r, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(g_counter, 1ull, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
  }
}

__device__ void Node::all_equal(void* arg) {
  __shared__ int go;
  if (threadIdx.x == 0) {
    fired_eq++;
    unsigned long long bad = 0;
    for (int k = 1; k < n; k++)
      bad += __hip_atomic_load(&g_payload[k], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT) !=
             (uint64_t)round * 1000003ull + (uint64_t)k;
    errors += bad;
    printf("Round %d: all %d ivars equal, payload errors %llu\n", round, n - 1,
           bad);
    go = fired_ge > 0;
  }
  __syncthreads();
  if (go) {
    if (round == rounds) {
      if (threadIdx.x == 0)
        printf("Async wait done: %d rounds, errors %llu\n", rounds, errors);
      charm::end();
    } else {
      start_round();
    }
  }
}

__device__ void Node::counted(void* arg) {
  __shared__ int go;
  if (threadIdx.x == 0) {
    fired_ge++;
    printf("Round %d: counter reached %llu\n", round,
           (unsigned long long)__hip_atomic_load(g_counter, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT));
    go = fired_eq > 0;
  }
  __syncthreads();
  if (go) {
    if (round == rounds) {
      if (threadIdx.x == 0)
        printf("Async wait done: %d rounds, errors %llu\n", rounds, errors);
      charm::end();
    } else {
      start_round();
    }
  }
}

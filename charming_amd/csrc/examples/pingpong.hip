// Pingpong between two chares (the reference's examples/pingpong,
// pingpong.cu:17-217): payload sizes double from min to max, `iters` timed
// round trips after `warmup`, one-way latency = elapsed / 2 / iters.  Added
// here: payloads are seeded bytes that the receiver checks, and every
// message carries refnum = its round-trip sequence number, which the
// receiver's refnum gate must match exactly.
//
// Like the reference it builds in both modes:
//   one PE per workgroup (default, the reference's SM_LEVEL): chares on PE 0
//   and PE n_pes/2, entry methods run by one workgroup;
//   GPU-level (built with CHARMING_GPU_LEVEL, the reference's default build,
//   pingpong.cu:9-15): one PE spans every workgroup, so on one GPU both
//   chares live on PE 0 and their messages go through its own inbox; entry
//   methods and invoke run on every thread of the launch, and the payload
//   buffers come from malloc_user.
//
// argv: [min_size [max_size [iters [warmup [check]]]]]
#include <hip/hip_runtime.h>
#include <stdio.h>

#include <charming.h>
#include <charming_chrono.h>

#ifdef SM_LEVEL
#define PP_TID (threadIdx.x)
#define PP_NT (blockDim.x)
#define BARRIER_LOCAL __syncthreads()
#else
#define PP_TID (GID)
#define PP_NT (gridDim.x * blockDim.x)
#define BARRIER_LOCAL charm::barrier_local()
#endif

struct Comm : charm::chare {
  int index, peer, init_cnt;
  size_t min_size, max_size, cur_size;
  int n_iters, warmup, iter, check;
  int seq;             // round trips started (chare 0) / answered (chare 1)
  int finished;
  unsigned long long errors;
  char* data;
  size_t data_off;     // GPU-level: offset of data in the user arena
  charm::chrono::time_point<charm::chrono::system_clock> start_tp;

  __device__ Comm() {}
  __device__ void init(void* arg);
  __device__ void init_done(void* arg);
  __device__ void send();
  __device__ void recv(void* arg);
};

__device__ void entry_init(Comm& c, void* arg) { c.init(arg); }
__device__ void entry_init_done(Comm& c, void* arg) { c.init_done(arg); }
__device__ void entry_recv(Comm& c, void* arg) { c.recv(arg); }

#ifdef SM_LEVEL
__shared__ charm::chare_proxy<Comm>* comm_proxy;
__device__ char* g_pp_buffers[2];  // set by main_host

static size_t g_host_max = 1 << 20;

void charm::main_host(int argc, char** argv) {
  if (argc >= 3) g_host_max = strtoull(argv[2], nullptr, 10);
  char* bufs[2] = {nullptr, nullptr};
  for (int i = 0; i < 2; i++)
    if (hipMalloc(&bufs[i], g_host_max + 64) != hipSuccess) bufs[i] = nullptr;
  (void)hipMemcpyToSymbol(HIP_SYMBOL(g_pp_buffers), bufs, sizeof(bufs));
}
#else
__device__ charm::chare_proxy<Comm>* comm_proxy;
__device__ size_t g_params[5];
__device__ unsigned long long g_bad;
#endif

// Byte k of a size-s payload travelling in direction dir.
__device__ inline unsigned char pattern(size_t k, size_t s, int dir) {
  unsigned long long z = (k + 1) * 0x9E3779B97F4A7C15ull ^ (s * 0xBF58476D1CE4E5B9ull) ^
                         (unsigned long long)(dir + 1);
  z ^= z >> 29;
  return (unsigned char)(z * 0x94D049BB133111EBull >> 56);
}

__device__ void fill(char* p, size_t s, int dir) {
  for (size_t k = PP_TID; k < s; k += PP_NT) p[k] = (char)pattern(k, s, dir);
  BARRIER_LOCAL;
}

// Mismatched bytes of a received payload (collective).
__device__ unsigned long long verify(const char* p, size_t s, int dir) {
  unsigned long long mine = 0;
  for (size_t k = PP_TID; k < s; k += PP_NT)
    mine += ((unsigned char)p[k] != pattern(k, s, dir));
#ifdef SM_LEVEL
  __shared__ unsigned long long bad;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  if (mine) atomicAdd(&bad, mine);
  __syncthreads();
  return bad;
#else
  if (GID == 0) g_bad = 0;
  BARRIER_LOCAL;
  if (mine) atomicAdd(&g_bad, mine);
  BARRIER_LOCAL;
  return __hip_atomic_load(&g_bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#endif
}

__device__ void charm::main(int argc, char** argv, size_t* argvs, int pe) {
  if (PP_TID == 0) {
    comm_proxy = new charm::chare_proxy<Comm>();
    comm_proxy->add_entry_method<&entry_init>();
    comm_proxy->add_entry_method<&entry_init_done>();
    comm_proxy->add_entry_method<&entry_recv>();
    const int P = charm::n_pes();
    int* block_map = new int[P];
    for (int i = 0; i < P; i++) block_map[i] = 0;
    block_map[0] += 1;
    block_map[P / 2] += 1;
    comm_proxy->create(2, block_map);
    delete[] block_map;
  }
  BARRIER_LOCAL;
  barrier();
  if (pe == 0) {
#ifdef SM_LEVEL
    __shared__ size_t params[5];
#else
    size_t* params = g_params;
#endif
    if (PP_TID == 0) {
      params[0] = 8;
      params[1] = 1048576;
      params[2] = 1000;
      params[3] = 10;
      params[4] = 1;
      for (int a = 1; a <= 5 && a < argc; a++)
        params[a - 1] = charm::device_atoi(argv[a], argvs[a]);
      printf("Pingpong\nPEs: %d\nSize: %llu - %llu\nIterations: %d (Warmup: %d)\n",
             charm::n_pes(), (unsigned long long)params[0],
             (unsigned long long)params[1], (int)params[2], (int)params[3]);
    }
    BARRIER_LOCAL;
    comm_proxy->invoke_all(0, params, sizeof(size_t) * 5);
  }
}

__device__ void Comm::init(void* arg) {
  if (PP_TID == 0) {
    const size_t* p = (const size_t*)arg;
    min_size = p[0];
    max_size = p[1];
    n_iters = (int)p[2];
    warmup = (int)p[3];
    check = (int)p[4];
    cur_size = min_size;
    iter = 0;
    seq = 0;
    errors = 0;
    init_cnt = 0;
    finished = 0;
    index = i;
    peer = (index == 0) ? 1 : 0;
#ifdef SM_LEVEL
    data = g_pp_buffers[index];
#else
    data = (char*)charm::malloc_user(max_size, data_off);
    if (!data) charm::abort();
#endif
    comm_proxy->set_refnum(index, 0);
  }
  BARRIER_LOCAL;
  comm_proxy->invoke(0, 1);
}

__device__ void Comm::init_done(void* arg) {
  if (PP_TID == 0) init_cnt++;
  BARRIER_LOCAL;
  if (init_cnt == 2) send();
}

__device__ void Comm::send() {
  // only chare 0 starts round trips
  if (iter == 0) fill(data, cur_size, 0);
  if (PP_TID == 0 && iter == warmup)
    start_tp = charm::chrono::system_clock::now();
  BARRIER_LOCAL;
  comm_proxy->invoke(peer, 2, data, cur_size, seq);
}

__device__ void Comm::recv(void* arg) {
  const size_t sz = cur_size;  // size of the message being handled
  const int my_seq = seq;
  const int dir_in = (index == 0) ? 1 : 0;
  BARRIER_LOCAL;
  const unsigned long long bad =
      check ? verify((const char*)arg, sz, dir_in) : 0;
  if (index == 1) {
    // echo side: answer with this chare's own seeded buffer of the same size
    if (iter == 0) fill(data, sz, 1);
    if (PP_TID == 0) {
      errors += bad;
      seq++;
      comm_proxy->set_refnum(index, seq);
      finished = 0;
      if (++iter == n_iters + warmup) {
        iter = 0;
        cur_size *= 2;
        if (cur_size > max_size) finished = 1;
      }
    }
    BARRIER_LOCAL;
    comm_proxy->invoke(peer, 2, data, sz, my_seq);
    if (finished && PP_TID == 0) {
      printf("Pingpong peer done: payload errors %llu\n", errors);
#ifndef SM_LEVEL
      charm::free_user(max_size, data_off);
#endif
    }
    return;
  }
  if (PP_TID == 0) {
    errors += bad;
    finished = 0;
    seq++;
    comm_proxy->set_refnum(index, seq);
    if (++iter == n_iters + warmup) {
      auto end_tp = charm::chrono::system_clock::now();
      charm::chrono::duration<double> diff = end_tp - start_tp;
      printf("Size %llu took %.3lf us\n", (unsigned long long)sz,
             (diff.count() / 2 / n_iters) * 1000000);
      cur_size *= 2;
      iter = 0;
      if (cur_size > max_size) finished = 1;
    }
  }
  BARRIER_LOCAL;
  if (finished) {
    if (PP_TID == 0) {
      printf("Pingpong done: round trips %d, payload errors %llu\n", seq,
             errors);
#ifndef SM_LEVEL
      charm::free_user(max_size, data_off);
#endif
    }
    charm::end();
  } else {
    send();
  }
}

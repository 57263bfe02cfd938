// Hello chain: 2 x n_pes chares pass an incrementing integer along the chain
// and the last one ends the run.  Same program as the reference's
// examples/hello (hello.cu:6-52), including its buffer contract: only the
// first thread allocates the integer it sends (a device `new`), every other
// thread passes a null pointer, and invoke copies from the first thread's
// buffer (get.cu:866-881).  Delivery count and order are checked by reading
// the printed lines.
//
// Built in both modes: one PE per workgroup (default), and GPU-level
// (CHARMING_GPU_LEVEL: one PE spanning every workgroup, so on one GPU the
// two chares of the one PE pass the integer through its own inbox, and GID,
// a __device__ proxy and grid barriers replace threadIdx.x, the __shared__
// proxy and __syncthreads, as in the reference's GPU-level programs).
#include <stdio.h>

#include <charming.h>

struct Hello : charm::chare {
  __device__ Hello() {}
  __device__ void greet(void* arg);
};

__device__ void entry_greet(Hello& h, void* arg) { h.greet(arg); }

#ifdef SM_LEVEL
#define HELLO_TID (threadIdx.x)
#define BARRIER_LOCAL __syncthreads()
__shared__ charm::chare_proxy<Hello>* hello_proxy;
#else
#define HELLO_TID (GID)
#define BARRIER_LOCAL charm::barrier_local()
__device__ charm::chare_proxy<Hello>* hello_proxy;
#endif

__device__ void charm::main(int argc, char** argv, size_t* argvs, int pe) {
  if (HELLO_TID == 0) {
    hello_proxy = new charm::chare_proxy<Hello>();
    hello_proxy->add_entry_method<&entry_greet>();
    hello_proxy->create(charm::n_pes() * 2);
  }
  BARRIER_LOCAL;

  barrier();

  if (pe == 0) {
    int* send_int_p = nullptr;
    if (HELLO_TID == 0) {
      send_int_p = new int;
      *send_int_p = 0;
    }
    BARRIER_LOCAL;
    hello_proxy->invoke(0, 0, send_int_p, sizeof(int));
  }
}

__device__ void Hello::greet(void* arg) {
  const int recv_int = *(int*)arg;
  int* send_int_p = nullptr;
  if (HELLO_TID == 0)
    printf("Hello I'm %d of %d! Received %d\n", i, n, recv_int);
  if (i == n - 1) {
    charm::end();
  } else {
    if (HELLO_TID == 0) {
      send_int_p = new int;
      *send_int_p = recv_int + 1;
    }
    BARRIER_LOCAL;
    hello_proxy->invoke(i + 1, 0, send_int_p, sizeof(int));
  }
}

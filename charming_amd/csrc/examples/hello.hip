// Hello chain: 2 x n_pes chares pass an incrementing integer along the chain
// and the last one ends the run.  Same program as the reference's
// examples/hello (hello.cu:6-52), including its buffer contract: only
// thread 0 allocates the integer it sends (a device `new`), every other
// thread passes a null pointer, and invoke copies from thread 0's buffer
// (get.cu:866-881).  Delivery count and order are checked by reading the
// printed lines.
#include <stdio.h>

#include <charming.h>

struct Hello : charm::chare {
  __device__ Hello() {}
  __device__ void greet(void* arg);
};

__device__ void entry_greet(Hello& h, void* arg) { h.greet(arg); }

__shared__ charm::chare_proxy<Hello>* hello_proxy;

__device__ void charm::main(int argc, char** argv, size_t* argvs, int pe) {
  if (threadIdx.x == 0) {
    hello_proxy = new charm::chare_proxy<Hello>();
    hello_proxy->add_entry_method<&entry_greet>();
    hello_proxy->create(charm::n_pes() * 2);
  }
  __syncthreads();

  barrier();

  if (pe == 0) {
    int* send_int_p = nullptr;
    if (threadIdx.x == 0) {
      send_int_p = new int;
      *send_int_p = 0;
    }
    __syncthreads();
    hello_proxy->invoke(0, 0, send_int_p, sizeof(int));
  }
}

__device__ void Hello::greet(void* arg) {
  const int recv_int = *(int*)arg;
  int* send_int_p = nullptr;
  if (threadIdx.x == 0)
    printf("Hello I'm %d of %d! Received %d\n", i, n, recv_int);
  if (i == n - 1) {
    charm::end();
  } else {
    if (threadIdx.x == 0) {
      send_int_p = new int;
      *send_int_p = recv_int + 1;
    }
    __syncthreads();
    hello_proxy->invoke(i + 1, 0, send_int_p, sizeof(int));
  }
}

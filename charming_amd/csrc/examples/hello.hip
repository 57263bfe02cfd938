// Hello chain: 2 x n_pes chares pass an incrementing integer along the chain
// and the last one ends the run.  Same program shape as the reference's
// examples/hello (hello.cu:6-52): delivery count and order are checked by
// reading the printed lines.
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
    __shared__ int first;
    if (threadIdx.x == 0) first = 0;
    __syncthreads();
    hello_proxy->invoke(0, 0, &first, sizeof(int));
  }
}

__device__ void Hello::greet(void* arg) {
  const int got = *(int*)arg;
  if (threadIdx.x == 0)
    printf("Hello I'm %d of %d! Received %d\n", i, n, got);
  if (i == n - 1) {
    charm::end();
  } else {
    __shared__ int next;
    if (threadIdx.x == 0) next = got + 1;
    __syncthreads();
    hello_proxy->invoke(i + 1, 0, &next, sizeof(int));
  }
}

// Refnum matching: chare 1 (refnum 1) first receives a message tagged 0,
// which must be deferred, then one tagged 1, which runs and moves the
// chare's refnum to 0, reviving the deferred message.  Program shape of
// the reference's examples/refnum (refnum.cu:6-53), with the send order
// chosen so deferral is certain under FIFO delivery, and with termination.
#include <stdio.h>

#include <charming.h>

struct Comm : charm::chare {
  int received;
  __device__ Comm() : received(0) {}
  __device__ void send(void* arg);
  __device__ void recv(void* arg);
};

__device__ void entry_send(Comm& c, void* arg) { c.send(arg); }
__device__ void entry_recv(Comm& c, void* arg) { c.recv(arg); }

__shared__ charm::chare_proxy<Comm>* comm_proxy;

__device__ void charm::main(int argc, char** argv, size_t* argvs, int pe) {
  if (threadIdx.x == 0) {
    comm_proxy = new charm::chare_proxy<Comm>();
    comm_proxy->add_entry_method<&entry_send>();
    comm_proxy->add_entry_method<&entry_recv>();
    const int P = charm::n_pes();
    int* block_map = new int[P];
    for (int i = 0; i < P; i++) block_map[i] = 0;
    block_map[0] += 1;
    block_map[P / 2] += 1;
    comm_proxy->create(2, block_map);
    delete[] block_map;
    if (pe == 0) comm_proxy->set_refnum(0, 0);
    if (pe == P / 2) comm_proxy->set_refnum(1, 1);
  }
  __syncthreads();
  barrier();
  if (pe == 0) comm_proxy->invoke(0, 0, nullptr, 0, 0);
}

__device__ void Comm::send(void* arg) {
  __shared__ int tag;
  if (threadIdx.x == 0) tag = 0;
  __syncthreads();
  comm_proxy->invoke(1, 1, &tag, sizeof(int), 0);  // deferred: refnum is 1
  if (threadIdx.x == 0) tag = 1;
  __syncthreads();
  comm_proxy->invoke(1, 1, &tag, sizeof(int), 1);  // runs, then revives 0
}

__device__ void Comm::recv(void* arg) {
  const int tag = *(int*)arg;
  if (threadIdx.x == 0) {
    printf("Received message with refnum %d\n", tag);
    received++;
    if (tag == 1) comm_proxy->set_refnum(1, 0);
  }
  __syncthreads();
  if (received == 2) charm::end();
}

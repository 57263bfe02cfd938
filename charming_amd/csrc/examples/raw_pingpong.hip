// The latency floor the runtime pingpong is compared to: two workgroups
// exchange a flag through per-PE mailboxes in HBM, with no runtime at all
// (examples/baseline/baseline.cu:14-160 measures the same on an SM pair).
//
// Mailbox layout as in the reference: PE p's slots are mbox[p][src][k],
// kSlots per sender.  send() claims a free slot of (dst, me) and writes a
// non-zero word with an agent-scope compare-and-swap; recv() scans all of
// its slots, takes the lowest occupied one (a workgroup minimum) and clears
// it.  No payload moves, so the round trip is the poll, the atomics and the
// coherence path between the two CUs.  Every wait is bounded: a message that
// never arrives ends the run with an error instead of a hang.
//
// usage: raw_pingpong [iters=1000] [warmup=10]
// prints one line per placement: the peer on another XCD (PE 1) and on the
// same XCD (PE n/2, as the runtime pingpong's PEs 0 and n/2).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

namespace {

constexpr int kSlots = 4;                  // per sender, as LOCAL_MAX
constexpr uint64_t kSpinLimit = 1ull << 20;  // polls before giving up

__device__ inline uint64_t ld(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline bool cas(uint64_t* p, uint64_t expect, uint64_t v) {
  return __hip_atomic_compare_exchange_strong(
      p, &expect, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
      __HIP_MEMORY_SCOPE_AGENT);
}

// Workgroup-collective: claim a free slot of (dst, me) and publish the
// message word there.  Returns false when the spin limit ran out.
__device__ bool send(uint64_t* mbox, int n_pes, int me, int dst,
                     uint64_t word) {
  uint64_t* slots = mbox + ((size_t)dst * n_pes + me) * kSlots;
  __shared__ int ok;
  if (threadIdx.x == 0) {
    ok = 0;
    for (uint64_t spin = 0; spin < kSpinLimit && !ok; spin++)
      for (int k = 0; k < kSlots && !ok; k++)
        if (ld(&slots[k]) == 0 && cas(&slots[k], 0, word)) ok = 1;
  }
  __syncthreads();
  return ok != 0;
}

// Workgroup-collective: one scan of my slots; take the lowest occupied one.
// Returns its word (0: none).
__device__ uint64_t recv_once(uint64_t* mbox, int n_pes, int me) {
  uint64_t* slots = mbox + (size_t)me * n_pes * kSlots;
  __shared__ int idx;
  __shared__ uint64_t word;
  if (threadIdx.x == 0) {
    idx = 0x7fffffff;
    word = 0;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n_pes * kSlots; i += blockDim.x)
    if (ld(&slots[i]) != 0) atomicMin(&idx, i);
  __syncthreads();
  if (threadIdx.x == 0 && idx != 0x7fffffff) {
    word = ld(&slots[idx]);
    cas(&slots[idx], word, 0);
  }
  __syncthreads();
  return word;
}

// out[0] = elapsed wall-clock ticks over the timed round trips (PE 0),
// out[1] = errors (lost messages, wrong words).
__global__ void pingpong_kernel(uint64_t* mbox, int n_pes, int peer,
                                int iters, int warmup, uint64_t* out) {
  const int me = blockIdx.x;
  if (me != 0 && me != peer) return;
  const int other = me == 0 ? peer : 0;
  uint64_t t0 = 0, errors = 0;
  for (int i = 0; i < iters + warmup && errors == 0; i++) {
    if (i == warmup) t0 = wall_clock64();
    const uint64_t word = (uint64_t)i + 1;  // never 0
    if (me == 0) {
      if (!send(mbox, n_pes, me, other, word)) errors++;
    }
    uint64_t got = 0;
    for (uint64_t spin = 0; got == 0 && spin < kSpinLimit; spin++)
      got = recv_once(mbox, n_pes, me);
    if (got != word) errors++;
    if (me != 0 && errors == 0 && !send(mbox, n_pes, me, other, word))
      errors++;
  }
  if (threadIdx.x == 0) {
    if (me == 0) out[0] = wall_clock64() - t0;
    if (errors) atomicAdd((unsigned long long*)&out[1], errors);
  }
}

#define CHECK(x)                                                         \
  do {                                                                   \
    hipError_t e_ = (x);                                                 \
    if (e_ != hipSuccess) {                                              \
      fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));            \
      exit(1);                                                           \
    }                                                                    \
  } while (0)

}  // namespace

int main(int argc, char** argv) {
  const int iters = argc > 1 ? atoi(argv[1]) : 1000;
  const int warmup = argc > 2 ? atoi(argv[2]) : 10;
  if (iters < 1 || warmup < 0) {
    fprintf(stderr, "usage: raw_pingpong [iters >= 1] [warmup >= 0]\n");
    return 2;
  }
  int dev = 0, n_cu = 0, khz = 0;
  CHECK(hipGetDevice(&dev));
  CHECK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount,
                              dev));
  CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  const int n_pes = n_cu;  // one workgroup per CU, all resident together
  uint64_t *mbox = nullptr, *out = nullptr;
  const size_t mbox_bytes = (size_t)n_pes * n_pes * kSlots * sizeof(uint64_t);
  CHECK(hipMalloc(&mbox, mbox_bytes));
  CHECK(hipMalloc(&out, 2 * sizeof(uint64_t)));
  int status = 0;
  const int peers[2] = {1, n_pes / 2};
  for (int p = 0; p < 2; p++) {
    CHECK(hipMemset(mbox, 0, mbox_bytes));
    CHECK(hipMemset(out, 0, 2 * sizeof(uint64_t)));
    pingpong_kernel<<<n_pes, 256>>>(mbox, n_pes, peers[p], iters, warmup, out);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    uint64_t h[2];
    CHECK(hipMemcpy(h, out, sizeof(h), hipMemcpyDeviceToHost));
    const double us = (double)h[0] / (khz * 1e-3) / 2.0 / iters;
    printf("raw mailbox pingpong: PEs 0 and %d of %d, %d round trips: "
           "%.3f us per one-way trip, errors %llu\n",
           peers[p], n_pes, iters, us, (unsigned long long)h[1]);
    if (h[1]) status = 1;
  }
  CHECK(hipFree(mbox));
  CHECK(hipFree(out));
  return status;
}

// Public runtime API of the charm:: interface (reference include/charming.h:
// 8-37): user hooks main_host/main, end/abort, PE queries, barriers and
// device_atoi.  Including this header in a chare program's translation unit
// also instantiates the persistent scheduler kernel there and registers it
// with the host library (libcharming_amd.so) at load time.
#ifndef _CHARMING_H_
#define _CHARMING_H_

#include "chare.h"
#include "message.h"
#include "kernel.h"

namespace charm {

// Main function executed on the host before the launch (charming.cu:73).
// Optional: a program that does not define it gets none.  Hidden, so that a
// program library's reference resolves to its own definition (or to none)
// and never to another program library's.
__attribute__((weak, visibility("hidden"))) void main_host(int argc,
                                                            char** argv);

// Main function executed on all PEs (by every thread of each PE's workgroup,
// or of every workgroup in GPU-level mode)
__device__ void main(int argc, char** argv, size_t* argvs, int pe);

// Runtime functions that can be called by the user
__device__ void end();
__device__ void abort();

__device__ int my_pe();
__device__ int n_pes();
__device__ int my_pe_node();
__device__ int n_pes_node();
__device__ int n_nodes();

// Element-level barrier (requires all PEs to participate)
__device__ void barrier();
__device__ void barrier_local();

__device__ int device_atoi(const char* str, int strlen);

#ifndef SM_LEVEL
// GPU-level mode only (charming.h:31-34 of the reference): memory from the
// PE's user arena; single-threaded.
__device__ void* malloc_user(size_t size, size_t& offset);
__device__ void free_user(size_t size, size_t offset);
#endif

// Seconds on the device wall clock (constant rate).
__device__ double wall_time();

// Messages sent by an entry method are published together when it returns
// (one cache writeback for all of them).  flush_sends() publishes the ones
// sent so far right away, for an entry method that keeps computing after
// sending.  Workgroup-collective.  (No counterpart in the reference, whose
// sends are published one by one.)
__device__ void flush_sends();

// Ask the scheduler to call the program's work hook (CHARMING_WORK_HOOK,
// defined before including this header) with `token` once the current entry
// method returns.  The hook is called directly, so the compiler gives it the
// scheduler kernel's register budget instead of an entry method's.  (No
// counterpart in the reference.)
__device__ void request_work(void* token);

}  // namespace charm

#ifndef CHARMING_NO_RUNTIME
#include "charming_detail/device_rt.h"
#endif

#endif  // _CHARMING_H_

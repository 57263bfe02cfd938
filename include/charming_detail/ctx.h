// Runtime context shared by the host library (libcharming_amd.so) and the
// device runtime instantiated in the user's translation unit.  Plain data
// only: the host fills it, copies it to a __constant__ symbol of the user TU
// and every workgroup reads it with scalar loads.
//
// Memory plan in HBM (all one-GPU, sized for 288 GB):
//   rings      n_pes x (ring_bytes + max_msg_bytes)  inbox byte rings; a message
//                                                    that starts before the end
//                                                    runs into the overflow tail
//                                                    so payloads stay contiguous
//   ready      n_pes x (ring_bytes / 64) u64          publish words, one per
//                                                    64-byte granule: tag =
//                                                    logical granule index + 1
//   tails      n_pes x 128 B                          producers' reservation
//                                                    counter (fetch-add)
//   heads      n_pes x 128 B                          consumer release counter
//   pes        n_pes x PeState                        owner-only state
//   deferred   n_pes x deferred_bytes                 refnum-mismatched copies
//   heap       n_pes x heap_bytes                     chare objects, tables
#ifndef CHARMING_DETAIL_CTX_H_
#define CHARMING_DETAIL_CTX_H_

#include <stddef.h>
#include <stdint.h>

#define CHARE_TYPE_CNT_MAX 1024  // chare.h:9 of the reference
#define EM_CNT_MAX 128           // chare.h:10
#define ASYNC_WAIT_MAX 16        // chare.h:11
#define MISMATCH_MAX 128         // chare.h:12 (per proxy per PE there)
// Deferred-message slots per PE here (all proxies together).
#ifndef CHARMING_DEFERRED_MAX
#define CHARMING_DEFERRED_MAX 1024
#endif

namespace charm {
struct chare_proxy_base;

namespace rt {

constexpr uint64_t kGranule = 64;  // ring allocation granule (bytes)

// device error codes (first one wins, see Ctx::ctl)
enum : uint32_t {
  ERR_NONE = 0,
  ERR_MSG_TOO_LARGE = 1,
  ERR_SELF_RING_FULL = 2,
  ERR_DEFERRED_FULL = 3,
  ERR_BAD_MESSAGE = 4,
  ERR_HEAP_FULL = 5,
  ERR_HOST_ABORT = 6,
  ERR_USER_ABORT = 7,
  ERR_SPIN_LIMIT = 8,
  ERR_BAD_CHARE = 9,
  ERR_UNSUPPORTED = 10,
  ERR_WAITS_FULL = 11,
};

// control words: each on its own 128-byte line
enum : int {
  CTL_TERM = 0,        // termination flag (charm::end)
  CTL_ERROR = 32,      // first error code
  CTL_ERROR_PE = 33,   // PE that raised it
  CTL_ERROR_AUX = 34,  // extra detail
  CTL_BAR_COUNT = 64,  // grid barrier arrivals
  CTL_BAR_GEN = 96,    // grid barrier generation
  CTL_WORDS = 128
};

struct DeferredEntry {
  uint32_t offset;  // byte offset in the PE's deferred arena
  uint32_t bytes;   // envelope alloc size
  int chare_id;
  int chare_idx;
  int refnum;
  int state;        // 0 free, 1 deferred, 2 revived
  uint32_t seq;     // arrival order
  uint32_t pad;
};

struct ProxyTable {
  chare_proxy_base* proxies[CHARE_TYPE_CNT_MAX];
  int count;
};

struct Stats {
  unsigned long long delivered;  // entry methods run (incl. revived)
  unsigned long long sent;
  unsigned long long deferred;
  unsigned long long revived;
  unsigned long long send_spins;  // polls while a ring was full
  unsigned long long idle_polls;
  unsigned long long overhead_ticks;  // wall-clock ticks in dispatch around
                                      // entry methods (decode, acquire, release)
  unsigned long long pad;
};

// A pending async_wait (reference chare.h:41-49): when every ivars[k],
// k < nelems, compares true against cmp_value, entry method ep of local
// chare idx of proxy `proxy` runs once with a null argument.
struct AsyncWait {
  uint64_t* ivars;
  uint64_t nelems;
  uint64_t cmp_value;
  int cmp, proxy, idx, ep;
  int valid;
};

// A free extent of the PE's user arena (malloc_user / free_user).
struct UserExtent {
  uint64_t offset;
  uint64_t size;
};
constexpr int kUserExtents = 64;

struct PeState {
  ProxyTable table;
  int n_user_free;     // free extents of the user arena, sorted by offset
  UserExtent user_free[kUserExtents];
  int n_waits;         // valid entries of waits[]
  AsyncWait waits[ASYNC_WAIT_MAX];
  uint64_t read_pos;   // logical position of the next inbox message
  uint64_t heap_used;  // bump allocator in the PE's heap slice
  uint32_t arena_used; // deferred arena bump pointer
  uint32_t seq;
  int n_deferred;      // entries in state 1 or 2
  int n_revived;       // entries in state 2
  int ended;           // charm::end() already called on this PE
  int hw;              // deferred entries in use lie below this index
  DeferredEntry deferred[CHARMING_DEFERRED_MAX];
  Stats stats;
};

struct Ctx {
  int n_pes;
  int threads_per_pe;
  uint64_t ring_bytes;      // power of two
  uint64_t ring_stride;     // ring_bytes + max_msg_bytes
  uint64_t max_msg_bytes;   // largest message slot (multiple of kGranule)
  char* rings;
  uint64_t* ready;          // n_pes * (ring_bytes / kGranule)
  uint64_t* tails;          // stride 16 u64 per PE
  uint64_t* heads;          // stride 16 u64 per PE
  PeState* pes;
  char* deferred;           // n_pes * deferred_bytes
  uint64_t deferred_bytes;
  char* heap;               // n_pes * heap_bytes
  uint64_t heap_bytes;
  uint32_t* ctl;            // CTL_WORDS control words
  volatile uint32_t* host_abort;  // host-coherent word set by the watchdog
  uint64_t spin_limit;      // 0 = unbounded waits
  int argc;
  char** argv;
  size_t* argvs;
  void* user;               // opaque pointer for the application
  double wall_hz;           // frequency of the device wall clock
  // GPU-level mode (one PE spanning every workgroup of the launch)
  uint64_t* bcast;          // leader -> all workgroups: decisions of the
                            // scheduler and of sends (kBcastWords words)
  char* user_arena;         // n_pes x user_bytes (malloc_user)
  uint64_t user_bytes;
};

constexpr int kBcastWords = 32;

}  // namespace rt
}  // namespace charm

#endif  // CHARMING_DETAIL_CTX_H_

// Message wire format and send API of the charm:: interface.
//
// Same layout as the reference (include/message.h:5-114): a 16-byte
// envelope {type, size}, a 16-byte regular_msg header {chare_id, chare_idx,
// ep_id, refnum}, then the payload; alloc_size = round16(16 + header +
// payload).  In the inbox ring each message occupies alloc_size rounded up
// to the 64-byte ring granule.
#ifndef _MESSAGE_H_
#define _MESSAGE_H_

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

// Alignment of every message and header (message.h:5)
#define ALIGN_SIZE 16

namespace charm {

// Message types (message.h:10-17)
enum class msgtype : int {
  regular,
  request,
  forward,
  user,
  begin_terminate,
  do_terminate
};

// Regular message header between chares (message.h:20-28)
struct alignas(ALIGN_SIZE) regular_msg {
  int chare_id;
  int chare_idx;
  int ep_id;
  int refnum;

  __device__ regular_msg(int chare_id_, int chare_idx_, int ep_id_, int refnum_)
      : chare_id(chare_id_), chare_idx(chare_idx_), ep_id(ep_id_),
        refnum(refnum_) {}
};

// Request sent from PE to CE (message.h:31-41).  Kept for the wire-size
// table; this one-GPU runtime has no CEs and never sends it.
struct alignas(ALIGN_SIZE) request_msg : regular_msg {
  msgtype type;
  void* buf;
  size_t payload_size;
  int dst_pe;

  __device__ request_msg(int chare_id_, int chare_idx_, int ep_id_,
                         msgtype type_, void* buf_, size_t payload_size_,
                         int dst_pe_, int refnum_)
      : regular_msg(chare_id_, chare_idx_, ep_id_, refnum_), type(type_),
        buf(buf_), payload_size(payload_size_), dst_pe(dst_pe_) {}
};

// Message forwarded between CEs (message.h:44-50); wire-size table only.
struct alignas(ALIGN_SIZE) forward_msg : regular_msg {
  int dst_pe;

  __device__ forward_msg(int chare_id_, int chare_idx_, int ep_id_,
                         int dst_pe_, int refnum_)
      : regular_msg(chare_id_, chare_idx_, ep_id_, refnum_), dst_pe(dst_pe_) {}
};

struct alignas(ALIGN_SIZE) envelope {
  msgtype type;
  size_t size;  // alloc_size of the whole message

  __device__ envelope(msgtype type_, size_t size_) : type(type_), size(size_) {}

  // Total size of a message for allocation (message.h:61-81)
  __host__ __device__ inline static size_t alloc_size(msgtype type,
                                                      size_t payload_size) {
    size_t type_size = 0;
    if (type == msgtype::regular || type == msgtype::user) {
      type_size += sizeof(regular_msg);
    } else if (type == msgtype::request) {
      type_size += sizeof(request_msg);
    } else if (type == msgtype::forward) {
      type_size += sizeof(forward_msg);
    }
    size_t s = sizeof(envelope) + type_size + payload_size;
    size_t rem = s % ALIGN_SIZE;
    if (rem != 0) s = s + ALIGN_SIZE - rem;
    return s;
  }
};

static_assert(sizeof(envelope) == 16, "envelope must stay 16 bytes");
static_assert(sizeof(regular_msg) == 16, "regular_msg must stay 16 bytes");

// User message handle (message.h:85-95).  The reference sketches a
// user-message API and leaves it commented out (message.h:84-95, 113-118,
// get.cu:1362-1398); here it is the zero-copy send: the message is
// allocated directly in the destination PE's inbox ring, the workgroup
// writes the payload in place, and send_user_msg publishes it.  Between
// alloc and send the destination consumes nothing behind this message, so
// the sender must not wait on other PEs in between.
struct alignas(ALIGN_SIZE) message {
  envelope* env;  // the message in the destination ring (nullptr: none)
  size_t offset;  // its position in that ring
  int dst_pe;

  __device__ message() : env(nullptr), offset(0), dst_pe(-1) {}
  // payload of the allocated message
  __device__ void* payload() const {
    return env ? (char*)env + sizeof(envelope) + sizeof(regular_msg) : nullptr;
  }
};

// Send API (message.h:97-114).  All are workgroup-collective: every thread
// of the calling PE's workgroup must call them together.
__device__ void send_reg_msg(int chare_id, int chare_idx, int ep_id, void* buf,
                             size_t payload_size, int dst_pe, int refnum);
__device__ void send_begin_term_msg();
// Zero-copy send (user messages): reserve `payload_size` bytes for entry
// method ep_id of chare chare_idx (proxy chare_id) in dst_pe's inbox and
// write the headers; returns the payload pointer (nullptr when the runtime
// is stopping).  The workgroup fills the payload, then send_user_msg
// publishes it (batched with the entry method's other sends).
__device__ void* alloc_user_msg(message& msg, int chare_id, int chare_idx,
                                int ep_id, size_t payload_size, int dst_pe,
                                int refnum);
__device__ void send_user_msg(message& msg);
__device__ void revive_mismatches(int chare_id, int chare_idx, int refnum);

}  // namespace charm

#endif  // _MESSAGE_H_

// Shared (host + device) definitions of the Jacobi2D task engine: the
// batched and served forms of the solver.
//
// Many independent Jacobi2D problems share one persistent launch.  Each
// problem is cut into square chare tiles (side chosen per problem); one task
// is one round of one tile: two sweeps fused in one pass (temporal
// blocking), or one sweep when a single one is left.  A tile's completion is
// the message that releases its neighbours: per tile and per round parity, a
// counter of unmet dependencies (the tile's own previous round, each of its
// up to eight neighbours' previous round - a fused round reads two cells
// beyond the tile, corners included - and, when the problem stops on a
// residual, the stop decision two rounds back) is decremented by every
// producer; whoever takes it to zero enqueues the task.  Halos are read in
// place from the neighbours' cells (the zero-copy halo mode: the message
// carries no payload).  Ready tasks go to the completing workgroup's own
// queue; idle workgroups steal from the others.
#pragma once
#include <stdint.h>

namespace chm {
namespace eng {

constexpr uint32_t kMaxSlots = 4096;        // problems in flight
constexpr uint32_t kMaxTiles = 1u << 18;    // tiles per problem
// task kinds: seed a tile, one round of a tile, hash a tile of the result,
// redo the first sweep of the stop round alone (a stop after an odd sweep
// count, whose result the fused round did not store)
enum : uint32_t { K_INIT = 0, K_ITER = 1, K_HASH = 2, K_FIX = 3 };

__host__ __device__ inline uint32_t task_word(uint32_t kind, uint32_t slot,
                                              uint32_t tile) {
  return (kind << 30) | (slot << 18) | tile;
}
__host__ __device__ inline uint32_t task_kind(uint32_t w) { return w >> 30; }
__host__ __device__ inline uint32_t task_slot(uint32_t w) {
  return (w >> 18) & (kMaxSlots - 1);
}
__host__ __device__ inline uint32_t task_tile(uint32_t w) {
  return w & (kMaxTiles - 1);
}

// A request as the host publishes it.  Three arrays with halos two deep
// (rows -1 .. H+2), pointers at row 0.  Round r reads array r % 3 and
// writes array (r + 1) % 3, so the input of round r survives round r+1:
// a stop decided on round r while round r+1 runs can still redo a sweep
// from it.
struct Request {
  double* a;
  double* b;
  double* c;
  uint32_t* pend;     // 3 words per tile: dependency counters of the next
                      // even and odd round, rounds done
  int64_t pitch;
  int W, H, bs, tr, tc, max_iters;
  int init_on_device; // 1: the engine seeds the grids (init tasks)
  int init_mode;
  double boundary;
  double tol2;        // < 0: run exactly max_iters sweeps
  uint64_t seed;
  uint64_t ticket;
  uint32_t slot;
  int rounds;         // ceil(max_iters / depth)
  int depth;          // sweeps per round = halo depth of the arrays (rows
                      // 1-depth .. H+depth): 2 with a tolerance, else 8
};

// Device state of a problem slot (HBM, written only by the engine).
struct Prob {
  Request r;
  uint32_t inflight;      // tasks enqueued and not finished
  uint32_t state;         // 0 running, 1 stopped, 4 redoing sweeps, 2
                          // hashing, 5 recomputing a round's residuals
  uint32_t stop_iters;    // sweeps in the final result
  uint32_t fix_round;     // round whose first fix_sweeps sweeps are redone
                          // (~0u: none)
  uint32_t round_tiles[2];  // tiles that finished round r (slot r & 1)
  uint32_t fix_sweeps;
  uint32_t pad0;
  uint64_t res[2][16][2];  // [r & 1][sweep j of round r][lo, hi] residual sums
  double resd[2][16];      // [r & 1][sweep j]: FP64 sums of squared changes
                          // (the stop filter of rounds between the first
                          // and the last)
  uint32_t resid_round;   // round recomputed with exact residuals (state 5)
  uint32_t resid_tiles;   // tiles done with that
  uint32_t held_tile;     // tile whose release of round resid_round+1 waits
  uint32_t pad1;
  uint64_t s0[2];         // residual of the first sweep
  uint64_t s_last[2];     // residual of the last sweep of the result
  uint64_t hash;
  uint64_t t_admit;       // device wall clock
};

// A completion record (host-coherent memory); `valid` = ticket + 1 is
// written last.
struct Completion {
  uint64_t ticket;
  uint64_t hash;
  uint64_t s_last_lo, s_last_hi;
  uint64_t s0_lo, s0_hi;
  uint64_t t_admit, t_done;  // device wall-clock ticks
  uint32_t iters;
  uint32_t slot;
  uint64_t valid;
};

// Device control words, one 128-byte line each.
enum : int {
  CW_REQ_HEAD = 0,   // requests claimed
  CW_REQ_TAIL = 16,  // requests published (mirror of the host word)
  CW_STOP = 32,      // 1: exit, 2: abort (mirror of the host word)
  CW_DONE = 48,      // completions written
  CW_ACTIVE = 64,    // problems admitted and not done
  CW_STATS = 80,     // [0] tasks [1] steals [2] idle polls [3] cell updates
  CW_WORDS = 96
};

struct EngineDesc {
  Prob* probs;              // kMaxSlots
  uint64_t* qcells;         // n_wg x qcap
  uint64_t* qhead;          // n_wg, stride 16
  uint64_t* qtail;          // n_wg, stride 16
  uint32_t qcap;            // power of two
  int n_wg;
  const Request* reqs;      // request ring (req_cap entries)
  uint32_t req_cap;         // power of two
  int n_pollers;            // workgroups that poll the host words
  volatile uint64_t* host_tail;  // host-written request count (or null)
  volatile uint32_t* host_stop;  // host-written stop word (or null)
  Completion* dones;        // host-coherent completion ring (req_cap entries)
  uint64_t* ctl;            // CW_WORDS device control words
  uint64_t expected;        // batch mode: exit after this many completions
  int nc, u;                // sweep tile shape
};

}  // namespace eng
}  // namespace chm

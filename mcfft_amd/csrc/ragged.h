// ragged.h -- one persistent launch for a ragged batch: thousands of transforms
// of different lengths and ranks, each split into per-pass tile tasks on a
// device-side work queue.
//
// Workgroups take tasks in queue order with one atomic per task.  A task of a
// transform's second (third) pass first waits until every tile of the earlier
// passes of the same transform has been counted; the queue lists all first-pass
// tasks before any second-pass task, so every awaited tile was taken earlier by
// a resident workgroup that never waits on a later task -- the wait always ends.
// It is bounded anyway (spin_limit naps), after which a status bit is set and
// the task runs on whatever data is there (the result is then wrong, never a hang).
//
// No reference counterpart: the reference runs one size per launch sequence
// (source/mcfft_1d_device.cpp:596-678).
#pragma once
#include "fft_pass.h"

namespace mcfft {

// A pass of one transform.  args.in / args.out hold BYTE OFFSETS from the
// buffer selected by in_buf / out_buf (0 = exec input, 1 = exec output,
// 2 = workspace); the table pointers are absolute.
struct RaggedPass {
  PassArgs args;
  int in_buf, out_buf;
  int variant;     // index into the ragged kernel list
  int counter;     // per-transform tile counter
  int wait_for;    // tiles of earlier passes that must be counted first (0 = none)
  int signal;      // count this pass's tiles
};

struct RaggedTask {
  int pass;
  int tile;
};

// One launch runs tasks [task_begin, task_end).  Phased execution launches the
// kernel once per pass phase (kernel boundaries order the phases, no waits);
// with use_deps a single launch runs every phase and second-phase tasks wait on
// their transform's counter.
struct RaggedArgs {
  const char* in_base;
  char* out_base;
  char* ws_base;
  const RaggedPass* passes;
  const RaggedTask* tasks;
  int task_begin, task_end;
  int queue;            // index of this launch's queue-head counter
  int use_deps;
  unsigned* counters;   // [0..3] queue heads, [4] status, [5 + i] transform-group i
  unsigned spin_limit;
  const u32x4* aops;
};
constexpr int kRaggedCounterBase = 5;

// ragged kernel variants: X(id, S, T, Q, NW, CM, KIND); all share NW = 8 and
// keep LDS <= 64 KiB (two workgroups per CU)
#define MCFFT_RAGGED_LIST_F(X, CM, QS)     \
  X(0, 1, 1, 256, 8, CM, 0)                \
  X(1, 1, 2, 128, 8, CM, 0)                \
  X(2, 1, 4, 64, 8, CM, 0)                 \
  X(3, 1, 8, 32, 8, CM, 0)                 \
  X(4, 2, 1, 16, 8, CM, 0)                 \
  X(5, 2, 2, 8, 8, CM, 0)                  \
  X(6, 2, 4, 4, 8, CM, 0)                  \
  X(7, 2, 8, 2, 8, CM, 0)                  \
  X(8, 3, 1, 1, 8, CM, 0)                  \
  X(9, 1, 1, 32, 8, CM, 1)                 \
  X(10, 1, 2, 32, 8, CM, 1)                \
  X(11, 1, 4, 32, 8, CM, 1)                \
  X(12, 1, 8, 32, 8, CM, 1)                \
  X(13, 2, 1, 32 / QS, 8, CM, 1)           \
  X(14, 2, 2, 16 / QS, 8, CM, 1)           \
  X(15, 1, 1, 32, 8, CM, 2)                \
  X(16, 1, 2, 32, 8, CM, 2)                \
  X(17, 1, 4, 32, 8, CM, 2)                \
  X(18, 1, 8, 32, 8, CM, 2)                \
  X(19, 2, 1, 32 / QS, 8, CM, 2)           \
  X(20, 2, 2, 32 / QS, 8, CM, 2)           \
  X(21, 2, 4, 16 / QS, 8, CM, 2)           \
  X(22, 2, 8, 8 / QS, 8, CM, 2)            \
  X(23, 3, 1, 4 / QS, 8, CM, 2)            \
  X(24, 1, 1, 32, 8, CM, 3)                \
  X(25, 1, 2, 32, 8, CM, 3)                \
  X(26, 1, 4, 32, 8, CM, 3)                \
  X(27, 1, 8, 32, 8, CM, 3)                \
  X(28, 2, 1, 32 / QS, 8, CM, 3)           \
  X(29, 2, 2, 32 / QS, 8, CM, 3)

constexpr int kRaggedVariants = 30;
constexpr int kRaggedNW = 8;

template <int CM>
constexpr int ragged_max_lds() {
  int m = 0;
#define MCFFT_RAGGED_MAXLDS(ID, S, T, Q, NW, CMx, K) \
  if (PassCfg<S, T, Q, NW, CMx, K>::LDS_BYTES > m) m = PassCfg<S, T, Q, NW, CMx, K>::LDS_BYTES;
  MCFFT_RAGGED_LIST_F(MCFFT_RAGGED_MAXLDS, CM, (CM == CM_SPLIT ? 2 : 1))
#undef MCFFT_RAGGED_MAXLDS
  return m;
}
// dynamic LDS of the ragged kernel: the largest variant plus the task slot
template <int CM>
constexpr int ragged_lds_bytes() {
  return ragged_max_lds<CM>() + 16;
}

// Every variant is inlined into the task loop (an out-of-line call would keep a
// call frame in scratch memory and raise the register count to the calling
// convention's); fft_tile recomputes its index arithmetic per tile, so one
// variant's addresses are not hoisted next to another's.  The pass arguments are
// read from the descriptor in global memory through a wave-uniform pointer
// (scalar loads).
template <int S, int T, int Q, int NW, int CM, int KIND>
MCFFT_DEV void ragged_tile(const PassArgs* gpa, const void* in, void* out, int tile, AOps A) {
  PassArgs a = *(const PassArgs*)uniform_ptr(gpa);
  a.in = uniform_ptr(in);
  a.out = (void*)uniform_ptr(out);
  fft_tile<S, T, Q, NW, CM, KIND, false>(a, uniform(tile), A);  // statically scaled
}

template <int CM>
MCFFT_DEV void ragged_dispatch(int variant, const PassArgs* gpa, const void* in, void* out, int tile, const AOps& A) {
  switch (variant) {
#define MCFFT_RAGGED_CASE(ID, S, T, Q, NW, CMx, K)         \
  case ID:                                                 \
    ragged_tile<S, T, Q, NW, CMx, K>(gpa, in, out, tile, A); \
    break;
    MCFFT_RAGGED_LIST_F(MCFFT_RAGGED_CASE, CM, (CM == CM_SPLIT ? 2 : 1))
#undef MCFFT_RAGGED_CASE
    default:
      break;
  }
}

template <int CM>
MCFFT_DEV void ragged_body(const RaggedArgs& r, int task_slot_offset) {
  // queue fetches, waits and counts are done by wave 0 (see wave_id in dev_prims.h)
  const int wid = wave_id();
  const unsigned one = lane_id() == 0 ? 1u : 0u;
  const AOps A = load_aops(r.aops, lane_id());
  MCFFT_DYN_LDS(unsigned char, ldsb);
  int* slot = (int*)(ldsb + task_slot_offset);
  unsigned* ctr = r.counters + kRaggedCounterBase;
  for (;;) {
    if (wid == 0) lds_st(slot, 0, r.task_begin + (int)wave_first(atomic_add_dev(&r.counters[r.queue], one)));
    block_barrier();
    const int ti = uniform(lds_ld(slot, 0));
    if (ti >= r.task_end) break;  // every workgroup leaves through here
    const RaggedTask t = r.tasks[ti];
    const RaggedPass* rp = (const RaggedPass*)uniform_ptr(r.passes + t.pass);
    // the descriptor's fields are the same for every lane: read as wave-uniform
    // values, so the branches on them are scalar branches
    const int wait_for = uniform(rp->wait_for), counter = uniform(rp->counter), signal = uniform(rp->signal);
    if (r.use_deps && wait_for > 0) {
      if (wid == 0) {
        unsigned it = 0;
        while (wave_first(atomic_load_dev(&ctr[counter])) < (unsigned)wait_for) {
          if (++it > r.spin_limit) {
            atomic_or_dev(&r.counters[4], one);
            break;
          }
          nap();
        }
      }
      block_barrier();
      acquire_dev();
    }
    const int ib = rp->in_buf, ob = rp->out_buf;
    const char* in = (ib == 0 ? r.in_base : ib == 1 ? (const char*)r.out_base : (const char*)r.ws_base) +
                     (size_t)rp->args.in;
    char* out = (ob == 1 ? r.out_base : r.ws_base) + (size_t)rp->args.out;
    ragged_dispatch<CM>(rp->variant, &rp->args, in, out, t.tile, A);  // ends with a barrier
    if (r.use_deps && signal) {
      release_dev();
      block_barrier();
      if (wid == 0) atomic_add_dev(&ctr[counter], one);
    }
  }
}

}  // namespace mcfft

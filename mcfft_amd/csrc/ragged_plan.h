// ragged_plan.h -- pure C++ planning of a ragged batch (no HIP calls): per-item
// passes built from the ragged kernel variants, shared twiddle tables per
// distinct shape, and the task queue of the persistent ragged kernel.
#pragma once
#include <vector>

#include "plan_core.h"

namespace mcfft {

struct RaggedItemSpec {
  int rank;
  long long dims[3];
  long long in_off, out_off;  // element offsets of the transform in the exec input / output
};

struct RaggedPassHost {
  PassPlan pp;            // tile structure, strides, tables (offsets into tw)
  int variant = 0;        // index into kRaggedVariantInfo[cm]
  long long in_off = 0;   // element offset inside the source buffer
  long long out_off = 0;  // element offset inside the destination buffer
  long long in_elems = 0, out_elems = 0;  // extent of the buffer views
  int counter = 0;        // transform-group counter
  int wait_for = 0;
  int signal = 0;
};

struct RaggedPlanCore {
  int cm = 0, in_fmt = 0, out_fmt = 0, scaling = 1;
  int n_items = 0;
  int n_counters = 0;               // transform groups with more than one pass
  std::vector<RaggedPassHost> passes;
  std::vector<int> task_pass, task_tile;
  std::vector<int> phase_begin;     // task index where each pass phase starts (+ end)
  std::vector<TwSpec> tw;
  std::vector<int> exponent;        // per item
  long long ws_elems = 0;
  long long in_extent = 0, out_extent = 0;  // elements the exec buffers must hold
  std::string describe() const;
};

// Items whose 1D length is at most 4096 and that sit back to back in both
// buffers with the same length are fused into one multi-row pass.
int ragged_build(RaggedPlanCore& r, int n_items, const RaggedItemSpec* items, int in_fmt, int out_fmt, int scaling);
// ST_OK if the queue's pass variants can run a transform of this shape (the
// others -- 1D above 2^18 points, outer axes narrower than a variant's column
// tile -- need a batched plan)
int ragged_shape_supported(int rank, const long long* dims, int in_fmt, int out_fmt, int scaling);

}  // namespace mcfft

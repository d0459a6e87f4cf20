// exec_args.h -- the block-floating-point and bookkeeping part of a pass
// launch's arguments (bfp.h), shared by the runtime and the CPU emulation build
// so both launch the kernels identically.
#pragma once
#include "fft_pass.h"
#include "plan_core.h"

namespace mcfft {

// a plan's device bookkeeping buffers (slots and tile_e only for AUTO multi-pass plans)
struct BfpBuffers {
  ExecState* state;
  unsigned long long* slots;
  signed char* tile_e;
  int* out_exp;
};

// bfp_* arguments of pass `pi` of plan `c` for a launch over transforms
// [t0, t0 + nt); c2r: the plan is the complex part of a C2R transform, whose
// pre-pass halves the input (one more power of two in the output exponent).
inline void fill_bfp(PassArgs& a, const PlanCore& c, int pi, long long t0, long long nt, bool c2r,
                     const int* in_exp, const BfpBuffers& b) {
  const int np = (int)c.passes.size();
  const PassPlan& pp = c.passes[pi];
  const bool last = pi == np - 1;
  int flags = 0;
  a.bfp_e_base = 0;
  if (c.scaling == SCALE_AUTO) {
    flags |= BFP_AUTO;
    if (pi == 0) flags |= BFP_FIRST;
    if (pi == 1) flags |= BFP_RECONCILE;
    if (last && np > 1) flags |= BFP_EPOCH;
    int e = 0;  // stage shifts of pass 0 (first pass), or of passes 1..pi
    for (int q = pi == 0 ? 0 : 1; q <= pi; ++q)
      for (int sh : c.passes[q].shifts) e += sh;
    a.bfp_e_base = e;
  } else if (last) {
    flags |= BFP_CHECK;
    a.bfp_e_base = c.out_exponent;
  }
  if (last) flags |= BFP_LAST;
  if (last && c2r) a.bfp_e_base += 1;
  if (np == 1 && pp.groups_per_transform == 0) flags |= BFP_SINGLE;
  a.bfp_flags = flags;
  a.bfp_quantum = np > 1 ? kBfpQuantumMultiPass : 1;
  a.bfp_tr0 = (int)t0;
  a.bfp_nt = (int)nt;
  a.bfp_gpt = pp.groups_per_transform > 0 ? (int)pp.groups_per_transform : 1;
  a.bfp_q0_lg = ilog2(c.passes[0].k->Q);
  a.bfp_tpg0 = c.passes[0].tiles_per_group;
  a.bfp_rg = pp.rc_g < 0 ? (int)pp.L : pp.rc_g;
  a.bfp_rm = pp.rc_m;
  a.bfp_rq = pp.rc_q;
  a.bfp_state = b.state;
  a.bfp_slots = b.slots;
  a.bfp_tile_e = b.tile_e;
  a.bfp_out_exp = b.out_exp;
  a.bfp_in_exp = in_exp;
}

// sizes of the bookkeeping buffers of plan c: slots (64-bit) and tile
// exponents (bytes: one per first-pass instance of a chunk, at least as many
// as the first pass's tiles whatever tile width it runs)
inline bool bfp_uses_slots(const PlanCore& c) { return c.scaling == SCALE_AUTO && c.passes.size() > 1; }
inline size_t bfp_slot_count(const PlanCore& c) { return bfp_uses_slots(c) ? 2 * (size_t)c.batch : 0; }
inline size_t bfp_tile_bytes(const PlanCore& c) {
  return bfp_uses_slots(c) ? (size_t)(c.chunk * (c.n / c.passes[0].L)) : 0;
}

}  // namespace mcfft

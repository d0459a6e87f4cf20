// exec_args.h -- a pass launch's arguments (PassArgs) from its plan, including
// the block-floating-point and bookkeeping part (bfp.h), shared by the runtime
// and the CPU emulation build so both launch the kernels identically.
#pragma once
#include "fft_pass.h"
#include "plan_core.h"

#include <cstring>

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
  // first-pass tiles per transform (its groups per transform times tiles per group)
  a.bfp_tpg0 = c.passes[0].tiles_per_group *
               (c.passes[0].groups_per_transform > 1 ? (int)c.passes[0].groups_per_transform : 1);
  a.bfp_rg = pp.rc_g < 0 ? (int)pp.L : pp.rc_g;
  a.bfp_rm = pp.rc_m;
  a.bfp_rq = pp.rc_q;
  a.bfp_state = b.state;
  a.bfp_slots = b.slots;
  a.bfp_tile_e = b.tile_e;
  a.bfp_out_exp = b.out_exp;
  a.bfp_in_exp = in_exp;
}

// Everything of pass pp's launch arguments but the bfp_* fields and n_tiles:
// `in`/`out` already point at the chunk's data, in_elems/out_elems are the
// element counts addressable from them, `tw` is the plan's twiddle table for
// the direction and `aops` its DFT-matrix operands.
inline void fill_pass_args(PassArgs& a, const PlanCore& c, const PassPlan& pp, const void* in, void* out,
                           long long in_elems, long long out_elems, int inverse, const f32x2* tw,
                           const u32x4* aops, const uint32_t* dft256) {
  memset(&a, 0, sizeof(a));
  a.in = in;
  a.out = out;
  a.stride_in = pp.stride_in;
  a.stride_out = pp.stride_out;
  a.in_sg = pp.in_sg;
  a.in_st = pp.in_st;
  a.out_sg = pp.out_sg;
  a.out_st = pp.out_st;
  a.n_inst = (int)pp.n_inst;
  a.tiles_per_group = pp.tiles_per_group;
  a.in_fmt = pp.src_fmt;
  a.in_elems = in_elems;
  a.out_elems = out_elems;
  a.stage_tw_entries = (int)((long long)c.tw.size() - pp.stage_tw_off);
  a.inverse = inverse;
  a.out_scale = pp.out_scale;
  for (size_t si = 0; si < 6; ++si) a.st_scale[si] = si < pp.shifts.size() ? pow2(-pp.shifts[si]) : 1.f;
  a.stage_tw = tw + pp.stage_tw_off;
  a.aops = aops;
  a.dft256 = dft256;
  if (pp.fourstep) {
    a.fs_tw = tw + pp.fs_tw_off;
    a.fs_tw_entries = (int)((long long)c.tw.size() - pp.fs_tw_off);
    a.fs_cols = pp.fs_cols;
    a.fs_lo_bits = pp.fs_lo_bits;
  } else if (pp.x_tw_off >= 0) {  // split ROWX: its cross-row twiddles
    a.fs_tw = tw + pp.x_tw_off;
    a.fs_tw_entries = (int)((long long)c.tw.size() - pp.x_tw_off);
  }
}

// A gathered batch (a ragged plan's same-shape items at arbitrary offsets): the
// pass's sides on the caller's buffers take transform t of the launch from
// tab[side][t0 + t] (tab[0]: input offsets, tab[1]: output offsets, elements);
// `in`/`out` given to fill_pass_args are then the buffers' bases.
inline void gather_args(PassArgs& a, const PassPlan& pp, const long long* const tab[2], long long t0) {
  auto side = [&](int b) -> const long long* { return b == BUF_WS ? nullptr : tab[b == BUF_IN ? 0 : 1] + t0; };
  a.gather_in = side(pp.src);
  a.gather_out = side(pp.dst);
  a.gather_gpt = (int)pp.groups_per_transform;
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

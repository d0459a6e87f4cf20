// plan_core.h -- pure C++ planning (no HIP calls): pass decomposition, strides,
// per-stage scale shifts and the twiddle-table specification of a plan.
//
// Reference counterpart: mcfftCreate's size switch
// (source/mcfft_1d_device.cpp:680-790, source/mcfft_2d_device.cpp:625-696),
// which hard-codes radices/mergings for ten 1D sizes and three 2D sides.  Here
// any power-of-two length 16..2^23 (1D) or side 16..4096 (2D/3D) is decomposed
// into passes of the kernel family in fft_pass.h.
#pragma once
#include <stdint.h>

#include <map>
#include <string>
#include <vector>

#include "kernel_list.h"

namespace mcfft {

enum Buf { BUF_IN = 0, BUF_OUT = 1, BUF_WS = 2 };
// SCALE_AUTO: block floating point following the data (bfp.h); its stages
// divide by their radix as SCALE_FULL's do, and each transform's exponent is
// found on the device
enum Scaling { SCALE_NONE = 0, SCALE_BALANCED = 1, SCALE_FULL = 2, SCALE_AUTO = 3 };
// first-pass normalisation step of multi-pass AUTO plans: tiles whose input
// maxima lie within one step of 2^4 agree on their exponent, so the pass after
// the first has nothing to reconcile for such data
constexpr int kBfpQuantumMultiPass = 4;
// 2D plans split an outer axis of at least this many points (ROWX + COL), so
// that its column pass keeps 128-byte row segments (MCFFT_2D_OUTER_RADIX
// overrides the radix: 0 or 1 off, 2/4/8)
constexpr long long kOuterSplitMin = 1024;
// 1D lengths: single pass up to 2^16 (FP16/BF16), two-pass four-step up to
// 2^23, three passes from 2^24 to 2^kMaxLg1D
constexpr int kThreePassMinLg = 24;
constexpr int kMaxLg1D = 27;

// one twiddle-table entry: exp(sign * 2 pi i * e / M) * scale (sign = -1 forward)
struct TwSpec {
  long long e, M;
  float scale;
};

struct PassPlan {
  const KernelInfo* k = nullptr;
  int kind = 0;                 // 0 ROW, 1 ROWT, 2 COL, 3 COLF, 4 COLM, 5 ROWX
  long long L = 0;              // points per instance
  int src = BUF_IN, dst = BUF_OUT;
  int src_fmt = 0;
  int stride_in = 0, stride_out = 0;  // strided-axis strides (see PassArgs)
  long long in_sg = 0, in_st = 0, out_sg = 0, out_st = 0;
  long long n_inst = 0;         // valid instances per group
  int tiles_per_group = 1;
  long long groups = 1;         // groups when the whole batch runs as one chunk
  long long groups_per_transform = 0;  // >0: groups = transforms in the chunk * this
  float out_scale = 1.f;
  long long stage_tw_off = 0;   // first entry of this pass's stage tables
  bool fourstep = false;
  long long fs_tw_off = 0;      // four-step table W_N^(k1*n2): N1 x fs_cols entries
  int fs_cols = 0;
  int fs_lo_bits = 0;           // > 0: the table in two factors (fft_pass.h, PassArgs::fs_tw)
  long long x_tw_off = -1;      // split ROWX: W_NX^(r1*k2) table (NX entries), else -1
  // reconcile pass (AUTO): element (q, m) of group g came from first-pass instance
  // (g % groups_per_transform) * rc_g + m * rc_m + q * rc_q (rc_g < 0: L)
  int rc_g = -1, rc_m = 1, rc_q = 0;
  std::vector<int> shifts;      // per stage (ROWX: the row stages, then its cross-row butterfly)
};

// Plan structure: the choices the autotuner enumerates (plan_structures) and
// records; -1 = the planner's default rule.
struct PlanOpts {
  int outer_radix = -1;   // 2D: radix R of the outer-axis split (0: none; 2, 4, 8)
  int fourstep_lg2 = -1;  // 1D two-pass four-step: log2 N2 (points of the second pass's rows);
                          // > 0 on a length with a single-pass kernel (>= 2^12): run it as two passes
  // gathered batch (transforms at arbitrary offsets, PassArgs::gather_*): every
  // pass has per-transform groups (a 1D single-pass plan one row per group)
  bool gathered = false;
};

struct PlanCore {
  int rank = 0;
  PlanOpts opts;                // the structure this plan was built with (defaults resolved)
  long long dims[3] = {1, 1, 1};
  long long batch = 0;
  long long n = 0;              // elements per transform
  int in_fmt = 0, out_fmt = 0, cm = 0;
  int scaling = SCALE_BALANCED;
  int out_exponent = 0;         // true value = stored * 2^out_exponent (AUTO: the stage shifts only)
  long long chunk = 0;          // transforms per chunk
  long long n_chunks = 0;
  long long ws_elems = 0;       // workspace elements in the compute format
  std::vector<PassPlan> passes;
  std::vector<TwSpec> tw;       // twiddle entries (same spec for both directions)
  std::string describe() const;
};

// status codes shared with the C ABI (include/mcfft.h)
enum {
  ST_OK = 0,
  ST_INVALID_VALUE = 1,
  ST_UNSUPPORTED = 2,
  ST_ALLOC_FAILED = 3,
  ST_HIP_ERROR = 4,
  ST_MISALIGNED = 5,
  ST_INVALID_PLAN = 6,
  ST_CACHE_ERROR = 7,
};

int elem_bytes(int fmt);
// every compiled kernel that can run pass `pp` (same kind, length and compute
// mode; the first is the planner's default); used by the tuner
std::vector<const KernelInfo*> pass_candidates(const PassPlan& pp, int cm, int scaling);
// whether kernel k may run a pass of a plan with this scaling mode
bool kernel_serves(const KernelInfo& k, int scaling);
// bind kernel `k` to pass `pp`, recomputing the tile-size dependent fields;
// false if its tile does not divide the pass
bool set_pass_kernel(PassPlan& pp, const KernelInfo* k);
// "rank|dims|batch|in|out|scaling" -- tuning cache key of a plan
std::string plan_signature(const PlanCore& p);
// helpers shared with the ragged planner
int ilog2(long long v);                          // log2 of a power of two, -1 otherwise
bool split_len(long long L, int& S, int& T);     // L = 16^S * T, S >= 1, T in {1,2,4,8}
float pow2(int e);                               // 2^e exactly
std::vector<int> add_stage_tables(std::vector<TwSpec>& tw, PassPlan& pp, int S, int T, int scaling);
// compute mode from the input/output formats; -1 if the pair is not supported
int compute_mode(int in_fmt, int out_fmt);
// ws_limit_bytes: upper bound of the four-step workspace (0 = default);
// opt: structure (defaults: the planner's rules)
int plan_build(PlanCore& p, int rank, const long long* dims, long long batch, int in_fmt, int out_fmt, int scaling,
               long long ws_limit_bytes, const PlanOpts& opt = PlanOpts());
// the distinct structures a plan of this shape can take (the default first;
// each resolved as plan_build would record it)
std::vector<PlanOpts> plan_structures(int rank, const long long* dims, long long batch, int in_fmt, int out_fmt,
                                      int scaling, long long ws_limit_bytes);
// "R=<outer radix>;L2=<four-step log2 N2>" and its inverse (false if malformed)
std::string opts_string(const PlanOpts& o);
bool parse_opts(const std::string& s, PlanOpts& o);

// Tuning cache (one file per GPU architecture, kept by the runtime): a header
// line "mcfft-tune-cache 2 <build id>", then one line per tuned plan,
//     <arch>|<plan signature>=R=<r>;L2=<l>;K=<kernel>,<kernel>,...
// A file whose header names another format or another build of the library
// (whose kernels of the same name may compute differently) yields no entries.
struct TuneChoice {
  PlanOpts opts;
  std::vector<std::string> kernels;  // one per pass
};
std::map<std::string, std::string> tune_cache_parse(const std::string& text, const std::string& build_id);
std::string tune_cache_format(const std::map<std::string, std::string>& entries, const std::string& build_id);
bool parse_tune_choice(const std::string& value, TuneChoice& c);
std::string format_tune_choice(const TuneChoice& c);

}  // namespace mcfft

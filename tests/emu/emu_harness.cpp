// emu_harness.cpp -- runs the library's planner and pass kernels on the CPU
// (test infrastructure).  The kernel sources are compiled unchanged with
// MCFFT_EMU: every LDS and global access is bounds-checked, so a wrong index
// aborts here instead of faulting a GPU.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fft_big.h"
#include "fft_pass.h"
#include "plan_core.h"
#include "ragged.h"
#include "ragged_plan.h"
#include "real_pass.h"
#include "layout_pass.h"
#include "exec_args.h"

using namespace mcfft;

static void twiddle_host(const TwSpec& s, int sign, f32x2& out) {
  long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)s.e / (long double)s.M;
  long double c = cosl(a), sn = sinl(a);
  if ((4 * s.e) % s.M == 0) {
    long long q = ((4 * s.e) / s.M) % 4;
    c = (q == 0) ? 1 : (q == 2) ? -1 : 0;
    sn = (q == 1) ? 1 : (q == 3) ? -1 : 0;
  }
  out.x = (float)((double)c * s.scale);
  out.y = (float)((double)(sign * sn) * s.scale);
}

static uint16_t f16b(double v) { return __builtin_bit_cast(uint16_t, (_Float16)v); }
static uint16_t bf16b(double v) { return __builtin_bit_cast(uint16_t, (__bf16)(float)v); }

static void dft_table_host(int sign, uint16_t* t) {
  for (int e = 0; e < 16; ++e) {
    TwSpec s{e, 16, 1.f};
    long double a = 2.0L * 3.14159265358979323846264338327950288L * e / 16.0L;
    double re = (double)cosl(a), im = (double)(sign * sinl(a));
    if ((4 * e) % 16 == 0) {
      int q = (4 * e) / 16;
      re = (q == 0) ? 1 : (q == 2) ? -1 : 0;
      im = sign * ((q == 1) ? 1 : (q == 3) ? -1 : 0);
    }
    (void)s;
    uint16_t rh = f16b(re), ih = f16b(im);
    t[e] = rh;
    t[16 + e] = ih;
    t[32 + e] = f16b((re - (double)__builtin_bit_cast(_Float16, rh)) * 2048.0);
    t[48 + e] = f16b((im - (double)__builtin_bit_cast(_Float16, ih)) * 2048.0);
    t[64 + e] = bf16b(re);
    t[80 + e] = bf16b(im);
  }
}

// the 65536-point kernel's row-scaled DFT matrices (fft_big.h), as the runtime
// makes them (FP64 values rounded once; exact at quarter turns)
static void dft256_table_host(int sign, int cm, uint32_t* t) {
  for (int w = 0; w < kDft256Words; ++w) {
    const Dft256Word d = dft256_word(w);
    uint32_t word = 0;
    for (int j = 0; j < 2; ++j) {
      const int e = dft256_exponent(d.r2, d.i, d.k + j);
      long double a = 2.0L * 3.14159265358979323846264338327950288L * e / 256.0L;
      double re = (double)cosl(a), im = (double)(sign * sinl(a));
      if (e % 64 == 0) {
        const int q = e / 64;
        re = (q == 0) ? 1 : (q == 2) ? -1 : 0;
        im = sign * ((q == 1) ? 1 : (q == 3) ? -1 : 0);
      }
      const double v = d.b == 0 ? re : d.b == 1 ? im : -im;
      word |= (uint32_t)(cm == 1 ? bf16b(v) : f16b(v)) << (16 * j);
    }
    t[w] = word;
  }
}

// the structures the autotuner would try for this plan shape, one
// opts_string per line
extern "C" int emu_plan_structures(int rank, const long long* dims, long long batch, int in_fmt, int out_fmt,
                                   int scaling, char* buf, int buflen) {
  std::string s;
  for (const PlanOpts& o : plan_structures(rank, dims, batch, in_fmt, out_fmt, scaling, 0)) s += opts_string(o) + "\n";
  snprintf(buf, buflen, "%s", s.c_str());
  return 0;
}

// the tuning cache's parser: the entry for `key` in cache file `text` as the
// runtime (built as `build_id`) would see it, re-formatted; 0 if found
extern "C" int emu_tune_cache_lookup(const char* text, const char* build_id, const char* key, char* buf, int buflen) {
  const auto m = tune_cache_parse(text, build_id);
  const auto it = m.find(key);
  if (it == m.end()) return 1;
  TuneChoice c;
  if (!parse_tune_choice(it->second, c)) return 2;
  snprintf(buf, buflen, "%s", format_tune_choice(c).c_str());
  return 0;
}

// a plan built with an explicit structure (the autotuner's candidates)
extern "C" int emu_plan_describe_opts(int rank, const long long* dims, long long batch, int in_fmt, int out_fmt,
                                      int scaling, int outer_radix, int fourstep_lg2, char* buf, int buflen) {
  PlanCore p;
  PlanOpts o;
  o.outer_radix = outer_radix;
  o.fourstep_lg2 = fourstep_lg2;
  const int st = plan_build(p, rank, dims, batch, in_fmt, out_fmt, scaling, 0, o);
  if (st) return st;
  snprintf(buf, buflen, "%s", p.describe().c_str());
  return 0;
}

extern "C" int emu_plan_describe(int rank, const long long* dims, long long batch, int in_fmt, int out_fmt,
                                 int scaling, char* buf, int buflen) {
  PlanCore p;
  int st = plan_build(p, rank, dims, batch, in_fmt, out_fmt, scaling, 0);
  if (st) return st;
  std::string s = p.describe();
  snprintf(buf, buflen, "%s", s.c_str());
  return 0;
}

// Run a whole plan on host buffers.  ws_limit bounds the four-step workspace.
// Kernel override for tests: passes whose candidates include a kernel of this
// name run that kernel instead of the planner's default (empty: no override).
static std::string g_kernel_override, g_last_kernels;
extern "C" void emu_set_kernel_override(const char* name) { g_kernel_override = name ? name : ""; }
extern "C" const char* emu_last_kernels() { return g_last_kernels.c_str(); }

// Per-transform output exponents and status word of the last emu_exec, and
// input exponents for the next ones (empty: none) -- the runtime's
// mcfft_get_transform_exponents / mcfft_get_status_flags / mcfft_exec_exp.
static std::vector<int> g_texps, g_in_exps;
static unsigned g_status = 0;
extern "C" int emu_transform_exponents(int* out, long long n) {
  if ((size_t)n != g_texps.size()) return -1;
  for (long long i = 0; i < n; ++i) out[i] = g_texps[(size_t)i];
  return 0;
}
extern "C" unsigned emu_status() { return g_status; }
extern "C" void emu_set_input_exponents(const int* e, long long n) { g_in_exps.assign(e, e + (e ? n : 0)); }

// Restrict emu_exec to pass `pass` of the plan, launched with one workgroup per
// tile, of which only workgroups [blk_first, blk_first + blk_count) run: one
// part of a plan too large to run whole on the CPU.  pass < 0 lifts it.
static int g_only_pass = -1;
static long long g_blk_first = 0, g_blk_count = 0;
extern "C" void emu_restrict(int pass, long long blk_first, long long blk_count) {
  g_only_pass = pass;
  g_blk_first = blk_first;
  g_blk_count = blk_count;
}

// plan structure of the next emu_exec calls (the autotuner's candidates; -1: defaults)
static PlanOpts g_opts;
// gathered batch for the next emu_exec calls (empty: transforms back to back):
// per-transform element offsets into the input and output buffers, and the
// buffers' element counts
static std::vector<long long> g_gather[2];
static long long g_gather_ext[2] = {0, 0};
extern "C" void emu_set_gather(const long long* in_off, const long long* out_off, long long n, long long ext_in,
                               long long ext_out) {
  g_gather[0].assign(in_off, in_off + (in_off ? n : 0));
  g_gather[1].assign(out_off, out_off + (out_off ? n : 0));
  g_gather_ext[0] = ext_in;
  g_gather_ext[1] = ext_out;
}
extern "C" void emu_set_structure(int outer_radix, int fourstep_lg2) {
  g_opts.outer_radix = outer_radix;
  g_opts.fourstep_lg2 = fourstep_lg2;
}

extern "C" int emu_exec(int rank, const long long* dims, long long batch, int in_fmt, int out_fmt, int scaling,
                        int inverse, const void* in, void* out, long long ws_limit, int* out_exponent) {
  PlanCore p;
  const bool gathered = !g_gather[0].empty();
  if (gathered && (long long)g_gather[0].size() != batch) return ST_INVALID_VALUE;
  PlanOpts opts = g_opts;
  opts.gathered = gathered;
  int st = plan_build(p, rank, dims, batch, in_fmt, out_fmt, scaling, ws_limit, opts);
  if (st) return st;
  if (!g_kernel_override.empty())
    for (PassPlan& pp : p.passes)
      for (const KernelInfo* k : pass_candidates(pp, p.cm, p.scaling))
        if (g_kernel_override == k->name) set_pass_kernel(pp, k);
  g_last_kernels.clear();
  for (const PassPlan& pp : p.passes) g_last_kernels += std::string(pp.k->name) + " ";
  if (out_exponent) *out_exponent = p.out_exponent;
  const int sign = inverse ? 1 : -1;
  std::vector<f32x2> tw(p.tw.size() + 1);
  for (size_t i = 0; i < p.tw.size(); ++i) twiddle_host(p.tw[i], sign, tw[i]);
  std::vector<uint16_t> dft(kDftTableEntries);
  dft_table_host(sign, dft.data());
  std::vector<u32x4> aops(64 * 4);
  for (int lane = 0; lane < 64; ++lane)
    aops_words(dft.data(), p.cm, lane, [&](int k, uint32_t w) { aops[lane * 4 + k / 4][k % 4] = w; });
  std::vector<uint32_t> dft256(kDft256Words);
  dft256_table_host(sign, p.cm, dft256.data());
  const int ws_fmt = (p.cm == 2) ? 2 : in_fmt;
  std::vector<unsigned char> ws((size_t)(p.ws_elems * elem_bytes(ws_fmt)) + 16);
  const size_t in_bytes = (size_t)((gathered ? g_gather_ext[0] : p.n * batch) * elem_bytes(in_fmt));
  const size_t out_bytes = (size_t)((gathered ? g_gather_ext[1] : p.n * batch) * elem_bytes(out_fmt));
  // bookkeeping buffers, as the runtime allocates them (all zero)
  ExecState state{0ull, 0u, 0u};
  std::vector<int> out_exp((size_t)batch, 0);
  // slots persist across emulated execs of the same plan shape, as they do in a plan
  static std::vector<unsigned long long> slots;
  static unsigned long long epoch = 0;
  static std::string slot_key;
  const std::string key = plan_signature(p);
  if (key != slot_key || slots.size() != bfp_slot_count(p) + 1) {
    slots.assign(bfp_slot_count(p) + 1, 0ull);
    epoch = 0;
    slot_key = key;
  }
  state.epoch = epoch;
  std::vector<signed char> tile_e(bfp_tile_bytes(p) + 1, 0);
  if (!g_in_exps.empty() && (long long)g_in_exps.size() != batch) return ST_INVALID_VALUE;
  const int* in_exp = g_in_exps.empty() ? nullptr : g_in_exps.data();
  const BfpBuffers bufs{&state, slots.data(), tile_e.data(), out_exp.data()};
  emu::clear_ranges();
  emu::register_range(in, in_bytes);
  emu::register_range(out, out_bytes);
  emu::register_range(ws.data(), (size_t)(p.ws_elems * elem_bytes(ws_fmt)));
  emu::register_range(tw.data(), p.tw.size() * sizeof(f32x2));
  emu::register_range(aops.data(), aops.size() * sizeof(u32x4));
  emu::register_range(dft256.data(), dft256.size() * sizeof(uint32_t));
  emu::register_range(&state, sizeof(state));
  emu::register_range(out_exp.data(), out_exp.size() * sizeof(int));
  emu::register_range(slots.data(), bfp_slot_count(p) * sizeof(unsigned long long));
  emu::register_range(tile_e.data(), bfp_tile_bytes(p));
  if (in_exp) emu::register_range(in_exp, g_in_exps.size() * sizeof(int));
  for (int d = 0; d < 2 && gathered; ++d) emu::register_range(g_gather[d].data(), g_gather[d].size() * sizeof(long long));
  for (long long c = 0; c < p.n_chunks; ++c) {
    const long long t0 = c * p.chunk;
    const long long nt = (t0 + p.chunk <= batch) ? p.chunk : batch - t0;
    for (size_t pi = 0; pi < p.passes.size(); ++pi) {
      const PassPlan& pp = p.passes[pi];
      if (g_only_pass >= 0 && (int)pi != g_only_pass) continue;
      auto bufptr = [&](int b, int fmt) -> char* {
        const size_t skip = gathered ? 0 : (size_t)(t0 * p.n * elem_bytes(fmt));
        if (b == BUF_IN) return (char*)in + skip;
        if (b == BUF_OUT) return (char*)out + skip;
        return (char*)ws.data();
      };
      auto view_elems = [&](int b) -> long long {
        if (b == BUF_WS) return p.ws_elems;
        return gathered ? g_gather_ext[b == BUF_IN ? 0 : 1] : (p.batch - t0) * p.n;
      };
      const int dst_fmt = (pp.dst == BUF_OUT) ? out_fmt : ws_fmt;
      PassArgs a;
      fill_pass_args(a, p, pp, bufptr(pp.src, pp.src_fmt), bufptr(pp.dst, dst_fmt), view_elems(pp.src),
                     view_elems(pp.dst), inverse, tw.data(), aops.data(), dft256.data());
      if (gathered) {
        const long long* const tab[2] = {g_gather[0].data(), g_gather[1].data()};
        gather_args(a, pp, tab, t0);
      }
      fill_bfp(a, p, (int)pi, t0, nt, false, in_exp, bufs);
      const long long groups = pp.groups_per_transform > 0 ? nt * pp.groups_per_transform : pp.groups;
      const long long tiles = groups * pp.tiles_per_group;
      a.n_tiles = (int)tiles;
      void (*fn)(PassArgs) = (void (*)(PassArgs))pp.k->fn;
      if (g_only_pass >= 0) {
        emu::launch((int)tiles, pp.k->nw * 64, (size_t)pp.k->lds_bytes, [&] { fn(a); }, (int)g_blk_first,
                    (int)(g_blk_first + g_blk_count));
        continue;
      }
      // fewer workgroups than tiles exercises the persistent tile loop
      const long long grid = tiles > 3 ? (tiles + 1) / 2 : tiles;
      emu::launch((int)grid, pp.k->nw * 64, (size_t)pp.k->lds_bytes, [&] { fn(a); });
    }
  }
  g_texps = out_exp;
  g_status = state.status;
  // an AUTO multi-pass exec advances the epoch once per chunk, and every slot
  // of the batch carries a tag of this exec
  if (bfp_uses_slots(p) && g_only_pass < 0) {
    if (state.epoch != epoch + (unsigned long long)p.n_chunks) return -100;
    for (size_t i = 0; i < bfp_slot_count(p); ++i)
      if ((slots[i] >> 16) < epoch || (slots[i] >> 16) >= state.epoch) return -101;
  }
  epoch = state.epoch;
  return 0;
}

extern "C" void mcfft_ragged_0(RaggedArgs r);
extern "C" void mcfft_ragged_1(RaggedArgs r);
extern "C" void mcfft_ragged_2(RaggedArgs r);

// Ragged batch through the persistent ragged kernel, `nblocks` workgroups run
// concurrently.  items: n x {rank, d0, d1, d2, in_off, out_off} (long long).
extern "C" int emu_ragged_exec(int n_items, const long long* spec, int in_fmt, int out_fmt, int scaling, int inverse,
                               const void* in, void* out, int nblocks, int* exponents, unsigned* status, int single) {
  std::vector<RaggedItemSpec> items(n_items);
  for (int i = 0; i < n_items; ++i) {
    const long long* q = spec + 6 * i;
    items[i].rank = (int)q[0];
    items[i].dims[0] = q[1];
    items[i].dims[1] = q[2];
    items[i].dims[2] = q[3];
    items[i].in_off = q[4];
    items[i].out_off = q[5];
  }
  RaggedPlanCore r;
  int st = ragged_build(r, n_items, items.data(), in_fmt, out_fmt, scaling);
  if (st) return st;
  for (int i = 0; i < n_items; ++i) exponents[i] = r.exponent[i];
  const int sign = inverse ? 1 : -1;
  std::vector<f32x2> tw(r.tw.size() + 1);
  for (size_t i = 0; i < r.tw.size(); ++i) twiddle_host(r.tw[i], sign, tw[i]);
  std::vector<uint16_t> dft(kDftTableEntries);
  dft_table_host(sign, dft.data());
  std::vector<u32x4> aops(64 * 4);
  for (int lane = 0; lane < 64; ++lane)
    aops_words(dft.data(), r.cm, lane, [&](int k, uint32_t w) { aops[lane * 4 + k / 4][k % 4] = w; });
  const int ws_fmt = (r.cm == 2) ? 2 : in_fmt;
  std::vector<unsigned char> ws((size_t)(r.ws_elems * elem_bytes(ws_fmt)) + 16);
  std::vector<RaggedPass> dp(r.passes.size());
  for (size_t i = 0; i < r.passes.size(); ++i) {
    const RaggedPassHost& h = r.passes[i];
    RaggedPass& d = dp[i];
    memset(&d, 0, sizeof(d));
    PassArgs& a = d.args;
    const int src_esz = elem_bytes(h.pp.src_fmt);
    const int dst_esz = elem_bytes(h.pp.dst == BUF_WS ? ws_fmt : out_fmt);
    a.in = (const void*)(size_t)(h.in_off * src_esz);
    a.out = (void*)(size_t)(h.out_off * dst_esz);
    a.in_sg = h.pp.in_sg; a.in_st = h.pp.in_st; a.out_sg = h.pp.out_sg; a.out_st = h.pp.out_st;
    a.stride_in = h.pp.stride_in; a.stride_out = h.pp.stride_out;
    a.n_inst = (int)h.pp.n_inst;
    a.tiles_per_group = h.pp.tiles_per_group;
    a.n_tiles = (int)(h.pp.groups * h.pp.tiles_per_group);
    a.in_fmt = h.pp.src_fmt;
    a.inverse = inverse;
    a.out_scale = h.pp.out_scale;
    for (size_t si = 0; si < 6; ++si) a.st_scale[si] = si < h.pp.shifts.size() ? pow2(-h.pp.shifts[si]) : 1.f;
    a.stage_tw = tw.data() + h.pp.stage_tw_off;
    a.stage_tw_entries = (int)((long long)r.tw.size() - h.pp.stage_tw_off);
    a.aops = aops.data();
    a.in_elems = h.in_elems;
    a.out_elems = h.out_elems;
    if (h.pp.fourstep) {
      a.fs_tw = tw.data() + h.pp.fs_tw_off;
      a.fs_tw_entries = (int)((long long)r.tw.size() - h.pp.fs_tw_off);
      a.fs_cols = h.pp.fs_cols;
    }
    d.in_buf = h.pp.src;
    d.out_buf = h.pp.dst;
    d.variant = h.variant;
    d.counter = h.counter;
    d.wait_for = h.wait_for;
    d.signal = h.signal;
  }
  std::vector<RaggedTask> tasks(r.task_pass.size());
  for (size_t i = 0; i < tasks.size(); ++i) tasks[i] = RaggedTask{r.task_pass[i], r.task_tile[i]};
  std::vector<unsigned> counters(kRaggedCounterBase + r.n_counters + 1, 0u);
  emu::clear_ranges();
  emu::register_range(in, (size_t)(r.in_extent * elem_bytes(in_fmt)));
  emu::register_range(out, (size_t)(r.out_extent * elem_bytes(out_fmt)));
  emu::register_range(ws.data(), (size_t)(r.ws_elems * elem_bytes(ws_fmt)));
  emu::register_range(tw.data(), r.tw.size() * sizeof(f32x2));
  emu::register_range(aops.data(), aops.size() * sizeof(u32x4));
  RaggedArgs ra;
  ra.in_base = (const char*)in;
  ra.out_base = (char*)out;
  ra.ws_base = (char*)ws.data();
  ra.passes = dp.data();
  ra.tasks = tasks.data();
  ra.counters = counters.data();
  ra.spin_limit = 1u << 22;
  ra.aops = aops.data();
  void (*fn)(RaggedArgs) = r.cm == 0 ? mcfft_ragged_0 : r.cm == 1 ? mcfft_ragged_1 : mcfft_ragged_2;
  const int lds = kRaggedKernels[r.cm].lds_bytes;
  const int n_phase = (int)r.phase_begin.size() - 1;
  for (int ph = 0; ph < (single ? 1 : n_phase); ++ph) {
    ra.task_begin = single ? 0 : r.phase_begin[ph];
    ra.task_end = single ? r.phase_begin.back() : r.phase_begin[ph + 1];
    ra.queue = single ? 3 : ph;
    ra.use_deps = single ? 1 : 0;
    emu::launch_concurrent(nblocks, kRaggedNW * 64, (size_t)lds, [&] { fn(ra); });
  }
  *status = counters[4];
  return 0;
}

// Structural checks of a ragged plan (no kernel runs): every dependency wait is
// satisfiable by tiles queued before it, tile indices are in range, buffer
// views stay inside the exec buffers / workspace, and phases hold one pass
// level each.  Returns the number of violations (first few printed).
extern "C" long long emu_ragged_check(int n_items, const long long* spec, int in_fmt, int out_fmt, int scaling,
                                      long long* stats) {
  std::vector<RaggedItemSpec> items(n_items);
  for (int i = 0; i < n_items; ++i) {
    const long long* q = spec + 6 * i;
    items[i] = RaggedItemSpec{(int)q[0], {q[1], q[2], q[3]}, q[4], q[5]};
  }
  RaggedPlanCore r;
  int st = ragged_build(r, n_items, items.data(), in_fmt, out_fmt, scaling);
  if (st) return -st;
  const int ws_fmt = (r.cm == 2) ? 2 : in_fmt;
  long long bad = 0;
  auto fail = [&](const char* what, long long a, long long b) {
    if (bad++ < 8) fprintf(stderr, "ragged check: %s (%lld, %lld)\n", what, a, b);
  };
  const size_t nt = r.task_pass.size();
  std::vector<long long> signalled(r.n_counters + 1, 0), needed(r.n_counters + 1, 0);
  // level of each pass inside its unit = number of earlier passes with the same counter
  for (size_t i = 0; i < nt; ++i) {
    const int pi = r.task_pass[i];
    if (pi < 0 || pi >= (int)r.passes.size()) { fail("pass index", (long long)i, pi); continue; }
    const RaggedPassHost& h = r.passes[pi];
    const long long tiles = h.pp.groups * h.pp.tiles_per_group;
    if (r.task_tile[i] < 0 || r.task_tile[i] >= tiles) fail("tile index", (long long)i, r.task_tile[i]);
    if (h.wait_for > 0 && signalled[h.counter] < h.wait_for) fail("wait not satisfiable", (long long)i, h.wait_for);
    if (h.signal) ++signalled[h.counter];
    needed[h.counter] = std::max(needed[h.counter], (long long)h.wait_for);
  }
  for (int c = 0; c < r.n_counters; ++c)
    if (signalled[c] != needed[c]) fail("counter total", c, signalled[c] - needed[c]);
  for (size_t pi = 0; pi < r.passes.size(); ++pi) {
    const RaggedPassHost& h = r.passes[pi];
    const long long lim_in = h.pp.src == BUF_IN ? r.in_extent : h.pp.src == BUF_OUT ? r.out_extent : r.ws_elems;
    const long long lim_out = h.pp.dst == BUF_OUT ? r.out_extent : r.ws_elems;
    if (h.in_off < 0 || h.in_off + h.in_elems > lim_in) fail("input view", (long long)pi, h.in_off + h.in_elems - lim_in);
    if (h.out_off < 0 || h.out_off + h.out_elems > lim_out) fail("output view", (long long)pi, h.out_off + h.out_elems - lim_out);
    if (h.variant < 0 || h.variant >= kRaggedVariants) fail("variant", (long long)pi, h.variant);
    const long long esz = h.pp.src == BUF_WS ? elem_bytes(ws_fmt) : elem_bytes(h.pp.src_fmt);
    if (h.in_elems * esz > 0xffffffffLL) fail("input view above 4 GiB", (long long)pi, h.in_elems);
  }
  for (size_t k = 0; k + 1 < r.phase_begin.size(); ++k)
    if (r.phase_begin[k] > r.phase_begin[k + 1]) fail("phase order", (long long)k, r.phase_begin[k]);
  if (r.phase_begin.back() != (int)nt) fail("phase end", r.phase_begin.back(), (long long)nt);
  stats[0] = (long long)nt;
  stats[1] = (long long)r.passes.size();
  stats[2] = (long long)r.phase_begin.size() - 1;
  stats[3] = r.n_counters;
  stats[4] = r.ws_elems;
  stats[5] = (long long)r.tw.size();
  return bad;
}

// Real transforms through the emulated kernels, mirroring exec_real in the
// runtime: kind 1 = R2C (complex FFT of length n/2, then the post-pass), kind 2 =
// C2R (pre-pass, then the inverse complex FFT).  Returns the output exponent.
extern "C" void mcfft_r2c_post_0(RealArgs a);
extern "C" void mcfft_r2c_post_1(RealArgs a);
extern "C" void mcfft_r2c_post_2(RealArgs a);
extern "C" void mcfft_c2r_pre_0(RealArgs a);
extern "C" void mcfft_c2r_pre_1(RealArgs a);
extern "C" void mcfft_c2r_pre_2(RealArgs a);

// rank 1..3, dims: the real shape (last dimension n even); the complex plan runs
// over (outer dims, n/2)
extern "C" int emu_real_exec_nd(int kind, int rank, const long long* dims, long long batch, int in_fmt, int out_fmt,
                                int scaling, const void* in, void* out, int* out_exponent) {
  if (rank < 1 || rank > 3) return ST_INVALID_VALUE;
  const long long n = dims[rank - 1], h = n / 2;
  long long rows = 1, cdims[3] = {1, 1, 1};
  for (int d = 0; d < rank; ++d) {
    cdims[d] = dims[d];
    if (d < rank - 1) rows *= dims[d];
  }
  cdims[rank - 1] = h;
  const int cm = compute_mode(in_fmt, out_fmt);
  if (cm < 0 || n < 32 || (n & (n - 1))) return ST_UNSUPPORTED;
  const int tmp_fmt = kind == 1 ? out_fmt : (cm == 2 ? 2 : in_fmt);
  const long long tmp_elems = rows * h * batch;
  std::vector<unsigned char> tmp((size_t)(tmp_elems * elem_bytes(tmp_fmt)) + 16);
  const int sign = kind == 1 ? -1 : 1;
  std::vector<f32x2> tw((size_t)h + 1);
  for (long long k = 0; k <= h; ++k) twiddle_host(TwSpec{k, n, 1.0f}, sign, tw[(size_t)k]);
  void (*post[3])(RealArgs) = {mcfft_r2c_post_0, mcfft_r2c_post_1, mcfft_r2c_post_2};
  void (*pre[3])(RealArgs) = {mcfft_c2r_pre_0, mcfft_c2r_pre_1, mcfft_c2r_pre_2};
  RealArgs ra;
  ra.h = h;
  ra.batch = batch;
  ra.rows = rows;
  ra.ox = rank > 1 ? dims[0] : 1;
  ra.oy = rank > 2 ? dims[1] : 1;
  ra.tw = tw.data();
  const long long total = (kind == 1 ? (h / 2 + 1) : h) * rows * batch;
  const int grid = (int)std::min<long long>((total + 1023) / 1024, 3);
  int e = 0;
  auto regs = [&](const void* a, size_t abytes, const void* b, size_t bbytes) {
    emu::clear_ranges();
    emu::register_range(a, abytes);
    emu::register_range(b, bbytes);
    emu::register_range(tw.data(), tw.size() * sizeof(f32x2));
  };
  const size_t spec_bytes_in = (size_t)(rows * (h + 1) * batch * elem_bytes(in_fmt));
  const size_t spec_bytes_out = (size_t)(rows * (h + 1) * batch * elem_bytes(out_fmt));
  if (kind == 1) {
    int st = emu_exec(rank, cdims, batch, in_fmt, out_fmt, scaling, 0, in, tmp.data(), 0, &e);
    if (st) return st;
    regs(tmp.data(), (size_t)(tmp_elems * elem_bytes(tmp_fmt)), out, spec_bytes_out);
    ra.in = tmp.data();
    ra.out = out;
    ra.in_fmt = out_fmt;
    emu::launch(grid, 1024, 0, [&] { post[cm](ra); });
  } else {
    regs(in, spec_bytes_in, tmp.data(), (size_t)(tmp_elems * elem_bytes(tmp_fmt)));
    ra.in = in;
    ra.out = tmp.data();
    ra.in_fmt = in_fmt;
    emu::launch(grid, 1024, 0, [&] { pre[cm](ra); });
    int st = emu_exec(rank, cdims, batch, tmp_fmt, out_fmt, scaling, 1, tmp.data(), out, 0, &e);
    if (st) return st;
    e += 1;
    for (int& v : g_texps) v += 1;  // the pre-pass halves its input
  }
  if (out_exponent) *out_exponent = e;
  return 0;
}

extern "C" int emu_real_exec(int kind, long long n, long long batch, int in_fmt, int out_fmt, int scaling,
                             const void* in, void* out, int* out_exponent) {
  return emu_real_exec_nd(kind, 1, &n, batch, in_fmt, out_fmt, scaling, in, out, out_exponent);
}

// ---- advanced layouts: gather -> transform -> scatter, as the runtime's
// exec_layout (staging shared between the sides when both stage and the
// formats match) ----
extern "C" void mcfft_layout_gather(LayoutArgs a);
extern "C" void mcfft_layout_scatter(LayoutArgs a);

// test hooks: force the one-element-per-thread path; whether the last exec used
// the four-element path on either side
static bool g_layout_scalar_only = false;
static int g_layout_vec_used = 0;
extern "C" void emu_layout_set_scalar_only(int v) { g_layout_scalar_only = v != 0; }
extern "C" int emu_layout_vec_used() { return g_layout_vec_used; }

extern "C" int emu_layout_exec(int rank, const long long* dims, long long batch, const long long* in_strides,
                               long long in_dist, int in_planar, const long long* out_strides, long long out_dist,
                               int out_planar, int in_fmt, int out_fmt, int scaling, int inverse, const void* in_re,
                               const void* in_im, void* out_re, void* out_im, int* out_exponent) {
  const long long* st[2] = {in_strides, out_strides};
  const long long dist[2] = {in_dist, out_dist};
  const int planar[2] = {in_planar, out_planar}, fmt[2] = {in_fmt, out_fmt};
  for (int sd = 0; sd < 2; ++sd) {
    const int v = layout_validate(rank, dims, batch, st[sd], dist[sd], planar[sd]);
    if (v) return v;
  }
  if ((in_im != nullptr) != (in_planar != 0) || (out_im != nullptr) != (out_planar != 0)) return ST_INVALID_VALUE;
  g_layout_vec_used = 0;
  long long n = 1;
  for (int d = 0; d < rank; ++d) n *= dims[d];
  bool dense[2];
  for (int sd = 0; sd < 2; ++sd) dense[sd] = layout_is_dense(rank, dims, st[sd], dist[sd], planar[sd]);
  const bool share = !dense[0] && !dense[1] && in_fmt == out_fmt;
  std::vector<unsigned char> buf[2];
  for (int sd = 0; sd < 2; ++sd)
    if (!dense[sd] && !(sd == 1 && share)) buf[sd].resize((size_t)(n * batch * elem_bytes(fmt[sd])));
  void* stage[2] = {buf[0].data(), share ? buf[0].data() : buf[1].data()};
  auto run = [&](int sd, bool gather, const void* s0, const void* s1, void* d0, void* d1) {
    LayoutArgs la{};
    layout_fill(la, rank, dims, batch, st[sd], dist[sd], planar[sd], fmt[sd]);
    la.src[0] = s0;
    la.src[1] = s1;
    la.dst[0] = d0;
    la.dst[1] = d1;
    la.vec = g_layout_scalar_only ? 0 : (layout_vec_ok(la, gather ? s0 : d0, gather ? s1 : d1) ? 1 : 0);
    g_layout_vec_used |= la.vec;
    const long long ext = layout_extent(rank, dims, batch, st[sd], dist[sd]);
    const size_t side_bytes = (size_t)ext * (size_t)(elem_bytes(fmt[sd]) / (planar[sd] ? 2 : 1));
    const size_t dense_bytes = (size_t)(n * batch * elem_bytes(fmt[sd]));
    emu::clear_ranges();
    if (gather) {
      emu::register_range(s0, side_bytes);
      if (s1) emu::register_range(s1, side_bytes);
      emu::register_range(d0, dense_bytes);
    } else {
      emu::register_range(s0, dense_bytes);
      emu::register_range(d0, side_bytes);
      if (d1) emu::register_range(d1, side_bytes);
    }
    const long long need = ((la.vec ? la.total / 4 : la.total) + kLayoutThreads - 1) / kLayoutThreads;
    const int grid = (int)std::min<long long>(need, 5);  // fewer blocks than elements: grid-stride loop
    emu::launch(grid, kLayoutThreads, 0, [&] {
      if (gather) mcfft_layout_gather(la);
      else mcfft_layout_scatter(la);
    });
  };
  const void* src = in_re;
  void* dst = out_re;
  if (!dense[0]) {
    run(0, true, in_re, in_im, stage[0], nullptr);
    src = stage[0];
  }
  if (!dense[1]) dst = stage[1];
  const int rc = emu_exec(rank, dims, batch, in_fmt, out_fmt, scaling, inverse, src, dst, 0, out_exponent);
  if (rc) return rc;
  if (!dense[1]) run(1, false, stage[1], nullptr, out_re, out_im);
  return 0;
}

// plan_core.cpp -- see plan_core.h
#include "plan_core.h"

#include <cstdio>
#include <cstdlib>
#include <sstream>

namespace mcfft {

int elem_bytes(int fmt) { return fmt == 2 ? 8 : 4; }

int compute_mode(int in_fmt, int out_fmt) {
  if (in_fmt == 0 && out_fmt == 0) return 0;        // FP16 -> FP16
  if (in_fmt == 1 && out_fmt == 1) return 1;        // BF16 -> BF16
  if ((in_fmt == 0 || in_fmt == 2) && out_fmt == 2) return 2;  // FP16 or split -> split
  return -1;
}

int ilog2(long long v) {
  int r = 0;
  while ((1LL << r) < v) ++r;
  return ((1LL << r) == v) ? r : -1;
}

// L = 16^S * T with S >= 1, T in {1,2,4,8}
bool split_len(long long L, int& S, int& T) {
  int lg = ilog2(L);
  if (lg < 4 || lg > 16) return false;
  S = lg / 4;
  T = 1 << (lg % 4);
  return true;
}

enum { KIND_ROW = 0, KIND_ROWT = 1, KIND_COL = 2, KIND_COLF = 3, KIND_COLM = 4, KIND_ROWX = 5 };
constexpr int kKinds = 6;
static const KernelInfo* const kLists[kKinds] = {kRowKernels, kRowTKernels, kColKernels,
                                                 kColFKernels, kColMKernels, kRowXKernels};
static int kind_count(int kind) {
  const int counts[kKinds] = {kRowKernelCount, kRowTKernelCount, kColKernelCount,
                              kColFKernelCount, kColMKernelCount, kRowXKernelCount};
  return counts[kind];
}

// Default tile shapes that measured fastest on MI355X (autotuner sweep over
// batched 1D/2D/3D plans); other combinations use the first compiled variant.
struct PreferredShape {
  int kind, S, T, cm, Q, nw;
};
static const PreferredShape kPreferred[] = {
    {KIND_ROWT, 2, 2, 0, 16, 8}, {KIND_ROWT, 2, 2, 1, 16, 8}, {KIND_ROWT, 2, 2, 2, 16, 8},
    // 256-point rows, FP16/BF16: 16 rows on 4 waves (1D 256 x 131072 0.060 -> 0.049 ms)
    {KIND_ROW, 2, 1, 0, 16, 4},  {KIND_ROW, 2, 1, 1, 16, 4},
    // 512-point rows: 8 rows on 2 waves (1D 512 x 131072 FP16 0.117 -> 0.106 ms, split 0.209 -> 0.186)
    {KIND_ROW, 2, 2, 0, 8, 2},   {KIND_ROW, 2, 2, 1, 8, 2},   {KIND_ROW, 2, 2, 2, 8, 2},
    // 1024/2048-point rows: four MFMA tiles per wave (2048^2 x 8 row pass 256 -> ~110 us)
    {KIND_ROW, 2, 4, 0, 4, 2},   {KIND_ROW, 2, 4, 1, 4, 2},   {KIND_ROW, 2, 4, 2, 4, 2},
    {KIND_ROW, 2, 8, 0, 4, 4},   {KIND_ROW, 2, 8, 1, 4, 4},   {KIND_ROW, 2, 8, 2, 4, 4},
    {KIND_COL, 2, 1, 0, 64, 8},  {KIND_COL, 2, 1, 1, 64, 8},
    {KIND_COL, 2, 4, 0, 32, 16}, {KIND_COL, 2, 4, 1, 32, 16},  // 1024 points: 128-byte segments
    // 4096-point rows, FP16/BF16: 2 rows per tile (1D 4096 x 16384: 0.139 -> 0.118 ms)
    {KIND_ROW, 3, 1, 0, 2, 8},   {KIND_ROW, 3, 1, 1, 2, 8},
    // 8192-point rows, FP16/BF16: 2 rows per tile (1D 8192 x 4096: 0.101 -> 0.084 ms)
    {KIND_ROW, 3, 2, 0, 2, 16},  {KIND_ROW, 3, 2, 1, 2, 16},
    {KIND_ROW, 4, 1, 0, 1, 16},  {KIND_ROW, 4, 1, 1, 1, 16},   // single-pass 65536: 16 waves
};

// A column tile's buffer view spans (L - 1) * stride + Q elements: its byte
// offsets must fit the 32-bit buffer addressing.
static bool col_view_fits(long long L, long long stride, long long Q, int cm) {
  return ((L - 1) * stride + Q) * (cm == 2 ? 8 : 4) <= (1LL << 32);
}

// The preferred tile shape of (S, T, cm, kind), else the first compiled one.
// `ninst` > 0 (column and transposed-row kinds): the pass has that many
// instances per group, which the tile's Q must divide, so a narrow outer axis
// falls back to a variant with a smaller tile.
// a kernel built for the other scaling family (KernelInfo::mode) cannot serve the plan
bool kernel_serves(const KernelInfo& k, int scaling) {
  return k.mode == 0 || (k.mode == 2) == (scaling == SCALE_AUTO);
}

// ROWX (`qx` > 0): the tile must hold exactly qx rows (the outer radix).
static const KernelInfo* find_kernel(int S, int T, int cm, int kind, int scaling, long long ninst = 0, int qx = 0) {
  const KernelInfo* list = kLists[kind];
  const int count = kind_count(kind);
  auto divides = [&](const KernelInfo& k) {
    return (ninst <= 0 || ninst % k.Q == 0) && (qx <= 0 || k.Q == qx) && kernel_serves(k, scaling);
  };
  for (const PreferredShape& p : kPreferred) {
    if (p.kind != kind || p.S != S || p.T != T || p.cm != cm) continue;
    for (int i = 0; i < count; ++i) {
      const KernelInfo& k = list[i];
      if (k.S == S && k.T == T && k.cm == cm && k.Q == p.Q && k.nw == p.nw && divides(k)) return &k;
    }
  }
  for (int i = 0; i < count; ++i) {
    const KernelInfo& k = list[i];
    if (k.S == S && k.T == T && k.cm == cm && divides(k)) return &k;
  }
  return nullptr;
}

static int stage_shift(int radix, int scaling) {
  int lg = ilog2(radix);
  if (scaling == SCALE_FULL || scaling == SCALE_AUTO) return lg;
  if (scaling == SCALE_BALANCED) return lg / 2;
  return 0;
}

float pow2(int e) {  // 2^e, exact
  float s = 1.f;
  for (int i = 0; i < e; ++i) s *= 2.f;
  for (int i = 0; i > e; --i) s *= 0.5f;
  return s;
}

// Append the stage post-twiddle tables of one pass of length L (stages 16^S then T)
// and return the per-stage shifts.  Table of stage P, entry j*16+rr, is the
// input twiddle of stage P+1 for the element stage P writes at
// m' = (j/Ns)*Ns*16 + j%Ns + rr*Ns, times 2^-shift_P.
std::vector<int> add_stage_tables(std::vector<TwSpec>& tw, PassPlan& pp, int S, int T, int scaling) {
  const long long L = (1LL << (4 * S)) * T;
  std::vector<int> radix;
  for (int i = 0; i < S; ++i) radix.push_back(16);
  if (T > 1) radix.push_back(T);
  std::vector<int> shifts;
  for (int r : radix) shifts.push_back(stage_shift(r, scaling));
  pp.stage_tw_off = (long long)tw.size();
  long long ns = 1;
  for (size_t P = 0; P + 1 < radix.size(); ++P) {
    const int R = radix[P];  // always 16 (the tail is last)
    const long long nsn = ns * R;
    const int Rn = radix[P + 1];
    for (long long j = 0; j < L / R; ++j)
      for (long long rr = 0; rr < R; ++rr) {
        const long long mo = (j / ns) * ns * R + (j % ns) + rr * ns;
        const long long r2 = mo / (L / Rn), j2 = mo % (L / Rn), k2 = j2 % nsn;
        tw.push_back({(r2 * k2) % (nsn * Rn), nsn * Rn, pow2(-shifts[P])});
      }
    ns = nsn;
  }
  return shifts;
}

// Radix R of a 2D plan's outer-axis split (NX = R * NX/R, ROWX + COL), 0 for
// the plain row + column plan.  A column tile of NX points holds 128 KiB / NX /
// (4 or 8 bytes) columns: under 32 (FP16/BF16) or 16 (split) columns, i.e.
// row segments under 128 bytes, from 1024 (FP16/BF16) or 2048 (split) points.
// Measured on MI355X (NX, NY in 1024..4096, all formats): R = 4 is the fastest
// or within 4 % of it everywhere except NX = NY = 4096, where the 4-row tiles of
// 4096-point rows (164-194 VGPRs FP16/BF16, 139 KiB of LDS split: one
// workgroup per CU) meet the 1024-point column pass (128 KiB tiles): R = 2 there
// (4096^2 x 4 FP16 0.290 -> 0.281 ms, split 0.587 -> 0.533; 1024 x 4096 stays at
// R = 4, 0.259 against 0.305 ms).  ROWX rows need at least 512 points.  `forced` >= 0 (a tuned or requested structure) replaces
// the choice; MCFFT_2D_OUTER_RADIX overrides both.
static int outer_radix(long long NX, long long NY, int forced) {
  int r = 0;
  if (NX >= kOuterSplitMin && NY >= 512) r = (NX >= 4096 && NY >= 4096) ? 2 : 4;
  if (forced >= 0) r = forced;
  if (const char* e = getenv("MCFFT_2D_OUTER_RADIX")) {
    const int v = atoi(e);
    if (v == 0 || v == 1) r = 0;
    if (v == 2 || v == 4 || v == 8) r = v;
  }
  if (r < 2 || NY < 512 || NX / r < 16) return 0;
  return r > 8 ? 0 : r;
}

// 1D lengths of at least 2^this run as three passes (MCFFT_THREE_PASS_MIN_LG
// lowers it, down to 2^15, so that the structure is exercised at small sizes)
static int three_pass_min_lg() {
  int v = kThreePassMinLg;
  if (const char* e = getenv("MCFFT_THREE_PASS_MIN_LG")) {
    const int w = atoi(e);
    if (w >= 15 && w <= kThreePassMinLg) v = w;
  }
  return v;
}

// Three-pass 1D plan, one transform per chunk: N = N1 * N2 * N3 (n = n1*N2*N3 +
// n2*N3 + n3, output k = k1 + N1*k2 + N1*N2*k3).
//   A (COLF, in -> out):  N1-point columns (stride M = N2*N3) times W_N^(k1*m),
//                         m = n2*N3 + n3, in place
//   B (COLM, out -> ws):  N2-point columns (stride N3) of each row k1 times
//                         W_M^(k2*n3), stored at (k2*N1 + k1)*N3 + n3
//   C (ROWT, ws -> out):  N3-point rows (k2*N1 + k1) stored transposed at
//                         (k2*N1 + k1) + N1*N2*k3
// Pass B's store puts the rows pass C needs in the order of its transposed
// store, so every pass is one of the kernel family's addressing modes.
static int plan_three_pass(PlanCore& p, long long N, int cm, int scaling, int ws_fmt, int& total_shift) {
  const int lg = ilog2(N);
  if (lg > kMaxLg1D) return ST_UNSUPPORTED;
  const int lg3 = lg / 3, lg1 = (lg - lg3 + 1) / 2, lg2 = lg - lg3 - lg1;
  const long long N1 = 1LL << lg1, N2 = 1LL << lg2, N3 = 1LL << lg3, M = N2 * N3;
  int S1, T1, S2, T2, S3, T3;
  if (!split_len(N1, S1, T1) || !split_len(N2, S2, T2) || !split_len(N3, S3, T3)) return ST_UNSUPPORTED;
  const KernelInfo* ka = find_kernel(S1, T1, cm, KIND_COLF, scaling, M);
  const KernelInfo* kb = find_kernel(S2, T2, cm, KIND_COLM, scaling, N3);
  const KernelInfo* kc = find_kernel(S3, T3, cm, KIND_ROWT, scaling, N1 * N2);
  if (!ka || !kb || !kc) return ST_UNSUPPORTED;
  p.chunk = 1;
  p.n_chunks = p.batch;
  p.ws_elems = N;
  // A: columns of length N1 at stride M, one group per transform
  PassPlan a;
  a.k = ka;
  a.kind = KIND_COLF;
  a.L = N1;
  a.src = BUF_IN;
  a.dst = BUF_OUT;
  a.src_fmt = p.in_fmt;
  a.stride_in = a.stride_out = (int)M;
  a.in_sg = a.out_sg = N;
  a.in_st = a.out_st = ka->Q;
  a.n_inst = M;
  a.tiles_per_group = (int)(M / ka->Q);
  a.groups = 1;
  a.groups_per_transform = 1;
  a.shifts = add_stage_tables(p.tw, a, S1, T1, scaling);
  a.fourstep = true;
  a.out_scale = pow2(-a.shifts.back());
  a.fs_cols = (int)M;
  a.fs_tw_off = (long long)p.tw.size();
  if (cm == 2) {
    // W_N^(k1*m) = W_N^(k1*(m % 2^b)) * W_N^(k1*(m - m % 2^b)): two tables of
    // N1 * 2^b and N1 * M / 2^b entries instead of one of N
    const int b = (lg2 + lg3 + 1) / 2;
    a.fs_lo_bits = b;
    for (long long k1 = 0; k1 < N1; ++k1)
      for (long long c = 0; c < (1LL << b); ++c) p.tw.push_back({(k1 * c) % N, N, a.out_scale});
    for (long long k1 = 0; k1 < N1; ++k1)
      for (long long c = 0; c < (M >> b); ++c) p.tw.push_back({(k1 * (c << b)) % N, N, 1.f});
  }
  for (int s : a.shifts) total_shift += s;
  p.passes.push_back(a);
  // B: group k1 = row of M elements; N3 columns of length N2 (stride N3), stored
  // at stride N1*N3 from k1*N3
  PassPlan b;
  b.k = kb;
  b.kind = KIND_COLM;
  b.L = N2;
  b.src = BUF_OUT;
  b.dst = BUF_WS;
  b.src_fmt = p.out_fmt;
  b.stride_in = (int)N3;
  b.stride_out = (int)(N1 * N3);
  b.in_sg = M;
  b.out_sg = N3;
  b.in_st = b.out_st = kb->Q;
  b.n_inst = N3;
  b.tiles_per_group = (int)(N3 / kb->Q);
  b.groups = N1;
  b.groups_per_transform = N1;
  b.shifts = add_stage_tables(p.tw, b, S2, T2, scaling);
  b.fourstep = true;
  b.out_scale = pow2(-b.shifts.back());
  b.fs_cols = (int)N3;
  b.fs_tw_off = (long long)p.tw.size();
  if (cm == 2)
    for (long long k2 = 0; k2 < N2; ++k2)
      for (long long n3 = 0; n3 < N3; ++n3) p.tw.push_back({(k2 * n3) % M, M, b.out_scale});
  // reconcile: element (q = n3, m = n2) came from pass A's instance n2*N3 + n3
  b.rc_g = 0;
  b.rc_m = (int)N3;
  b.rc_q = 1;
  for (int s : b.shifts) total_shift += s;
  p.passes.push_back(b);
  // C: N1*N2 rows of N3 points, transposed store at stride N1*N2
  PassPlan c;
  c.k = kc;
  c.kind = KIND_ROWT;
  c.L = N3;
  c.src = BUF_WS;
  c.dst = BUF_OUT;
  c.src_fmt = ws_fmt;
  c.stride_out = (int)(N1 * N2);
  c.in_sg = c.out_sg = N;
  c.in_st = (long long)kc->Q * N3;
  c.out_st = kc->Q;
  c.n_inst = N1 * N2;
  c.tiles_per_group = (int)(N1 * N2 / kc->Q);
  c.groups = 1;
  c.groups_per_transform = 1;
  c.shifts = add_stage_tables(p.tw, c, S3, T3, scaling);
  c.out_scale = pow2(-c.shifts.back());
  for (int s : c.shifts) total_shift += s;
  p.passes.push_back(c);
  return ST_OK;
}

int plan_build(PlanCore& p, int rank, const long long* dims, long long batch, int in_fmt, int out_fmt, int scaling,
               long long ws_limit_bytes, const PlanOpts& opt) {
  p = PlanCore();
  p.opts.outer_radix = 0;
  p.opts.fourstep_lg2 = 0;
  p.opts.gathered = opt.gathered;
  if (rank < 1 || rank > 3 || batch < 1 || !dims) return ST_INVALID_VALUE;
  if (in_fmt < 0 || in_fmt > 2 || out_fmt < 0 || out_fmt > 2) return ST_INVALID_VALUE;
  if (scaling < 0 || scaling > 3) return ST_INVALID_VALUE;
  const int cm = compute_mode(in_fmt, out_fmt);
  if (cm < 0) return ST_UNSUPPORTED;
  p.rank = rank;
  p.batch = batch;
  p.in_fmt = in_fmt;
  p.out_fmt = out_fmt;
  p.cm = cm;
  p.scaling = scaling;
  p.n = 1;
  for (int d = 0; d < rank; ++d) {
    if (dims[d] < 1) return ST_INVALID_VALUE;
    if (ilog2(dims[d]) < 0) return ST_UNSUPPORTED;  // powers of two only
    p.dims[d] = dims[d];
    p.n *= dims[d];
  }
  const int ws_fmt = (cm == 2) ? 2 : in_fmt;
  int total_shift = 0;

  // contiguous rows of length L: `n_rows` rows per group; with per_transform
  // the groups are the chunk's transforms (rows = n_rows per transform)
  auto add_row = [&](long long L, long long n_rows, int src, int dst, int src_fmt, bool per_transform) -> int {
    int S, T;
    if (!split_len(L, S, T)) return ST_UNSUPPORTED;
    // one row per group (a gathered batch: each transform its own group, or a
    // single transform): a tile of several rows would run mostly empty
    const KernelInfo* k = n_rows == 1 ? find_kernel(S, T, cm, KIND_ROW, scaling, 0, 1) : nullptr;
    if (!k) k = find_kernel(S, T, cm, KIND_ROW, scaling);
    if (!k) return ST_UNSUPPORTED;
    PassPlan pp;
    pp.k = k;
    pp.kind = KIND_ROW;
    pp.L = L;
    pp.src = src;
    pp.dst = dst;
    pp.src_fmt = src_fmt;
    pp.in_sg = pp.out_sg = per_transform ? p.n : 0;
    pp.in_st = pp.out_st = (long long)k->Q * L;
    pp.n_inst = n_rows;
    pp.tiles_per_group = (int)((n_rows + k->Q - 1) / k->Q);
    pp.groups = per_transform ? batch : 1;
    pp.groups_per_transform = per_transform ? 1 : 0;
    pp.shifts = add_stage_tables(p.tw, pp, S, T, scaling);
    pp.out_scale = pow2(-pp.shifts.back());
    for (int s : pp.shifts) total_shift += s;
    p.passes.push_back(pp);
    return ST_OK;
  };
  // strided columns: L points at stride `stride`, `ncols` consecutive columns per
  // group, groups `groups` apart by `gstride` elements
  auto add_col = [&](long long L, long long stride, long long ncols, long long groups, long long gstride, int src,
                     int dst, int src_fmt) -> int {
    int S, T;
    if (!split_len(L, S, T)) return ST_UNSUPPORTED;
    const KernelInfo* k = find_kernel(S, T, cm, KIND_COL, scaling, ncols);
    if (!k || ncols % k->Q != 0 || !col_view_fits(L, stride, k->Q, cm)) return ST_UNSUPPORTED;
    PassPlan pp;
    pp.k = k;
    pp.kind = KIND_COL;
    pp.L = L;
    pp.src = src;
    pp.dst = dst;
    pp.src_fmt = src_fmt;
    pp.stride_in = pp.stride_out = (int)stride;
    pp.in_sg = pp.out_sg = gstride;
    pp.in_st = pp.out_st = k->Q;
    pp.n_inst = ncols;
    pp.tiles_per_group = (int)(ncols / k->Q);
    pp.groups = groups;
    pp.groups_per_transform = groups / batch;
    pp.shifts = add_stage_tables(p.tw, pp, S, T, scaling);
    pp.out_scale = pow2(-pp.shifts.back());
    for (int s : pp.shifts) total_shift += s;
    p.passes.push_back(pp);
    return ST_OK;
  };

  // Two-pass four-step FFT of every row (length N) of the chunk's transforms,
  // `rows` rows per transform (1D: 1; 2D/3D rows too long for one pass: the
  // product of the outer sides): columns of length N1 = N/N2 (stride N2) with
  // twiddle W_N^(k1*n2) (in -> ws), then rows of length N2 = 2^lg2 stored
  // transposed X[k1 + N1*k2] (ws -> out).  The caller sets the chunk and a
  // workspace of one chunk.
  auto add_fourstep = [&](long long N, long long rows, int lg2) -> int {
    const int lg = ilog2(N), lg1 = lg - lg2;
    p.opts.fourstep_lg2 = lg2;
    if (lg1 > 12 || lg2 > 12) return ST_UNSUPPORTED;
    const long long N1 = 1LL << lg1, N2 = 1LL << lg2;
    int S1, T1, S2, T2;
    if (!split_len(N1, S1, T1) || !split_len(N2, S2, T2)) return ST_UNSUPPORTED;
    const KernelInfo* kc = find_kernel(S1, T1, cm, KIND_COLF, scaling, N2);
    const KernelInfo* kr = find_kernel(S2, T2, cm, KIND_ROWT, scaling, N1);
    if (!kc || !kr || N2 % kc->Q != 0 || N1 % kr->Q != 0) return ST_UNSUPPORTED;
    // pass A (columns + four-step twiddle): in -> ws, one group per row
    PassPlan a;
    a.k = kc;
    a.kind = KIND_COLF;
    a.L = N1;
    a.src = BUF_IN;
    a.dst = BUF_WS;
    a.src_fmt = in_fmt;
    a.stride_in = a.stride_out = (int)N2;
    a.in_sg = a.out_sg = N;
    a.in_st = a.out_st = kc->Q;
    a.n_inst = N2;
    a.tiles_per_group = (int)(N2 / kc->Q);
    a.groups = p.chunk * rows;
    a.groups_per_transform = rows;
    a.shifts = add_stage_tables(p.tw, a, S1, T1, scaling);
    a.fourstep = true;
    // four-step twiddle W_N^(k1*n2) times the last stage's scale: computed in the
    // kernel for FP16/BF16, a direct table (k1 < N1, n2 < N2) for split
    const float sl = pow2(-a.shifts.back());
    a.out_scale = sl;
    a.fs_cols = (int)N2;
    a.fs_tw_off = (long long)p.tw.size();
    if (cm == 2)
      for (long long k1 = 0; k1 < N1; ++k1)
        for (long long n2 = 0; n2 < N2; ++n2) p.tw.push_back({(k1 * n2) % N, N, sl});
    for (int sh : a.shifts) total_shift += sh;
    p.passes.push_back(a);
    // pass B (rows, transposed store): ws -> out; AUTO reconcile: element (q, m)
    // of row g came from pass A's instance m of that row ((g % rows) * N2 + m
    // of the transform)
    PassPlan b;
    b.k = kr;
    b.kind = KIND_ROWT;
    b.L = N2;
    b.src = BUF_WS;
    b.dst = BUF_OUT;
    b.src_fmt = ws_fmt;
    b.stride_out = (int)N1;
    b.in_sg = b.out_sg = N;
    b.in_st = (long long)kr->Q * N2;
    b.out_st = kr->Q;
    b.n_inst = N1;
    b.tiles_per_group = (int)(N1 / kr->Q);
    b.groups = p.chunk * rows;
    b.groups_per_transform = rows;
    b.rc_g = (int)N2;
    b.shifts = add_stage_tables(p.tw, b, S2, T2, scaling);
    b.out_scale = pow2(-b.shifts.back());
    for (int sh : b.shifts) total_shift += sh;
    p.passes.push_back(b);
    return ST_OK;
  };
  // default split of a two-pass row FFT of length N: N2 = 2^ceil(lg/2), at most 2^11
  auto fourstep_lg2 = [](long long N) {
    const int lg = ilog2(N);
    return (lg + 1) / 2 > 11 ? 11 : (lg + 1) / 2;
  };
  // whether rows of length L run in one pass (a ROW kernel exists)
  auto row_single = [&](long long L) {
    int S, T;
    return split_len(L, S, T) && find_kernel(S, T, cm, KIND_ROW, scaling) != nullptr;
  };

  int st = ST_OK;
  if (rank == 1) {
    const long long N = p.dims[0];
    const int lg = ilog2(N);
    int S0, T0;
    // single pass (up to 2^15 / 2^16 in FP16/BF16), unless a two-pass split is requested
    const bool two_pass = opt.fourstep_lg2 > 0 && lg >= 12;
    if (!two_pass && split_len(N, S0, T0) && find_kernel(S0, T0, cm, KIND_ROW, scaling)) {
      p.chunk = batch;
      p.n_chunks = 1;
      st = add_row(N, opt.gathered ? 1 : batch, BUF_IN, BUF_OUT, in_fmt, opt.gathered);
    } else if (lg >= three_pass_min_lg()) {
      st = plan_three_pass(p, N, cm, scaling, ws_fmt, total_shift);
    } else {
      // four-step: N = N1 * N2; columns of length N1 (stride N2) with twiddle
      // W_N^(k1*n2), then rows of length N2 stored transposed X[k1 + N1*k2]
      int lg2 = fourstep_lg2(N);
      // a requested split: N2 = 2^4 .. 2^11 (ROWT), N1 = 2^4 .. 2^12 (COLF)
      if (opt.fourstep_lg2 >= 4 && opt.fourstep_lg2 <= 11 && lg - opt.fourstep_lg2 >= 4 &&
          lg - opt.fourstep_lg2 <= 12)
        lg2 = opt.fourstep_lg2;
      long long limit = ws_limit_bytes > 0 ? ws_limit_bytes : (256LL << 20);
      long long chunk = limit / (N * elem_bytes(ws_fmt));
      if (chunk < 1) chunk = 1;
      if (chunk > batch) chunk = batch;
      p.chunk = chunk;
      p.n_chunks = (batch + chunk - 1) / chunk;
      p.ws_elems = chunk * N;
      st = add_fourstep(N, 1, lg2);
    }
  } else if (rank == 2) {
    // transforms run in chunks whose data fits the Infinity Cache, so the
    // column pass reads what the row pass just wrote from the cache
    const long long NX = p.dims[0], NY = p.dims[1];
    const long long limit = ws_limit_bytes > 0 ? ws_limit_bytes : (128LL << 20);
    long long chunk = limit / (p.n * elem_bytes(p.out_fmt));
    if (chunk < 1) chunk = 1;
    if (chunk > batch) chunk = batch;
    p.chunk = chunk;
    p.n_chunks = (batch + chunk - 1) / chunk;
    const int R = outer_radix(NX, NY, opt.outer_radix);
    int SY, TY, SX, TX;
    const KernelInfo* kx = nullptr;
    const KernelInfo* kcx = nullptr;
    if (R > 1 && split_len(NY, SY, TY) && split_len(NX / R, SX, TX)) {
      kx = find_kernel(SY, TY, cm, KIND_ROWX, scaling, 0, R);
      kcx = find_kernel(SX, TX, cm, KIND_COL, scaling, R * NY);
    }
    if (kx && kcx && col_view_fits(NX / R, R * NY, kcx->Q, cm)) {
      p.opts.outer_radix = R;
      // Outer axis split NX = R * (NX/R), rows r = r1 + (NX/R)*r2:
      //   ROWX (in -> ws): tile r1 takes rows r1 + (NX/R)*r2 (r2 < R), their
      //     R-point DFT times W_NX^(r1*k2), then the NY-point row FFTs; result k2
      //     stored as row R*r1 + k2
      //   COL  (ws -> out): NX/R-point columns at stride R rows (column
      //     k2*NY + c of the R*NY columns is rows R*r1 + k2), output row k2 + R*k1
      // The tiles of ROWX write rows other tiles read, so it runs out of place
      // through a workspace (one chunk).
      p.ws_elems = chunk * p.n;
      PassPlan a;
      a.k = kx;
      a.kind = KIND_ROWX;
      a.L = NY;
      a.src = BUF_IN;
      a.dst = BUF_WS;
      a.src_fmt = in_fmt;
      a.stride_in = (int)(NX / R * NY);
      a.stride_out = (int)NY;
      a.in_sg = a.out_sg = p.n;
      a.in_st = NY;
      a.out_st = R * NY;
      a.n_inst = NX;
      a.tiles_per_group = (int)(NX / R);
      a.groups = batch;
      a.groups_per_transform = 1;
      a.shifts = add_stage_tables(p.tw, a, SY, TY, scaling);
      a.out_scale = pow2(-a.shifts.back());
      a.shifts.push_back(stage_shift(R, scaling));
      if (cm == 2) {
        a.x_tw_off = (long long)p.tw.size();
        for (long long r1 = 0; r1 < NX / R; ++r1)
          for (long long k2 = 0; k2 < R; ++k2) p.tw.push_back({(r1 * k2) % NX, NX, 1.f});
      }
      for (int sh : a.shifts) total_shift += sh;
      p.passes.push_back(a);
      st = add_col(NX / R, R * NY, R * NY, batch, p.n, BUF_WS, BUF_OUT, ws_fmt);
      if (st == ST_OK) {
        // AUTO reconcile: element (q, m) of the column pass is row R*m + k2, made
        // by ROWX tile m (first-pass tile = (m*R) >> log2(R))
        PassPlan& c = p.passes.back();
        c.rc_g = 0;
        c.rc_m = R;
        c.rc_q = 0;
      }
    } else if (row_single(NY)) {
      st = add_row(NY, NX, BUF_IN, BUF_OUT, in_fmt, true);
      if (st == ST_OK) st = add_col(NX, NY, NY, batch, NX * NY, BUF_OUT, BUF_OUT, p.out_fmt);
    } else {
      // rows too long for one pass (split 2^15, 2^16; FP16/BF16 above 2^16): a
      // two-pass four-step over every row, then the columns
      p.ws_elems = chunk * p.n;
      st = add_fourstep(NY, NX, fourstep_lg2(NY));
      if (st == ST_OK) st = add_col(NX, NY, NY, batch, NX * NY, BUF_OUT, BUF_OUT, p.out_fmt);
    }
  } else {
    const long long NX = p.dims[0], NY = p.dims[1], NZ = p.dims[2];
    p.chunk = batch;
    p.n_chunks = 1;
    if (row_single(NZ)) {
      st = add_row(NZ, NX * NY, BUF_IN, BUF_OUT, in_fmt, true);
    } else {  // innermost rows too long for one pass: four-step over every row
      p.ws_elems = batch * p.n;
      st = add_fourstep(NZ, NX * NY, fourstep_lg2(NZ));
    }
    if (st == ST_OK) st = add_col(NY, NZ, NZ, batch * NX, NY * NZ, BUF_OUT, BUF_OUT, p.out_fmt);
    if (st == ST_OK) st = add_col(NX, NY * NZ, NY * NZ, batch, NX * NY * NZ, BUF_OUT, BUF_OUT, p.out_fmt);
  }
  if (st != ST_OK) return st;
  p.out_exponent = total_shift;
  return ST_OK;
}

std::string opts_string(const PlanOpts& o) {
  return "R=" + std::to_string(o.outer_radix) + ";L2=" + std::to_string(o.fourstep_lg2);
}

bool parse_opts(const std::string& s, PlanOpts& o) {
  return sscanf(s.c_str(), "R=%d;L2=%d", &o.outer_radix, &o.fourstep_lg2) == 2;
}

static const char kTuneCacheMagic[] = "mcfft-tune-cache 2";

std::map<std::string, std::string> tune_cache_parse(const std::string& text, const std::string& build_id) {
  std::map<std::string, std::string> out;
  std::istringstream in(text);
  std::string line;
  if (!std::getline(in, line) || line != std::string(kTuneCacheMagic) + " " + build_id) return out;
  while (std::getline(in, line)) {
    const size_t eq = line.find('=');
    if (eq != std::string::npos && eq > 0) out[line.substr(0, eq)] = line.substr(eq + 1);
  }
  return out;
}

std::string tune_cache_format(const std::map<std::string, std::string>& entries, const std::string& build_id) {
  std::string s = std::string(kTuneCacheMagic) + " " + build_id + "\n";
  for (const auto& kv : entries) s += kv.first + "=" + kv.second + "\n";
  return s;
}

bool parse_tune_choice(const std::string& value, TuneChoice& c) {
  const size_t k = value.find(";K=");
  if (k == std::string::npos || !parse_opts(value.substr(0, k), c.opts)) return false;
  c.kernels.clear();
  std::stringstream ss(value.substr(k + 3));
  std::string n;
  while (std::getline(ss, n, ',')) c.kernels.push_back(n);
  return !c.kernels.empty();
}

std::string format_tune_choice(const TuneChoice& c) {
  std::string s = opts_string(c.opts) + ";K=";
  for (size_t i = 0; i < c.kernels.size(); ++i) s += (i ? "," : "") + c.kernels[i];
  return s;
}

std::vector<PlanOpts> plan_structures(int rank, const long long* dims, long long batch, int in_fmt, int out_fmt,
                                      int scaling, long long ws_limit_bytes) {
  std::vector<PlanOpts> out, tries;
  tries.push_back(PlanOpts());
  if (rank == 2) {
    for (int r : {0, 2, 4, 8}) {
      PlanOpts o;
      o.outer_radix = r;
      tries.push_back(o);
    }
  } else if (rank == 1 && dims[0] > 0) {
    // the default split and its neighbours (N1 = N/N2 one factor of 2 either way);
    // from 2^12 up, lengths with a single-pass kernel may run as two passes too
    const int lg = ilog2(dims[0]), d = (lg + 1) / 2 > 11 ? 11 : (lg + 1) / 2;
    for (int l2 = d - 1; l2 <= d + 1; ++l2) {
      PlanOpts o;
      o.fourstep_lg2 = l2;
      tries.push_back(o);
    }
  }
  for (const PlanOpts& t : tries) {
    PlanCore c;
    if (plan_build(c, rank, dims, batch, in_fmt, out_fmt, scaling, ws_limit_bytes, t) != ST_OK) continue;
    bool seen = false;
    for (const PlanOpts& o : out) seen |= o.outer_radix == c.opts.outer_radix && o.fourstep_lg2 == c.opts.fourstep_lg2;
    if (!seen) out.push_back(c.opts);
  }
  return out;
}

std::vector<const KernelInfo*> pass_candidates(const PassPlan& pp, int cm, int scaling) {
  std::vector<const KernelInfo*> out;
  if (!pp.k) return out;
  for (int i = 0; i < kind_count(pp.kind); ++i) {
    const KernelInfo& k = kLists[pp.kind][i];
    // a ROWX tile's row count is the plan's outer radix: only its wave counts vary
    if (pp.kind == KIND_ROWX && k.Q != pp.k->Q) continue;
    if (k.S == pp.k->S && k.T == pp.k->T && k.cm == cm && kernel_serves(k, scaling)) out.push_back(&k);
  }
  return out;
}

bool set_pass_kernel(PassPlan& pp, const KernelInfo* k) {
  const long long Q = k->Q;
  switch (pp.kind) {
    case KIND_ROW:
      pp.in_st = pp.out_st = Q * pp.L;
      pp.tiles_per_group = (int)((pp.n_inst + Q - 1) / Q);
      break;
    case KIND_ROWX:
      if (Q != pp.k->Q) return false;  // the tile's rows are the plan's outer radix
      break;
    case KIND_ROWT:
      if (pp.n_inst % Q) return false;
      pp.in_st = Q * pp.L;
      pp.out_st = Q;
      pp.tiles_per_group = (int)(pp.n_inst / Q);
      break;
    default:
      if (pp.n_inst % Q || !col_view_fits(pp.L, pp.stride_in, Q, k->cm)) return false;
      pp.in_st = pp.out_st = Q;
      pp.tiles_per_group = (int)(pp.n_inst / Q);
      break;
  }
  pp.k = k;
  return true;
}

std::string plan_signature(const PlanCore& p) {
  std::ostringstream o;
  o << p.rank << "|" << p.dims[0];
  for (int d = 1; d < p.rank; ++d) o << "x" << p.dims[d];
  o << "|" << p.batch << "|" << p.in_fmt << "|" << p.out_fmt << "|" << p.scaling;
  return o.str();
}

std::string PlanCore::describe() const {
  std::ostringstream o;
  o << "rank=" << rank << " dims=" << dims[0];
  for (int d = 1; d < rank; ++d) o << "x" << dims[d];
  o << " batch=" << batch << " cm=" << cm << " in=" << in_fmt << " out=" << out_fmt << " exp=" << out_exponent
    << " chunk=" << chunk << "x" << n_chunks << " ws=" << ws_elems << " struct=" << opts_string(opts) << "\n";
  for (const auto& pp : passes) {
    o << "  pass " << pp.k->name << " src=" << pp.src << " dst=" << pp.dst << " groups=" << pp.groups
      << " tiles/group=" << pp.tiles_per_group << " n_inst=" << pp.n_inst << " stride(in,out)=(" << pp.stride_in
      << "," << pp.stride_out << ") lds=" << pp.k->lds_bytes
      << (pp.fourstep ? " fourstep" : "") << " shifts=";
    for (int s : pp.shifts) o << s << ",";
    o << "\n";
  }
  return o.str();
}

}  // namespace mcfft

// ragged_plan.cpp -- see ragged_plan.h
#include "ragged_plan.h"

#include <algorithm>
#include <map>
#include <sstream>
#include <tuple>

namespace mcfft {

namespace {

enum { K_ROW = 0, K_ROWT = 1, K_COL = 2, K_COLF = 3 };

int find_variant(int cm, int S, int T, int kind) {
  for (int i = 0; i < 30; ++i) {
    const KernelInfo& k = kRaggedVariantInfo[cm][i];
    if (k.S == S && k.T == T && k.kind == kind) return i;
  }
  return -1;
}

// Passes of one transform shape, with tables appended to r.tw once per shape.
struct ShapePasses {
  std::vector<RaggedPassHost> passes;  // offsets relative to the transform
  int exponent = 0;
  long long ws = 0;                    // workspace elements per transform
  int n_rows_fusable = 0;              // > 0: single ROW pass over rows of this length
};

int build_shape(RaggedPlanCore& r, int rank, const long long* dims, long long rows, ShapePasses& out) {
  const int cm = r.cm;
  int total = 0;
  auto row_pass = [&](long long L, long long n_rows, int src, int dst, int src_fmt) -> int {
    int S, T;
    if (!split_len(L, S, T)) return ST_UNSUPPORTED;
    const int v = find_variant(cm, S, T, K_ROW);
    if (v < 0) return ST_UNSUPPORTED;
    const KernelInfo& k = kRaggedVariantInfo[cm][v];
    RaggedPassHost h;
    h.variant = v;
    h.pp.src = src;
    h.pp.dst = dst;
    h.pp.src_fmt = src_fmt;
    h.pp.in_st = h.pp.out_st = (long long)k.Q * L;
    h.pp.n_inst = n_rows;
    h.pp.tiles_per_group = (int)((n_rows + k.Q - 1) / k.Q);
    h.pp.shifts = add_stage_tables(r.tw, h.pp, S, T, r.scaling);
    h.pp.out_scale = pow2(-h.pp.shifts.back());
    for (int s : h.pp.shifts) total += s;
    out.passes.push_back(h);
    return ST_OK;
  };
  auto col_pass = [&](long long L, long long stride, long long ncols, long long groups, long long gstride, int src,
                      int dst, int src_fmt) -> int {
    int S, T;
    if (!split_len(L, S, T)) return ST_UNSUPPORTED;
    const int v = find_variant(cm, S, T, K_COL);
    if (v < 0) return ST_UNSUPPORTED;
    const KernelInfo& k = kRaggedVariantInfo[cm][v];
    // the tile's view spans (L - 1) * stride + Q elements of 32-bit byte offsets
    if (ncols % k.Q != 0 || ((L - 1) * stride + k.Q) * (cm == 2 ? 8 : 4) > (1LL << 32)) return ST_UNSUPPORTED;
    RaggedPassHost h;
    h.variant = v;
    h.pp.src = src;
    h.pp.dst = dst;
    h.pp.src_fmt = src_fmt;
    h.pp.stride_in = h.pp.stride_out = (int)stride;
    h.pp.in_sg = h.pp.out_sg = gstride;
    h.pp.in_st = h.pp.out_st = k.Q;
    h.pp.n_inst = ncols;
    h.pp.tiles_per_group = (int)(ncols / k.Q);
    h.pp.groups = groups;
    h.pp.shifts = add_stage_tables(r.tw, h.pp, S, T, r.scaling);
    h.pp.out_scale = pow2(-h.pp.shifts.back());
    for (int s : h.pp.shifts) total += s;
    out.passes.push_back(h);
    return ST_OK;
  };
  const int ws_fmt = (cm == 2) ? 2 : r.in_fmt;  // split compute keeps hi+lo in the workspace
  int st = ST_OK;
  if (rank == 1 && dims[0] <= 4096) {
    st = row_pass(dims[0], rows, BUF_IN, BUF_OUT, r.in_fmt);
    out.n_rows_fusable = (int)dims[0];
  } else if (rank == 1) {
    const long long N = dims[0];
    const int lg = ilog2(N);
    if (lg < 0) return ST_UNSUPPORTED;
    int lg2 = (lg + 1) / 2;
    if (lg2 > 9) lg2 = 9;
    const int lg1 = lg - lg2;
    const long long N1 = 1LL << lg1, N2 = 1LL << lg2;
    int S1, T1, S2, T2;
    if (lg1 > 9 || !split_len(N1, S1, T1) || !split_len(N2, S2, T2)) return ST_UNSUPPORTED;
    const int vc = find_variant(cm, S1, T1, K_COLF), vr = find_variant(cm, S2, T2, K_ROWT);
    if (vc < 0 || vr < 0) return ST_UNSUPPORTED;
    const KernelInfo& kc = kRaggedVariantInfo[cm][vc];
    const KernelInfo& kr = kRaggedVariantInfo[cm][vr];
    if (N2 % kc.Q != 0 || N1 % kr.Q != 0) return ST_UNSUPPORTED;
    out.ws = N;
    RaggedPassHost a;
    a.variant = vc;
    a.pp.src = BUF_IN;
    a.pp.dst = BUF_WS;
    a.pp.src_fmt = r.in_fmt;
    a.pp.stride_in = a.pp.stride_out = (int)N2;
    a.pp.in_st = a.pp.out_st = kc.Q;
    a.pp.n_inst = N2;
    a.pp.tiles_per_group = (int)(N2 / kc.Q);
    a.pp.shifts = add_stage_tables(r.tw, a.pp, S1, T1, r.scaling);
    a.pp.fourstep = true;
    // four-step twiddle W_N^(k1*n2) times the last stage's scale: computed in the
    // kernel for FP16/BF16, a direct table (k1 < N1, n2 < N2) for split
    const float sl = pow2(-a.pp.shifts.back());
    a.pp.out_scale = sl;
    a.pp.fs_cols = (int)N2;
    a.pp.fs_tw_off = (long long)r.tw.size();
    if (cm == 2)
      for (long long k1 = 0; k1 < N1; ++k1)
        for (long long n2 = 0; n2 < N2; ++n2) r.tw.push_back({(k1 * n2) % N, N, sl});
    for (int s : a.pp.shifts) total += s;
    out.passes.push_back(a);
    RaggedPassHost b;
    b.variant = vr;
    b.pp.src = BUF_WS;
    b.pp.dst = BUF_OUT;
    b.pp.src_fmt = ws_fmt;
    b.pp.stride_out = (int)N1;
    b.pp.in_st = (long long)kr.Q * N2;
    b.pp.out_st = kr.Q;
    b.pp.n_inst = N1;
    b.pp.tiles_per_group = (int)(N1 / kr.Q);
    b.pp.shifts = add_stage_tables(r.tw, b.pp, S2, T2, r.scaling);
    b.pp.out_scale = pow2(-b.pp.shifts.back());
    for (int s : b.pp.shifts) total += s;
    out.passes.push_back(b);
  } else if (rank == 2) {
    const long long NX = dims[0], NY = dims[1];
    st = row_pass(NY, NX, BUF_IN, BUF_OUT, r.in_fmt);
    if (st == ST_OK) st = col_pass(NX, NY, NY, 1, NX * NY, BUF_OUT, BUF_OUT, r.out_fmt);
  } else if (rank == 3) {
    const long long NX = dims[0], NY = dims[1], NZ = dims[2];
    st = row_pass(NZ, NX * NY, BUF_IN, BUF_OUT, r.in_fmt);
    if (st == ST_OK) st = col_pass(NY, NZ, NZ, NX, NY * NZ, BUF_OUT, BUF_OUT, r.out_fmt);
    if (st == ST_OK) st = col_pass(NX, NY * NZ, NY * NZ, 1, NX * NY * NZ, BUF_OUT, BUF_OUT, r.out_fmt);
  } else {
    return ST_INVALID_VALUE;
  }
  out.exponent = total;
  return st;
}

}  // namespace

int ragged_build(RaggedPlanCore& r, int n_items, const RaggedItemSpec* items, int in_fmt, int out_fmt, int scaling) {
  r = RaggedPlanCore();
  if (n_items < 1 || !items) return ST_INVALID_VALUE;
  if (in_fmt < 0 || in_fmt > 2 || out_fmt < 0 || out_fmt > 2 || scaling < 0 || scaling > 3) return ST_INVALID_VALUE;
  if (scaling == SCALE_AUTO) return ST_UNSUPPORTED;  // ragged items are statically scaled
  const int cm = compute_mode(in_fmt, out_fmt);
  if (cm < 0) return ST_UNSUPPORTED;
  r.cm = cm;
  r.in_fmt = in_fmt;
  r.out_fmt = out_fmt;
  r.scaling = scaling;
  r.n_items = n_items;
  r.exponent.assign(n_items, 0);

  // 1. units: runs of back-to-back 1D items of one length <= 4096 become one unit
  struct Unit {
    int first, count, rank;
    long long dims[3];
    long long n;
  };
  std::vector<Unit> units;
  for (int i = 0; i < n_items; ++i) {
    const RaggedItemSpec& it = items[i];
    if (it.rank < 1 || it.rank > 3 || it.in_off < 0 || it.out_off < 0) return ST_INVALID_VALUE;
    long long n = 1;
    for (int d = 0; d < it.rank; ++d) {
      if (it.dims[d] < 1) return ST_INVALID_VALUE;
      if (ilog2(it.dims[d]) < 0) return ST_UNSUPPORTED;
      n *= it.dims[d];
    }
    r.in_extent = std::max(r.in_extent, it.in_off + n);
    r.out_extent = std::max(r.out_extent, it.out_off + n);
    if (!units.empty()) {
      Unit& u = units.back();
      const RaggedItemSpec& last = items[u.first + u.count - 1];
      if (u.rank == 1 && it.rank == 1 && u.dims[0] == it.dims[0] && it.dims[0] <= 4096 &&
          it.in_off == last.in_off + it.dims[0] && it.out_off == last.out_off + it.dims[0]) {
        ++u.count;
        continue;
      }
    }
    Unit u{i, 1, it.rank, {it.dims[0], it.rank > 1 ? it.dims[1] : 1, it.rank > 2 ? it.dims[2] : 1}, n};
    units.push_back(u);
  }

  // 2. passes per unit; tables shared by units of one shape (fused row counts differ
  //    only in tile counts, so rows are patched per unit)
  std::map<std::tuple<int, long long, long long, long long>, ShapePasses> shapes;
  std::vector<std::vector<RaggedPassHost>> unit_passes(units.size());
  std::vector<long long> unit_cost(units.size());
  for (size_t ui = 0; ui < units.size(); ++ui) {
    const Unit& u = units[ui];
    const auto key = std::make_tuple(u.rank, u.dims[0], u.dims[1], u.dims[2]);
    auto itS = shapes.find(key);
    if (itS == shapes.end()) {
      ShapePasses sp;
      int st = build_shape(r, u.rank, u.dims, 1, sp);
      if (st != ST_OK) return st;
      itS = shapes.emplace(key, sp).first;
    }
    const ShapePasses& sp = itS->second;
    std::vector<RaggedPassHost> ps = sp.passes;
    const RaggedItemSpec& it0 = items[u.first];
    const long long ws_off = r.ws_elems;
    r.ws_elems += sp.ws;
    const int counter = (ps.size() > 1) ? r.n_counters++ : -1;
    int done_tiles = 0;
    for (size_t k = 0; k < ps.size(); ++k) {
      RaggedPassHost& h = ps[k];
      if (sp.n_rows_fusable > 0) {  // rows = items in the unit
        const KernelInfo& kv = kRaggedVariantInfo[r.cm][h.variant];
        h.pp.n_inst = u.count;
        h.pp.tiles_per_group = (int)((u.count + kv.Q - 1) / kv.Q);
      }
      auto off_of = [&](int buf) -> long long { return buf == BUF_IN ? it0.in_off : buf == BUF_OUT ? it0.out_off : ws_off; };
      auto ext_of = [&](int buf) -> long long { return buf == BUF_WS ? sp.ws : u.n * u.count; };
      h.in_off = off_of(h.pp.src);
      h.out_off = off_of(h.pp.dst);
      h.in_elems = ext_of(h.pp.src);
      h.out_elems = ext_of(h.pp.dst);
      h.counter = counter < 0 ? 0 : counter;
      h.wait_for = done_tiles;
      const int tiles = (int)(h.pp.groups * h.pp.tiles_per_group);
      h.signal = (counter >= 0 && k + 1 < ps.size()) ? 1 : 0;
      done_tiles += tiles;
    }
    unit_cost[ui] = u.n * u.count;
    for (int k = 0; k < u.count; ++k) r.exponent[u.first + k] = sp.exponent;
    unit_passes[ui] = ps;
  }

  // 3. tasks: pass 0 of every unit (largest first), then pass 1, then pass 2
  std::vector<size_t> order(units.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return unit_cost[a] > unit_cost[b]; });
  std::vector<std::vector<int>> pass_index(units.size());
  for (size_t ui = 0; ui < units.size(); ++ui)
    for (const RaggedPassHost& h : unit_passes[ui]) {
      pass_index[ui].push_back((int)r.passes.size());
      r.passes.push_back(h);
    }
  for (int phase = 0; phase < 3; ++phase) {
    r.phase_begin.push_back((int)r.task_pass.size());
    for (size_t ui : order) {
      if ((size_t)phase >= pass_index[ui].size()) continue;
      const int pi = pass_index[ui][phase];
      const RaggedPassHost& h = r.passes[pi];
      const int tiles = (int)(h.pp.groups * h.pp.tiles_per_group);
      for (int t = 0; t < tiles; ++t) {
        r.task_pass.push_back(pi);
        r.task_tile.push_back(t);
      }
    }
  }
  r.phase_begin.push_back((int)r.task_pass.size());
  while (r.phase_begin.size() > 2 && r.phase_begin[r.phase_begin.size() - 2] == r.phase_begin.back())
    r.phase_begin.pop_back();  // drop empty trailing phases
  return ST_OK;
}

int ragged_shape_supported(int rank, const long long* dims, int in_fmt, int out_fmt, int scaling) {
  RaggedPlanCore r;
  r.cm = compute_mode(in_fmt, out_fmt);
  if (r.cm < 0 || rank < 1 || rank > 3) return ST_UNSUPPORTED;
  r.in_fmt = in_fmt;
  r.out_fmt = out_fmt;
  r.scaling = scaling;
  const long long d[3] = {dims[0], rank > 1 ? dims[1] : 1, rank > 2 ? dims[2] : 1};
  for (int i = 0; i < rank; ++i)
    if (ilog2(d[i]) < 0) return ST_UNSUPPORTED;
  ShapePasses sp;
  return build_shape(r, rank, d, 1, sp);
}

std::string RaggedPlanCore::describe() const {
  std::ostringstream o;
  o << "ragged items=" << n_items << " passes=" << passes.size() << " tasks=" << task_pass.size()
    << " counters=" << n_counters << " ws=" << ws_elems << " tw=" << tw.size() << " cm=" << cm << "\n";
  return o.str();
}

}  // namespace mcfft

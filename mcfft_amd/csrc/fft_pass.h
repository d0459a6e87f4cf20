// fft_pass.h -- the one kernel family of the library: a batched, strided,
// LDS-resident Stockham FFT pass whose radix-16 stages run as complex 16x16
// DFT-matrix GEMMs on v_mfma_f32_32x32x16_{f16,bf16} with FP32 accumulation.
//
// A pass computes length-L sub-FFTs (L = 16^S * T, T in {1,2,4,8}) on a tile of
// Q instances per workgroup.  Element m of instance q of the tile lives at
//     in  + base_in  + q*in_sq  + m*in_sm
//     out + base_out + q*out_sq + m*out_sm
// which covers contiguous rows (in_sm = 1), strided columns (in_sq = 1) and the
// transposed store of a four-step decomposition (out_sq = 1, out_sm = N1).
//
// Stage p (radix R, Ns = product of earlier radices) follows the Stockham
// autosort recurrence: it reads m = j + r*(L/R) and writes
// m' = (j/Ns)*Ns*R + j%Ns + r'*Ns, so input and output are in natural order and
// every stage reads 32 consecutive j (or q) per MFMA column block.  The twiddle a
// stage's input needs is applied by the previous stage to its FP32 accumulators
// before they are rounded (one rounding per stage); per-stage power-of-two
// scale factors are folded into those tables.
//
// Complex GEMM on real MFMA: with F = Fr + i Fi and X = Xr + i Xi,
//     [Yr; Yi] (32 x 32 cols) = [Fr; Fi] * Xr + [-Fi; Fr] * Xi
// i.e. two 32x32x16 MFMAs per 32 sixteen-point DFTs.  Lane l holds the eight
// inputs r = 8*(l/32)+t of column l%32 and receives outputs
// r' = 4*(l/32) + s%4 + 8*(s/4) (s = 0..7) as accumulators s (re) and 8+s (im).
//
// Reference counterpart: the layer_{256,512,1024}_0 / layer_{256,512}_1 family
// of source/mcfft_1d_device.cpp:58-593 (rocWMMA, digit-reversed input,
// per-lane multiplicative twiddle recurrence).  This design differs on purpose:
// natural-order input, table twiddles generated in FP64, explicit MFMA lane
// layout, one kernel template for rows, columns and four-step passes.
#pragma once
#include "dev_prims.h"

namespace mcfft {

enum { CM_F16 = 0, CM_BF16 = 1, CM_SPLIT = 2 };   // compute / storage mode
enum { FMT_F16 = 0, FMT_BF16 = 1, FMT_SPLIT = 2 }; // global element formats

// Split FP16 element: x = hi + lo * 2^-11 (lo holds the rounding residual of hi,
// scaled so that it stays in FP16's normal range wherever hi does).
constexpr float kSplitLoScale = 2048.0f;
constexpr float kSplitLoInv = 1.0f / 2048.0f;

struct PassArgs {
  const void* in;
  void* out;
  long long in_sq, in_sm, out_sq, out_sm;   // element strides: instance, position
  long long in_sg, in_st, out_sg, out_st;   // block -> base: group stride, tile stride
  long long n_inst;           // valid instances per group (partial last tile)
  int tiles_per_group;
  int in_fmt;                 // split mode only: FMT_F16 or FMT_SPLIT input
  int inverse;                // 0 forward (exp(-i...)), 1 inverse
  float out_scale;            // last-stage scale when there is no four-step table
  const f32x2* stage_tw;      // (nstages-1) post-twiddle tables of L entries each
  const uint16_t* dft;        // W16^e tables for this direction (see DftTable)
  const f32x2* fs_hi;         // four-step twiddle W_N^e = fs_hi[e>>fs_shift] * fs_lo[e&fs_mask]
  const f32x2* fs_lo;
  int fs_shift;
  unsigned fs_mask;
  unsigned long long fs_nmask;  // N-1 (N a power of two)
};

// Layout of the 16-entry DFT root tables (uint16 bit patterns), per direction:
//   [0..15] re_hi f16, [16..31] im_hi f16, [32..47] re_lo f16 (x 2^11),
//   [48..63] im_lo f16 (x 2^11), [64..79] re bf16, [80..95] im bf16
// for W16^e = exp(-+ 2 pi i e / 16), e = 0..15.
constexpr int kDftTableEntries = 96;

MCFFT_DEV f32x2 cmul(f32x2 a, f32x2 b) {
  f32x2 r;
  r.x = __builtin_fmaf(a.x, b.x, -a.y * b.y);
  r.y = __builtin_fmaf(a.x, b.y, a.y * b.x);
  return r;
}

template <int CM> struct Traits;
template <> struct Traits<CM_F16> {
  typedef uint32_t T;
  static MCFFT_DEV T pack(f32x2 v) { return pack2(f16_bits(v.x), f16_bits(v.y)); }
  static MCFFT_DEV f32x2 unpack(T e) { f32x2 r; r.x = f16_val(e & 0xffffu); r.y = f16_val(e >> 16); return r; }
  static MCFFT_DEV T load(const void* p, long long i, int) { return g_ld((const uint32_t*)p, i); }
  static MCFFT_DEV void store(void* p, long long i, T v) { g_st((uint32_t*)p, i, v); }
};
template <> struct Traits<CM_BF16> {
  typedef uint32_t T;
  static MCFFT_DEV T pack(f32x2 v) { return pack2(bf16_bits(v.x), bf16_bits(v.y)); }
  static MCFFT_DEV f32x2 unpack(T e) { f32x2 r; r.x = bf16_val(e & 0xffffu); r.y = bf16_val(e >> 16); return r; }
  static MCFFT_DEV T load(const void* p, long long i, int) { return g_ld((const uint32_t*)p, i); }
  static MCFFT_DEV void store(void* p, long long i, T v) { g_st((uint32_t*)p, i, v); }
};
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
template <> struct Traits<CM_SPLIT> {
  typedef u32x2 T;  // .x = hi pair (re, im), .y = lo pair
  static MCFFT_DEV T pack(f32x2 v) {
    uint32_t hr = f16_bits(v.x), hi = f16_bits(v.y);
    uint32_t lr = f16_bits((v.x - f16_val(hr)) * kSplitLoScale);
    uint32_t li = f16_bits((v.y - f16_val(hi)) * kSplitLoScale);
    T e;
    e.x = pack2(hr, hi);
    e.y = pack2(lr, li);
    return e;
  }
  static MCFFT_DEV f32x2 unpack(T e) {
    f32x2 r;
    r.x = __builtin_fmaf(f16_val(e.y & 0xffffu), kSplitLoInv, f16_val(e.x & 0xffffu));
    r.y = __builtin_fmaf(f16_val(e.y >> 16), kSplitLoInv, f16_val(e.x >> 16));
    return r;
  }
  static MCFFT_DEV T load(const void* p, long long i, int fmt) {
    if (fmt == FMT_SPLIT) return g_ld((const u32x2*)p, i);
    T e;
    e.x = g_ld((const uint32_t*)p, i);
    e.y = 0u;
    return e;
  }
  static MCFFT_DEV void store(void* p, long long i, T v) { g_st((u32x2*)p, i, v); }
};

// DFT-matrix operands of one lane: A1 = [Fr; Fi], A2 = [-Fi; Fr] (hi and lo parts)
struct AOps {
  u32x4 a1, a2, a1lo, a2lo;
};

template <int CM>
MCFFT_DEV AOps load_dft_operands(const uint16_t* tab, int lane) {
  AOps A;
  const int i = lane & 31, h = lane >> 5, row = i & 15;
  const bool top = i < 16;
  const int ore = (CM == CM_BF16) ? 64 : 0, oim = (CM == CM_BF16) ? 80 : 16;
  uint32_t w1[4] = {0, 0, 0, 0}, w2[4] = {0, 0, 0, 0}, w1l[4] = {0, 0, 0, 0}, w2l[4] = {0, 0, 0, 0};
  for (int t = 0; t < 8; ++t) {
    const int k = 8 * h + t, e = (row * k) & 15;
    uint32_t re = g_ld(tab, (long long)(ore + e)), im = g_ld(tab, (long long)(oim + e));
    uint32_t v1 = top ? re : im;
    uint32_t v2 = top ? (im ^ 0x8000u) : re;
    w1[t >> 1] |= v1 << (16 * (t & 1));
    w2[t >> 1] |= v2 << (16 * (t & 1));
    if (CM == CM_SPLIT) {
      uint32_t rl = g_ld(tab, (long long)(32 + e)), il = g_ld(tab, (long long)(48 + e));
      uint32_t l1 = top ? rl : il;
      uint32_t l2 = top ? (il ^ 0x8000u) : rl;
      w1l[t >> 1] |= l1 << (16 * (t & 1));
      w2l[t >> 1] |= l2 << (16 * (t & 1));
    }
  }
  for (int w = 0; w < 4; ++w) {
    A.a1[w] = w1[w];
    A.a2[w] = w2[w];
    A.a1lo[w] = w1l[w];
    A.a2lo[w] = w2l[w];
  }
  return A;
}

// 16-point DFTs of 32 columns: x[t] = input r = 8*(lane/32)+t of this lane's column.
// Returns the FP32 outputs (re in [0..7], im in [8..15] of the slot order above).
template <int CM>
MCFFT_DEV f32x16 dft16(const typename Traits<CM>::T (&x)[8], const AOps& A) {
  f32x16 acc = {};
  if constexpr (CM == CM_SPLIT) {
    u32x4 bre, bim, lre, lim;
    for (int w = 0; w < 4; ++w) {
      bre[w] = lo_halves(x[2 * w].x, x[2 * w + 1].x);
      bim[w] = hi_halves(x[2 * w].x, x[2 * w + 1].x);
      lre[w] = lo_halves(x[2 * w].y, x[2 * w + 1].y);
      lim[w] = hi_halves(x[2 * w].y, x[2 * w + 1].y);
    }
    f32x16 cor = {};
    acc = mfma_f16(A.a1, bre, acc);
    acc = mfma_f16(A.a2, bim, acc);
    cor = mfma_f16(A.a1, lre, cor);
    cor = mfma_f16(A.a2, lim, cor);
    cor = mfma_f16(A.a1lo, bre, cor);
    cor = mfma_f16(A.a2lo, bim, cor);
    for (int v = 0; v < 16; ++v) acc[v] = __builtin_fmaf(cor[v], kSplitLoInv, acc[v]);
  } else {
    u32x4 bre, bim;
    for (int w = 0; w < 4; ++w) {
      bre[w] = lo_halves(x[2 * w], x[2 * w + 1]);
      bim[w] = hi_halves(x[2 * w], x[2 * w + 1]);
    }
    if constexpr (CM == CM_BF16) {
      acc = mfma_bf16(A.a1, bre, acc);
      acc = mfma_bf16(A.a2, bim, acc);
    } else {
      acc = mfma_f16(A.a1, bre, acc);
      acc = mfma_f16(A.a2, bim, acc);
    }
  }
  return acc;
}

// radix-T butterflies (T = 2, 4, 8) on FP32 values, sign: forward exp(-i), inverse exp(+i)
MCFFT_DEV f32x2 mul_i(f32x2 z, bool inv) {  // multiply by -i (forward) or +i (inverse)
  f32x2 r;
  if (inv) { r.x = -z.y; r.y = z.x; } else { r.x = z.y; r.y = -z.x; }
  return r;
}
template <int T>
MCFFT_DEV void dft_small(f32x2 (&v)[T], bool inv) {
  if constexpr (T == 2) {
    f32x2 a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  } else if constexpr (T == 4) {
    f32x2 s0 = v[0] + v[2], d0 = v[0] - v[2], s1 = v[1] + v[3], d1 = mul_i(v[1] - v[3], inv);
    v[0] = s0 + s1;
    v[2] = s0 - s1;
    v[1] = d0 + d1;
    v[3] = d0 - d1;
  } else if constexpr (T == 8) {
    f32x2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    dft_small<4>(e, inv);
    dft_small<4>(o, inv);
    const float c = 0.70710678118654752f;
    f32x2 w1, w3;
    w1.x = c; w1.y = inv ? c : -c;     // W8^1
    w3.x = -c; w3.y = inv ? c : -c;    // W8^3
    f32x2 t1 = cmul(o[1], w1), t2 = mul_i(o[2], inv), t3 = cmul(o[3], w3);
    v[0] = e[0] + o[0];
    v[4] = e[0] - o[0];
    v[1] = e[1] + t1;
    v[5] = e[1] - t1;
    v[2] = e[2] + t2;
    v[6] = e[2] - t2;
    v[3] = e[3] + t3;
    v[7] = e[3] - t3;
  }
}

constexpr int ipow16(int s) { return s == 0 ? 1 : 16 * ipow16(s - 1); }

template <int V> struct IC { static constexpr int value = V; };
template <int P, int E> struct For {
  template <class F> MCFFT_DEV static void run(F&& f) {
    f(IC<P>{});
    For<P + 1, E>::run(f);
  }
};
template <int E> struct For<E, E> {
  template <class F> MCFFT_DEV static void run(F&&) {}
};

// ---------------------------------------------------------------------------
// Pass configuration (compile time)
//   S, T      L = 16^S * T
//   Q         instances per workgroup tile
//   QFAST     MFMA columns and LDS rows run over instances (strided columns)
//             instead of positions (contiguous rows)
//   NW        waves per workgroup
//   CM        compute/storage mode
//   STAGED    the last stage writes LDS and a copy loop stores instance-fastest
//             (coalesced when out_sq == 1, e.g. the transposed four-step store)
//   FOURSTEP  multiply outputs by W_N^(m * column) before storing
// ---------------------------------------------------------------------------
template <int S, int T, int Q, bool QFAST, int NW, int CM, bool STAGED, bool FOURSTEP>
struct PassCfg {
  static constexpr int L = ipow16(S) * T;
  static constexpr int NT = NW * 64;
  static constexpr int NSTAGE = S + (T > 1 ? 1 : 0);
  static constexpr int ROW = L + L / 16 + 1;  // padded LDS row (contiguous-row layout)
  static constexpr int LDS_ELEMS = QFAST ? Q * L : Q * ROW;
  typedef typename Traits<CM>::T El;
  static constexpr int LDS_BYTES = LDS_ELEMS * (int)sizeof(El);
  static_assert(S >= 1, "at least one radix-16 stage");
  static_assert((Q * L) % 512 == 0, "tile must hold a multiple of 32 sixteen-point columns");
  static_assert(T == 1 || T == 2 || T == 4 || T == 8, "tail radix");

  static MCFFT_DEV int lidx(int q, int m) {
    if constexpr (QFAST) return m * Q + q;
    else return q * ROW + m + (m >> 4);
  }
  template <int R>
  static MCFFT_DEV void colmap(int c, int& q, int& j) {
    if constexpr (QFAST) { q = c % Q; j = c / Q; }
    else { q = c / (L / R); j = c % (L / R); }
  }
};

template <int S, int T, int Q, bool QFAST, int NW, int CM, bool STAGED, bool FOURSTEP>
MCFFT_DEV void fft_pass_body(const PassArgs& a) {
  typedef PassCfg<S, T, Q, QFAST, NW, CM, STAGED, FOURSTEP> C;
  typedef Traits<CM> Tr;
  typedef typename C::El El;
  constexpr int L = C::L;
  MCFFT_DYN_LDS(El, lds);

  const int tid = thread_id(), lane = lane_id(), wave = tid >> 6, h = lane >> 5;
  const int blk = block_id();
  const int grp = blk / a.tiles_per_group, til = blk - grp * a.tiles_per_group;
  const long long ib = (long long)grp * a.in_sg + (long long)til * a.in_st;
  const long long ob = (long long)grp * a.out_sg + (long long)til * a.out_st;
  const bool inv = a.inverse != 0;
  const long long inst0 = (long long)til * Q;
  auto valid = [&](int q) -> bool { return inst0 + q < a.n_inst; };

  const AOps A = load_dft_operands<CM>(a.dft, lane);

  // output post-processing of the pass's last stage: four-step twiddle or scale
  auto finish = [&](f32x2 v, int q, int m) -> f32x2 {
    if constexpr (FOURSTEP) {
      const unsigned long long col = (unsigned long long)(inst0 + q);
      const unsigned long long e = ((unsigned long long)m * col) & a.fs_nmask;
      const f32x2 w = cmul(g_ld(a.fs_hi, (long long)(e >> a.fs_shift)), g_ld(a.fs_lo, (long long)(e & a.fs_mask)));
      return cmul(v, w);
    } else {
      return v * a.out_scale;
    }
  };
  auto store_out = [&](int q, int m, El e) {
    if (valid(q)) Tr::store(a.out, ob + (long long)q * a.out_sq + (long long)m * a.out_sm, e);
  };

  // ---------------- radix-16 MFMA stages ----------------
  For<0, S>::run([&](auto pc) {
    constexpr int P = decltype(pc)::value;
    constexpr int NS = ipow16(P);
    constexpr bool LAST = (P == C::NSTAGE - 1);
    constexpr int NCOLT = Q * L / 512;             // 32-column MFMA tiles
    constexpr int TPW = (NCOLT + NW - 1) / NW;     // tiles per wave
    El x[TPW][8];
    for (int k = 0; k < TPW; ++k) {
      const int ti = wave + k * NW;
      if (ti < NCOLT) {
        int q, j;
        C::template colmap<16>(ti * 32 + (lane & 31), q, j);
        for (int t = 0; t < 8; ++t) {
          const int m = j + (8 * h + t) * (L / 16);
          if constexpr (P == 0) {
            if (valid(q)) x[k][t] = Tr::load(a.in, ib + (long long)q * a.in_sq + (long long)m * a.in_sm, a.in_fmt);
            else x[k][t] = El{};
          } else
            x[k][t] = lds_ld(lds, C::lidx(q, m));
        }
      }
    }
    // P > 0: all reads of LDS before any overwrite.  A single-stage pass
    // (L = 16) reads and writes global memory in one stage: the barrier keeps an
    // in-place pass from overwriting elements another wave has not read yet.
    if constexpr (P > 0 || (LAST && !STAGED)) block_barrier();
    for (int k = 0; k < TPW; ++k) {
      const int ti = wave + k * NW;
      if (ti < NCOLT) {
        int q, j;
        C::template colmap<16>(ti * 32 + (lane & 31), q, j);
        const f32x16 y = dft16<CM>(x[k], A);
        for (int s = 0; s < 8; ++s) {
          const int rr = 4 * h + (s & 3) + 8 * (s >> 2);
          const int mo = (j / NS) * NS * 16 + (j % NS) + rr * NS;
          f32x2 v;
          v.x = y[s];
          v.y = y[8 + s];
          if constexpr (!LAST) {
            v = cmul(v, g_ld(a.stage_tw, (long long)(P * L + rr * (L / 16) + j)));
            lds_st(lds, C::lidx(q, mo), Tr::pack(v));
          } else {
            v = finish(v, q, mo);
            if constexpr (STAGED) lds_st(lds, C::lidx(q, mo), Tr::pack(v));
            else store_out(q, mo, Tr::pack(v));
          }
        }
      }
    }
    if constexpr (!LAST || STAGED) block_barrier();
  });

  // ---------------- radix-T tail stage (VALU, FP32) ----------------
  if constexpr (T > 1) {
    constexpr int NCOL = Q * L / T;
    constexpr int CPT = (NCOL + C::NT - 1) / C::NT;
    El x[CPT][T];
    for (int k = 0; k < CPT; ++k) {
      const int c = tid + k * C::NT;
      if (c < NCOL) {
        int q, j;
        C::template colmap<T>(c, q, j);
        for (int r = 0; r < T; ++r) x[k][r] = lds_ld(lds, C::lidx(q, j + r * (L / T)));
      }
    }
    if constexpr (STAGED) block_barrier();
    for (int k = 0; k < CPT; ++k) {
      const int c = tid + k * C::NT;
      if (c < NCOL) {
        int q, j;
        C::template colmap<T>(c, q, j);
        f32x2 v[T];
        for (int r = 0; r < T; ++r) v[r] = Tr::unpack(x[k][r]);
        dft_small<T>(v, inv);
        for (int r = 0; r < T; ++r) {
          const int mo = j + r * (L / T);
          const f32x2 w = finish(v[r], q, mo);
          if constexpr (STAGED) lds_st(lds, C::lidx(q, mo), Tr::pack(w));
          else store_out(q, mo, Tr::pack(w));
        }
      }
    }
    if constexpr (STAGED) block_barrier();
  }

  // ---------------- staged store: instance-fastest order ----------------
  if constexpr (STAGED) {
    for (int e = tid; e < Q * L; e += C::NT) {
      const int q = e % Q, m = e / Q;
      store_out(q, m, lds_ld(lds, C::lidx(q, m)));
    }
  }
}

}  // namespace mcfft

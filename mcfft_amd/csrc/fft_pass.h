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
// before they are rounded (one rounding per stage), with the stage's
// power-of-two scale factor.  The element written at m' is input
// r2 = m' / (L/Rn) of column j2 = m' % (L/Rn) of the next stage (radix Rn), so
// its twiddle is W_M^(r2 * (m' % 16Ns)), M = 16 Ns Rn: an exact rational angle.
// FP16/BF16 passes evaluate it with the hardware sin/cos (argument in
// revolutions, about one FP32 ulp); split passes, whose results carry about 22
// bits, read FP32 tables generated in FP64 (scale folded in).
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
#include "bfp.h"
#include "dev_prims.h"

namespace mcfft {

enum { CM_F16 = 0, CM_BF16 = 1, CM_SPLIT = 2 };   // compute / storage mode
enum { FMT_F16 = 0, FMT_BF16 = 1, FMT_SPLIT = 2 }; // global element formats

// Split FP16 element: x = hi + lo * 2^-11 (lo holds the rounding residual of hi,
// scaled so that it stays in FP16's normal range wherever hi does).
constexpr float kSplitLoScale = 2048.0f;
constexpr float kSplitLoInv = 1.0f / 2048.0f;
// Split passes take their twiddles from FP32 tables generated in FP64 (true) or
// evaluate them in the kernel like FP16/BF16 passes (false).
constexpr bool kSplitTableTwiddles = true;

enum { KIND_ROW = 0, KIND_ROWT = 1, KIND_COL = 2, KIND_COLF = 3, KIND_COLM = 4, KIND_ROWX = 5 };

// Element (q, m) of a tile (instance q, position m) is at, in elements,
//   ROW : in/out  base + q*L + m                  (contiguous rows)
//   ROWT: in      base + q*L + m,  out  base + q + m*stride_out  (transposed store)
//   COL, COLF, COLM: in base + q + m*stride_in, out base + q + m*stride_out (columns)
//   ROWX: in  base + q*stride_in + m,  out  base + q*stride_out + m  (rows, instance strides)
// with base = group*sg + tile*st for the tile's group/tile indices.
//
// ROWX is the row pass of a 2D plan whose outer axis (NX points) is split as
// NX = R * (NX/R): its tile holds the R rows r1 + (NX/R)*r2 (r2 < R) of tile
// r1, and before the row FFT it takes the R-point DFT across them (the outer
// axis's first factor) times W_NX^(r1*k2), storing result k2 as row R*r1 + k2.
// The column pass then has NX/R-point columns at stride R rows.  Q = R and the
// waves' MFMA tiles cover a row in C = L/(512 NW) sets of 32 columns per wave,
// so a thread's tiles k = C*q + c are row q at the same positions for every c
// and the butterfly needs no exchange.
struct PassArgs {
  const void* in;
  void* out;
  long long in_sg, in_st, out_sg, out_st;   // tile -> base: group stride, tile stride
  int stride_in, stride_out;  // runtime strides of the strided axis (elements)
  int n_inst;                 // valid instances per group (partial last tile)
  int tiles_per_group;
  int n_tiles;                // total tiles; workgroups loop over them
  int in_fmt;                 // split mode only: FMT_F16 or FMT_SPLIT input
  int inverse;                // 0 forward (exp(-i...)), 1 inverse
  float out_scale;            // last-stage scale when there is no four-step table
  const f32x2* stage_tw;      // post-twiddle tables [stage][j][16] (see plan_core)
  const u32x4* aops;          // DFT-matrix MFMA operands [64 lanes][4]: A1, A2, A1lo, A2lo
  const uint32_t* dft256;     // 65536-point kernel: 16-point DFT rows scaled by W_256 powers (fft_big.h)
  const f32x2* fs_tw;         // four-step twiddles W_N^(m*column), N1 x fs_cols entries, or (fs_lo_bits > 0)
                              // (split ROWX: W_NX^(r1*k2) at r1*R + k2, unscaled)
                              // W_N^(m*(column % 2^b)) [N1][2^b] then W_N^(m*(column - column % 2^b)) [N1][fs_cols >> b]
  int fs_cols;                // N2 (columns of the four-step decomposition)
  int fs_lo_bits;             // b: split four-step table factorised in two (0: one table)
  long long in_elems, out_elems;  // elements addressable from in / out (bounds of the buffer views)
  int stage_tw_entries, fs_tw_entries;
  float st_scale[6];          // 2^-shift of stage P (twiddles computed in the kernel carry it);
                              // ROWX: [NSTAGE] is its cross-row butterfly's
  // block floating point and exec bookkeeping (bfp.h); zero flags: static
  // scaling, nothing recorded
  int bfp_flags;
  int bfp_quantum;            // AUTO first pass: step of the normalisation exponent
  int bfp_e_base;             // exponent the pass's stores add (first pass: its stage shifts; later
                              // passes: the shifts of passes 1..this; static last pass: the plan's)
  int bfp_tr0;                // global index of the launch's first transform
  int bfp_gpt;                // groups of this pass per transform
  int bfp_nt;                 // transforms of the launch
  int bfp_q0_lg, bfp_tpg0;    // first pass: log2 of its tile width, tiles per group
  int bfp_rg, bfp_rm, bfp_rq; // reconcile pass: first-pass instance of element (q, m) of group g
                              // (see rescale_stage0)
  ExecState* bfp_state;
  unsigned long long* bfp_slots;  // [2t]: max e, [2t+1]: min e of t's first-pass tiles (see bfp.h)
  signed char* bfp_tile_e;    // exponents of the launch's first-pass tiles
  int* bfp_out_exp;           // per transform
  const int* bfp_in_exp;      // per transform (caller's input exponents), or null
  // gathered batch: element offset of each transform of the launch from `in` /
  // `out` (null: transforms follow each other, group g at g*sg); group g of
  // transform t = g / gather_gpt starts at off[t] + (g % gather_gpt) * sg
  const long long* gather_in;
  const long long* gather_out;
  int gather_gpt;
  // start stagger of the single-pass 65536-point kernel: workgroup b waits
  // ((b / 8) % stagger_phases) * stagger_ticks ticks of the 100 MHz clock before
  // its first load, so that its memory phases fall into other workgroups'
  // compute phases instead of all workgroups loading, then computing, together
  // (0: no wait)
  int stagger_ticks, stagger_phases;
};

// first element of group `grp` of a side with group stride sg (see PassArgs::gather_*).
// grp is wave-uniform at every call and the offset table is read-only, so the
// offset is a scalar load: a buffer descriptor built from a vector register would
// wrap every load and store of the tile in a waterfall loop, and a vector load's
// wait would also wait for the previous tile's stores.
MCFFT_DEV long long group_base(const long long* off, int gpt, long long sg, int grp) {
  if (!off) return (long long)grp * sg;
  const int t = grp / gpt;
  return s_ld(off, (long long)t) + (long long)(grp - t * gpt) * sg;
}

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
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
// RNE conversions of a pair (one v_cvt_pk_{f16,bf16}_f32 on gfx950)
MCFFT_DEV uint32_t pk_f16(f32x2 v) { return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, f16x2)); }
MCFFT_DEV uint32_t pk_bf16(f32x2 v) { return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2)); }
MCFFT_DEV f32x2 up_f16(uint32_t e) { return __builtin_convertvector(__builtin_bit_cast(f16x2, e), f32x2); }
MCFFT_DEV f32x2 up_bf16(uint32_t e) { return __builtin_convertvector(__builtin_bit_cast(bf16x2, e), f32x2); }
// product of two packed binary16 pairs (v_pk_mul_f16, RNE)
MCFFT_DEV uint32_t pk_mul_f16(uint32_t a, uint32_t b) {
  return __builtin_bit_cast(uint32_t, __builtin_bit_cast(f16x2, a) * __builtin_bit_cast(f16x2, b));
}

// Byte offset of element i (esz bytes each) inside a buffer view.  Views span up
// to 2^32 bytes, so the product is formed in unsigned 32-bit arithmetic (element
// indices stay below 2^31; their byte offsets need not) and handed to the
// buffer instruction, whose offsets are unsigned, as the same 32 bits.
MCFFT_DEV int boff(int i, unsigned esz) { return (int)((unsigned)i * esz); }

template <> struct Traits<CM_F16> {
  typedef uint32_t T;
  static MCFFT_DEV T pack(f32x2 v) { return pk_f16(v); }
  static MCFFT_DEV f32x2 unpack(T e) { return up_f16(e); }
  static MCFFT_DEV T load(GBuf b, int i, int) { return buf_ld32(b, boff(i, 4)); }
  static MCFFT_DEV void store(GBuf b, int i, T v) { buf_st32(b, boff(i, 4), v); }
  static MCFFT_DEV void store_s(GBuf b, int i, int si, T v) { buf_st32s(b, boff(i, 4), boff(si, 4), v); }
};
template <> struct Traits<CM_BF16> {
  typedef uint32_t T;
  static MCFFT_DEV T pack(f32x2 v) { return pk_bf16(v); }
  static MCFFT_DEV f32x2 unpack(T e) { return up_bf16(e); }
  static MCFFT_DEV T load(GBuf b, int i, int) { return buf_ld32(b, boff(i, 4)); }
  static MCFFT_DEV void store(GBuf b, int i, T v) { buf_st32(b, boff(i, 4), v); }
  static MCFFT_DEV void store_s(GBuf b, int i, int si, T v) { buf_st32s(b, boff(i, 4), boff(si, 4), v); }
};
template <> struct Traits<CM_SPLIT> {
  typedef u32x2 T;  // .x = hi pair (re, im), .y = lo pair
  static MCFFT_DEV T pack(f32x2 v) {
    T e;
    e.x = pk_f16(v);
    e.y = pk_f16((v - up_f16(e.x)) * kSplitLoScale);
    return e;
  }
  static MCFFT_DEV f32x2 unpack(T e) { return up_f16(e.y) * kSplitLoInv + up_f16(e.x); }
  static MCFFT_DEV T load(GBuf b, int i, int fmt) {
    if (fmt == FMT_SPLIT) return buf_ld64(b, boff(i, 8));
    T e;
    e.x = buf_ld32(b, boff(i, 4));
    e.y = 0u;
    return e;
  }
  static MCFFT_DEV void store(GBuf b, int i, T v) { buf_st64(b, boff(i, 8), v); }
  static MCFFT_DEV void store_s(GBuf b, int i, int si, T v) { buf_st64s(b, boff(i, 8), boff(si, 8), v); }
};

// DFT-matrix operands of one lane: A1 = [Fr; Fi], A2 = [-Fi; Fr] (hi and lo parts)
struct AOps {
  u32x4 a1, a2, a1lo, a2lo;
};

// Build lane `lane`'s 16 operand words (A1, A2, A1lo, A2lo; 4 words each) from a
// W16 table (layout above).  Lane l holds A[l%32][8*(l/32)+t], t = 0..7, two
// halves per word; row i < 16 is the real part of DFT row i, row i >= 16 the
// imaginary part of row i-16.
template <class Fn>
MCFFT_DEV void aops_words(const uint16_t* tab, int cm, int lane, Fn&& put) {
  const int i = lane & 31, h = lane >> 5, row = i & 15;
  const bool top = i < 16;
  const int ore = (cm == CM_BF16) ? 64 : 0, oim = (cm == CM_BF16) ? 80 : 16;
  uint32_t w[16] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int t = 0; t < 8; ++t) {
    const int k = 8 * h + t, e = (row * k) & 15;
    const uint32_t re = tab[ore + e], im = tab[oim + e];
    const uint32_t rl = tab[32 + e], il = tab[48 + e];
    const uint32_t sh = 16 * (t & 1);
    w[0 + (t >> 1)] |= (top ? re : im) << sh;
    w[4 + (t >> 1)] |= (top ? (im ^ 0x8000u) : re) << sh;
    if (cm == CM_SPLIT) {
      w[8 + (t >> 1)] |= (top ? rl : il) << sh;
      w[12 + (t >> 1)] |= (top ? (il ^ 0x8000u) : rl) << sh;
    }
  }
  for (int k = 0; k < 16; ++k) put(k, w[k]);
}

MCFFT_DEV AOps load_aops(const u32x4* tab, int lane) {
  AOps A;
  A.a1 = g_ld(tab, (long long)(lane * 4 + 0));
  A.a2 = g_ld(tab, (long long)(lane * 4 + 1));
  A.a1lo = g_ld(tab, (long long)(lane * 4 + 2));
  A.a2lo = g_ld(tab, (long long)(lane * 4 + 3));
  return A;
}

// Row-scaled DFT matrices G_r2[i][k] = W_256^(i (16 k + r2)) = W_16^(i k) W_256^(r2 i)
// (direction's sign; r2 = 0 is the plain 16-point DFT matrix), stored as MFMA A
// operand words (see aops_words): 16-byte group ((b * 256 + r2 * 16 + i) * 2 + h)
// holds entries k = 8h..8h+7 of row i, two per word (k even in the low half), as
// their real parts (block b = 0), imaginary parts (b = 1) or negated imaginary
// parts (b = 2), in the plan's FP16 or BF16 format.
constexpr int kDft256Words = 3 * 256 * 2 * 4;
// the two entries of word w: row i, columns k and k + 1, block b, scale r2
struct Dft256Word {
  int b, r2, i, k;
};
MCFFT_DEV Dft256Word dft256_word(int w) {
  const int q = w >> 2, rem = q & 511;
  Dft256Word d;
  d.b = q >> 9;
  d.r2 = rem >> 5;
  d.i = (rem >> 1) & 15;
  d.k = 8 * (rem & 1) + 2 * (w & 3);
  return d;
}
MCFFT_DEV int dft256_exponent(int r2, int i, int k) { return (i * (16 * k + r2)) & 255; }

// MFMA A operands of lane `lane` for the DFT matrix with rows scaled by
// W_256^(r2 rr), read from the table in LDS: rows < 16 (real parts of the
// outputs) take (re, -im), rows >= 16 (imaginary parts) take (im, re).  NB = 2:
// a table of blocks 0 and 1 only (where LDS is short), -im made by a sign flip.
template <int NB = 3>
MCFFT_DEV AOps dft256_aops(const uint32_t* tab, int r2, int lane) {
  const int i = lane & 15, h = lane >> 5;
  const bool top = (lane & 16) == 0;
  AOps A;
  A.a1 = lds_ld((const u32x4*)tab, ((top ? 0 : 1) * 256 + r2 * 16 + i) * 2 + h);
  if constexpr (NB == 3) {
    A.a2 = lds_ld((const u32x4*)tab, ((top ? 2 : 0) * 256 + r2 * 16 + i) * 2 + h);
  } else {
    const u32x4 w = lds_ld((const u32x4*)tab, ((top ? 1 : 0) * 256 + r2 * 16 + i) * 2 + h);
    const uint32_t neg = top ? 0x80008000u : 0u;
    for (int c = 0; c < 4; ++c) A.a2[c] = w[c] ^ neg;
  }
  A.a1lo = A.a1;
  A.a2lo = A.a2;
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
constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v / 2); }

// W_M^e * scale (forward exp(-2 pi i e / M), inverse exp(+...)); inv_m = 1/M
// is a power of two, so e * inv_m is exact for e < 2^24 (larger e, from the
// first pass of a 1D plan above 2^24 points, round to 24 bits: an angle error
// below 2^-24 of a revolution)
MCFFT_DEV f32x2 twiddle_rev(int e, float inv_m, bool inv, float scale) {
  f32x2 w = cis_rev((float)e * inv_m);
  w.x *= scale;
  w.y *= inv ? scale : -scale;
  return w;
}

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
//   NW        waves per workgroup
//   CM        compute/storage mode
//   KIND      ROW / ROWT / COL / COLF / COLM (addressing above).  The column
//             kinds run the MFMA columns and LDS rows over instances (q
//             fastest); COLF (a plan's first pass) and COLM (the middle pass of a
//             three-pass 1D plan) multiply their outputs by W_N^(m * column);
//             ROWT's last stage writes LDS and a copy loop stores
//             instance-fastest (coalesced for the transposed four-step store).
// ---------------------------------------------------------------------------
template <int S, int T, int Q, int NW, int CM, int KIND>
struct PassCfg {
  static constexpr int L = ipow16(S) * T;
  static constexpr int Q_ = Q, NW_ = NW, CM_ = CM, KIND_ = KIND;
  static constexpr int NT = NW * 64;
  static constexpr int NSTAGE = S + (T > 1 ? 1 : 0);
  static constexpr bool QFAST = (KIND == KIND_COL || KIND == KIND_COLF || KIND == KIND_COLM);
  static constexpr bool XROWS = (KIND == KIND_ROWX);
  static constexpr bool STAGED = (KIND == KIND_ROWT);
  static constexpr bool FOURSTEP = (KIND == KIND_COLF || KIND == KIND_COLM);
  static constexpr int ROW = L + L / 16 + 1;  // padded LDS row (contiguous-row layout)
  static constexpr int LDS_ELEMS = QFAST ? Q * L : Q * ROW;
  typedef typename Traits<CM>::T El;
  // the tile's data, then one word per wave for block-wide reductions
  static constexpr int RED_OFF = LDS_ELEMS * (int)sizeof(El);
  // Row twiddles folded into the DFT matrix (dft256_aops): FP16/BF16 row tiles
  // of 8192, 16384 and 32768 points, where the 32 columns of an MFMA tile share
  // r2 = j >> (lg(L/Rn) - 4) at every twiddled stage; the matrices sit in LDS
  // after the tile (these kernels run one workgroup per CU, held by registers),
  // all three blocks where they fit, else the first two (TAB_BLOCKS).
  static constexpr bool FOLD_TW = KIND == KIND_ROW && CM != CM_SPLIT && S == 3 && T >= 2;
  static constexpr int TAB_OFF = (RED_OFF + 64 + 15) / 16 * 16;
  static constexpr int TAB_BLOCKS = TAB_OFF + kDft256Words * 4 <= 160 * 1024 ? 3 : 2;
  static constexpr int LDS_BYTES = FOLD_TW ? TAB_OFF + TAB_BLOCKS * (kDft256Words / 3) * 4 : RED_OFF + 64;
  static constexpr int NCOLT = Q * L / 512;          // 32-column MFMA tiles of the tile
  static constexpr int TPW = (NCOLT + NW - 1) / NW;  // MFMA tiles per wave
  // L = 512: the radix-2 tail is done right after the last radix-16 stage
  // instead of in another pass through LDS.  Its partners (positions m and
  // m + 256) come from columns j and j + 16 of that stage, which the stage
  // places in the same lane of two MFMA tiles of one wave (needs an even number
  // of tiles per wave); in the row layout with an odd number, in lanes l and
  // l ^ 16 of one tile (exchanged with ds_swizzle).
  static constexpr bool FUSE2 =
      S == 2 && T == 2 && NCOLT % NW == 0 && (!QFAST || (NCOLT / NW) % 2 == 0);
  static_assert(S >= 1, "at least one radix-16 stage");
  static_assert((Q * L) % 512 == 0, "tile must hold a multiple of 32 sixteen-point columns");
  static_assert(T == 1 || T == 2 || T == 4 || T == 8, "tail radix");
  // ROWX: wave w's MFMA tile k = XC*q + c is row q at columns 32 (w + NW c) .. +31
  static constexpr int XC = L / (512 * NW);
  static_assert(!XROWS || ((Q == 2 || Q == 4 || Q == 8) && XC >= 1 && XC * 512 * NW == L && TPW == Q * XC),
                "ROWX tile shape");

  static MCFFT_DEV int lidx(int q, int m) {
    if constexpr (QFAST) return m * Q + q;
    else return q * ROW + m + (m >> 4);
  }
  // lidx(q, mb + d) where d is a multiple of 16, or mb is and 0 <= d < 16: the
  // padding term then splits, (mb + d) >> 4 == (mb >> 4) + (d >> 4), so a lane's
  // slots are one per-lane base plus offsets the compiler can fold into the LDS
  // instructions instead of a shift and add per element
  static MCFFT_DEV int lidx_sum(int q, int mb, int d) {
    if constexpr (QFAST) return (mb + d) * Q + q;
    else return q * ROW + mb + (mb >> 4) + d + (d >> 4);
  }
  template <int R>
  static MCFFT_DEV void colmap(int c, int& q, int& j) {  // c >= 0: unsigned division is a shift
    const unsigned u = (unsigned)c;
    if constexpr (QFAST) { q = (int)(u % (unsigned)Q); j = (int)(u / (unsigned)Q); }
    else { q = (int)(u / (unsigned)(L / R)); j = (int)(u % (unsigned)(L / R)); }
  }
  static MCFFT_DEV int in_off(int q, int m, int stride) {
    if constexpr (QFAST) return q + m * stride;
    else if constexpr (XROWS) return q * stride + m;
    else return q * L + m;
  }
  static MCFFT_DEV int out_off(int q, int m, int stride) {
    if constexpr (KIND == KIND_ROW) return q * L + m;
    else if constexpr (XROWS) return q * stride + m;
    else return q + m * stride;
  }
};

MCFFT_DEV GBuf clamp_view(const void* base, long long first, long long total, int esz) {
  const long long rem = (total - first) * esz;
  return make_gbuf((const char*)base + first * esz, (unsigned)(rem > 0xffffffffLL ? 0xffffffffLL : rem));
}

// Stage-0 inputs of one tile: MFMA tile k = wave + k*NW, lane inputs r = 8h + t.
template <class C, class Tr>
MCFFT_DEV void load_stage0(const PassArgs& a, GBuf gin, int wave, int lane, typename C::El (*x)[8]) {
  constexpr int L = C::L;
  const int h = lane >> 5;
  for (int k = 0; k < C::TPW; ++k) {
    const int ti = wave + k * C::NW_;
    if (ti < C::NCOLT) {
      int q, j;
      C::template colmap<16>(ti * 32 + (lane & 31), q, j);
      const int o0 = C::in_off(q, j + 8 * h * (L / 16), a.stride_in);
      const int ostep = C::in_off(0, L / 16, a.stride_in);
      for (int t = 0; t < 8; ++t) x[k][t] = Tr::load(gin, o0 + t * ostep, a.in_fmt);
    }
  }
}

// AUTO first pass: `acc` holds this thread's largest |component| of the tile's
// input (packed, see pk_absmax).  Returns the tile's normalisation 2^n, sets
// e_out (the stored output's exponent) and, unless the pass is also the last,
// records the tile's exponent (byte) and folds it into its transform's slots.
// An all-zero tile keeps 2^0 and records no exponent; an Inf/NaN input sets the
// status bit and is treated like it.  Ends after a barrier (block_max; its
// per-wave words at red[wave * red_stride]).
MCFFT_DEV float bfp_first_tile(const PassArgs& a, uint32_t acc, int cm, unsigned* red, int nw, int tile, int t_loc,
                               int& e_out, SlotAcc& sacc, int red_stride = 1) {
  const int tid = thread_id();
  const uint32_t mb = block_max(pk_fold(acc), red, nw, tid >> 6, tid & 63, red_stride);
  const bool nonfinite = bits_nonfinite(mb, cm), zero = nonfinite || mb == 0;
  if (nonfinite && tid == 0) atomic_or_dev(&a.bfp_state->status, kStatusNonFinite);
  const int n = zero ? 0 : bfp_norm(mb, cm, a.bfp_quantum);
  e_out = a.bfp_e_base - n;
  if (!(a.bfp_flags & BFP_LAST)) {
    if (tid == 0) g_st(a.bfp_tile_e, (long long)tile, (signed char)(zero ? kBfpZeroTile : e_out));
    slot_note(sacc, a.bfp_slots, (long long)a.bfp_tr0 + t_loc, e_out, zero);
  }
  return exp2i(n);
}

// Last pass, end of a tile: the per-transform output exponents (a single-pass
// tile's instances are transforms, q < Q; otherwise the first tile of the
// transform's first group writes it) and, for a statically scaled plan, the
// status bit when one of this thread's stores was Inf/NaN.
MCFFT_DEV void bfp_last_tile(const PassArgs& a, int Q, int inst0, int grp, int til, int t_loc, int e_out,
                             bool stored_nonfinite) {
  const int tid = thread_id();
  const bool single = (a.bfp_flags & BFP_SINGLE) != 0;
  if (single ? (tid < Q && inst0 + tid < a.n_inst) : (tid == 0 && til == 0 && grp == t_loc * a.bfp_gpt)) {
    const long long t = (long long)a.bfp_tr0 + (single ? inst0 + tid : t_loc);
    // the caller's input exponent: a scalar load where t is wave-uniform (a
    // vector load here would wait for all of the tile's stores); each entry is
    // read before its one writer (this thread) may overwrite it in place
    int e_in = 0;
    if (a.bfp_in_exp) e_in = (!single || Q == 1) ? s_ld(a.bfp_in_exp, t) : g_ld(a.bfp_in_exp, t);
    g_st(a.bfp_out_exp, t, e_out + e_in);
  }
  if ((a.bfp_flags & BFP_CHECK) && stored_nonfinite) atomic_or_dev(&a.bfp_state->status, kStatusNonFinite);
}

// AUTO scaling, pass after the first, when the first pass's tiles of this
// transform chose different exponents: bring every stage-0 input to the common
// exponent E (the largest), multiplying by 2^(e_tile - E) <= 1.  Element (q, m) of
// group `grp` (tile instances from inst0) is instance
//   i0 = (grp % gpt) * rg + m * rm + (inst0 + q) * rq
// of first-pass group grp / gpt -- 2D and four-step 1D: gpt = 1, i0 = m; 3D
// y-axis pass: gpt = NX, rg = L; middle pass of a three-pass 1D plan
// N1 x N2 x N3: rg = 0, rm = N3, rq = 1 -- made by first-pass tile
// (grp / gpt) * tpg0 + (i0 >> q0_lg).
template <class C, class Tr>
MCFFT_DEV void rescale_stage0(const PassArgs& a, int wave, int lane, int grp, int inst0, int E,
                              typename C::El (*x)[8]) {
  constexpr int L = C::L;
  const int h = lane >> 5;
  const int g0 = grp / a.bfp_gpt, ibase = (grp - g0 * a.bfp_gpt) * a.bfp_rg;
  const long long tbase = (long long)g0 * a.bfp_tpg0;
  for (int k = 0; k < C::TPW; ++k) {
    const int ti = wave + k * C::NW_;
    if (ti < C::NCOLT) {
      int q, j;
      C::template colmap<16>(ti * 32 + (lane & 31), q, j);
      const int iq = ibase + (inst0 + q) * a.bfp_rq;
      for (int t = 0; t < 8; ++t) {
        const int m = j + (8 * h + t) * (L / 16);
        const int e = (int)g_ld(a.bfp_tile_e, tbase + ((iq + m * a.bfp_rm) >> a.bfp_q0_lg));
        const int d = e - E;
        if constexpr (C::CM_ == CM_F16) {
          // 2^d as binary16 (normal down to 2^-14, subnormal to 2^-24), both
          // halves; one packed FP16 multiply per pair, rounded as a repack of
          // the FP32 product would be
          const uint32_t f16 = (e == kBfpZeroTile || d < -24) ? 0u
                               : d >= -14 ? (uint32_t)(15 + d) << 10 : 1u << (24 + d);
          x[k][t] = pk_mul_f16(x[k][t], f16 | (f16 << 16));
        } else {
          // BF16; split through FP32, so that bits a subnormal hi part loses move
          // into the lo part
          const float f = (e == kBfpZeroTile || d < -126) ? 0.0f : exp2i(d);
          x[k][t] = Tr::pack(Tr::unpack(x[k][t]) * f);
        }
      }
    }
  }
}

// ROWX, after the stage-0 loads: the radix-Q DFT across the tile's rows (the
// outer axis's first factor, in FP32), result k2 times W_NX^(r1*k2), the
// butterfly's stage scale and the AUTO normalisation `nsc`, rounded once.  The
// thread's MFMA tiles XC*k + c hold row k at the same positions (PassCfg::XROWS),
// so the butterfly stays in registers.  Row r1 = the tile's index in its group;
// NX = Q * tiles_per_group.
template <class C, class Tr>
MCFFT_DEV void cross_rows(const PassArgs& a, int r1, bool inv, float nsc, typename C::El (*x)[8]) {
  constexpr int R = C::Q_;
  const float sc = a.st_scale[C::NSTAGE] * nsc;
  f32x2 w[R];
  if constexpr (C::CM_ == CM_SPLIT && kSplitTableTwiddles) {
    const GBuf g = make_gbuf(a.fs_tw, (unsigned)a.fs_tw_entries * 8u);
    for (int k = 0; k < R; ++k) w[k] = __builtin_bit_cast(f32x2, buf_ld64(g, (r1 * R + k) * 8)) * sc;
  } else {
    const float inv_nx = 1.0f / (float)(R * a.tiles_per_group);
    for (int k = 0; k < R; ++k) w[k] = twiddle_rev(r1 * k, inv_nx, inv, sc);
  }
  for (int c = 0; c < C::XC; ++c)
    for (int t = 0; t < 8; ++t) {
      f32x2 v[R];
      for (int k = 0; k < R; ++k) v[k] = Tr::unpack(x[C::XC * k + c][t]);
      dft_small<R>(v, inv);
      for (int k = 0; k < R; ++k) x[C::XC * k + c][t] = Tr::pack(cmul(v[k], w[k]));
    }
}

// The input view of `tile`: its group/tile base and, for ROW tiles (contiguous
// instances), an end at the last valid instance, so that loads beyond read 0.
// Column kinds always run full tiles (the planner guarantees it).
template <class C>
MCFFT_DEV GBuf tile_in_view(const PassArgs& a, int tile, int esz_in) {
  const int grp = uniform(tile / a.tiles_per_group), til = uniform(tile - grp * a.tiles_per_group);
  const long long ib = group_base(a.gather_in, a.gather_gpt, a.in_sg, grp) + (long long)til * a.in_st;
  long long in_end = a.in_elems;
  if constexpr (C::KIND_ == KIND_ROW) {
    const int inst0 = til * C::Q_;
    const long long vq = (long long)(a.n_inst - inst0 < C::Q_ ? a.n_inst - inst0 : C::Q_);
    in_end = ib + vq * C::L < in_end ? ib + vq * C::L : in_end;
  }
  return clamp_view(a.in, ib, in_end, esz_in);
}

// Stage-0 inputs of `tile` (what fft_tile loads first), issued ahead of the
// tile so that they arrive while the workgroup computes the one before it.
template <int S, int T, int Q, int NW, int CM, int KIND>
MCFFT_DEV void tile_prefetch(const PassArgs& a, int tile, typename Traits<CM>::T (*x)[8]) {
  typedef PassCfg<S, T, Q, NW, CM, KIND> C;
  const int tid = opaque(thread_id());
  const int esz_in = (CM == CM_SPLIT && a.in_fmt == FMT_SPLIT) ? 8 : 4;
  load_stage0<C, Traits<CM>>(a, tile_in_view<C>(a, tile, esz_in), tid >> 6, tid & 63, x);
}

// One tile of a pass: stages on the tile's Q instances, ending with a barrier
// (LDS is reused by the workgroup's next tile).  BFP: the tile takes part in the
// block floating point and exec bookkeeping of bfp.h (off in the ragged kernel,
// whose passes are statically scaled).
// `pre`: the tile's stage-0 inputs, already loaded (tile_prefetch); null: load them here.
template <int S, int T, int Q, int NW, int CM, int KIND, bool BFP = true>
MCFFT_DEV void fft_tile(const PassArgs& a, int tile, const AOps& A, SlotAcc* sacc = nullptr,
                        const typename Traits<CM>::T (*pre)[8] = nullptr) {
  typedef PassCfg<S, T, Q, NW, CM, KIND> C;
  typedef Traits<CM> Tr;
  typedef typename C::El El;
  constexpr int L = C::L;
  constexpr bool STAGED = C::STAGED, FOURSTEP = C::FOURSTEP;
  MCFFT_DYN_LDS(El, lds);
  // LDS slot of (q, mb + d), see lidx_sum.  The ragged kernel's split variants
  // (BFP off, capped at 128 VGPRs) keep the per-element form: the per-lane bases
  // of the split one spilled 10 VGPRs there.
  const uint32_t* tabf = (const uint32_t*)((const unsigned char*)lds + C::TAB_OFF);  // FOLD_TW matrices
  auto lds_slot = [](int q, int mb, int d) {
    if constexpr (BFP || CM != CM_SPLIT) return C::lidx_sum(q, mb, d);
    else return C::lidx(q, mb + d);
  };

  // opaque per tile: index math is redone for each tile instead of being hoisted
  // out of the tile loop and held in registers (or spilled) across it
  const int tid = opaque(thread_id()), lane = tid & 63, wave = tid >> 6, h = lane >> 5;
  const bool inv = a.inverse != 0;
  const int sout = a.stride_out;
  const int esz_in = (CM == CM_SPLIT && a.in_fmt == FMT_SPLIT) ? 8 : 4;
  constexpr int esz_out = (int)sizeof(El);
  const GBuf gtw = make_gbuf(a.stage_tw, (unsigned)a.stage_tw_entries * 8u);
  GBuf gfs;
  if constexpr (FOURSTEP) gfs = make_gbuf(a.fs_tw, (unsigned)a.fs_tw_entries * 8u);

  {
    // wave-uniform (scalar registers): kept to the end of the tile by the bookkeeping
    const int grp = uniform(tile / a.tiles_per_group), til = uniform(tile - grp * a.tiles_per_group);
    const long long ob = group_base(a.gather_out, a.gather_gpt, a.out_sg, grp) + (long long)til * a.out_st;
    const int inst0 = til * Q;
    // A ROW tile's instances are contiguous, so its buffer views end at the last
    // valid instance: loads beyond read 0 and stores beyond are dropped.
    long long out_end = a.out_elems;
    if constexpr (KIND == KIND_ROW) {
      const long long vq = (long long)(a.n_inst - inst0 < Q ? a.n_inst - inst0 : Q);
      out_end = ob + vq * L < out_end ? ob + vq * L : out_end;
    }
    const GBuf gout = clamp_view(a.out, ob, out_end, esz_out);

    // ---- block floating point (bfp.h) ----
    // Row and four-step column passes are always a plan's first pass; column and
    // transposed-row passes never are (so only they can need reconciling).
    constexpr bool FIRST_KIND = (KIND == KIND_ROW || KIND == KIND_COLF || KIND == KIND_ROWX);
    const int bf = BFP ? a.bfp_flags : 0;
    const int t_loc = (BFP && a.bfp_gpt > 0) ? uniform(grp / a.bfp_gpt) : 0;  // transform of the tile in the launch
    float nsc = 1.0f;        // AUTO first pass: 2^n applied to the stage-0 results
    int e_out = BFP ? a.bfp_e_base : 0;  // exponent of the stored output (true = stored * 2^e_out)
    bool rescale = false;    // AUTO reconcile pass: the transform's first-pass tiles disagree
    uint32_t watch = ~0u;    // static last pass: Inf/NaN watch over this thread's stores (bfp.h)

    // Scale factors.  With computed twiddles the last stage's power-of-two scale
    // rides on the twiddle the stage before it writes (the stages are linear), so
    // the outputs need no extra multiply; twiddle tables carry each stage's own.
    // An AUTO first pass's normalisation nsc multiplies stage 0's results.
    constexpr bool TABLE_TW = (CM == CM_SPLIT) && kSplitTableTwiddles;
    constexpr bool FOLD = !TABLE_TW && C::NSTAGE >= 2;
    const float last_scale = FOLD ? 1.0f : a.out_scale;
    auto stage_scale = [&](int P) -> float {  // scale carried by the twiddle stage P writes
      const float s = (FOLD && P == C::NSTAGE - 2) ? a.st_scale[P] * a.st_scale[P + 1] : a.st_scale[P];
      return P == 0 ? s * nsc : s;
    };
    // scale of a single-stage pass's outputs (stage 0 is its last)
    auto single_scale = [&]() -> float { return C::NSTAGE == 1 ? a.out_scale * nsc : a.out_scale; };

    // output post-processing of the pass's last stage: four-step twiddle or scale
    // four-step twiddle W_N^(m * column), N = L * fs_cols (m < L, column < fs_cols)
    const float inv_n = 1.0f / ((float)L * (float)a.fs_cols);
    auto fs_twiddle = [&](int q, int m) -> f32x2 {
      if constexpr (TABLE_TW) {
        const int col = inst0 + q, b = a.fs_lo_bits;
        if (b == 0) return __builtin_bit_cast(f32x2, buf_ld64(gfs, (m * a.fs_cols + col) * 8));
        const int lo = (m << b) + (col & ((1 << b) - 1));
        const int hi = (L << b) + m * (a.fs_cols >> b) + (col >> b);
        return cmul(__builtin_bit_cast(f32x2, buf_ld64(gfs, lo * 8)), __builtin_bit_cast(f32x2, buf_ld64(gfs, hi * 8)));
      } else {
        return twiddle_rev(m * (inst0 + q), inv_n, inv, last_scale);
      }
    };
    auto finish = [&](f32x2 v, int q, int m) -> f32x2 {
      if constexpr (FOURSTEP) {
        return cmul(v, fs_twiddle(q, m));
      } else if constexpr (FOLD) {
        return v;
      } else {
        return v * single_scale();
      }
    };
    const bool check = BFP && (bf & BFP_CHECK);
    auto emit = [&](int off, El e) {
      if (check) watch = nonfinite_fold(watch, e, CM);
      Tr::store(gout, off, e);
    };
    auto store_out = [&](int q, int m, El e) { emit(C::out_off(q, m, sout), e); };

    // ---------------- radix-16 MFMA stages ----------------
    For<0, S>::run([&](auto pc) {
      constexpr int P = decltype(pc)::value;
      constexpr int NS = ipow16(P);
      constexpr bool LAST = (P == C::NSTAGE - 1);
      constexpr int NCOLT = C::NCOLT;
      constexpr int TPW = C::TPW;
      // Fused radix-2 tail with the partners in one lane: in this (last radix-16)
      // stage MFMA tiles 2i and 2i+1 of a wave take the same 32 (instance, column)
      // slots with column j < 16 and j + 16 (row layout: lanes 0-15 instance 2p,
      // 16-31 instance 2p+1; column layout: 32 instances of one column), so the
      // outputs at m and m + 256 meet in the same thread.
      constexpr bool PAIR = C::FUSE2 && P == S - 1 && TPW % 2 == 0;
      auto stage_col = [&](int k, int& q, int& j) {
        if constexpr (PAIR) {
          const unsigned cp = (unsigned)((wave + (k >> 1) * NW) * 32 + (lane & 31));
          if constexpr (C::QFAST) {
            q = (int)(cp % (unsigned)Q);
            j = (int)(cp / (unsigned)Q) + 16 * (k & 1);
          } else {
            q = (int)(cp >> 4);
            j = (int)(cp & 15u) + 16 * (k & 1);
          }
        } else {
          C::template colmap<16>((wave + k * NW) * 32 + (lane & 31), q, j);
        }
      };
      El x[TPW][8];
      if constexpr (P == 0) {
        if (pre) {
          for (int k = 0; k < TPW; ++k)
            for (int t = 0; t < 8; ++t) x[k][t] = pre[k][t];
        } else {
          load_stage0<C, Tr>(a, tile_in_view<C>(a, tile, esz_in), wave, lane, x);
        }
        if (BFP && !FIRST_KIND && (bf & BFP_AUTO)) {
          // AUTO later pass: the transform's exponent slots (scalar loads), read
          // once the tile's loads are in flight so that the two latencies overlap
          int emax, emin;
          slot_read(a.bfp_slots, (long long)a.bfp_tr0 + t_loc, emax, emin);
          e_out = uniform(e_out + emax);
          rescale = (bf & BFP_RECONCILE) && uniform(emin) != uniform(emax);
        }
        if (BFP && FIRST_KIND && (bf & BFP_FIRST)) {
          // AUTO first pass: the tile's largest |component| sets its normalisation
          uint32_t acc = 0;
          for (int k = 0; k < TPW; ++k)
            if (wave + k * NW < NCOLT)
              for (int t = 0; t < 8; ++t) acc = pk_absmax(acc, x[k][t]);
          nsc = bfp_first_tile(a, acc, CM, (unsigned*)((unsigned char*)lds + C::RED_OFF), NW, tile, t_loc, e_out,
                               *sacc);
        } else if (BFP && !FIRST_KIND && rescale) {
          rescale_stage0<C, Tr>(a, wave, lane, grp, inst0, e_out - a.bfp_e_base, x);
        }
        if constexpr (C::XROWS) {
          // the normalisation rides on the cross-row butterfly instead of stage 0
          cross_rows<C, Tr>(a, til, inv, nsc, x);
          nsc = 1.0f;
        }
      } else {
        for (int k = 0; k < TPW; ++k) {
          const int ti = wave + k * NW;
          if (ti < NCOLT) {
            int q, j;
            stage_col(k, q, j);
            for (int t = 0; t < 8; ++t) x[k][t] = lds_ld(lds, lds_slot(q, j, (8 * h + t) * (L / 16)));
          }
        }
      }
      // P > 0: all reads of LDS before any overwrite.  A single-stage pass
      // (L = 16) reads and writes global memory in one stage: the barrier keeps an
      // in-place pass from overwriting elements another wave has not read yet.
      if constexpr (P > 0 || (LAST && !STAGED)) block_barrier();
      if constexpr (PAIR) {
        for (int k = 0; k < TPW; k += 2) {
          int q, j0;
          stage_col(k, q, j0);                            // j0 < 16; partner column j0 + 16
          const f32x16 y0 = dft16<CM>(x[k], A);
          const f32x16 y1 = dft16<CM>(x[k + 1], A);
          // outputs of column j0 land at m = j0 + 16 rr < 256 (tail input 0, twiddle 1),
          // those of j0 + 16 at m + 256 (tail input 1, twiddle W_512^m)
          const float sc = stage_scale(P);
          f32x2 d;  // W_512^16; split: table entry (column 16, rr = 1) without its scale
          if constexpr (!TABLE_TW) d = twiddle_rev(16, 1.0f / 512.0f, inv, 1.0f);
          else d = __builtin_bit_cast(f32x2, buf_ld64(gtw, (P * L + 16 * 16 + 1) * 8)) * (1.0f / a.st_scale[P]);
          for (int g = 0; g < 2; ++g) {
            const int m0 = j0 + (4 * h + 8 * g) * 16;
            f32x2 tw[4];
            if constexpr (TABLE_TW) {
              tw[0] = __builtin_bit_cast(f32x2, buf_ld64(gtw, (P * L + (j0 + 16) * 16 + 4 * h + 8 * g) * 8));
            } else {
              tw[0] = twiddle_rev(m0, 1.0f / 512.0f, inv, sc);
            }
            for (int u = 1; u < 4; ++u) tw[u] = cmul(tw[u - 1], d);
            for (int u = 0; u < 4; ++u) {
              const int s = 4 * g + u, m = m0 + 16 * u;
              f32x2 v0, v1;
              v0.x = y0[s];
              v0.y = y0[8 + s];
              v1.x = y1[s];
              v1.y = y1[8 + s];
              // split tables carry the scale in both entries (W^0 * scale for column j0)
              const f32x2 w0 = TABLE_TW ? v0 * a.st_scale[P] : v0 * sc;
              const f32x2 w1 = cmul(v1, tw[u]);
              const f32x2 lo = finish(w0 + w1, q, m), hi = finish(w0 - w1, q, m + 256);
              if constexpr (STAGED) {
                // m = j0 + 16 (4h + 8g + u)
                lds_st(lds, lds_slot(q, j0, m - j0), Tr::pack(lo));
                lds_st(lds, lds_slot(q, j0, m - j0 + 256), Tr::pack(hi));
              } else {
                store_out(q, m, Tr::pack(lo));
                store_out(q, m + 256, Tr::pack(hi));
              }
            }
          }
        }
      } else
      for (int k = 0; k < TPW; ++k) {
        const int ti = wave + k * NW;
        if constexpr (C::FOLD_TW && BFP && !LAST) {
          // twiddle W_M^(r2 (m' % 16Ns)) = W_M^(r2 (j % Ns)) (column factor) *
          // W_(16 Rn)^(r2 rr) (in the tile's DFT matrix, table entry r2 * 16 / Rn)
          constexpr int RN = (P + 1 < S) ? 16 : T;
          constexpr int SH = ilog2c(L / RN) - 4;
          constexpr float INV_M = 1.0f / (float)(16 * NS * RN);
          static_assert(SH >= 5 && (L / 16) % 32 == 0, "the 32 columns of an MFMA tile share r2");
          if (ti < NCOLT) {
            int q, j;
            stage_col(k, q, j);
            const int r2 = j >> SH;
            const f32x16 y = dft16<CM>(x[k], dft256_aops<C::TAB_BLOCKS>(tabf, r2 * (16 / RN), lane));
            const int jb = (int)(((unsigned)j / NS) * NS * 16 + ((unsigned)j % NS));
            f32x2 cf;
            if constexpr (P == 0) {
              cf.x = stage_scale(0);
              cf.y = 0.0f;
            } else {
              cf = twiddle_rev(r2 * (j & (NS - 1)), INV_M, inv, stage_scale(P));
            }
            for (int s = 0; s < 8; ++s) {
              const int rr = 4 * h + (s & 3) + 8 * (s >> 2);
              f32x2 v;
              v.x = y[s];
              v.y = y[8 + s];
              lds_st(lds, lds_slot(q, jb, rr * NS), Tr::pack(P == 0 ? v * cf.x : cmul(v, cf)));
            }
          }
        } else
        if (ti < NCOLT) {
          int q, j;
          stage_col(k, q, j);
          const f32x16 y = dft16<CM>(x[k], A);
          const int jb = (int)(((unsigned)j / NS) * NS * 16 + ((unsigned)j % NS));
          // Computed twiddles of the lane's outputs rr = 4h + 8g + u, written at
          // m = jb + rr*NS: r2 = m >> lg(L/RN) is the same for all eight and
          // m % 16NS = j % NS + rr*NS, so the exponent is E + u*D with
          // D = r2*NS: W^E (one sin/cos per g) times powers of W^D.
          constexpr int RN = (P + 1 < S) ? 16 : T;        // radix of the next stage
          constexpr int LGQ2 = ilog2c(L / (RN > 1 ? RN : 1));
          constexpr int NSN = 16 * NS;
          constexpr float INV_M = 1.0f / (float)(NSN * (RN > 1 ? RN : 1));
          f32x2 tw_d;
          int r2 = 0;
          if constexpr (!LAST) {
            r2 = (jb + 4 * h * NS) >> LGQ2;
            if constexpr (!TABLE_TW) {
              tw_d = twiddle_rev(r2 * NS, INV_M, inv, 1.0f);
            } else {
              // split: W^D is the table entry (column r2 * L/(16 RN), rr = 1), whose
              // column has j % NS = 0 and the same r2; the table carries the scale
              constexpr int JD = L / (16 * (RN > 1 ? RN : 1));
              tw_d = __builtin_bit_cast(f32x2, buf_ld64(gtw, (P * L + r2 * JD * 16 + 1) * 8)) * (1.0f / a.st_scale[P]);
            }
          }
          // four-step twiddles of the last stage: W_N^(m * column), m = mo0 + u*NS
          f32x2 fs_d;
          if constexpr (LAST && !STAGED && FOURSTEP) {
            if constexpr (!TABLE_TW) fs_d = twiddle_rev(NS * (inst0 + q), inv_n, inv, 1.0f);
            else fs_d = fs_twiddle(q, NS) * (1.0f / a.out_scale);
          }
          for (int g = 0; g < 2; ++g) {
            f32x2 tw[4];
            const int mo0 = jb + (4 * h + 8 * g) * NS;
            if constexpr (!LAST) {
              if constexpr (TABLE_TW) {
                // W^E: table entry rr = 4h + 8g of row j (FP64-generated, scale folded in)
                tw[0] = __builtin_bit_cast(f32x2, buf_ld64(gtw, (P * L + j * 16 + 4 * h + 8 * g) * 8));
                if constexpr (P == 0) tw[0] = tw[0] * nsc;
              } else {
                tw[0] = twiddle_rev(r2 * (mo0 & (NSN - 1)), INV_M, inv, stage_scale(P));
              }
              for (int u = 1; u < 4; ++u) tw[u] = cmul(tw[u - 1], tw_d);
            }
            for (int u = 0; u < 4; ++u) {
              const int s = 4 * g + u;
              const int mo = mo0 + u * NS;
              // mo = jb + rr * NS: jb a multiple of 16 when NS = 1 (rr < 16), rr * NS one when NS >= 16
              const int rr = 4 * h + 8 * g + u;
              f32x2 v;
              v.x = y[s];
              v.y = y[8 + s];
              if constexpr (C::FUSE2 && P == S - 1) {
                // radix-2 tail: lane j < 16 holds x0 = X[m], lane j + 16 holds
                // x1 W = X[m + 256] W; outputs x0 + x1 W at m, x0 - x1 W at m + 256
                const f32x2 w = cmul(v, tw[u]);
                f32x2 o;
                o.x = lane_xor16(w.x);
                o.y = lane_xor16(w.y);
                const float sg = (lane & 16) ? -1.0f : 1.0f;
                f32x2 r;
                r.x = __builtin_fmaf(sg, w.x, o.x);
                r.y = __builtin_fmaf(sg, w.y, o.y);
                if constexpr (STAGED) lds_st(lds, lds_slot(q, jb, rr * NS), Tr::pack(finish(r, q, mo)));
                else store_out(q, mo, Tr::pack(finish(r, q, mo)));
              } else if constexpr (!LAST) {
                lds_st(lds, lds_slot(q, jb, rr * NS), Tr::pack(cmul(v, tw[u])));
              } else if constexpr (STAGED) {
                lds_st(lds, lds_slot(q, jb, rr * NS), Tr::pack(finish(v, q, mo)));
              }
            }
            if constexpr (LAST && !STAGED) {
              // direct store: offsets and four-step exponents advance linearly in u
              {
                const int o0 = C::out_off(q, mo0, sout), ostep = C::out_off(0, NS, sout);
                f32x2 fw;
                if constexpr (FOURSTEP) fw = fs_twiddle(q, mo0);
                for (int u = 0; u < 4; ++u) {
                  const int s = 4 * g + u;
                  f32x2 v;
                  v.x = y[s];
                  v.y = y[8 + s];
                  if constexpr (FOURSTEP) {
                    if (u > 0) fw = cmul(fw, fs_d);
                    v = cmul(v, fw);
                  } else if constexpr (!FOLD) {
                    v = v * single_scale();
                  }
                  emit(o0 + u * ostep, Tr::pack(v));
                }
              }
            }
          }
        }
      }
      if constexpr (!LAST || STAGED) block_barrier();
    });

    // ---------------- radix-T tail stage (VALU, FP32) ----------------
    // Column (q, j) reads and writes the same T positions j + r*L/T, so with
    // the staged store the results go back in place without a barrier.
    if constexpr (T > 1 && !C::FUSE2) {
      constexpr int NCOL = Q * L / T;
      constexpr int CPT = (NCOL + C::NT - 1) / C::NT;
      for (int k = 0; k < CPT; ++k) {
        const int c = tid + k * C::NT;
        if (c < NCOL) {
          int q, j;
          C::template colmap<T>(c, q, j);
          f32x2 v[T];
          for (int r = 0; r < T; ++r) v[r] = Tr::unpack(lds_ld(lds, lds_slot(q, j, r * (L / T))));
          dft_small<T>(v, inv);
          for (int r = 0; r < T; ++r) {
            const int mo = j + r * (L / T);
            const f32x2 w = finish(v[r], q, mo);
            if constexpr (STAGED) lds_st(lds, lds_slot(q, j, r * (L / T)), Tr::pack(w));
            else store_out(q, mo, Tr::pack(w));
          }
        }
      }
      if constexpr (STAGED) block_barrier();
    }

    // ---------------- staged store: instance-fastest order ----------------
    if constexpr (STAGED) {
      // element e = tid + i*NT is (q, m) = (tid % Q, tid / Q + i*NT/Q): one
      // per-lane offset plus a wave-uniform one (scalar register) per i
      static_assert(C::NT % Q == 0 && (Q * L) % C::NT == 0, "staged store covers the tile exactly");
      constexpr int MSTEP = C::NT / Q;
      const int q = (int)((unsigned)tid % Q), m0 = (int)((unsigned)tid / Q);
      const int vo = C::out_off(q, m0, sout);
      for (int i = 0; i < Q * L / C::NT; ++i)
      {
        const El e = lds_ld(lds, (C::QFAST || MSTEP % 16 == 0) ? lds_slot(q, m0, i * MSTEP) : C::lidx(q, m0 + i * MSTEP));
        if (check) watch = nonfinite_fold(watch, e, CM);
        Tr::store_s(gout, vo, i * MSTEP * sout, e);
      }
    }
    if (BFP && (bf & BFP_LAST)) bfp_last_tile(a, Q, inst0, grp, til, t_loc, e_out, nonfinite_seen(watch));
    block_barrier();  // LDS is reused by the next tile
  }
}

// A pass kernel: the workgroups loop over the pass's tiles.
template <int S, int T, int Q, int NW, int CM, int KIND>
MCFFT_DEV void fft_pass_body(const PassArgs& a) {
  const AOps A = load_aops(a.aops, lane_id());
  if constexpr (PassCfg<S, T, Q, NW, CM, KIND>::FOLD_TW) {
    // the row-scaled DFT matrices into LDS, read by every tile's stages
    typedef PassCfg<S, T, Q, NW, CM, KIND> C;
    MCFFT_DYN_LDS(uint32_t, lds);
    u32x4* tab = (u32x4*)(lds + C::TAB_OFF / 4);
    for (int i = thread_id(); i < C::TAB_BLOCKS * (kDft256Words / 3) / 4; i += C::NT)
      lds_st(tab, i, g_ld((const u32x4*)a.dft256, (long long)i));
    block_barrier();
  }
  // Column tiles narrower than a 128-byte line share lines with their
  // neighbours.  Workgroups are dealt to the 8 XCDs round robin (block b on XCD
  // b % 8), each XCD with its own L2, so within every aligned group of 32 blocks
  // the 4 blocks b, b+8, b+16, b+24 (one XCD) take 4 adjacent tiles.  This only
  // changes which block runs which tile.
  constexpr bool XCD_GROUP = PassCfg<S, T, Q, NW, CM, KIND>::QFAST && Q * (int)sizeof(typename Traits<CM>::T) < 128;
  int b = block_id();
  const int g = grid_size();
  if constexpr (XCD_GROUP) {
    if (b < (g & ~31)) b = (b & ~31) + 4 * (b & 7) + ((b >> 3) & 3);
  }
  SlotAcc sacc;
  if (a.bfp_flags & BFP_FIRST) sacc.epoch = (unsigned long long)uniform64(g_ld(&a.bfp_state->epoch, 0));
  // the last pass of an AUTO multi-pass exec advances the epoch (nothing in it
  // reads the epoch; the next first pass is a later launch)
  if ((a.bfp_flags & BFP_EPOCH) && block_id() == 0 && thread_id() == 0) atomic_add_dev64(&a.bfp_state->epoch, 1);
  // Row kinds in FP16/BF16: a software pipeline over the workgroup's tiles, the
  // next tile's stage-0 loads issued before the current tile is computed (a
  // workgroup barrier waits for LDS only, so they stay in flight through the
  // tile's stages).  Measured on MI355X: 3-10 % faster row passes; column and
  // split passes ran 4-90 % slower with it (the extra registers cost occupancy).
  constexpr bool PREFETCH = (KIND == KIND_ROW || KIND == KIND_ROWX) && CM != CM_SPLIT;
  if constexpr (PREFETCH) {
    typedef typename Traits<CM>::T El;
    constexpr int TPW = PassCfg<S, T, Q, NW, CM, KIND>::TPW;
    El nxt[TPW][8];
    if (b < a.n_tiles) tile_prefetch<S, T, Q, NW, CM, KIND>(a, b, nxt);
    for (int tile = b; tile < a.n_tiles; tile += g) {
      El cur[TPW][8];
      for (int k = 0; k < TPW; ++k)
        for (int t = 0; t < 8; ++t) cur[k][t] = nxt[k][t];
      if (tile + g < a.n_tiles) tile_prefetch<S, T, Q, NW, CM, KIND>(a, tile + g, nxt);
      fft_tile<S, T, Q, NW, CM, KIND>(a, tile, A, &sacc, cur);
    }
  } else {
    for (int tile = b; tile < a.n_tiles; tile += g) fft_tile<S, T, Q, NW, CM, KIND>(a, tile, A, &sacc);
  }
  slot_flush(sacc, a.bfp_slots);
}

}  // namespace mcfft

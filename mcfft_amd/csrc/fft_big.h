// fft_big.h -- single-pass 65536-point FFT: one workgroup (8 waves with 128
// elements per lane, or 16 waves with 64) holds a whole transform (256 KiB in
// FP16/BF16) in registers and exchanges it between its four radix-16 MFMA
// stages through LDS in two rounds of half the data.
//
// The exchange needs each lane to free, in each round, registers for the inputs
// it receives.  With the Stockham indexing used everywhere in fft_pass.h, bit 13
// of an element's position m decides its round: for a writer (stage P, MFMA
// tile k of the wave, column j = 32*(wave + NW k) + lane%32) that bit is bit 9
// of j (bit 1 of k for 8 waves, bit 0 for 16), for a reader (stage P+1, input
// slot t) it is bit 1 of t -- for every P.  So round 0 moves the outputs of the
// "early" tiles and delivers inputs t in {0,1,4,5}; inputs that belong to the
// other tiles wait in spare registers until round 1 frees their home.  Peak
// live data per lane: 1.25x the lane's share (8 waves: 128 + 32 packed
// elements, 2 waves per SIMD; 16 waves: 64 + 16, 4 waves per SIMD).
//
// Twiddles.  Between stages the element of output row rr (0..15) of column j
// needs W_M^(r2 * (j % NS)) * W_256^(r2 * rr), M = 256 NS, r2 = j >> 8.  A
// thread's MFMA tile k holds 32 columns with one r2, so the second factor scales
// the DFT matrix's rows: the kernel multiplies by a DFT matrix with rows scaled
// by W_256^(r2 rr) (16 of them, read from a table in LDS), and only the first
// factor, the same for all outputs of a lane's column, is applied to the FP32
// accumulators (stage 0: none, only the stage's scale).
//
// No reference counterpart: the reference has no 65536-point plan
// (source/mcfft_1d_device.cpp:695-775 falls to `default:` for it).
#pragma once
#include "fft_pass.h"

namespace mcfft {

template <int NW_>
struct BigCfg {
  static constexpr int L = 65536, LQ = 4096, NW = NW_, NT = 64 * NW_, TPW = 128 / NW_;
  static_assert(NW_ == 8 || NW_ == 16, "8 or 16 waves");
  static constexpr int HALF = 32768;
  static constexpr int LDS_ELEMS = HALF + HALF / 16;
  // block reductions: one word per wave in the exchange's padding slots
  // 17 q + 16, which no exchange access touches (see lidx below)
  static constexpr int RED_SLOT0 = 16, RED_STRIDE = 17;
  static constexpr int TAB_OFF = LDS_ELEMS * 4;  // then the row-scaled DFT matrices (kDft256Words)
  static constexpr int LDS_BYTES = TAB_OFF + kDft256Words * 4;
  static_assert(LDS_BYTES <= 160 * 1024, "LDS of one workgroup");
  // LDS slot of position m inside its round: drop bit 13, pad every 16 elements
  //   lidx(m) = c + (c >> 4),  c = (m & 0x1fff) | ((m >> 14) << 13).
  // Column j of MFMA tile k is low + 32 NW k (low = 32 wave + lane%32 < 32 NW).
  // Every position below is a per-lane part below 8192 plus a compile-time part
  // that is a multiple of 4096, so lidx splits into lidx(lane part) + lconst:
  //   reader (input r = 8h + t of column j):  rbase + 34 NW k + roff(t)
  //   writer (stage P, output rr of column j): wbase_P + lconst(512 NW k) + wmul_P * rr
  // and bit 13 (the exchange round) comes from the compile-time part alone.
  static constexpr int lconst(int m) {
    const int c = (m & 0x1fff) | ((m >> 14) << 13);
    return c + (c >> 4);
  }
  static constexpr int kpos(int k) { return 512 * NW * k; }         // writer: compile-time part of m'
  static constexpr bool late(int k) { return (kpos(k) >> 13) & 1; }  // writer round of tile k
  static constexpr int roff(int t) { return (t & 1) * 4352 + (t >> 2) * 8704; }
  static constexpr int rrc(int s) { return (s & 3) + 8 * (s >> 2); }  // rr - 4h of MFMA output slot s
  static constexpr int NLATE = TPW / 2;                               // tiles written in round 1
  // index among the round-1 tiles
  static constexpr int late_idx(int k) {
    int n = 0;
    for (int i = 0; i < k; ++i) n += late(i) ? 1 : 0;
    return n;
  }
};

// GBuf of transform `tile` (ROW layout, one transform per tile)
MCFFT_DEV GBuf big_view(const void* p, const long long* off, int gpt, long long sg, long long st, int tpg, int tile) {
  const int grp = tile / tpg, til = tile - grp * tpg;
  const long long base = group_base(off, gpt, sg, grp) + (long long)til * st;
  return make_gbuf((const char*)p + base * 4, 65536u * 4u);
}

// One transform whose stage-0 inputs are already in x.  The last stage stores
// MFMA tile k and, if next >= 0, immediately loads tile k of transform `next`
// into the registers it just freed, so the next transform's loads overlap this
// one's stores.
// AUTO: the kernel serves MCFFT_SCALE_AUTO plans (first-pass normalisation and
// output exponents); otherwise statically scaled plans (output exponents).
template <int CM, int NW, bool AUTO>
MCFFT_DEV void fft_big_tile(const PassArgs& a, int tile, int next,
                            typename Traits<CM>::T (&x)[BigCfg<NW>::TPW][8], SlotAcc& sacc) {
  static_assert(CM == CM_F16 || CM == CM_BF16, "4-byte elements only");
  typedef Traits<CM> Tr;
  typedef typename Tr::T El;
  typedef BigCfg<NW> B;
  MCFFT_DYN_LDS(El, lds);
  const uint32_t* tab = (const uint32_t*)((const unsigned char*)lds + B::TAB_OFF);
  // The per-lane bases below do not depend on the tile; an opaque copy of the
  // thread index keeps the compiler from hoisting them (and everything derived
  // from them) out of the workgroup's tile loop.
  const int tid = opaque(thread_id());
  const int lane = tid & 63, wave = tid >> 6, h = lane >> 5, c5 = lane & 31;
  const int low = 32 * wave + c5;  // column j of MFMA tile k is low + 32 NW k
  const GBuf gout = big_view(a.out, a.gather_out, a.gather_gpt, a.out_sg, a.out_st, a.tiles_per_group, tile);
  const GBuf gnext =
      big_view(a.in, a.gather_in, a.gather_gpt, a.in_sg, a.in_st, a.tiles_per_group, next < 0 ? tile : next);
  const int rbase = low + (low >> 4) + 2 * h * 8704;
  const int vin = (low + 8 * h * B::LQ) * 4;  // stage-0 input r = 8h + t of column j: element j + r * 4096

  El tmp[B::NLATE][4];
  constexpr int T0[4] = {0, 1, 4, 5}, T1[4] = {2, 3, 6, 7};

  // block floating point (bfp.h): an AUTO first pass normalises the transform
  // from its largest input |component| (stage 0's scale carries 2^n)
  const int bf = a.bfp_flags;
  int e_out = a.bfp_e_base;
  float nsc = 1.0f;
  if (AUTO && (bf & BFP_FIRST)) {
    uint32_t acc = 0;
#pragma unroll
    for (int k = 0; k < B::TPW; ++k)
#pragma unroll
      for (int t = 0; t < 8; ++t) acc = pk_absmax(acc, x[k][t]);
    const int grp = tile / a.tiles_per_group;
    nsc = bfp_first_tile(a, acc, CM, (unsigned*)lds + B::RED_SLOT0, NW, tile,
                         a.bfp_gpt > 0 ? grp / a.bfp_gpt : 0, e_out, sacc, B::RED_STRIDE);
  }

  For<0, 4>::run([&](auto pc) __attribute__((always_inline)) {
    constexpr int P = decltype(pc)::value;
    // per-lane part of the position written by output rr (the k part is kpos(k))
    const int hi8 = low >> 8, lo8 = low & 255;
    const int wbase = P == 0 ? 17 * low + 4 * h
                    : P == 1 ? (low >> 4) * 272 + (low & 15) + 68 * h
                             : hi8 * 4352 + lo8 + (lo8 >> 4) + 272 * 4 * h;
    constexpr int wmul = P == 0 ? 1 : P == 1 ? 17 : 272;
    const int vout = (low + 4 * h * B::LQ) * 4;   // last stage: element j + rr * 4096
    // The element written at m' = (j / NS) * 16 NS + j % NS + rr * NS is input
    // r2 = m' >> 12 = j >> 8 of the next stage's column m' & 4095, and its
    // twiddle W_{256 NS}^(r2 * (m' % 16 NS)) splits into W_256^(r2 rr) (in the
    // tile's DFT matrix) and the column factor cf = W_{256 NS}^(r2 * (j % NS)).
    // Tile k's r2 is r20 + DR * k, so cf advances by W_{256 NS}^(DR * (j % NS))
    // per tile.  The stage's scale rides on cf (stage 0: on the accumulators);
    // the last stage's scale rides on stage 2's.
    constexpr int NS = ipow16(P < 3 ? P : 0);
    constexpr float INV_M = 1.0f / (float)(256 * NS);
    constexpr int DR = (32 * NW) >> 8;
    const int r20 = low >> 8;
    const float sc = P == 2 ? a.st_scale[2] * a.st_scale[3] : P == 0 ? a.st_scale[0] * nsc : a.st_scale[P];
    f32x2 cf = {sc, 0.0f}, cf_step = {1.0f, 0.0f};
    if constexpr (P == 1 || P == 2) {
      const int jm = low & (NS - 1);
      cf = twiddle_rev(r20 * jm, INV_M, a.inverse != 0, sc);
      cf_step = twiddle_rev(DR * jm, INV_M, a.inverse != 0, 1.0f);
    }
    AOps A0;
    if constexpr (P == 3) A0 = dft256_aops(tab, 0, lane);
#pragma unroll
    for (int k = 0; k < B::TPW; ++k) {
      f32x16 y;
      if constexpr (P < 3) y = dft16<CM>(x[k], dft256_aops(tab, r20 + DR * k, lane));
      else y = dft16<CM>(x[k], A0);
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        f32x2 v;
        v.x = y[s];
        v.y = y[8 + s];
        if constexpr (P == 0) {
          x[k][s] = Tr::pack(v * cf.x);
        } else if constexpr (P < 3) {
          x[k][s] = Tr::pack(cmul(v, cf));
        } else {
          const El e = Tr::pack(v);
          buf_st32s(gout, vout, (k * 32 * NW + B::rrc(s) * B::LQ) * 4, e);
        }
      }
      if constexpr (P == 1 || P == 2) cf = cmul(cf, cf_step);
      if constexpr (P == 3) {
        if (next >= 0) {
#pragma unroll
          for (int t = 0; t < 8; ++t) x[k][t] = buf_ld32s(gnext, vin, k * 128 * NW + t * 16384);
        }
      }
    }
    if constexpr (P < 3) {
      // ---- round 0: outputs of the early tiles out, inputs t in T0 in ----
#pragma unroll
      for (int k = 0; k < B::TPW; ++k) {
        if (B::late(k)) continue;
#pragma unroll
        for (int s = 0; s < 8; ++s) lds_st(lds, wbase + B::lconst(B::kpos(k)) + wmul * B::rrc(s), x[k][s]);
      }
      block_barrier();
#pragma unroll
      for (int k = 0; k < B::TPW; ++k) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const El v = lds_ld(lds, rbase + 34 * NW * k + B::roff(T0[i]));
          if (!B::late(k)) x[k][T0[i]] = v;
          else tmp[B::late_idx(k)][i] = v;
        }
      }
      block_barrier();
      // ---- round 1: outputs of the late tiles out, inputs t in T1 in ----
#pragma unroll
      for (int k = 0; k < B::TPW; ++k) {
        if (!B::late(k)) continue;
#pragma unroll
        for (int s = 0; s < 8; ++s) lds_st(lds, wbase + B::lconst(B::kpos(k)) + wmul * B::rrc(s), x[k][s]);
      }
      block_barrier();
#pragma unroll
      for (int k = 0; k < B::TPW; ++k) {
#pragma unroll
        for (int i = 0; i < 4; ++i) x[k][T1[i]] = lds_ld(lds, rbase + 34 * NW * k + B::roff(T1[i]));
        if (B::late(k)) {
#pragma unroll
          for (int i = 0; i < 4; ++i) x[k][T0[i]] = tmp[B::late_idx(k)][i];
        }
      }
      block_barrier();
    }
  });
  if (bf & BFP_LAST) {
    const int grp = tile / a.tiles_per_group, til = tile - grp * a.tiles_per_group;
    // no Inf/NaN watch over the stores: every form tried (a fold of the packed
    // values, a maximum or compares of the FP32 results) made the static kernel
    // spill hundreds of bytes per lane; AUTO plans cannot overflow
    bfp_last_tile(a, 1, til, grp, til, a.bfp_gpt > 0 ? grp / a.bfp_gpt : 0, e_out, false);
  }
}

template <int CM, int NW, bool AUTO>
MCFFT_DEV void fft_big_body(const PassArgs& a) {
  typedef BigCfg<NW> B;
  typename Traits<CM>::T x[B::TPW][8];
  {
    // the row-scaled DFT matrices into LDS (read by every stage; the first
    // barrier of stage 0 comes after its first use, hence the one here)
    MCFFT_DYN_LDS(uint32_t, lds);
    u32x4* tab = (u32x4*)(lds + B::TAB_OFF / 4);
    for (int i = thread_id(); i < kDft256Words / 4; i += B::NT) lds_st(tab, i, g_ld((const u32x4*)a.dft256, (long long)i));
    block_barrier();
  }
  const int g = grid_size();
  int tile = block_id();
  // start stagger (PassArgs::stagger_*), only when every workgroup has several
  // transforms: with one each, the late half's wait would add to the kernel time
  if (a.stagger_ticks > 0 && a.stagger_phases > 1 && a.n_tiles >= 4 * g)
    stall_ticks(((tile >> 3) % a.stagger_phases) * a.stagger_ticks);
  if (tile < a.n_tiles) {
    const int tid = thread_id(), lane = tid & 63, low = 32 * (tid >> 6) + (lane & 31);
    const int vin = (low + 8 * (lane >> 5) * B::LQ) * 4;
    const GBuf gin = big_view(a.in, a.gather_in, a.gather_gpt, a.in_sg, a.in_st, a.tiles_per_group, tile);
#pragma unroll
    for (int k = 0; k < B::TPW; ++k)
#pragma unroll
      for (int t = 0; t < 8; ++t) x[k][t] = buf_ld32s(gin, vin, k * 128 * NW + t * 16384);
  }
  SlotAcc sacc;
  if (AUTO && (a.bfp_flags & BFP_FIRST)) sacc.epoch = (unsigned long long)uniform64(g_ld(&a.bfp_state->epoch, 0));
  for (; tile < a.n_tiles; tile += g) {
    const int next = tile + g < a.n_tiles ? tile + g : -1;
    fft_big_tile<CM, NW, AUTO>(a, tile, next, x, sacc);
  }
  if constexpr (AUTO) slot_flush(sacc, a.bfp_slots);
}

}  // namespace mcfft

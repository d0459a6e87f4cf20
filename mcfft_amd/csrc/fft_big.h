// fft_big.h -- single-pass 65536-point FFT: one workgroup (8 waves, 128
// elements per lane) holds a whole transform (256 KiB in FP16/BF16) in
// registers and exchanges it between its four radix-16 MFMA stages through LDS
// in two rounds of half the data.
//
// The exchange needs each lane to free, in each round, registers for the inputs
// it receives.  With the Stockham indexing used everywhere in fft_pass.h, bit 13
// of an element's position m decides its round: for a writer (stage P, MFMA
// tile k of the wave, column j = 32*(wave + 8k) + lane%32) that bit is bit 9 of
// j = bit 1 of k, for a reader (stage P+1, input slot t) it is bit 1 of t -- for
// every P.  So round 0 moves the outputs of tiles k with bit 1 clear and
// delivers inputs t in {0,1,4,5}; inputs that belong to the other tiles wait in
// 32 spare registers until round 1 frees their home.  Peak live data per lane:
// 128 + 32 packed elements (2 waves per SIMD, up to 256 VGPRs each).
//
// No reference counterpart: the reference has no 65536-point plan
// (source/mcfft_1d_device.cpp:695-775 falls to `default:` for it).
#pragma once
#include "fft_pass.h"

namespace mcfft {

struct BigCfg {
  static constexpr int L = 65536, LQ = 4096, NW = 8, NT = 512, TPW = 16;
  static constexpr int HALF = 32768;
  static constexpr int LDS_ELEMS = HALF + HALF / 16;
  static constexpr int LDS_BYTES = LDS_ELEMS * 4;
  // LDS slot of position m inside its round: drop bit 13, pad every 16 elements
  //   lidx(m) = c + (c >> 4),  c = (m & 0x1fff) | ((m >> 14) << 13).
  // With column j = low + 256 k (low = 32 wave + lane%32) this is linear in the
  // tile k and the slot, so every access below is one per-lane base plus a
  // compile-time constant:
  //   reader (input r = 8h + t of column j):   rbase + 272 k + roff(t)
  //   writer (stage P, output rr of column j): wbase_P + kpart(k) + wmul_P * rr
  static constexpr int kpart(int k) { return (k & 1) * 4352 + (k >> 2) * 8704; }
  static constexpr int roff(int t) { return (t & 1) * 4352 + (t >> 2) * 8704; }
  static constexpr int rrc(int s) { return (s & 3) + 8 * (s >> 2); }  // rr - 4h of MFMA output slot s
};

// GBuf of transform `tile` (ROW layout, one transform per tile)
MCFFT_DEV GBuf big_view(const void* p, long long sg, long long st, int tpg, int tile) {
  const int grp = tile / tpg, til = tile - grp * tpg;
  const long long base = (long long)grp * sg + (long long)til * st;
  return make_gbuf((const char*)p + base * 4, (unsigned)BigCfg::L * 4u);
}

// One transform whose stage-0 inputs are already in x.  The last stage stores
// MFMA tile k and, if next >= 0, immediately loads tile k of transform `next`
// into the registers it just freed, so the next transform's loads overlap this
// one's stores (one workgroup per CU leaves no other overlap).
template <int CM>
MCFFT_DEV void fft_big_tile(const PassArgs& a, int tile, int next, const AOps& A,
                            typename Traits<CM>::T (&x)[BigCfg::TPW][8]) {
  static_assert(CM == CM_F16 || CM == CM_BF16, "4-byte elements only");
  typedef Traits<CM> Tr;
  typedef typename Tr::T El;
  typedef BigCfg B;
  MCFFT_DYN_LDS(El, lds);
  // The per-lane bases below do not depend on the tile; an opaque copy of the
  // thread index keeps the compiler from hoisting them (and everything derived
  // from them) out of the workgroup's tile loop.
  const int tid = opaque(thread_id());
  const int lane = tid & 63, wave = tid >> 6, h = lane >> 5, c5 = lane & 31;
  const int low = 32 * wave + c5;  // column j of MFMA tile k is low + 256 k
  const GBuf gout = big_view(a.out, a.out_sg, a.out_st, a.tiles_per_group, tile);
  const GBuf gnext = big_view(a.in, a.in_sg, a.in_st, a.tiles_per_group, next < 0 ? tile : next);
  const int rbase = low + (low >> 4) + 2 * h * 8704;
  const int vin = (low + 8 * h * B::LQ) * 4;  // stage-0 input r = 8h + t of column j: element j + r * 4096

  El tmp[8][4];
  constexpr int T0[4] = {0, 1, 4, 5}, T1[4] = {2, 3, 6, 7};

  For<0, 4>::run([&](auto pc) __attribute__((always_inline)) {
    constexpr int P = decltype(pc)::value;
    const int wbase = P == 0 ? 17 * low + 4 * h
                    : P == 1 ? (low >> 4) * 272 + (low & 15) + 68 * h
                             : low + (low >> 4) + 272 * 4 * h;
    constexpr int wmul = P == 0 ? 1 : P == 1 ? 17 : 272;
    const int vout = (low + 4 * h * B::LQ) * 4;   // last stage: element j + rr * 4096
#pragma unroll
    for (int k = 0; k < B::TPW; ++k) {
      const f32x16 y = dft16<CM>(x[k], A);
#pragma unroll
      for (int g = 0; g < 2; ++g) {
        f32x2 tw[4];
        if constexpr (P < 3) {
          // element written at m' = jb + rr * NS is input m' >> 12 of the next
          // stage's column m' & 4095: twiddle W_{256 NS}^((m' >> 12) * (m' % 16NS))
          // r2 = m' >> 12 is the same for the lane's outputs and m' % 16NS grows
          // by NS per u: W^E (one sin/cos) times powers of W^(r2 NS).  The last
          // stage's scale rides on the twiddles of stage 2.
          constexpr int NS = ipow16(P);
          constexpr float INV_M = 1.0f / (float)(256 * NS);
          const int j = low + 256 * k;
          const int mo0 = (int)(((unsigned)j / NS) * NS * 16 + ((unsigned)j % NS)) + (4 * h + 8 * g) * NS;
          const int r2 = mo0 >> 12;
          const float sc = P == 2 ? a.st_scale[2] * a.st_scale[3] : a.st_scale[P];
          const f32x2 d = twiddle_rev(r2 * NS, INV_M, a.inverse != 0, 1.0f);
          tw[0] = twiddle_rev(r2 * (mo0 & (16 * NS - 1)), INV_M, a.inverse != 0, sc);
#pragma unroll
          for (int u = 1; u < 4; ++u) tw[u] = cmul(tw[u - 1], d);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int s = 4 * g + u;
          f32x2 v;
          v.x = y[s];
          v.y = y[8 + s];
          if constexpr (P < 3) {
            x[k][s] = Tr::pack(cmul(v, tw[u]));
          } else {
            buf_st32s(gout, vout, (k * 256 + B::rrc(s) * B::LQ) * 4, Tr::pack(v));
          }
        }
      }
      if constexpr (P == 3) {
        if (next >= 0) {
#pragma unroll
          for (int t = 0; t < 8; ++t) x[k][t] = buf_ld32s(gnext, vin, k * 1024 + t * 16384);
        }
      }
    }
    if constexpr (P < 3) {
      // ---- round 0: outputs of tiles with k bit 1 clear out, inputs t in T0 in ----
#pragma unroll
      for (int k = 0; k < B::TPW; ++k) {
        if (k & 2) continue;
#pragma unroll
        for (int s = 0; s < 8; ++s) lds_st(lds, wbase + B::kpart(k) + wmul * B::rrc(s), x[k][s]);
      }
      block_barrier();
#pragma unroll
      for (int k = 0; k < B::TPW; ++k) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const El v = lds_ld(lds, rbase + 272 * k + B::roff(T0[i]));
          if (!(k & 2)) x[k][T0[i]] = v;
          else tmp[(k >> 2) * 2 + (k & 1)][i] = v;
        }
      }
      block_barrier();
      // ---- round 1: outputs of the other tiles out, inputs t in T1 in ----
#pragma unroll
      for (int k = 0; k < B::TPW; ++k) {
        if (!(k & 2)) continue;
#pragma unroll
        for (int s = 0; s < 8; ++s) lds_st(lds, wbase + B::kpart(k) + wmul * B::rrc(s), x[k][s]);
      }
      block_barrier();
#pragma unroll
      for (int k = 0; k < B::TPW; ++k) {
#pragma unroll
        for (int i = 0; i < 4; ++i) x[k][T1[i]] = lds_ld(lds, rbase + 272 * k + B::roff(T1[i]));
        if (k & 2) {
#pragma unroll
          for (int i = 0; i < 4; ++i) x[k][T0[i]] = tmp[(k >> 2) * 2 + (k & 1)][i];
        }
      }
      block_barrier();
    }
  });
}

template <int CM>
MCFFT_DEV void fft_big_body(const PassArgs& a) {
  typedef BigCfg B;
  const AOps A = load_aops(a.aops, lane_id());
  typename Traits<CM>::T x[B::TPW][8];
  const int g = grid_size();
  int tile = block_id();
  if (tile < a.n_tiles) {
    const int tid = thread_id(), lane = tid & 63, low = 32 * (tid >> 6) + (lane & 31);
    const int vin = (low + 8 * (lane >> 5) * B::LQ) * 4;
    const GBuf gin = big_view(a.in, a.in_sg, a.in_st, a.tiles_per_group, tile);
#pragma unroll
    for (int k = 0; k < B::TPW; ++k)
#pragma unroll
      for (int t = 0; t < 8; ++t) x[k][t] = buf_ld32s(gin, vin, k * 1024 + t * 16384);
  }
  for (; tile < a.n_tiles; tile += g) {
    const int next = tile + g < a.n_tiles ? tile + g : -1;
    fft_big_tile<CM>(a, tile, next, A, x);
  }
}

}  // namespace mcfft

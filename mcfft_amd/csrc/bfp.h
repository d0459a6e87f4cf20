// bfp.h -- data-following block floating point (MCFFT_SCALE_AUTO) and the
// per-exec bookkeeping the pass kernels share: per-transform output exponents
// and a status word.
//
// AUTO scaling of transform t:
//   * First pass: each tile finds M, the largest |re| or |im| of its input, and
//     multiplies its stage-0 results by 2^n with n chosen so that M 2^n < 2^13
//     (n rounded down to a multiple of `quantum`); every radix-R stage of every
//     pass divides by R.  |DFT_L x| <= L max|x| with |x| <= sqrt(2) M, so no
//     stored value exceeds 2^13.5 (FP16 holds 65504 ~ 2^16): no finite input
//     overflows, and small inputs (generate_data's ~1e-4 values, partly FP16
//     subnormals) are scaled up instead of underflowing.  The tile's exponent e
//     (true value = stored * 2^e) goes to a per-tile byte array, and into the
//     transform's two slots: max and min of e over its tiles.
//   * Later passes: the transform's data carries E = max e.  If every tile agreed
//     (min = max) the data is used as it is; otherwise the pass right after the
//     first multiplies each element by 2^(e_tile - E) (exact, <= 1) as it loads
//     it.  The output exponent is E plus the later passes' stage shifts.
//   * The last pass writes out_exp[t] (plus the caller's input exponent of t when
//     one is given).
// A single-pass plan's tiles hold whole transforms: their exponent is final.
//
// Slots are 64-bit: (epoch << 16) | value, value = kBfpBias + e (max slot) or
// kBfpBias - e (min slot), combined with atomic max.  The first pass tags its
// values with the plan's current epoch, so they override whatever an earlier
// exec left; every tile reports (an all-zero tile with sentinel values that no
// real exponent loses to), so after the first pass every slot of the launch
// carries the current epoch, and later passes read the value bits only.  One
// thread of the last pass advances the epoch (48 bits: no wrap in practice).
// Nothing is cleared between execs, which keeps an exec self-contained under
// hipGraph replay.
//
// Status: bit kStatusNonFinite is set (sticky until the host reads and clears
// it) when an AUTO first pass sees an Inf/NaN input, or when a statically
// scaled plan's last pass stores an Inf/NaN (overflow of FP16 included) --
// except the single-pass 65536-point kernel, which has no registers to spare
// for watching its stores.
//
// No reference counterpart: the reference chains its passes with no scaling at
// all (source/mcfft_1d_device.cpp:596-678).
#pragma once
#include "dev_prims.h"

namespace mcfft {

// PassArgs::bfp_flags
enum : int {
  BFP_AUTO = 1,       // data-following scaling (MCFFT_SCALE_AUTO)
  BFP_FIRST = 2,      // AUTO first pass: measure, normalise and record each tile
  BFP_RECONCILE = 4,  // AUTO pass after the first: tiles' exponents may differ
  BFP_LAST = 8,       // last pass: write per-transform output exponents
  BFP_EPOCH = 16,     // AUTO multi-pass last pass: advance the slot epoch
  BFP_CHECK = 32,     // static scaling, last pass: flag non-finite outputs
  BFP_SINGLE = 64,    // each instance of the pass is a transform (single-pass 1D)
};

constexpr unsigned kStatusNonFinite = 2u;
constexpr int kBfpTarget = 12;          // normalised maximum in [2^12, 2^13) at quantum 1
constexpr int kBfpBias = 0x8000;        // slot value = kBfpBias +- e (16 bits)
constexpr int kBfpZeroSentinel = 30000; // an all-zero tile reports e = -30000 / +30000
constexpr int kBfpZeroTile = -128;      // tile byte of an all-zero tile

struct ExecState {                 // device, one per plan
  unsigned long long epoch;        // AUTO multi-pass plans: tag of the current exec's slot values
  unsigned status;                 // sticky status bits
  unsigned pad;
};

// largest |component| of packed 16-bit pairs, kept as packed 16-bit maxima of the
// sign-cleared bit patterns (FP16 and BF16 magnitudes order like their bits)
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
MCFFT_DEV uint32_t pk_absmax(uint32_t acc, uint32_t pair) {
  const u16x2 a = __builtin_bit_cast(u16x2, pair & 0x7fff7fffu);
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(u16x2, acc), a));
}
MCFFT_DEV uint32_t pk_absmax(uint32_t acc, u32x2 split) { return pk_absmax(acc, split.x); }  // hi parts
MCFFT_DEV uint32_t pk_fold(uint32_t acc) {
  const uint32_t lo = acc & 0xffffu, hi = acc >> 16;
  return lo > hi ? lo : hi;
}
MCFFT_DEV bool bits_nonfinite(uint32_t bits, int cm) { return bits >= (cm == 1 ? 0x7f80u : 0x7c00u); }
// Inf/NaN watch over stored pairs: acc starts at all ones; each pair folds in
// ~e & (exponent field of both halves) with a packed 16-bit minimum, so a half
// of acc is zero iff some stored half had an all-ones exponent (Inf or NaN;
// split elements: their hi pair).  Branch-free, one register.
MCFFT_DEV uint32_t expmask2(int cm) { return cm == 1 ? 0x7f807f80u : 0x7c007c00u; }
MCFFT_DEV uint32_t nonfinite_fold(uint32_t acc, uint32_t e, int cm) {
  const u16x2 v = __builtin_bit_cast(u16x2, ~e & expmask2(cm));
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(u16x2, acc), v));
}
MCFFT_DEV uint32_t nonfinite_fold(uint32_t acc, u32x2 e, int cm) { return nonfinite_fold(acc, e.x, cm); }
MCFFT_DEV bool nonfinite_seen(uint32_t acc) { return (acc & 0xffffu) == 0 || (acc >> 16) == 0; }

// 2^n as a float, |n| <= 126
MCFFT_DEV float exp2i(int n) { return __builtin_bit_cast(float, (uint32_t)(n + 127) << 23); }

// Normalisation exponent n of a tile whose largest |component| has the 16-bit
// pattern `bits` (FP16, or BF16 when cm == 1): M 2^n in [2^(13-q), 2^13).
MCFFT_DEV int bfp_norm(uint32_t bits, int cm, int quantum) {
  const float m = cm == 1 ? bf16_val(bits) : f16_val(bits);
  const int ef = (int)((__builtin_bit_cast(uint32_t, m) >> 23) & 0xffu);
  const int k = ef - 127;  // floor(log2 m) for normal m (a subnormal m counts as 2^-127)
  int n = kBfpTarget - k;
  if (quantum > 1) n = (n >= 0 ? n / quantum : -((-n + quantum - 1) / quantum)) * quantum;
  return n < -100 ? -100 : n > 100 ? 100 : n;
}

// Largest |component| over the workgroup (each wave's lane 0 writes its wave's
// value to red[wave * stride]; every thread gets the maximum, as a wave-uniform
// value).  Ends after a barrier.
MCFFT_DEV uint32_t block_max(uint32_t v, unsigned* red, int nw, int wave, int lane, int stride = 1) {
  v = wave_max(v);
  if (lane == 0) lds_st(red, wave * stride, v);
  block_barrier();
  uint32_t m = 0;
  for (int w = 0; w < nw; ++w) {
    const uint32_t x = lds_ld(red, w * stride);
    m = x > m ? x : m;
  }
  return (uint32_t)uniform((int)m);
}

// A workgroup's pending slot update: the largest kBfpBias + e and kBfpBias - e
// of its first-pass tiles of transform t, since consecutive tiles of a
// workgroup often belong to one transform.  Flushed (two atomics whose results
// nothing waits for) when the transform changes and at the end of the launch.
struct SlotAcc {
  unsigned long long epoch = 0;  // the launch's tag (read once per workgroup)
  long long t = -1;
  unsigned vmax = 0, vmin = 0;
};
MCFFT_DEV void slot_flush(SlotAcc& s, unsigned long long* slots) {
  if (s.t >= 0 && thread_id() == 0) {
    atomic_max_dev64(slots + 2 * s.t, (s.epoch << 16) | s.vmax);
    atomic_max_dev64(slots + 2 * s.t + 1, (s.epoch << 16) | s.vmin);
  }
  s.t = -1;
}
// e: the tile's exponent; an all-zero tile passes zero = true
MCFFT_DEV void slot_note(SlotAcc& s, unsigned long long* slots, long long t, int e, bool zero) {
  if (t != s.t) {
    slot_flush(s, slots);
    s.t = t;
    s.vmax = s.vmin = 0;
  }
  const unsigned hi = (unsigned)(kBfpBias + (zero ? -kBfpZeroSentinel : e));
  const unsigned lo = (unsigned)(kBfpBias - (zero ? kBfpZeroSentinel : e));
  s.vmax = hi > s.vmax ? hi : s.vmax;
  s.vmin = lo > s.vmin ? lo : s.vmin;
}
// exponents a later pass reads from transform t's slots; emax = emin = 0 for a
// transform whose first-pass tiles all held zeros.  t is wave-uniform and the
// slots were written by an earlier launch, so they are scalar loads (whose wait
// does not cover the previous tile's stores, unlike a vector load's)
MCFFT_DEV void slot_read(const unsigned long long* slots, long long t, int& emax, int& emin) {
  const int vmax = (int)(s_ld(slots, 2 * t) & 0xffffu), vmin = (int)(s_ld(slots, 2 * t + 1) & 0xffffu);
  emax = vmax - kBfpBias;
  emin = kBfpBias - vmin;
  if (emax <= -kBfpZeroSentinel) emax = emin = 0;
}

}  // namespace mcfft

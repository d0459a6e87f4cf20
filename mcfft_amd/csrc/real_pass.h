// real_pass.h -- real-input forward (R2C) and real-output inverse (C2R) 1D
// transforms of even length N = 2H on top of the complex plans.
//
// R2C: the N reals of a transform, read as H complex values z[n] = x[2n] +
// i x[2n+1], go through a length-H complex FFT Z = E + iO (E, O: the FFTs of the
// even and odd samples).  This pass recovers, for k = 0..H,
//     E[k] = (Z[k] + conj Z[H-k]) / 2,   O[k] = (Z[k] - conj Z[H-k]) / (2i),
//     X[k] = E[k] + W_N^k O[k],           X[H-k] = conj(E[k] - W_N^k O[k]),
// one thread per pair (k, H-k), k = 0..H/2 (indices of Z taken mod H).
//
// C2R (unnormalised inverse of the Hermitian-extended spectrum): the pre-pass
// forms, for k = 0..H-1,
//     Z'[k] = ((X[k] + conj X[H-k]) + i (X[k] - conj X[H-k]) W_N^-k) / 2
// and a length-H inverse complex FFT of Z' gives (x[2n] + i x[2n+1]) / 2 (the
// plan reports one more power of two in its exponent).
//
// W_N^k comes from a table generated in FP64 on the GPU (H + 1 entries); the
// arithmetic is FP32, one rounding into the storage format.
//
// No reference counterpart: the reference has only the forward C2C transform.
#pragma once
#include "fft_pass.h"

namespace mcfft {

struct RealArgs {
  const void* in;
  void* out;
  const f32x2* tw;       // W_N^k, k = 0..H, forward sign (conjugated for the inverse)
  long long h;           // H = N / 2
  long long batch;
  int in_fmt;            // storage format of `in` (split compute mode: FP16 or split)
};

template <int CM>
MCFFT_DEV f32x2 real_ld(GBuf b, long long i, int fmt) {
  return Traits<CM>::unpack(Traits<CM>::load(b, (int)i, fmt));
}
MCFFT_DEV f32x2 conj2(f32x2 v) {
  v.y = -v.y;
  return v;
}
MCFFT_DEV f32x2 mul_by_i(f32x2 v) {  // i * v
  f32x2 r;
  r.x = -v.y;
  r.y = v.x;
  return r;
}

// R2C post-pass: Z (H complex per transform, contiguous) -> X (H + 1 per transform)
template <int CM>
MCFFT_DEV void r2c_post_body(const RealArgs& a) {
  typedef Traits<CM> Tr;
  const long long h = a.h, pairs = h / 2 + 1, total = pairs * a.batch;
  const long long stride = (long long)grid_size() * 1024;
  for (long long i = (long long)block_id() * 1024 + thread_id(); i < total; i += stride) {
    const long long b = i / pairs, k = i - b * pairs;
    const long long kk = (h - k) % h;  // index of Z[H - k] (mod H)
    const GBuf gz = make_gbuf((const char*)a.in + b * h * (long long)sizeof(typename Tr::T),
                              (unsigned)(h * (long long)sizeof(typename Tr::T)));
    const GBuf gx = make_gbuf((char*)a.out + b * (h + 1) * (long long)sizeof(typename Tr::T),
                              (unsigned)((h + 1) * (long long)sizeof(typename Tr::T)));
    const GBuf gt = make_gbuf(a.tw, (unsigned)((h + 1) * 8));
    const f32x2 zk = real_ld<CM>(gz, k % h, a.in_fmt), zr = conj2(real_ld<CM>(gz, kk, a.in_fmt));
    const f32x2 e = (zk + zr) * 0.5f;
    const f32x2 o = mul_by_i(zr - zk) * 0.5f;  // (Z[k] - conj Z[H-k]) / (2i)
    const f32x2 t = cmul(__builtin_bit_cast(f32x2, buf_ld64(gt, (int)(k * 8))), o);
    Tr::store(gx, (int)k, Tr::pack(e + t));
    if (k != h - k) Tr::store(gx, (int)(h - k), Tr::pack(conj2(e - t)));
  }
}

// C2R pre-pass: X (H + 1 complex per transform) -> Z' (H per transform)
template <int CM>
MCFFT_DEV void c2r_pre_body(const RealArgs& a) {
  typedef Traits<CM> Tr;
  const long long h = a.h, total = h * a.batch;
  const int esz_in = (CM == CM_SPLIT && a.in_fmt == FMT_SPLIT) ? 8 : 4;
  const long long stride = (long long)grid_size() * 1024;
  for (long long i = (long long)block_id() * 1024 + thread_id(); i < total; i += stride) {
    const long long b = i / h, k = i - b * h;
    const GBuf gx = make_gbuf((const char*)a.in + b * (h + 1) * esz_in, (unsigned)((h + 1) * esz_in));
    const GBuf gz = make_gbuf((char*)a.out + b * h * (long long)sizeof(typename Tr::T),
                              (unsigned)(h * (long long)sizeof(typename Tr::T)));
    const GBuf gt = make_gbuf(a.tw, (unsigned)((h + 1) * 8));
    const f32x2 xk = real_ld<CM>(gx, k, a.in_fmt), xr = conj2(real_ld<CM>(gx, h - k, a.in_fmt));
    const f32x2 w = __builtin_bit_cast(f32x2, buf_ld64(gt, (int)(k * 8)));  // W_N^-k (table of the inverse)
    const f32x2 z = ((xk + xr) + mul_by_i(cmul(xk - xr, w))) * 0.5f;
    Tr::store(gz, (int)k, Tr::pack(z));
  }
}

}  // namespace mcfft

// real_pass.h -- real-input forward (R2C) and real-output inverse (C2R)
// transforms (1D, 2D, 3D; last dimension N = 2H even) on top of the complex
// plans.
//
// R2C: the N reals of a transform, read as H complex values z[n] = x[2n] +
// i x[2n+1], go through a length-H complex FFT Z = E + iO (E, O: the FFTs of the
// even and odd samples).  This pass recovers, for k = 0..H,
//     E[k] = (Z[k] + conj Z[H-k]) / 2,   O[k] = (Z[k] - conj Z[H-k]) / (2i),
//     X[k] = E[k] + W_N^k O[k],           X[H-k] = conj(E[k] - W_N^k O[k]),
// one thread per pair (k, H-k), k = 0..H/2 (indices of Z taken mod H).
//
// 2D / 3D: each row r (outer indices m) of the last dimension is read the same
// way and the complex plan runs over (outer dims, H).  The spectra E and O of
// the even and odd samples are then Hermitian over all dimensions, so the
// identities above hold with conj Z[H-k] taken from the mirror row r' (outer
// indices -m mod each dimension):
//     E[r][k] = (Z[r][k] + conj Z[r'][H-k]) / 2, ..., X[r'][H-k] = conj(E - W^k O)
// and one thread serves the pair ((r, k), (r', H-k)), k = 0..H/2; at k = H/2
// only X[r][k] is written (the pair's other element is another thread's first).
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
  long long rows;        // rows of the last dimension per transform (1 in 1D)
  long long ox, oy;      // outer dimensions (1D: 1, 1; 2D: NX, 1; 3D: NX, NY)
};

// the row whose outer indices are the negatives of row r's
MCFFT_DEV long long mirror_row(long long r, long long ox, long long oy) {
  const long long m1 = r / oy, m2 = r - m1 * oy;
  return ((ox - m1) % ox) * oy + (oy - m2) % oy;
}

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

// R2C post-pass: Z (rows x H complex per transform, contiguous) -> X (rows x (H + 1))
template <int CM>
MCFFT_DEV void r2c_post_body(const RealArgs& a) {
  typedef Traits<CM> Tr;
  constexpr long long esz = (long long)sizeof(typename Tr::T);
  const long long h = a.h, pairs = h / 2 + 1, per = pairs * a.rows, total = per * a.batch;
  const long long stride = (long long)grid_size() * 1024;
  for (long long i = (long long)block_id() * 1024 + thread_id(); i < total; i += stride) {
    const long long b = i / per, rk = i - b * per, r = rk / pairs, k = rk - r * pairs;
    const long long rm = mirror_row(r, a.ox, a.oy);
    const long long kk = (h - k) % h;  // index of Z[H - k] (mod H)
    const GBuf gz = make_gbuf((const char*)a.in + b * a.rows * h * esz, (unsigned)(a.rows * h * esz));
    const GBuf gx = make_gbuf((char*)a.out + b * a.rows * (h + 1) * esz, (unsigned)(a.rows * (h + 1) * esz));
    const GBuf gt = make_gbuf(a.tw, (unsigned)((h + 1) * 8));
    const f32x2 zk = real_ld<CM>(gz, r * h + k % h, a.in_fmt), zr = conj2(real_ld<CM>(gz, rm * h + kk, a.in_fmt));
    const f32x2 e = (zk + zr) * 0.5f;
    const f32x2 o = mul_by_i(zr - zk) * 0.5f;  // (Z[k] - conj Z[H-k]) / (2i)
    const f32x2 t = cmul(__builtin_bit_cast(f32x2, buf_ld64(gt, (int)(k * 8))), o);
    Tr::store(gx, (int)(r * (h + 1) + k), Tr::pack(e + t));
    if (k != h - k) Tr::store(gx, (int)(rm * (h + 1) + h - k), Tr::pack(conj2(e - t)));
  }
}

// C2R pre-pass: X (rows x (H + 1) complex per transform) -> Z' (rows x H)
template <int CM>
MCFFT_DEV void c2r_pre_body(const RealArgs& a) {
  typedef Traits<CM> Tr;
  constexpr long long esz = (long long)sizeof(typename Tr::T);
  const long long h = a.h, per = h * a.rows, total = per * a.batch;
  const long long esz_in = (CM == CM_SPLIT && a.in_fmt == FMT_SPLIT) ? 8 : 4;
  const long long stride = (long long)grid_size() * 1024;
  for (long long i = (long long)block_id() * 1024 + thread_id(); i < total; i += stride) {
    const long long b = i / per, rk = i - b * per, r = rk / h, k = rk - r * h;
    const long long rm = mirror_row(r, a.ox, a.oy);
    const GBuf gx = make_gbuf((const char*)a.in + b * a.rows * (h + 1) * esz_in, (unsigned)(a.rows * (h + 1) * esz_in));
    const GBuf gz = make_gbuf((char*)a.out + b * a.rows * h * esz, (unsigned)(a.rows * h * esz));
    const GBuf gt = make_gbuf(a.tw, (unsigned)((h + 1) * 8));
    const f32x2 xk = real_ld<CM>(gx, r * (h + 1) + k, a.in_fmt);
    const f32x2 xr = conj2(real_ld<CM>(gx, rm * (h + 1) + h - k, a.in_fmt));
    const f32x2 w = __builtin_bit_cast(f32x2, buf_ld64(gt, (int)(k * 8)));  // W_N^-k (table of the inverse)
    const f32x2 z = ((xk + xr) + mul_by_i(cmul(xk - xr, w))) * 0.5f;
    Tr::store(gz, (int)(r * h + k), Tr::pack(z));
  }
}

}  // namespace mcfft

// compat_common.h -- helpers shared by the reference-compatible shim libraries.
#pragma once
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <utility>
#include <vector>

#include <hip/hip_runtime.h>

#include "mcfft.h"

namespace mcfft_compat {

inline mcfftStatus& last_error() {
  static mcfftStatus s = MCFFT_SUCCESS;
  return s;
}

inline bool report(mcfftStatus st, const char* what) {
  last_error() = st;
  if (st != MCFFT_SUCCESS) {
    fprintf(stderr, "mcfft: %s failed: %s\n", what, mcfft_status_string(st));
    return false;
  }
  return true;
}

// the plan of the last exec: mcfftGetLastError() also reports what its device
// status word saw (an unnormalised mcfftExec that overflowed FP16)
inline mcfftPlan& last_plan() {
  static mcfftPlan p = nullptr;
  return p;
}

// mcfftGetLastError(): the last host-side failure, else MCFFT_NONFINITE once an
// exec of the last plan has stored an Inf/NaN (kept until the next failure or
// success is reported).  Synchronises with the device.
inline mcfftStatus get_last_error() {
  if (last_error() == MCFFT_SUCCESS && last_plan()) {
    unsigned f = 0;
    if (mcfft_get_status_flags(last_plan(), &f) == MCFFT_SUCCESS && (f & MCFFT_STATUS_NONFINITE))
      last_error() = MCFFT_NONFINITE;
  }
  return last_error();
}

inline void forget_plan(mcfftPlan p) {
  if (last_plan() == p) last_plan() = nullptr;
}

// In-place forward exec on the null stream (the reference's mcfftExec contract).
// in_exp: the data's per-transform exponents, for data an earlier exec of an
// AUTO plan left in place (null: true units).
inline bool exec_in_place(mcfftPlan p, void* data, const int* in_exp, const char* what) {
  if (!p) return report(MCFFT_INVALID_PLAN, what);
  last_plan() = p;
  return report(mcfft_exec_exp(p, data, in_exp, data, MCFFT_FORWARD, nullptr), what);
}

// The utils layer's state, one per library (the reference keeps the same as
// globals, source/mcfft_1d_utils.cpp:3-6).  Its plan scales with
// MCFFT_SCALE_AUTO, so FP16 neither overflows nor underflows whatever the data;
// finalize() multiplies each transform by 2^(its exponent) in FP64.  Repeated
// execs (doit(iter)) work on the previous result in place and take its
// exponents as their input exponents, so finalize() returns the iterated
// transform in true units.
template <class Scalar>
struct UtilsState {
  std::vector<Scalar> host;   // interleaved re, im
  Scalar* dev = nullptr;
  long long batch = 0, per = 0;  // transforms, complex elements per transform
  int execs = 0;                 // execs since setup()
};

template <class Scalar>
inline bool upload(UtilsState<Scalar>& u, long long batch, long long per) {
  u.batch = batch;
  u.per = per;
  u.execs = 0;
  if (u.dev) (void)hipFree(u.dev);
  u.dev = nullptr;
  if (hipMalloc(&u.dev, sizeof(Scalar) * u.host.size()) != hipSuccess) return report(MCFFT_ALLOC_FAILED, "setup");
  if (hipMemcpy(u.dev, u.host.data(), sizeof(Scalar) * u.host.size(), hipMemcpyHostToDevice) != hipSuccess)
    return report(MCFFT_HIP_ERROR, "setup");
  return true;
}

template <class Scalar>
inline void run(UtilsState<Scalar>& u, mcfftPlan p, int iter) {
  const int* own = nullptr;
  if (p && mcfft_get_exponent_buffer(p, &own) != MCFFT_SUCCESS) own = nullptr;
  for (int t = 0; t < iter; ++t, ++u.execs)
    if (!exec_in_place(p, u.dev, u.execs > 0 ? own : nullptr, "doit")) return;
  if (hipDeviceSynchronize() != hipSuccess) report(MCFFT_HIP_ERROR, "doit");
}

template <class Scalar>
inline void download(UtilsState<Scalar>& u, mcfftPlan p, double* result) {
  if (!u.dev || !p) {
    report(MCFFT_INVALID_VALUE, "finalize");
    return;
  }
  std::vector<int> e((size_t)u.batch, 0);
  if (hipMemcpy(u.host.data(), u.dev, sizeof(Scalar) * u.host.size(), hipMemcpyDeviceToHost) != hipSuccess ||
      (u.execs > 0 && mcfft_get_transform_exponents(p, e.data()) != MCFFT_SUCCESS)) {
    report(MCFFT_HIP_ERROR, "finalize");
    return;
  }
  for (long long b = 0; b < u.batch; ++b) {
    const double s = std::ldexp(1.0, e[(size_t)b]);
    for (long long i = 2 * b * u.per; i < 2 * (b + 1) * u.per; ++i) result[i] = (double)(float)u.host[(size_t)i] * s;
  }
}

// Digit-reversal permutation for a radix list (same contract as the reference's
// gen_rev, source/mcfft_1d_utils.cpp:8-35: rev[i] is the source index placed at
// position i).  The compatible plans report one radix, for which it is the
// identity; it is kept so a caller that reads plan.radices sees a consistent
// answer.
inline void gen_rev(int n, int* rev, const int* radices, int n_radices) {
  std::vector<int> a(n), b(n);
  for (int i = 0; i < n; ++i) a[i] = i;
  int now = n;
  for (int i = n_radices - 1; i >= 0; --i) {
    const int r = radices[i];
    for (int j = 0; j < n; j += now)
      for (int k = 0; k < r; ++k)
        for (int l = 0; l < now / r; ++l) b[j + l + k * (now / r)] = a[j + l * r + k];
    now /= r;
    std::swap(a, b);
  }
  for (int i = 0; i < n; ++i) rev[i] = a[i];
}

// The reference generator (source/mcfft_1d_utils.cpp:88-100): srand(seed), then
// re and im of every element from 0.0001f * rand() / RAND_MAX in float.
inline void generate(double* data, int n, int n_batch, int seed) {
  srand((unsigned)seed);
  for (long long i = 0; i < n_batch; ++i)
    for (long long j = 0; j < n; ++j) {
      float a = 0.0001f * (float)rand();
      a = a / (float)RAND_MAX;
      data[0 + j * 2 + i * n * 2] = (double)a;
      float b = 0.0001f * (float)rand();
      b = b / (float)RAND_MAX;
      data[1 + j * 2 + i * n * 2] = (double)b;
    }
}

inline mcfftFormat storage_format() {
#ifdef MCFFT_COMPAT_BF16
  return MCFFT_FMT_BF16;
#else
  return MCFFT_FMT_FP16;
#endif
}

}  // namespace mcfft_compat

// compat_common.h -- helpers shared by the reference-compatible shim libraries.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <utility>
#include <vector>

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

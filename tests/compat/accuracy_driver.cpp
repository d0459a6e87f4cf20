// accuracy_driver.cpp -- a caller written against the reference's utils API
// (generate_data / setup / doit / finalize, as benchmark/accuracy_1d.cpp:58-97
// and benchmark/accuracy_2d.cpp use it), linked against the compatible shim
// library instead of the reference.  Prints one JSON line with the rel-L2,
// max-abs and the reference's clipped metric against the FP64 oracle.
// Built three times: -DDIM=1, 2, 3 (libmcfft_compat_{1,2,3}d); the 3D build calls
// setup(data, nx, ny, nz, n_batch) when -z is given and the reference's 2D-shaped
// setup(data, nx, ny, n_batch) otherwise.
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#if DIM == 1
#include "mcfft_1d_utils.h"
#elif DIM == 2
#include "mcfft_2d_utils.h"
#else
#include "mcfft_3d_utils.h"
#endif

extern "C" {
void orc_fft_many(double* data, int rank, const long* dims, long batch, int sign);
double orc_rel_l2(const double* a, const double* ref, long n_total);
double orc_max_abs(const double* a, const double* ref, long n_total);
double orc_ref_metric(const double* tested, const double* standard, long n_total);
void orc_round_half(double* x, long count);
}

int main(int argc, char** argv) {
  int n = 256, nx = 256, ny = 256, nz = 0, n_batch = 1, seed = 42, iters = 1;
  int c;
  while ((c = getopt(argc, argv, "n:x:y:z:b:s:i:")) != -1) {
    if (c == 'n') n = atoi(optarg);
    if (c == 'x') nx = atoi(optarg);
    if (c == 'y') ny = atoi(optarg);
    if (c == 'z') nz = atoi(optarg);
    if (c == 'b') n_batch = atoi(optarg);
    if (c == 's') seed = atoi(optarg);
    if (c == 'i') iters = atoi(optarg);
  }
#if DIM == 1
  const long total = (long)n * n_batch;
  std::vector<double> data(2 * total), tested(2 * total);
  generate_data(data.data(), n, n_batch, seed);
  setup(data.data(), n, n_batch);
#else
  const long nzz = nz > 0 ? nz : 1;
  const long total = (long)nx * ny * nzz * n_batch;
  std::vector<double> data(2 * total), tested(2 * total);
  generate_data(data.data(), (int)(nx * ny * nzz), n_batch, seed);
#if DIM == 3
  if (nz > 0)
    setup(data.data(), nx, ny, nz, n_batch);
  else
#endif
    setup(data.data(), nx, ny, n_batch);
#endif
  doit(iters);
  finalize(tested.data());
  // FP64 oracle on the FP16-quantised input (the metric of the reference also
  // charges the input rounding; both are printed)
  std::vector<double> q(data), standard(data);
  orc_round_half(q.data(), 2 * total);
#if DIM == 1
  const long dims[1] = {n};
  orc_fft_many(q.data(), 1, dims, n_batch, -1);
  orc_fft_many(standard.data(), 1, dims, n_batch, -1);
#else
  const long dims[3] = {nx, ny, nzz};
  const int rank = nz > 0 ? 3 : 2;
  orc_fft_many(q.data(), rank, dims, n_batch, -1);
  orc_fft_many(standard.data(), rank, dims, n_batch, -1);
#endif
  printf("{\"dim\": %d, \"total\": %ld, \"rel_l2\": %.4e, \"max_abs\": %.4e, \"ref_metric\": %.6e, "
         "\"rel_l2_vs_unquantised\": %.4e, \"status\": %d}\n",
         DIM, total, orc_rel_l2(tested.data(), q.data(), total), orc_max_abs(tested.data(), q.data(), total),
         orc_ref_metric(tested.data(), standard.data(), total), orc_rel_l2(tested.data(), standard.data(), total),
         (int)mcfftGetLastError());
  return mcfftGetLastError() == MCFFT_SUCCESS ? 0 : 1;
}

// compat_2d.cpp -- the reference's 2D API (include/mcfft_2d_device.h:37-38,
// include/mcfft_2d_utils.h:8-11) implemented on the C ABI of libmcfft.
#include "mcfft_2d_utils.h"

#include "compat_common.h"

using namespace mcfft_compat;

static void create(mcfftHandle* plan, int nx, int ny, int n_batch, mcfftScaling scaling) {
  plan->Nx = nx;
  plan->Ny = ny;
  plan->N_batch = n_batch;
  plan->radices_x[0] = nx;
  plan->radices_y[0] = ny;
  for (int i = 1; i < 3; ++i) plan->radices_x[i] = plan->radices_y[i] = 1;
  plan->n_radices_x = plan->n_radices_y = 1;
  plan->mergings[0] = plan->mergings[1] = 0;
  plan->impl = nullptr;
  const long long d[2] = {nx, ny};
  report(mcfft_plan_many(&plan->impl, 2, d, n_batch, storage_format(), storage_format(), scaling), "mcfftCreate");
}

// the reference's contract: unnormalised forward transform in place
void mcfftCreate(mcfftHandle* plan, int nx, int ny, int n_batch) { create(plan, nx, ny, n_batch, MCFFT_SCALE_NONE); }

void mcfftExec(mcfftHandle plan, MCTYPE* data) { exec_in_place(plan.impl, data, nullptr, "mcfftExec"); }

void mcfftDestroy(mcfftHandle* plan) {
  if (plan->impl) {
    forget_plan(plan->impl);
    mcfft_destroy(plan->impl);
  }
  plan->impl = nullptr;
}

mcfftStatus mcfftGetLastError(void) { return get_last_error(); }

// ---- utils layer (source/mcfft_2d_utils.cpp:33-79); one radix per axis, so
// the reference's digit reversal is the identity ----
static UtilsState<MCTYPE> g_u;
static mcfftHandle g_plan;

void setup(double* data, int nx, int ny, int n_batch) {
  mcfftDestroy(&g_plan);
  create(&g_plan, nx, ny, n_batch, MCFFT_SCALE_AUTO);
  const long long total = (long long)nx * ny * n_batch;
  g_u.host.resize((size_t)2 * total);
  for (long long i = 0; i < 2 * total; ++i) g_u.host[i] = (MCTYPE)(float)data[i];
  upload(g_u, n_batch, (long long)nx * ny);
}

void finalize(double* result) { download(g_u, g_plan.impl, result); }

void doit(int iter) { run(g_u, g_plan.impl, iter); }

void generate_data(double* data, int n, int n_batch, int seed) { generate(data, n, n_batch, seed); }

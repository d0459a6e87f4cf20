// compat_3d.cpp -- 3D API: the reference's (nx, ny, n_batch) signatures
// (include/mcfft_3d_device.h:37-38, include/mcfft_3d_utils.h:8) plan a 2D
// transform as the reference's stub does; the nz overloads plan a real 3D one.
#include "mcfft_3d_utils.h"

#include "compat_common.h"

using namespace mcfft_compat;

static void create(mcfftHandle* plan, int rank, int nx, int ny, int nz, int n_batch, mcfftScaling scaling) {
  plan->Nx = nx;
  plan->Ny = ny;
  plan->Nz = nz;
  plan->N_batch = n_batch;
  plan->radices_x[0] = nx;
  plan->radices_y[0] = ny;
  plan->radices_z[0] = nz;
  for (int i = 1; i < 3; ++i) plan->radices_x[i] = plan->radices_y[i] = plan->radices_z[i] = 1;
  plan->n_radices_x = plan->n_radices_y = plan->n_radices_z = 1;
  plan->mergings[0] = plan->mergings[1] = plan->mergings[2] = 0;
  plan->impl = nullptr;
  const long long d[3] = {nx, ny, nz};
  report(mcfft_plan_many(&plan->impl, rank, d, n_batch, storage_format(), storage_format(), scaling), "mcfftCreate");
}

// the reference's contract: unnormalised forward transform in place
void mcfftCreate(mcfftHandle* plan, int nx, int ny, int n_batch) {
  create(plan, 2, nx, ny, 1, n_batch, MCFFT_SCALE_NONE);
}
void mcfftCreate(mcfftHandle* plan, int nx, int ny, int nz, int n_batch) {
  create(plan, 3, nx, ny, nz, n_batch, MCFFT_SCALE_NONE);
}

void mcfftExec(mcfftHandle plan, MCTYPE* data) { exec_in_place(plan.impl, data, nullptr, "mcfftExec"); }

void mcfftDestroy(mcfftHandle* plan) {
  if (plan->impl) {
    forget_plan(plan->impl);
    mcfft_destroy(plan->impl);
  }
  plan->impl = nullptr;
}

mcfftStatus mcfftGetLastError(void) { return get_last_error(); }

// ---- utils layer: as 2D, plus the (nx, ny, nz, n_batch) overload ----
static UtilsState<MCTYPE> g_u;
static mcfftHandle g_plan;

static void setup_any(double* data, int rank, int nx, int ny, int nz, int n_batch) {
  mcfftDestroy(&g_plan);
  create(&g_plan, rank, nx, ny, nz, n_batch, MCFFT_SCALE_AUTO);
  const long long per = (long long)nx * ny * nz, total = per * n_batch;
  g_u.host.resize((size_t)2 * total);
  for (long long i = 0; i < 2 * total; ++i) g_u.host[i] = (MCTYPE)(float)data[i];  // identity digit reversal
  upload(g_u, n_batch, per);
}

void setup(double* data, int nx, int ny, int n_batch) { setup_any(data, 2, nx, ny, 1, n_batch); }
void setup(double* data, int nx, int ny, int nz, int n_batch) { setup_any(data, 3, nx, ny, nz, n_batch); }

void finalize(double* result) { download(g_u, g_plan.impl, result); }

void doit(int iter) { run(g_u, g_plan.impl, iter); }

void generate_data(double* data, int n, int n_batch, int seed) { generate(data, n, n_batch, seed); }

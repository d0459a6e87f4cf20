// compat_3d.cpp -- 3D API: the reference's (nx, ny, n_batch) signatures
// (include/mcfft_3d_device.h:37-38, include/mcfft_3d_utils.h:8) plan a 2D
// transform as the reference's stub does; the nz overloads plan a real 3D one.
#include "mcfft_3d_utils.h"

#include "compat_common.h"

using namespace mcfft_compat;

static void create(mcfftHandle* plan, int rank, int nx, int ny, int nz, int n_batch) {
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
  report(mcfft_plan_many(&plan->impl, rank, d, n_batch, storage_format(), storage_format(), MCFFT_SCALE_NONE),
         "mcfftCreate");
}

void mcfftCreate(mcfftHandle* plan, int nx, int ny, int n_batch) { create(plan, 2, nx, ny, 1, n_batch); }
void mcfftCreate(mcfftHandle* plan, int nx, int ny, int nz, int n_batch) { create(plan, 3, nx, ny, nz, n_batch); }

void mcfftExec(mcfftHandle plan, MCTYPE* data) {
  if (!plan.impl) {
    report(MCFFT_INVALID_PLAN, "mcfftExec");
    return;
  }
  report(mcfft_exec(plan.impl, data, data, MCFFT_FORWARD, nullptr), "mcfftExec");
}

void mcfftDestroy(mcfftHandle* plan) {
  if (plan->impl) mcfft_destroy(plan->impl);
  plan->impl = nullptr;
}

mcfftStatus mcfftGetLastError(void) { return last_error(); }

static std::vector<MCTYPE> g_host;
static MCTYPE* g_dev = nullptr;
static mcfftHandle g_plan;

static void setup_any(double* data, long long total) {
  g_host.resize((size_t)2 * total);
  for (long long i = 0; i < 2 * total; ++i) g_host[i] = (MCTYPE)(float)data[i];  // identity digit reversal
  if (g_dev) (void)hipFree(g_dev);
  g_dev = nullptr;
  if (hipMalloc(&g_dev, sizeof(MCTYPE) * g_host.size()) != hipSuccess) {
    report(MCFFT_ALLOC_FAILED, "setup");
    return;
  }
  if (hipMemcpy(g_dev, g_host.data(), sizeof(MCTYPE) * g_host.size(), hipMemcpyHostToDevice) != hipSuccess)
    report(MCFFT_HIP_ERROR, "setup");
}

void setup(double* data, int nx, int ny, int n_batch) {
  mcfftCreate(&g_plan, nx, ny, n_batch);
  setup_any(data, (long long)nx * ny * n_batch);
}

void setup(double* data, int nx, int ny, int nz, int n_batch) {
  mcfftCreate(&g_plan, nx, ny, nz, n_batch);
  setup_any(data, (long long)nx * ny * nz * n_batch);
}

void finalize(double* result) {
  if (!g_dev) {
    report(MCFFT_INVALID_VALUE, "finalize");
    return;
  }
  if (hipMemcpy(g_host.data(), g_dev, sizeof(MCTYPE) * g_host.size(), hipMemcpyDeviceToHost) != hipSuccess) {
    report(MCFFT_HIP_ERROR, "finalize");
    return;
  }
  for (size_t i = 0; i < g_host.size(); ++i) result[i] = (double)(float)g_host[i];
}

void doit(int iter) {
  for (int t = 0; t < iter; ++t) mcfftExec(g_plan, g_dev);
  if (hipDeviceSynchronize() != hipSuccess) report(MCFFT_HIP_ERROR, "doit");
}

void generate_data(double* data, int n, int n_batch, int seed) { generate(data, n, n_batch, seed); }

// compat_2d.cpp -- the reference's 2D API (include/mcfft_2d_device.h:37-38,
// include/mcfft_2d_utils.h:8-11) implemented on the C ABI of libmcfft.
#include "mcfft_2d_utils.h"

#include "compat_common.h"

using namespace mcfft_compat;

void mcfftCreate(mcfftHandle* plan, int nx, int ny, int n_batch) {
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
  report(mcfft_plan_many(&plan->impl, 2, d, n_batch, storage_format(), storage_format(), MCFFT_SCALE_NONE),
         "mcfftCreate");
}

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

static int g_nx, g_ny, g_batch;
static std::vector<MCTYPE> g_host;
static MCTYPE* g_dev = nullptr;
static mcfftHandle g_plan;

void setup(double* data, int nx, int ny, int n_batch) {
  g_nx = nx;
  g_ny = ny;
  g_batch = n_batch;
  mcfftCreate(&g_plan, nx, ny, n_batch);
  std::vector<int> rx(nx), ry(ny);
  gen_rev(nx, rx.data(), g_plan.radices_x, g_plan.n_radices_x);
  gen_rev(ny, ry.data(), g_plan.radices_y, g_plan.n_radices_y);
  g_host.resize((size_t)2 * nx * ny * n_batch);
  for (long long i = 0; i < n_batch; ++i)
    for (long long j = 0; j < nx; ++j)
      for (long long k = 0; k < ny; ++k)
        for (int c = 0; c < 2; ++c)
          g_host[2 * (i * nx * ny + j * ny + k) + c] = (MCTYPE)(float)data[2 * (i * nx * ny + rx[j] * ny + ry[k]) + c];
  if (g_dev) (void)hipFree(g_dev);
  g_dev = nullptr;
  if (hipMalloc(&g_dev, sizeof(MCTYPE) * g_host.size()) != hipSuccess) {
    report(MCFFT_ALLOC_FAILED, "setup");
    return;
  }
  if (hipMemcpy(g_dev, g_host.data(), sizeof(MCTYPE) * g_host.size(), hipMemcpyHostToDevice) != hipSuccess)
    report(MCFFT_HIP_ERROR, "setup");
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

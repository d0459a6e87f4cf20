// compat_1d.cpp -- the reference's 1D API (include/mcfft_1d_device.h:37-38,
// include/mcfft_1d_utils.h:8-11) implemented on the C ABI of libmcfft.
#include "mcfft_1d_utils.h"

#include "compat_common.h"

using namespace mcfft_compat;

void mcfftCreate(mcfftHandle* plan, int n, int n_batch) {
  plan->N = n;
  plan->N_batch = n_batch;
  for (int i = 0; i < 9; ++i) plan->radices[i] = 16;
  plan->radices[0] = n;
  plan->n_radices = 1;
  for (int i = 0; i < 3; ++i) plan->mergings[i] = 0;
  plan->impl = nullptr;
  const long long d = n;
  if (!report(mcfft_plan_many(&plan->impl, 1, &d, n_batch, storage_format(), storage_format(), MCFFT_SCALE_NONE),
              "mcfftCreate"))
    return;
  int np = 0;
  mcfft_get_pass_count(plan->impl, &np);
  plan->n_mergings = np;
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

// ---- utils layer: one global plan, as in source/mcfft_1d_utils.cpp:3-6 ----
static int g_n, g_batch;
static std::vector<MCTYPE> g_host;
static MCTYPE* g_dev = nullptr;
static mcfftHandle g_plan;

void setup(double* data, int n, int n_batch) {
  g_n = n;
  g_batch = n_batch;
  mcfftCreate(&g_plan, n, n_batch);
  std::vector<int> rev(n);
  gen_rev(n, rev.data(), g_plan.radices, g_plan.n_radices);
  g_host.resize((size_t)2 * n * n_batch);
  for (long long j = 0; j < n_batch; ++j)
    for (long long i = 0; i < n; ++i) {
      g_host[2 * n * j + 2 * i + 0] = (MCTYPE)(float)data[2 * n * j + 2 * rev[i] + 0];
      g_host[2 * n * j + 2 * i + 1] = (MCTYPE)(float)data[2 * n * j + 2 * rev[i] + 1];
    }
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

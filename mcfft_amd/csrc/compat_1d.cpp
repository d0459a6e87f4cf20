// compat_1d.cpp -- the reference's 1D API (include/mcfft_1d_device.h:37-38,
// include/mcfft_1d_utils.h:8-11) implemented on the C ABI of libmcfft.
#include "mcfft_1d_utils.h"

#include "compat_common.h"

using namespace mcfft_compat;

static void create(mcfftHandle* plan, int n, int n_batch, mcfftScaling scaling) {
  plan->N = n;
  plan->N_batch = n_batch;
  for (int i = 0; i < 9; ++i) plan->radices[i] = 16;
  plan->radices[0] = n;
  plan->n_radices = 1;
  for (int i = 0; i < 3; ++i) plan->mergings[i] = 0;
  plan->n_mergings = 0;
  plan->impl = nullptr;
  const long long d = n;
  if (!report(mcfft_plan_many(&plan->impl, 1, &d, n_batch, storage_format(), storage_format(), scaling),
              "mcfftCreate"))
    return;
  int np = 0;
  mcfft_get_pass_count(plan->impl, &np);
  plan->n_mergings = np;
}

// the reference's contract: unnormalised forward transform in place
void mcfftCreate(mcfftHandle* plan, int n, int n_batch) { create(plan, n, n_batch, MCFFT_SCALE_NONE); }

void mcfftExec(mcfftHandle plan, MCTYPE* data) { exec_in_place(plan.impl, data, nullptr, "mcfftExec"); }

void mcfftDestroy(mcfftHandle* plan) {
  if (plan->impl) {
    forget_plan(plan->impl);
    mcfft_destroy(plan->impl);
  }
  plan->impl = nullptr;
}

mcfftStatus mcfftGetLastError(void) { return get_last_error(); }

// ---- utils layer: one global plan, as in source/mcfft_1d_utils.cpp:3-6 ----
static UtilsState<MCTYPE> g_u;
static mcfftHandle g_plan;

void setup(double* data, int n, int n_batch) {
  mcfftDestroy(&g_plan);
  create(&g_plan, n, n_batch, MCFFT_SCALE_AUTO);
  std::vector<int> rev(n);
  gen_rev(n, rev.data(), g_plan.radices, g_plan.n_radices);
  g_u.host.resize((size_t)2 * n * n_batch);
  for (long long j = 0; j < n_batch; ++j)
    for (long long i = 0; i < n; ++i) {
      g_u.host[2 * n * j + 2 * i + 0] = (MCTYPE)(float)data[2 * n * j + 2 * rev[i] + 0];
      g_u.host[2 * n * j + 2 * i + 1] = (MCTYPE)(float)data[2 * n * j + 2 * rev[i] + 1];
    }
  upload(g_u, n_batch, n);
}

void finalize(double* result) { download(g_u, g_plan.impl, result); }

void doit(int iter) { run(g_u, g_plan.impl, iter); }

void generate_data(double* data, int n, int n_batch, int seed) { generate(data, n, n_batch, seed); }

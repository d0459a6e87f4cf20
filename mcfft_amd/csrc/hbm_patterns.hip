// hbm_patterns.hip -- HBM bandwidth of the access patterns the FFT passes use,
// as plain copies (no arithmetic), to give each pass a ceiling to be measured
// against:
//   contiguous   1 GiB copy, 16 B per lane: one vector per thread with a grid of
//                one workgroup per 4 KiB ("flat"), and persistent grid-stride
//                loops of several grid sizes
//   strided      4 x 4096 x 4096 matrices of 4-byte elements: each workgroup
//                copies all 4096 rows of a W-element column block (segments of
//                4W bytes, rows 16 KiB apart), loads first then stores, as a
//                column pass does; W = 8, 16, 32, 64, with and without the
//                XCD-grouped tile order of the column passes
// Prints one line per pattern: time per copy and GB/s (read + write bytes),
// after 200 ms of warm-up load.
#include <hip/hip_runtime.h>

#include <cstdio>

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

template <int UNROLL>
__global__ __launch_bounds__(1024) void copy_grid_stride(const u32x4* __restrict__ in, u32x4* __restrict__ out,
                                                         long long n16) {
  const long long nthreads = (long long)gridDim.x * blockDim.x;
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  for (long long base = tid; base < n16; base += nthreads * UNROLL) {
    u32x4 v[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
      const long long i = base + (long long)u * nthreads;
      if (i < n16) v[u] = in[i];
    }
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
      const long long i = base + (long long)u * nthreads;
      if (i < n16) out[i] = v[u];
    }
  }
}

constexpr int NX = 4096, NY = 4096, NB = 4;

template <int W, bool XCD>
__global__ __launch_bounds__(1024) void copy_column_tiles(const u32x4* __restrict__ in, u32x4* __restrict__ out,
                                                          int ntiles) {
  constexpr int QPR = W / 4;               // 16-byte vectors per row segment
  constexpr int ROWS_PER_IT = 1024 / QPR;  // rows one sweep of the workgroup covers
  constexpr int UNR = 8;
  constexpr int TPM = NY / W;              // tiles per matrix
  int b = blockIdx.x;
  const int g = gridDim.x;
  if (XCD && b < (g & ~31)) b = (b & ~31) + 4 * (b & 7) + ((b >> 3) & 3);
  for (int tile = b; tile < ntiles; tile += g) {
    const int mat = tile / TPM, cb = tile % TPM;
    const long long base = (long long)mat * NX * NY / 4 + cb * QPR;
    const int q = threadIdx.x % QPR, r0 = threadIdx.x / QPR;
    for (int r = r0; r < NX; r += ROWS_PER_IT * UNR) {
      u32x4 v[UNR];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const int rr = r + u * ROWS_PER_IT;
        if (rr < NX) v[u] = in[base + (long long)rr * (NY / 4) + q];
      }
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const int rr = r + u * ROWS_PER_IT;
        if (rr < NX) out[base + (long long)rr * (NY / 4) + q] = v[u];
      }
    }
  }
}

#define CHECK(x)                                                          \
  do {                                                                    \
    if ((x) != hipSuccess) {                                              \
      fprintf(stderr, "%s failed at %s:%d\n", #x, __FILE__, __LINE__);    \
      return 1;                                                           \
    }                                                                     \
  } while (0)

int main() {
  const long long bytes = 1ll << 30;  // >= NB * NX * NY * 4 (256 MiB) too
  u32x4 *a, *b;
  CHECK(hipMalloc(&a, bytes));
  CHECK(hipMalloc(&b, bytes));
  CHECK(hipMemset(a, 1, bytes));
  CHECK(hipMemset(b, 0, bytes));
  int ncu = 0;
  CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0));
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  int bad = 0;
  auto timeit = [&](const char* name, long long moved, auto launch) {
    for (int w = 0; w < 3; ++w) launch();
    const int reps = 20;
    bad |= hipEventRecord(e0, 0) != hipSuccess;
    for (int r = 0; r < reps; ++r) launch();
    bad |= hipEventRecord(e1, 0) != hipSuccess;
    bad |= hipEventSynchronize(e1) != hipSuccess;
    float ms = 0;
    bad |= hipEventElapsedTime(&ms, e0, e1) != hipSuccess;
    ms /= reps;
    printf("%-44s %8.4f ms %8.1f GB/s\n", name, ms, 2.0 * moved / (ms * 1e-3) / 1e9);
    fflush(stdout);
  };
  char nm[128];
  const long long n16 = bytes / 16;
  printf("CUs %d\n", ncu);
  // the GPU's clocks settle after about 40 ms of sustained load: copy for
  // 200 ms before the first measurement (the patterns then run back to back)
  {
    float spent = 0;
    while (spent < 200.0f && !bad) {
      bad |= hipEventRecord(e0, 0) != hipSuccess;
      for (int r = 0; r < 20; ++r) copy_grid_stride<1><<<(unsigned)(n16 / 256), 256>>>(a, b, n16);
      bad |= hipEventRecord(e1, 0) != hipSuccess;
      bad |= hipEventSynchronize(e1) != hipSuccess;
      float ms = 0;
      bad |= hipEventElapsedTime(&ms, e0, e1) != hipSuccess;
      spent += ms;
    }
  }
  timeit("contiguous 1 GiB, flat (4 KiB per workgroup)", bytes,
         [&] { copy_grid_stride<1><<<(unsigned)(n16 / 256), 256>>>(a, b, n16); });
  for (int threads : {256, 512, 1024})
    for (int per_cu : {1, 2, 4}) {
      const int grid = ncu * per_cu * (1024 / threads);
      snprintf(nm, sizeof nm, "contiguous 1 GiB, grid-stride %4d x %4d", grid, threads);
      timeit(nm, bytes, [&] { copy_grid_stride<1><<<grid, threads>>>(a, b, n16); });
    }
  const long long mbytes = (long long)NB * NX * NY * 4;
#define COLUMN_CASE(W, X)                                                                               \
  for (int per_cu : {1, 2}) {                                                                           \
    const int nt = NB * (NY / W), grid = ncu * per_cu;                                                  \
    snprintf(nm, sizeof nm, "column tiles %3d B segments%s, grid %d", 4 * W, X ? " xcd-grouped" : "", grid); \
    timeit(nm, mbytes, [&] { copy_column_tiles<W, X><<<grid, 1024>>>(a, b, nt); });                   \
  }
  COLUMN_CASE(8, false)
  COLUMN_CASE(8, true)
  COLUMN_CASE(16, false)
  COLUMN_CASE(16, true)
  COLUMN_CASE(32, false)
  COLUMN_CASE(64, false)
  CHECK(hipDeviceSynchronize());
  CHECK(hipFree(a));
  CHECK(hipFree(b));
  if (bad) {
    fprintf(stderr, "a HIP event call failed\n");
    return 1;
  }
  printf("done\n");
  return 0;
}

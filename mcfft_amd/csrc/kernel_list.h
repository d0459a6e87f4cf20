// kernel_list.h -- the compiled pass configurations (X-macro lists).
//
// X(S, T, Q, QFAST, NW, CM, STAGED, FOURSTEP): L = 16^S * T points, Q instances
// per workgroup, NW waves.  Tile sizes keep one workgroup's LDS at <= 128 KiB
// (gfx950 has 160 KiB per CU) and a tile at >= 32 sixteen-point MFMA columns.
//   ROW   contiguous rows (1D single pass, 2D/3D innermost axis)
//   ROWT  contiguous rows, transposed store (last pass of a four-step 1D plan)
//   COL   strided columns (2D/3D outer axes)
//   COLF  strided columns + four-step twiddle (first pass of a four-step plan)
#pragma once

// ---- ROW: tile of 2048 elements (4096 at L = 4096) ----
#define MCFFT_ROW_LIST_CM(X, CM)            \
  X(1, 1, 128, false, 4, CM, false, false)  \
  X(1, 2, 64, false, 4, CM, false, false)   \
  X(1, 4, 32, false, 4, CM, false, false)   \
  X(1, 8, 16, false, 4, CM, false, false)   \
  X(2, 1, 8, false, 4, CM, false, false)    \
  X(2, 2, 4, false, 4, CM, false, false)    \
  X(2, 4, 2, false, 4, CM, false, false)    \
  X(2, 8, 1, false, 4, CM, false, false)    \
  X(3, 1, 1, false, 8, CM, false, false)

#define MCFFT_ROW_LIST(X)        \
  MCFFT_ROW_LIST_CM(X, 0)        \
  MCFFT_ROW_LIST_CM(X, 1)        \
  MCFFT_ROW_LIST_CM(X, 2)

// ---- ROWT: Q rows stored transposed (segments of Q elements) ----
#define MCFFT_ROWT_LIST_CM(X, CM, QB)       \
  X(1, 1, 32, false, 4, CM, true, false)    \
  X(1, 2, 32, false, 4, CM, true, false)    \
  X(1, 4, 32, false, 4, CM, true, false)    \
  X(1, 8, 32, false, 4, CM, true, false)    \
  X(2, 1, 32, false, 8, CM, true, false)    \
  X(2, 2, 32, false, 8, CM, true, false)    \
  X(2, 4, 16, false, 8, CM, true, false)    \
  X(2, 8, QB, false, 16, CM, true, false)

#define MCFFT_ROWT_LIST(X)        \
  MCFFT_ROWT_LIST_CM(X, 0, 16)    \
  MCFFT_ROWT_LIST_CM(X, 1, 16)    \
  MCFFT_ROWT_LIST_CM(X, 2, 8)

// ---- COL / COLF: Q consecutive columns ----
#define MCFFT_COL_LIST_CM(X, CM, FS, Q512, Q1024, Q2048, Q4096) \
  X(1, 1, 32, true, 4, CM, false, FS)                          \
  X(1, 2, 32, true, 4, CM, false, FS)                          \
  X(1, 4, 32, true, 4, CM, false, FS)                          \
  X(1, 8, 32, true, 4, CM, false, FS)                          \
  X(2, 1, 32, true, 8, CM, false, FS)                          \
  X(2, 2, Q512, true, 8, CM, false, FS)                        \
  X(2, 4, Q1024, true, 8, CM, false, FS)                       \
  X(2, 8, Q2048, true, 16, CM, false, FS)                      \
  X(3, 1, Q4096, true, 16, CM, false, FS)

#define MCFFT_COL_LIST(X)                              \
  MCFFT_COL_LIST_CM(X, 0, false, 32, 16, 16, 8)        \
  MCFFT_COL_LIST_CM(X, 1, false, 32, 16, 16, 8)        \
  MCFFT_COL_LIST_CM(X, 2, false, 16, 16, 8, 4)

#define MCFFT_COLF_LIST(X)                             \
  MCFFT_COL_LIST_CM(X, 0, true, 32, 16, 16, 8)         \
  MCFFT_COL_LIST_CM(X, 1, true, 32, 16, 16, 8)         \
  MCFFT_COL_LIST_CM(X, 2, true, 16, 16, 8, 4)

namespace mcfft {
struct KernelInfo {
  const char* name;
  const void* fn;   // __global__ entry (host stub); launched with hipLaunchKernel
  int S, T, Q, qfast, nw, cm, staged, fourstep;
  int lds_bytes;
};
// tables defined in each kernel translation unit
extern const KernelInfo kRowKernels[];
extern const int kRowKernelCount;
extern const KernelInfo kRowTKernels[];
extern const int kRowTKernelCount;
extern const KernelInfo kColKernels[];
extern const int kColKernelCount;
extern const KernelInfo kColFKernels[];
extern const int kColFKernelCount;
}  // namespace mcfft

#define MCFFT_KNAME(S, T, Q, QF, NW, CM, ST, FS) mcfft_pass_##S##_##T##_##Q##_##QF##_##NW##_##CM##_##ST##_##FS

#define MCFFT_DEFINE_KERNEL(S, T, Q, QF, NW, CM, ST, FS)                                   \
  MCFFT_KERNEL(NW * 64) void MCFFT_KNAME(S, T, Q, QF, NW, CM, ST, FS)(mcfft::PassArgs a) { \
    mcfft::fft_pass_body<S, T, Q, QF, NW, CM, ST, FS>(a);                                  \
  }

#define MCFFT_TABLE_ENTRY(S, T, Q, QF, NW, CM, ST, FS)                                                      \
  {#S "_" #T "_" #Q "_" #QF "_" #NW "_" #CM "_" #ST "_" #FS, (const void*)&MCFFT_KNAME(S, T, Q, QF, NW, CM, ST, FS), \
   S, T, Q, QF, NW, CM, ST, FS, mcfft::PassCfg<S, T, Q, QF, NW, CM, ST, FS>::LDS_BYTES},

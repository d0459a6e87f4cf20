// kernel_list.h -- the compiled pass configurations (X-macro lists).
//
// X(S, T, Q, NW, CM, KIND): L = 16^S * T points, Q instances per workgroup,
// NW waves, compute mode CM, KIND 0 ROW / 1 ROWT / 2 COL / 3 COLF / 4 COLM / 5 ROWX.  Tile sizes keep one workgroup's LDS at <= 128 KiB
// (gfx950 has 160 KiB per CU) and a tile at >= 32 sixteen-point MFMA columns.
//   ROW   contiguous rows (1D single pass, 2D/3D innermost axis)
//   ROWT  contiguous rows, transposed store (last pass of a four-step 1D plan)
//   COL   strided columns (2D/3D outer axes)
//   COLF  strided columns + four-step twiddle (first pass of a four-step plan)
//   COLM  strided columns + four-step twiddle (middle pass of a three-pass 1D plan)
//   ROWX  Q rows NX/Q apart + radix-Q butterfly across them (row pass of a 2D plan
//         with a split outer axis of NX points)
#pragma once

// ---- ROW: tile of 2048 elements (4096 at L = 4096) ----
#define MCFFT_ROW_LIST_CM(X, CM)  \
  X(1, 1, 128, 4, CM, 0)          \
  X(1, 2, 64, 4, CM, 0)           \
  X(1, 4, 32, 4, CM, 0)           \
  X(1, 8, 16, 4, CM, 0)           \
  X(2, 1, 8, 4, CM, 0)            \
  X(2, 2, 4, 4, CM, 0)            \
  X(2, 4, 2, 4, CM, 0)            \
  X(2, 8, 1, 4, CM, 0)            \
  X(3, 1, 1, 8, CM, 0)            \
  X(3, 2, 1, 16, CM, 0)           \
  X(3, 4, 1, 16, CM, 0)

// alternative tile shapes the autotuner may pick (the first entry of a
// (S, T, CM, KIND) combination above is the planner's default)
#define MCFFT_ROW_TUNE_CM(X, CM)  \
  X(2, 1, 4, 2, CM, 0)            \
  X(2, 1, 16, 8, CM, 0)           \
  X(2, 1, 16, 4, CM, 0)           \
  X(2, 2, 8, 8, CM, 0)            \
  X(3, 1, 1, 4, CM, 0)            \
  X(3, 1, 2, 8, CM, 0)

// four MFMA tiles per wave (NW = Q*L/2048): a tile's fixed costs (barriers,
// load latency) spread over four times the data; 512..2048-point rows
#define MCFFT_ROW_WIDE_CM(X, CM) \
  X(2, 2, 8, 2, CM, 0)           \
  X(2, 4, 4, 2, CM, 0)           \
  X(2, 8, 4, 4, CM, 0)

// 8192-point rows, 2 per tile (70 KiB of LDS)
#define MCFFT_ROW_8K2_CM(X, CM) X(3, 2, 2, 16, CM, 0)

// split-mode alternatives for short rows (latency-bound small batches)
#define MCFFT_ROW_TUNE_SPLIT(X) \
  X(2, 1, 4, 2, 2, 0)           \
  X(2, 1, 4, 4, 2, 0)           \
  X(2, 1, 16, 8, 2, 0)          \
  X(2, 2, 2, 2, 2, 0)           \
  X(2, 2, 8, 8, 2, 0)

// single-pass 2^15 (LDS 139 KiB at 4 B per element; not for split)
#define MCFFT_ROW_LIST(X)   \
  MCFFT_ROW_LIST_CM(X, 0)   \
  MCFFT_ROW_LIST_CM(X, 1)   \
  MCFFT_ROW_LIST_CM(X, 2)   \
  X(3, 8, 1, 16, 0, 0)      \
  X(3, 8, 1, 16, 1, 0)      \
  MCFFT_ROW_TUNE_CM(X, 0)   \
  MCFFT_ROW_TUNE_CM(X, 1)   \
  MCFFT_ROW_WIDE_CM(X, 0)   \
  MCFFT_ROW_WIDE_CM(X, 1)   \
  MCFFT_ROW_WIDE_CM(X, 2)   \
  MCFFT_ROW_8K2_CM(X, 0)    \
  MCFFT_ROW_8K2_CM(X, 1)    \
  MCFFT_ROW_TUNE_SPLIT(X)

// ---- ROWT: Q rows stored transposed (segments of Q elements) ----
#define MCFFT_ROWT_LIST_CM(X, CM, QB)  \
  X(1, 1, 32, 4, CM, 1)                \
  X(1, 2, 32, 4, CM, 1)                \
  X(1, 4, 32, 4, CM, 1)                \
  X(1, 8, 32, 4, CM, 1)                \
  X(2, 1, 32, 8, CM, 1)                \
  X(2, 2, 32, 8, CM, 1)                \
  X(2, 4, 16, 8, CM, 1)                \
  X(2, 8, QB, 16, CM, 1)

#define MCFFT_ROWT_TUNE_CM(X, CM)  \
  X(2, 1, 16, 4, CM, 1)            \
  X(2, 2, 16, 8, CM, 1)            \
  X(2, 2, 16, 4, CM, 1)

// 4-byte elements only: 256-byte output segments per row of the transposed store
#define MCFFT_ROWT_TUNE_WIDE(X, CM) \
  X(2, 1, 64, 8, CM, 1)             \
  X(2, 2, 64, 8, CM, 1)             \
  X(2, 2, 64, 16, CM, 1)

#define MCFFT_ROWT_LIST(X)      \
  MCFFT_ROWT_LIST_CM(X, 0, 16)  \
  MCFFT_ROWT_LIST_CM(X, 1, 16)  \
  MCFFT_ROWT_LIST_CM(X, 2, 8)   \
  MCFFT_ROWT_TUNE_CM(X, 0)      \
  MCFFT_ROWT_TUNE_CM(X, 1)      \
  MCFFT_ROWT_TUNE_CM(X, 2)      \
  MCFFT_ROWT_TUNE_WIDE(X, 0)

// ---- COL / COLF: Q consecutive columns ----
#define MCFFT_COL_LIST_CM(X, CM, K, Q512, Q1024, Q2048, Q4096) \
  X(1, 1, 32, 4, CM, K)                                        \
  X(1, 2, 32, 4, CM, K)                                        \
  X(1, 4, 32, 4, CM, K)                                        \
  X(1, 8, 32, 4, CM, K)                                        \
  X(2, 1, 32, 8, CM, K)                                        \
  X(2, 2, Q512, 8, CM, K)                                      \
  X(2, 4, Q1024, 8, CM, K)                                     \
  X(2, 8, Q2048, 16, CM, K)                                    \
  X(3, 1, Q4096, 16, CM, K)

#define MCFFT_COL_TUNE_CM(X, CM, K)  \
  X(2, 1, 16, 4, CM, K)              \
  X(2, 1, 64, 8, CM, K)              \
  X(3, 1, 4, 8, CM, K)

// 16-column tiles of 32..128-point columns (outer axes over 16 columns)
#define MCFFT_COL_NARROW_CM(X, CM) \
  X(1, 2, 16, 4, CM, 2)            \
  X(1, 4, 16, 4, CM, 2)            \
  X(1, 8, 16, 4, CM, 2)

// 128-byte (FP16/BF16: 32 columns) and 256-byte row segments for the column
// pass of a 2D plan with a split outer axis (1024 and 512 points)
#define MCFFT_COL_WIDE_CM(X, CM) \
  X(2, 4, 32, 16, CM, 2)         \
  X(2, 4, 32, 8, CM, 2)          \
  X(2, 2, 64, 16, CM, 2)

#define MCFFT_COL_LIST(X)                          \
  MCFFT_COL_LIST_CM(X, 0, 2, 32, 16, 16, 8)        \
  MCFFT_COL_LIST_CM(X, 1, 2, 32, 16, 16, 8)        \
  MCFFT_COL_LIST_CM(X, 2, 2, 16, 16, 8, 4)         \
  MCFFT_COL_TUNE_CM(X, 0, 2)                       \
  MCFFT_COL_TUNE_CM(X, 1, 2)                       \
  MCFFT_COL_WIDE_CM(X, 0)                          \
  MCFFT_COL_WIDE_CM(X, 1)                          \
  X(2, 1, 16, 4, 2, 2)                             \
  X(2, 2, 16, 8, 0, 2)                             \
  X(2, 2, 16, 8, 1, 2)                             \
  MCFFT_COL_NARROW_CM(X, 0)                        \
  MCFFT_COL_NARROW_CM(X, 1)                        \
  MCFFT_COL_NARROW_CM(X, 2)

#define MCFFT_COLF_LIST(X)                         \
  MCFFT_COL_LIST_CM(X, 0, 3, 32, 16, 16, 8)        \
  MCFFT_COL_LIST_CM(X, 1, 3, 32, 16, 16, 8)        \
  MCFFT_COL_LIST_CM(X, 2, 3, 16, 16, 8, 4)         \
  MCFFT_COL_TUNE_CM(X, 0, 3)                       \
  X(2, 1, 16, 4, 2, 3)

// COLM: the middle pass of a three-pass 1D plan (N2 = 2^4..2^9 points, columns
// 2^5..2^9 apart)
#define MCFFT_COLM_LIST_CM(X, CM, Q512) \
  X(1, 1, 32, 4, CM, 4)                 \
  X(1, 2, 32, 4, CM, 4)                 \
  X(1, 4, 32, 4, CM, 4)                 \
  X(1, 8, 32, 4, CM, 4)                 \
  X(2, 1, 32, 8, CM, 4)                 \
  X(2, 2, Q512, 8, CM, 4)

#define MCFFT_COLM_LIST(X)         \
  MCFFT_COLM_LIST_CM(X, 0, 32)     \
  MCFFT_COLM_LIST_CM(X, 1, 32)     \
  MCFFT_COLM_LIST_CM(X, 2, 16)

// ---- ROWX: Q = R rows on NW <= L / 512 waves ----
#define MCFFT_ROWX_L_CM(X, S, T, NW, CM) \
  X(S, T, 2, NW, CM, 5)                  \
  X(S, T, 4, NW, CM, 5)

#define MCFFT_ROWX_CM(X, CM)          \
  MCFFT_ROWX_L_CM(X, 2, 2, 1, CM)     \
  MCFFT_ROWX_L_CM(X, 2, 4, 2, CM)     \
  MCFFT_ROWX_L_CM(X, 2, 8, 4, CM)     \
  MCFFT_ROWX_L_CM(X, 3, 1, 8, CM)

// R = 8 where the tile fits the LDS (<= 139 KiB); 4096-point rows also on 4
// waves (two column sets per wave: a tuner candidate, measured 9-17 % slower
// than 8 waves on 4096^2)
#define MCFFT_ROWX_LIST(X) \
  MCFFT_ROWX_CM(X, 0)      \
  MCFFT_ROWX_CM(X, 1)      \
  MCFFT_ROWX_CM(X, 2)      \
  X(3, 1, 2, 4, 0, 5)      \
  X(3, 1, 2, 4, 1, 5)      \
  X(3, 1, 2, 4, 2, 5)      \
  X(2, 8, 8, 4, 0, 5)      \
  X(3, 1, 8, 8, 0, 5)      \
  X(2, 8, 8, 4, 1, 5)      \
  X(3, 1, 8, 8, 1, 5)      \
  X(2, 8, 8, 4, 2, 5)

namespace mcfft {
struct KernelInfo {
  const char* name;
  const void* fn;   // __global__ entry (host stub); launched with hipLaunchKernel
  int S, T, Q, nw, cm, kind;
  int lds_bytes;
  int mode = 0;     // scaling it serves: 0 any, 1 static (NONE/BALANCED/FULL) only, 2 AUTO only
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
extern const KernelInfo kColMKernels[];
extern const int kColMKernelCount;
extern const KernelInfo kRowXKernels[];
extern const int kRowXKernelCount;
// ragged batches: pass variants per compute mode and the persistent kernels
extern const KernelInfo kRaggedVariantInfo[3][30];
extern const KernelInfo kRaggedKernels[3];
}  // namespace mcfft

#define MCFFT_KNAME(S, T, Q, NW, CM, K) mcfft_pass_##S##_##T##_##Q##_##NW##_##CM##_##K

#define MCFFT_DEFINE_KERNEL(S, T, Q, NW, CM, K)                                   \
  MCFFT_KERNEL(NW * 64) void MCFFT_KNAME(S, T, Q, NW, CM, K)(mcfft::PassArgs a) { \
    mcfft::fft_pass_body<S, T, Q, NW, CM, K>(a);                                   \
  }

#define MCFFT_TABLE_ENTRY(S, T, Q, NW, CM, K)                                                            \
  {"mcfft_pass_" #S "_" #T "_" #Q "_" #NW "_" #CM "_" #K, (const void*)&MCFFT_KNAME(S, T, Q, NW, CM, K), \
   S, T, Q, NW, CM, K, mcfft::PassCfg<S, T, Q, NW, CM, K>::LDS_BYTES},

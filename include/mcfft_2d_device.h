// mcfft_2d_device.h -- reference-compatible 2D plan/exec API (include/mcfft_2d_device.h:18-38
// of the reference).  Data is [batch][x][y] with y contiguous.  The handle
// reports radices_x = {Nx}, radices_y = {Ny} (identity digit reversal); sides
// are powers of two 16..4096.  See mcfft_1d_device.h for the other differences.
#ifndef MCFFT_2D_DEVICE_H
#define MCFFT_2D_DEVICE_H

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include "mcfft.h"

#ifdef MCFFT_COMPAT_BF16
typedef __hip_bfloat16 MCTYPE;
typedef __hip_bfloat162 MCTYPE2;
#else
typedef __half MCTYPE;
typedef __half2 MCTYPE2;
#endif

typedef struct {
  int Nx, Ny, N_batch;
  int radices_x[3];
  int radices_y[3];
  int n_radices_x, n_radices_y;
  int mergings[2];
  mcfftPlan impl;
} mcfftHandle;

void mcfftExec(mcfftHandle plan, MCTYPE* data);
void mcfftCreate(mcfftHandle* plan, int nx, int ny, int n_batch);
void mcfftDestroy(mcfftHandle* plan);
mcfftStatus mcfftGetLastError(void);

#endif

// mcfft_3d_device.h -- 3D plan/exec API.  The reference's 3D files are the 2D
// code under a 3D name (include/mcfft_3d_device.h:37-38 takes (nx, ny, n_batch));
// that signature is kept and plans a 2D transform exactly as the reference
// does, and a (nx, ny, nz, n_batch) overload plans a real 3D transform
// ([batch][x][y][z], z contiguous, sides 16..4096).  See mcfft_1d_device.h
// for the other differences.
#ifndef MCFFT_3D_DEVICE_H
#define MCFFT_3D_DEVICE_H

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
  int Nx, Ny, Nz, N_batch;
  int radices_x[3];
  int radices_y[3];
  int radices_z[3];
  int n_radices_x, n_radices_y, n_radices_z;
  int mergings[3];
  mcfftPlan impl;
} mcfftHandle;

void mcfftExec(mcfftHandle plan, MCTYPE* data);
void mcfftCreate(mcfftHandle* plan, int nx, int ny, int n_batch);
void mcfftCreate(mcfftHandle* plan, int nx, int ny, int nz, int n_batch);
void mcfftDestroy(mcfftHandle* plan);
mcfftStatus mcfftGetLastError(void);

#endif

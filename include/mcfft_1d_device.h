// mcfft_1d_device.h -- reference-compatible 1D plan/exec API on top of mcfft.h.
//
// Keeps the signatures of include/mcfft_1d_device.h:18-38 of the reference
// (mcfftHandle, mcfftCreate, mcfftExec) so the reference's utils layer and
// drivers compile unchanged.  Differences a caller can see:
//   * MCTYPE is binary16 (__half) and MCTYPE2 its pair type (__half2), in
//     place of the reference's double / double2 (include/common.h:4-5) -- the
//     FP16 MFMA path; define MCFFT_COMPAT_BF16 for bfloat16 storage instead.
//   * mcfftCreate reports radices = {N}, n_radices = 1, so the reference's
//     gen_rev permutation is the identity and data is in natural order.
//   * any power of two 16..2^23 is supported; an unsupported size or a HIP
//     failure is reported on stderr and by mcfftGetLastError() (the reference
//     leaves the plan uninitialised).
//   * mcfftExec transforms are unnormalised (MCFFT_SCALE_NONE), in place, on
//     the null stream, as in the reference; a result that overflowed FP16 is
//     reported by mcfftGetLastError() as MCFFT_NONFINITE (which waits for the
//     device).
//   * the utils layer (setup / doit / finalize, FP64 on the host) runs its plan
//     with block floating point (MCFFT_SCALE_AUTO) and returns true values:
//     no overflow or underflow whatever the data, and doit(iter) iterates the
//     transform exactly as the reference would at FP64.
#ifndef MCFFT_1D_DEVICE_H
#define MCFFT_1D_DEVICE_H

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
  int N, N_batch;
  int radices[9];
  int n_radices;
  int mergings[3];
  int n_mergings;
  mcfftPlan impl;
} mcfftHandle;

void mcfftExec(mcfftHandle plan, MCTYPE* data);
void mcfftCreate(mcfftHandle* plan, int n, int n_batch);
// additions (no reference counterpart)
void mcfftDestroy(mcfftHandle* plan);
mcfftStatus mcfftGetLastError(void);

#endif

// mcfft_1d_utils.h -- reference-compatible host utilities (include/mcfft_1d_utils.h:8-11
// of the reference): FP64 host data in, FP64 host data out, one global plan.
#ifndef MCFFT_1D_UTILS_H
#define MCFFT_1D_UTILS_H

#include "mcfft_1d_device.h"

void setup(double* data, int n, int n_batch);
void doit(int iter);
void finalize(double* result);
void generate_data(double* data, int n, int n_batch, int seed = 42);

#endif

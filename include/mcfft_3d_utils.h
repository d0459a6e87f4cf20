// mcfft_3d_utils.h -- host utilities for 3D (include/mcfft_3d_utils.h:8-11 of the
// reference, whose setup takes only (data, nx, ny, n_batch); kept, plus an
// (nx, ny, nz, n_batch) overload).
#ifndef MCFFT_3D_UTILS_H
#define MCFFT_3D_UTILS_H

#include "mcfft_3d_device.h"

void setup(double* data, int nx, int ny, int n_batch);
void setup(double* data, int nx, int ny, int nz, int n_batch);
void doit(int iter);
void finalize(double* result);
void generate_data(double* data, int n, int n_batch, int seed = 42);

#endif

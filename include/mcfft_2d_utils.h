// mcfft_2d_utils.h -- reference-compatible host utilities (include/mcfft_2d_utils.h:8-11).
#ifndef MCFFT_2D_UTILS_H
#define MCFFT_2D_UTILS_H

#include "mcfft_2d_device.h"

void setup(double* data, int nx, int ny, int n_batch);
void doit(int iter);
void finalize(double* result);
void generate_data(double* data, int n, int n_batch, int seed = 42);

#endif

// k_row.hip -- instantiations of the pass kernel for the MCFFT_ROW_LIST configurations.
#include "fft_pass.h"
#include "kernel_list.h"

MCFFT_ROW_LIST(MCFFT_DEFINE_KERNEL)

namespace mcfft {
const KernelInfo kRowKernels[] = {MCFFT_ROW_LIST(MCFFT_TABLE_ENTRY)};
const int kRowKernelCount = (int)(sizeof(kRowKernels) / sizeof(kRowKernels[0]));
}  // namespace mcfft

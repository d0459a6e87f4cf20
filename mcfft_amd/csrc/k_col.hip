// k_col.hip -- instantiations of the pass kernel for the MCFFT_COL_LIST configurations.
#include "fft_pass.h"
#include "kernel_list.h"

MCFFT_COL_LIST(MCFFT_DEFINE_KERNEL)

namespace mcfft {
const KernelInfo kColKernels[] = {MCFFT_COL_LIST(MCFFT_TABLE_ENTRY)};
const int kColKernelCount = (int)(sizeof(kColKernels) / sizeof(kColKernels[0]));
}  // namespace mcfft

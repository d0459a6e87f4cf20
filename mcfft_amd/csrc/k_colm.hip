// k_colm.hip -- instantiations of the pass kernel for the MCFFT_COLM_LIST configurations.
#include "fft_pass.h"
#include "kernel_list.h"

MCFFT_COLM_LIST(MCFFT_DEFINE_KERNEL)

namespace mcfft {
const KernelInfo kColMKernels[] = {MCFFT_COLM_LIST(MCFFT_TABLE_ENTRY)};
const int kColMKernelCount = (int)(sizeof(kColMKernels) / sizeof(kColMKernels[0]));
}  // namespace mcfft

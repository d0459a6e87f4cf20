// k_colf.hip -- instantiations of the pass kernel for the MCFFT_COLF_LIST configurations.
#include "fft_pass.h"
#include "kernel_list.h"

MCFFT_COLF_LIST(MCFFT_DEFINE_KERNEL)

namespace mcfft {
const KernelInfo kColFKernels[] = {MCFFT_COLF_LIST(MCFFT_TABLE_ENTRY)};
const int kColFKernelCount = (int)(sizeof(kColFKernels) / sizeof(kColFKernels[0]));
}  // namespace mcfft

// k_rowx.hip -- instantiations of the pass kernel for the MCFFT_ROWX_LIST configurations.
#include "fft_pass.h"
#include "kernel_list.h"

MCFFT_ROWX_LIST(MCFFT_DEFINE_KERNEL)

namespace mcfft {
const KernelInfo kRowXKernels[] = {MCFFT_ROWX_LIST(MCFFT_TABLE_ENTRY)};
const int kRowXKernelCount = (int)(sizeof(kRowXKernels) / sizeof(kRowXKernels[0]));
}  // namespace mcfft

// k_rowt.hip -- instantiations of the pass kernel for the MCFFT_ROWT_LIST configurations.
#include "fft_pass.h"
#include "kernel_list.h"

MCFFT_ROWT_LIST(MCFFT_DEFINE_KERNEL)

namespace mcfft {
const KernelInfo kRowTKernels[] = {MCFFT_ROWT_LIST(MCFFT_TABLE_ENTRY)};
const int kRowTKernelCount = (int)(sizeof(kRowTKernels) / sizeof(kRowTKernels[0]));
}  // namespace mcfft

// k_ragged.hip -- the persistent ragged-batch kernels (one per compute mode) and
// the table of their pass variants for the planner.
#include "kernel_list.h"
#include "ragged.h"

// two 8-wave workgroups per CU: at most 128 VGPRs (4 waves per SIMD)
#define MCFFT_RAGGED_KERNEL(CM)                                                                   \
  MCFFT_KERNEL(mcfft::kRaggedNW * 64) MCFFT_WAVES_PER_SIMD(4)                                       \
  void mcfft_ragged_##CM(mcfft::RaggedArgs r) {                                                   \
    mcfft::ragged_body<CM>(r, mcfft::ragged_max_lds<CM>());                              \
  }
MCFFT_RAGGED_KERNEL(0)
MCFFT_RAGGED_KERNEL(1)
MCFFT_RAGGED_KERNEL(2)

namespace mcfft {
#define MCFFT_RAGGED_INFO(ID, S, T, Q, NW, CM, K) \
  {"ragged_" #ID, nullptr, S, T, Q, NW, CM, K, PassCfg<S, T, Q, NW, CM, K>::LDS_BYTES},
const KernelInfo kRaggedVariantInfo[3][kRaggedVariants] = {
    {MCFFT_RAGGED_LIST_F(MCFFT_RAGGED_INFO, 0, 1)},
    {MCFFT_RAGGED_LIST_F(MCFFT_RAGGED_INFO, 1, 1)},
    {MCFFT_RAGGED_LIST_F(MCFFT_RAGGED_INFO, 2, 2)},
};
#undef MCFFT_RAGGED_INFO
const KernelInfo kRaggedKernels[3] = {
    {"mcfft_ragged_0", (const void*)&mcfft_ragged_0, 0, 0, 0, kRaggedNW, 0, -1, ragged_lds_bytes<0>()},
    {"mcfft_ragged_1", (const void*)&mcfft_ragged_1, 0, 0, 0, kRaggedNW, 1, -1, ragged_lds_bytes<1>()},
    {"mcfft_ragged_2", (const void*)&mcfft_ragged_2, 0, 0, 0, kRaggedNW, 2, -1, ragged_lds_bytes<2>()},
};
}  // namespace mcfft

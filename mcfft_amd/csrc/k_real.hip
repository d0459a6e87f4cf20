// k_real.hip -- R2C post-pass and C2R pre-pass kernels (one per compute mode).
#include "real_pass.h"

#define MCFFT_REAL_KERNELS(CM)                                                                                   \
  MCFFT_KERNEL(1024) void mcfft_r2c_post_##CM(mcfft::RealArgs a) { mcfft::r2c_post_body<CM>(a); }             \
  MCFFT_KERNEL(1024) void mcfft_c2r_pre_##CM(mcfft::RealArgs a) { mcfft::c2r_pre_body<CM>(a); }
MCFFT_REAL_KERNELS(0)
MCFFT_REAL_KERNELS(1)
MCFFT_REAL_KERNELS(2)

namespace mcfft {
extern const void* const kR2CPost[3] = {(const void*)&mcfft_r2c_post_0, (const void*)&mcfft_r2c_post_1,
                                 (const void*)&mcfft_r2c_post_2};
extern const void* const kC2RPre[3] = {(const void*)&mcfft_c2r_pre_0, (const void*)&mcfft_c2r_pre_1,
                                (const void*)&mcfft_c2r_pre_2};
}  // namespace mcfft

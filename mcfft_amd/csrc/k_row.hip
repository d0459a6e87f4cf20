// k_row.hip -- instantiations of the pass kernel for the MCFFT_ROW_LIST configurations.
#include "fft_pass.h"
#include "fft_big.h"
#include "kernel_list.h"

MCFFT_ROW_LIST(MCFFT_DEFINE_KERNEL)
// single-pass 65536 points (one workgroup per transform, two-round LDS exchange),
// in a statically scaled form and an AUTO (block floating point) form, each
// carrying only its own bookkeeping (the 16-wave kernel has no spare registers)
MCFFT_KERNEL(512) void mcfft_pass_big_0(mcfft::PassArgs a) { mcfft::fft_big_body<0, 8, false>(a); }
MCFFT_KERNEL(512) void mcfft_pass_big_1(mcfft::PassArgs a) { mcfft::fft_big_body<1, 8, false>(a); }
MCFFT_KERNEL(1024) void mcfft_pass_big16_0(mcfft::PassArgs a) { mcfft::fft_big_body<0, 16, false>(a); }
MCFFT_KERNEL(1024) void mcfft_pass_big16_1(mcfft::PassArgs a) { mcfft::fft_big_body<1, 16, false>(a); }
MCFFT_KERNEL(512) void mcfft_pass_bigauto_0(mcfft::PassArgs a) { mcfft::fft_big_body<0, 8, true>(a); }
MCFFT_KERNEL(512) void mcfft_pass_bigauto_1(mcfft::PassArgs a) { mcfft::fft_big_body<1, 8, true>(a); }
MCFFT_KERNEL(1024) void mcfft_pass_bigauto16_0(mcfft::PassArgs a) { mcfft::fft_big_body<0, 16, true>(a); }
MCFFT_KERNEL(1024) void mcfft_pass_bigauto16_1(mcfft::PassArgs a) { mcfft::fft_big_body<1, 16, true>(a); }

namespace mcfft {
const KernelInfo kRowKernels[] = {
    MCFFT_ROW_LIST(MCFFT_TABLE_ENTRY)
    {"mcfft_pass_big_0", (const void*)&mcfft_pass_big_0, 4, 1, 1, 8, 0, 0, BigCfg<8>::LDS_BYTES, 1},
    {"mcfft_pass_big_1", (const void*)&mcfft_pass_big_1, 4, 1, 1, 8, 1, 0, BigCfg<8>::LDS_BYTES, 1},
    {"mcfft_pass_big16_0", (const void*)&mcfft_pass_big16_0, 4, 1, 1, 16, 0, 0, BigCfg<16>::LDS_BYTES, 1},
    {"mcfft_pass_big16_1", (const void*)&mcfft_pass_big16_1, 4, 1, 1, 16, 1, 0, BigCfg<16>::LDS_BYTES, 1},
    {"mcfft_pass_bigauto_0", (const void*)&mcfft_pass_bigauto_0, 4, 1, 1, 8, 0, 0, BigCfg<8>::LDS_BYTES, 2},
    {"mcfft_pass_bigauto_1", (const void*)&mcfft_pass_bigauto_1, 4, 1, 1, 8, 1, 0, BigCfg<8>::LDS_BYTES, 2},
    {"mcfft_pass_bigauto16_0", (const void*)&mcfft_pass_bigauto16_0, 4, 1, 1, 16, 0, 0, BigCfg<16>::LDS_BYTES, 2},
    {"mcfft_pass_bigauto16_1", (const void*)&mcfft_pass_bigauto16_1, 4, 1, 1, 16, 1, 0, BigCfg<16>::LDS_BYTES, 2},
};
const int kRowKernelCount = (int)(sizeof(kRowKernels) / sizeof(kRowKernels[0]));
}  // namespace mcfft

// k_layout.hip -- gather (caller layout -> contiguous) and scatter (contiguous
// -> caller layout) kernels of the advanced-layout exec.
#include "layout_pass.h"

MCFFT_KERNEL(256) void mcfft_layout_gather(mcfft::LayoutArgs a) { mcfft::layout_body<true>(a); }
MCFFT_KERNEL(256) void mcfft_layout_scatter(mcfft::LayoutArgs a) { mcfft::layout_body<false>(a); }

namespace mcfft {
extern const void* const kLayoutGather = (const void*)&mcfft_layout_gather;
extern const void* const kLayoutScatter = (const void*)&mcfft_layout_scatter;
}  // namespace mcfft

// dev_prims.h -- the handful of device primitives the FFT kernels use.
//
// The kernels are written once against this header.  Built by hipcc for
// gfx950 the primitives are the CDNA4 instructions themselves (MFMA
// 32x32x16 f16/bf16, LDS, wave64 lane id, workgroup barrier).  Built by a
// host compiler with MCFFT_EMU defined (test infrastructure under tests/emu/)
// each workgroup runs as host threads: LDS and global accesses are
// bounds-checked and the MFMA is evaluated from the documented operand
// layout, so indexing mistakes surface on the CPU before any GPU run.
#pragma once
#include <stdint.h>

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

#ifndef MCFFT_EMU
#include <hip/hip_runtime.h>

#define MCFFT_DEV __device__ __forceinline__
#define MCFFT_NOINLINE __device__ __attribute__((noinline))
#define MCFFT_KERNEL(nt) extern "C" __global__ __launch_bounds__(nt)
// a kernel whose registers must leave room for w waves per SIMD
#define MCFFT_WAVES_PER_SIMD(w) __attribute__((amdgpu_waves_per_eu(w)))
#define MCFFT_TKERNEL(nt) __global__ __launch_bounds__(nt)
#define MCFFT_DYN_LDS(type, name) \
  extern __shared__ __attribute__((aligned(16))) unsigned char mcfft_lds_raw[]; \
  type* name = (type*)mcfft_lds_raw

namespace mcfft {
MCFFT_DEV int lane_id() { return __lane_id(); }
MCFFT_DEV int thread_id() { return threadIdx.x; }
MCFFT_DEV int block_id() { return blockIdx.x; }
MCFFT_DEV int grid_size() { return gridDim.x; }
MCFFT_DEV int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
MCFFT_DEV long long uniform64(unsigned long long v) {
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v), hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return (long long)(((unsigned long long)hi << 32) | lo);
}
MCFFT_DEV const void* uniform_ptr(const void* p) {
  const unsigned long long v = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v), hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return (const void*)(((unsigned long long)hi << 32) | lo);
}
// (cos 2 pi x, sin 2 pi x) from the hardware transcendental units, whose argument
// is in revolutions (v_cos_f32 / v_sin_f32; |error| <= 1.3e-7 over x = e / 2^k,
// e < 2^23, measured on MI355X -- about one FP32 ulp of 1)
MCFFT_DEV f32x2 cis_rev(float x) {
  f32x2 r;
  r.x = __builtin_amdgcn_cosf(x);
  r.y = __builtin_amdgcn_sinf(x);
  return r;
}
// value of lane (lane ^ 16) of the same wave (ds_swizzle bit mode: and 0x1f,
// xor 0x10 within each 32-lane half; all 64 lanes take part)
MCFFT_DEV float lane_xor16(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_ds_swizzle(__builtin_bit_cast(int, v), 0x401f));
}
// a copy of v the compiler cannot see through (keeps values derived from it
// from being hoisted out of a loop)
MCFFT_DEV int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}
MCFFT_DEV void block_barrier() { __syncthreads(); }
// the compiler may not move memory accesses across this point
MCFFT_DEV void mem_order_point() { asm volatile("" ::: "memory"); }

template <class T> MCFFT_DEV T lds_ld(const T* p, int i) { return p[i]; }
template <class T> MCFFT_DEV void lds_st(T* p, int i, T v) { p[i] = v; }
template <class T> MCFFT_DEV T g_ld(const T* p, long long i) { return p[i]; }
// wave-uniform read of memory that no kernel of the launch writes, through the
// scalar cache (constant address space: s_load): its wait covers earlier scalar
// loads only, where a vector load's wait would also cover the wave's pending
// vector loads and stores
template <class T> MCFFT_DEV T s_ld(const T* p, long long i) {
  typedef const __attribute__((address_space(4))) T* ConstPtr;
  return ((ConstPtr)p)[i];
}
template <class T> MCFFT_DEV void g_st(T* p, long long i, T v) { p[i] = v; }

// Buffer-resource addressing: a 64-bit base in scalar registers plus 32-bit byte
// offsets per lane (no 64-bit address arithmetic per access).  Loads outside
// [0, bytes) return 0 and stores there are dropped by the hardware.
struct GBuf {
  __amdgpu_buffer_rsrc_t r;
};
MCFFT_DEV GBuf make_gbuf(const void* p, unsigned bytes) {
  GBuf b;
  b.r = __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)bytes, 0x00020000);
  return b;
}
MCFFT_DEV uint32_t buf_ld32(GBuf b, int off) { return __builtin_amdgcn_raw_buffer_load_b32(b.r, off, 0, 0); }
MCFFT_DEV u32x2 buf_ld64(GBuf b, int off) { return __builtin_amdgcn_raw_buffer_load_b64(b.r, off, 0, 0); }
MCFFT_DEV u32x4 buf_ld128(GBuf b, int off) { return __builtin_amdgcn_raw_buffer_load_b128(b.r, off, 0, 0); }
MCFFT_DEV void buf_st32(GBuf b, int off, uint32_t v) { __builtin_amdgcn_raw_buffer_store_b32(v, b.r, off, 0, 0); }
MCFFT_DEV void buf_st64(GBuf b, int off, u32x2 v) { __builtin_amdgcn_raw_buffer_store_b64(v, b.r, off, 0, 0); }
// the same with a wave-uniform byte offset added in a scalar register
MCFFT_DEV uint32_t buf_ld32s(GBuf b, int voff, int soff) { return __builtin_amdgcn_raw_buffer_load_b32(b.r, voff, soff, 0); }
MCFFT_DEV u32x4 buf_ld128s(GBuf b, int voff, int soff) { return __builtin_amdgcn_raw_buffer_load_b128(b.r, voff, soff, 0); }
MCFFT_DEV void buf_st32s(GBuf b, int voff, int soff, uint32_t v) {
  __builtin_amdgcn_raw_buffer_store_b32(v, b.r, voff, soff, 0);
}
MCFFT_DEV void buf_st64s(GBuf b, int voff, int soff, u32x2 v) {
  __builtin_amdgcn_raw_buffer_store_b64(v, b.r, voff, soff, 0);
}

// device-scope (all XCDs) atomics and fences for the ragged work queue
MCFFT_DEV unsigned atomic_add_dev(unsigned* p, unsigned v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
MCFFT_DEV unsigned atomic_load_dev(const unsigned* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
MCFFT_DEV void atomic_or_dev(unsigned* p, unsigned v) {
  (void)__hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
MCFFT_DEV void atomic_max_dev64(unsigned long long* p, unsigned long long v) {
  (void)__hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
MCFFT_DEV void atomic_add_dev64(unsigned long long* p, unsigned long long v) {
  (void)__hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// maximum of v over the 64 lanes of the wave, as a wave-uniform value (every
// lane takes part): xor butterflies inside each 32-lane half with ds_swizzle
// (the pattern is an immediate, so no lane-address registers are needed, unlike
// ds_bpermute), then the two halves' results combined in scalar registers
MCFFT_DEV unsigned wave_max(unsigned v) {
#define MCFFT_SWZ_MAX(pattern)                                                  \
  {                                                                             \
    const unsigned w = (unsigned)__builtin_amdgcn_ds_swizzle((int)v, pattern);  \
    v = w > v ? w : v;                                                          \
  }
  MCFFT_SWZ_MAX(0x041f)  // lane ^ 1
  MCFFT_SWZ_MAX(0x081f)  // lane ^ 2
  MCFFT_SWZ_MAX(0x101f)  // lane ^ 4
  MCFFT_SWZ_MAX(0x201f)  // lane ^ 8
  MCFFT_SWZ_MAX(0x401f)  // lane ^ 16
#undef MCFFT_SWZ_MAX
  const unsigned a = (unsigned)__builtin_amdgcn_readlane((int)v, 0), b = (unsigned)__builtin_amdgcn_readlane((int)v, 32);
  return a > b ? a : b;
}
// device-scope release / acquire, one side each: a producer's release (writes back
// its XCD's dirty L2 lines) before it counts a tile, a consumer's acquire (drops
// stale L1/L2 lines) after its wait
MCFFT_DEV void release_dev() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
MCFFT_DEV void acquire_dev() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
MCFFT_DEV void nap() { __builtin_amdgcn_s_sleep(4); }
// wait n ticks of the constant 100 MHz clock (s_memrealtime); ends by itself
MCFFT_DEV void stall_ticks(int n) {
  const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
  while (__builtin_amdgcn_s_memrealtime() - t0 < (unsigned long long)n) __builtin_amdgcn_s_sleep(8);
}

// Work done for the whole workgroup by one wave inside a loop that also holds
// barriers: branch on wave_id() (wave-uniform, a scalar branch) and let every
// lane of that wave run the atomics, lane 0 with the real operand and the
// others with 0, reading results through wave_first.  A branch on the thread
// index instead (tid == 0) is divergent inside the wave, and the compiler may
// then turn the loop into a nest whose inner loop runs the other lanes through
// further iterations (and barriers) while lane 0 waits outside it.
MCFFT_DEV int wave_id() { return __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6); }
MCFFT_DEV unsigned wave_first(unsigned v) { return (unsigned)__builtin_amdgcn_readfirstlane((int)v); }

// pack the low halves / high halves of two packed pairs into one pair (v_perm_b32)
MCFFT_DEV uint32_t lo_halves(uint32_t e0, uint32_t e1) { return __builtin_amdgcn_perm(e1, e0, 0x05040100u); }
MCFFT_DEV uint32_t hi_halves(uint32_t e0, uint32_t e1) { return __builtin_amdgcn_perm(e1, e0, 0x07060302u); }

// D = A(32x16) * B(16x32) + C, f16 operands, fp32 accumulate (v_mfma_f32_32x32x16_f16).
// Lane l holds A[l%32][8*(l/32)+t] and B[8*(l/32)+t][l%32] (t = 0..7) and
// D[8*(v/4) + 4*(l/32) + v%4][l%32] in accumulator v = 0..15.
MCFFT_DEV f32x16 mfma_f16(u32x4 a, u32x4 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
}
MCFFT_DEV f32x16 mfma_bf16(u32x4 a, u32x4 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}
}  // namespace mcfft

#else  // ------------------------------ CPU emulation ---------------------------
#include "emu_runtime.h"
#endif

namespace mcfft {
// ---- scalar format conversions (RNE) ----
MCFFT_DEV uint32_t f16_bits(float x) { return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)x); }
MCFFT_DEV float f16_val(uint32_t bits) { return (float)__builtin_bit_cast(_Float16, (uint16_t)bits); }
MCFFT_DEV uint32_t bf16_bits(float x) { return (uint32_t)__builtin_bit_cast(uint16_t, (__bf16)x); }
MCFFT_DEV float bf16_val(uint32_t bits) { return (float)__builtin_bit_cast(__bf16, (uint16_t)bits); }
MCFFT_DEV uint32_t pack2(uint32_t lo, uint32_t hi) { return (lo & 0xffffu) | (hi << 16); }
}  // namespace mcfft

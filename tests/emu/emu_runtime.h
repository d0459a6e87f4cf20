// emu_runtime.h -- CPU emulation of the device primitives used by the FFT
// kernels (test infrastructure; included by mcfft_amd/csrc/dev_prims.h only
// when MCFFT_EMU is defined).  Each workgroup runs as blockDim host threads;
// LDS and global accesses are bounds-checked; the MFMA is evaluated from the
// operand layout documented in dev_prims.h.
#pragma once
#include <barrier>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <cmath>
#include <thread>
#include <vector>

#define MCFFT_DEV inline
#define MCFFT_NOINLINE __attribute__((noinline))
#define MCFFT_KERNEL(nt) extern "C"
#define MCFFT_WAVES_PER_SIMD(w)
#define MCFFT_TKERNEL(nt)
#define MCFFT_DYN_LDS(type, name) type* name = (type*)::emu::lds_base()
#define __restrict__
#define __device__
#define __global__
#define __host__

namespace emu {
struct Range { const char* lo; const char* hi; };
inline std::vector<Range>& ranges() { static std::vector<Range> r; return r; }
inline void register_range(const void* p, size_t bytes) { ranges().push_back({(const char*)p, (const char*)p + bytes}); }
inline void clear_ranges() { ranges().clear(); }
[[noreturn]] inline void fail(const char* what, long long i) {
  fprintf(stderr, "EMU BOUNDS ERROR: %s index %lld\n", what, i);
  fflush(stderr);
  abort();
}
inline void check_global(const void* p, size_t bytes, long long i) {
  const char* c = (const char*)p;
  for (auto& r : ranges()) if (c >= r.lo && c + bytes <= r.hi) return;
  fail("global", i);
}

struct Block {
  int nt = 0, bid = 0, nblocks = 0;
  std::vector<unsigned char> lds;
  std::barrier<>* bar = nullptr;
  std::vector<std::barrier<>*> wbar;
  // per-wave MFMA exchange buffers: [wave][lane][8] for A and B
  std::vector<float> xa, xb;
  std::vector<float> xs;  // per-wave lane exchange: [wave][lane]
};
inline thread_local Block* tl_blk = nullptr;
inline thread_local int tl_tid = 0;

inline void* lds_base() { return tl_blk->lds.data(); }
inline size_t lds_bytes() { return tl_blk->lds.size(); }

// launch all workgroups at once (each as nt host threads): needed by kernels whose
// workgroups wait on each other (the ragged work queue)
inline void launch_concurrent(int nblocks, int nt, size_t lds, const std::function<void()>& body) {
  std::vector<Block*> blks;
  std::vector<std::barrier<>*> bars;
  std::vector<std::thread> th;
  for (int b = 0; b < nblocks; ++b) {
    Block* blk = new Block();
    blk->nt = nt;
    blk->bid = b;
    blk->nblocks = nblocks;
    blk->lds.assign(lds, 0xCD);
    blk->bar = new std::barrier<>(nt);
    for (int w = 0; w < nt / 64; ++w) blk->wbar.push_back(new std::barrier<>(64));
    blk->xa.assign((size_t)(nt / 64) * 64 * 8, 0.f);
    blk->xb.assign((size_t)(nt / 64) * 64 * 8, 0.f);
    blk->xs.assign((size_t)nt, 0.f);
    blks.push_back(blk);
  }
  for (int b = 0; b < nblocks; ++b)
    for (int t = 0; t < nt; ++t) th.emplace_back([&, b, t] { tl_blk = blks[b]; tl_tid = t; body(); });
  for (auto& x : th) x.join();
  for (Block* blk : blks) {
    delete blk->bar;
    for (auto* w : blk->wbar) delete w;
    delete blk;
  }
}

// workgroups of a launch() that run at the same time (EMU_BLOCKS, default 4):
// the pass kernels' workgroups only meet through atomics, which are real ones
inline int concurrent_blocks() {
  static const int n = [] {
    const char* e = getenv("EMU_BLOCKS");
    const int v = e ? atoi(e) : 4;
    return v < 1 ? 1 : v;
  }();
  return n;
}

// launch `body` as a grid of nblocks workgroups of nt threads with `lds` bytes of
// LDS, running only workgroups [b0, b1) of it (default: all), a few at a time
inline void launch(int nblocks, int nt, size_t lds, const std::function<void()>& body, int b0 = 0, int b1 = -1) {
  if (b1 < 0 || b1 > nblocks) b1 = nblocks;
  const int nw = nt / 64;
  for (int g0 = b0; g0 < b1; g0 += concurrent_blocks()) {
    const int g1 = g0 + concurrent_blocks() < b1 ? g0 + concurrent_blocks() : b1;
    std::vector<Block> blks((size_t)(g1 - g0));
    std::vector<std::thread> th;
    for (int b = g0; b < g1; ++b) {
      Block& blk = blks[(size_t)(b - g0)];
      blk.nt = nt;
      blk.bid = b;
      blk.nblocks = nblocks;
      blk.lds.assign(lds, 0xCD);
      blk.bar = new std::barrier<>(nt);
      for (int w = 0; w < nw; ++w) blk.wbar.push_back(new std::barrier<>(64));
      blk.xa.assign((size_t)nw * 64 * 8, 0.f);
      blk.xb.assign((size_t)nw * 64 * 8, 0.f);
      blk.xs.assign((size_t)nt, 0.f);
    }
    for (int b = g0; b < g1; ++b)
      for (int t = 0; t < nt; ++t) {
        Block* bp = &blks[(size_t)(b - g0)];
        th.emplace_back([&body, bp, t] { tl_blk = bp; tl_tid = t; body(); });
      }
    for (auto& x : th) x.join();
    for (Block& blk : blks) {
      delete blk.bar;
      for (auto* p : blk.wbar) delete p;
    }
  }
}
}  // namespace emu

namespace mcfft {
inline int lane_id() { return emu::tl_tid % 64; }
inline int thread_id() { return emu::tl_tid; }
inline int block_id() { return emu::tl_blk->bid; }
inline int grid_size() { return emu::tl_blk->nblocks; }
inline int uniform(int v) { return v; }
inline long long uniform64(unsigned long long v) { return (long long)v; }
inline const void* uniform_ptr(const void* p) { return p; }
inline int opaque(int v) { return v; }
inline float lane_xor(float v, int mask);
// value of lane (lane ^ 16) of the same wave (all 64 lanes take part)
inline float lane_xor16(float v) { return lane_xor(v, 16); }
// value of lane (lane ^ mask) of the same wave (all 64 lanes take part)
inline float lane_xor(float v, int mask) {
  emu::Block* b = emu::tl_blk;
  const int t = emu::tl_tid, w = t / 64, l = t % 64;
  b->xs[(size_t)w * 64 + l] = v;
  b->wbar[w]->arrive_and_wait();
  const float r = b->xs[(size_t)w * 64 + (l ^ mask)];
  b->wbar[w]->arrive_and_wait();
  return r;
}
inline f32x2 cis_rev(float x) {
  f32x2 r;
  r.x = (float)std::cos(2.0 * M_PI * (double)x);
  r.y = (float)std::sin(2.0 * M_PI * (double)x);
  return r;
}
inline void block_barrier() { emu::tl_blk->bar->arrive_and_wait(); }
inline void mem_order_point() {}

template <class T> inline T lds_ld(const T* p, int i) {
  const char* base = (const char*)emu::lds_base();
  const char* a = (const char*)(p + i);
  if (i < 0 || a < base || a + sizeof(T) > base + emu::lds_bytes()) emu::fail("lds load", i);
  return p[i];
}
template <class T> inline void lds_st(T* p, int i, T v) {
  const char* base = (const char*)emu::lds_base();
  const char* a = (const char*)(p + i);
  if (i < 0 || a < base || a + sizeof(T) > base + emu::lds_bytes()) emu::fail("lds store", i);
  p[i] = v;
}
template <class T> inline T g_ld(const T* p, long long i) { emu::check_global(p + i, sizeof(T), i); return p[i]; }
template <class T> inline T s_ld(const T* p, long long i) { return g_ld(p, i); }
template <class T> inline void g_st(T* p, long long i, T v) { emu::check_global(p + i, sizeof(T), i); p[i] = v; }

// buffer-resource emulation with the hardware's range check: an access outside
// [0, bytes) of the view reads 0 / is dropped; an access inside it must hit
// registered memory (otherwise the view itself is wrong: abort)
struct GBuf {
  const char* p;
  unsigned bytes;
};
inline GBuf make_gbuf(const void* p, unsigned bytes) { return GBuf{(const char*)p, bytes}; }
// (the hardware reads a buffer offset as an unsigned 32-bit byte count)
template <class T> inline T buf_ld(GBuf b, int off) {
  T v;
  const size_t uo = (unsigned)off;
  if (uo + sizeof(T) > b.bytes) {
    memset(&v, 0, sizeof(T));
    return v;
  }
  emu::check_global(b.p + uo, sizeof(T), (long long)uo);
  memcpy(&v, b.p + uo, sizeof(T));
  return v;
}
template <class T> inline void buf_st(GBuf b, int off, T v) {
  const size_t uo = (unsigned)off;
  if (uo + sizeof(T) > b.bytes) return;
  emu::check_global(b.p + uo, sizeof(T), (long long)uo);
  memcpy((char*)b.p + uo, &v, sizeof(T));
}
inline uint32_t buf_ld32(GBuf b, int off) { return buf_ld<uint32_t>(b, off); }
inline u32x2 buf_ld64(GBuf b, int off) { return buf_ld<u32x2>(b, off); }
inline u32x4 buf_ld128(GBuf b, int off) { return buf_ld<u32x4>(b, off); }
inline void buf_st32(GBuf b, int off, uint32_t v) { buf_st<uint32_t>(b, off, v); }
inline void buf_st64(GBuf b, int off, u32x2 v) { buf_st<u32x2>(b, off, v); }
// vector + scalar offsets add as unsigned 32-bit values
inline int uadd(int a, int b) { return (int)((unsigned)a + (unsigned)b); }
inline uint32_t buf_ld32s(GBuf b, int voff, int soff) { return buf_ld<uint32_t>(b, uadd(voff, soff)); }
inline u32x4 buf_ld128s(GBuf b, int voff, int soff) { return buf_ld<u32x4>(b, uadd(voff, soff)); }
inline void buf_st32s(GBuf b, int voff, int soff, uint32_t v) { buf_st<uint32_t>(b, uadd(voff, soff), v); }
inline void buf_st64s(GBuf b, int voff, int soff, u32x2 v) { buf_st<u32x2>(b, uadd(voff, soff), v); }

inline unsigned atomic_add_dev(unsigned* p, unsigned v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline unsigned atomic_load_dev(const unsigned* p) { return __atomic_load_n(p, __ATOMIC_SEQ_CST); }
inline void atomic_or_dev(unsigned* p, unsigned v) { (void)__atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
inline void atomic_max_dev64(unsigned long long* p, unsigned long long v) {
  unsigned long long cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (cur < v && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
  }
}
inline void atomic_add_dev64(unsigned long long* p, unsigned long long v) { (void)__atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
// maximum of v over the 64 lanes of the wave (all 64 lanes take part): xor
// butterflies inside each 32-lane half, then lanes 0 and 32 combined
inline unsigned wave_max(unsigned v) {
  for (int o = 1; o <= 16; o <<= 1) {
    const unsigned w = __builtin_bit_cast(unsigned, lane_xor(__builtin_bit_cast(float, v), o));
    v = w > v ? w : v;
  }
  emu::Block* b = emu::tl_blk;
  const int t = emu::tl_tid, w = t / 64, l = t % 64;
  b->xs[(size_t)w * 64 + l] = __builtin_bit_cast(float, v);
  b->wbar[w]->arrive_and_wait();
  const unsigned lo = __builtin_bit_cast(unsigned, b->xs[(size_t)w * 64]);
  const unsigned hi = __builtin_bit_cast(unsigned, b->xs[(size_t)w * 64 + 32]);
  b->wbar[w]->arrive_and_wait();
  return lo > hi ? lo : hi;
}
inline void release_dev() { __atomic_thread_fence(__ATOMIC_RELEASE); }
inline void acquire_dev() { __atomic_thread_fence(__ATOMIC_ACQUIRE); }
inline void nap() { std::this_thread::yield(); }
inline void stall_ticks(int) {}
inline int wave_id() { return emu::tl_tid / 64; }
// lane 0's value, for every lane of the wave (all 64 lanes take part)
inline unsigned wave_first(unsigned v) {
  emu::Block* b = emu::tl_blk;
  const int t = emu::tl_tid, w = t / 64, l = t % 64;
  b->xs[(size_t)w * 64 + l] = __builtin_bit_cast(float, v);
  b->wbar[w]->arrive_and_wait();
  const unsigned r = __builtin_bit_cast(unsigned, b->xs[(size_t)w * 64]);
  b->wbar[w]->arrive_and_wait();
  return r;
}

inline uint32_t lo_halves(uint32_t e0, uint32_t e1) { return (e0 & 0xffffu) | (e1 << 16); }
inline uint32_t hi_halves(uint32_t e0, uint32_t e1) { return (e0 >> 16) | (e1 & 0xffff0000u); }

template <class Dec>
inline f32x16 emu_mfma(u32x4 a, u32x4 b, f32x16 c, Dec dec) {
  emu::Block* blk = emu::tl_blk;
  int l = emu::tl_tid % 64, w = emu::tl_tid / 64;
  float* xa = &blk->xa[(size_t)w * 512];
  float* xb = &blk->xb[(size_t)w * 512];
  for (int t = 0; t < 8; ++t) {
    uint32_t wa = a[t / 2], wbv = b[t / 2];
    xa[l * 8 + t] = dec((t & 1) ? (wa >> 16) : (wa & 0xffff));
    xb[l * 8 + t] = dec((t & 1) ? (wbv >> 16) : (wbv & 0xffff));
  }
  blk->wbar[w]->arrive_and_wait();
  f32x16 d;
  int h = l / 32, col = l % 32;
  for (int v = 0; v < 16; ++v) {
    int row = 8 * (v / 4) + 4 * h + v % 4;
    float s = c[v];
    for (int k = 0; k < 16; ++k) s += xa[(row + 32 * (k / 8)) * 8 + k % 8] * xb[(col + 32 * (k / 8)) * 8 + k % 8];
    d[v] = s;
  }
  blk->wbar[w]->arrive_and_wait();
  return d;
}
inline f32x16 mfma_f16(u32x4 a, u32x4 b, f32x16 c) {
  return emu_mfma(a, b, c, [](uint32_t x) { return (float)__builtin_bit_cast(_Float16, (uint16_t)x); });
}
inline f32x16 mfma_bf16(u32x4 a, u32x4 b, f32x16 c) {
  return emu_mfma(a, b, c, [](uint32_t x) { return (float)__builtin_bit_cast(__bf16, (uint16_t)x); });
}
}  // namespace mcfft

// layout_pass.h -- gather / scatter between a caller's strided, interleaved or
// planar data layout and the contiguous interleaved layout the FFT passes use.
//
// Caller side: element (b, i0, i1, i2) of transform b (dimensions padded to
// three at the front, innermost last) sits at b*dist + i0*s0 + i1*s1 + i2*s2,
// counted in elements of the side's form:
//   interleaved: one array of complex elements (FP16/BF16 4 B, split 8 B);
//   planar: a real-part array and an imaginary-part array of scalars
//     (FP16/BF16 2 B; split: binary16 (hi, lo), 4 B).
// Contiguous side: element e = b*n + row-major index, interleaved.  One thread
// per element, the contiguous side coalesced; pure data movement (HBM-bound).
//
// No reference counterpart: the reference takes contiguous interleaved data
// only (source/mcfft_1d_utils.cpp:37-66, source/mcfft_2d_utils.cpp:49-50).
#pragma once
#include "fft_pass.h"
#include "plan_core.h"

namespace mcfft {

struct LayoutArgs {
  const void* src[2];    // gather: caller's re (or interleaved) and im arrays; scatter: [0] contiguous
  void* dst[2];          // scatter: caller's re (or interleaved) and im arrays; gather: [0] contiguous
  long long strides[3];  // caller's element strides of dimensions 0..2 (innermost last)
  long long dist;        // caller's element distance between transforms
  long long total;       // elements of all transforms
  int lg[3];             // log2 of the (padded) dimensions
  int planar;
  int split;             // 1: 8-byte split elements / 4-byte planar scalars
  int vec;               // 1: four consecutive elements per thread (layout_vec_ok)
};

MCFFT_DEV long long layout_offset(const LayoutArgs& a, long long e) {
  const int lgn = a.lg[0] + a.lg[1] + a.lg[2];
  const long long b = e >> lgn;
  const long long i2 = e & ((1ll << a.lg[2]) - 1);
  const long long i1 = (e >> a.lg[2]) & ((1ll << a.lg[1]) - 1);
  const long long i0 = (e >> (a.lg[1] + a.lg[2])) & ((1ll << a.lg[0]) - 1);
  return b * a.dist + i0 * a.strides[0] + i1 * a.strides[1] + i2 * a.strides[2];
}

// FP16/BF16 element: re in bits 0-15, im in bits 16-31.  Split element
// (u32x2): .x = (re_hi, im_hi), .y = (re_lo, im_lo); a planar split scalar is
// hi in bits 0-15, lo in bits 16-31.
MCFFT_DEV void layout_gather_one(const LayoutArgs& a, long long e) {
  const long long s = layout_offset(a, e);
  if (!a.split) {
    uint32_t v;
    if (!a.planar) {
      v = g_ld((const uint32_t*)a.src[0], s);
    } else {
      const uint32_t re = g_ld((const uint16_t*)a.src[0], s), im = g_ld((const uint16_t*)a.src[1], s);
      v = re | (im << 16);
    }
    g_st((uint32_t*)a.dst[0], e, v);
  } else {
    u32x2 v;
    if (!a.planar) {
      v = g_ld((const u32x2*)a.src[0], s);
    } else {
      const uint32_t re = g_ld((const uint32_t*)a.src[0], s), im = g_ld((const uint32_t*)a.src[1], s);
      v.x = (re & 0xffffu) | (im << 16);
      v.y = (re >> 16) | (im & 0xffff0000u);
    }
    g_st((u32x2*)a.dst[0], e, v);
  }
}

MCFFT_DEV void layout_scatter_one(const LayoutArgs& a, long long e) {
  const long long d = layout_offset(a, e);
  if (!a.split) {
    const uint32_t v = g_ld((const uint32_t*)a.src[0], e);
    if (!a.planar) {
      g_st((uint32_t*)a.dst[0], d, v);
    } else {
      g_st((uint16_t*)a.dst[0], d, (uint16_t)(v & 0xffffu));
      g_st((uint16_t*)a.dst[1], d, (uint16_t)(v >> 16));
    }
  } else {
    const u32x2 v = g_ld((const u32x2*)a.src[0], e);
    if (!a.planar) {
      g_st((u32x2*)a.dst[0], d, v);
    } else {
      g_st((uint32_t*)a.dst[0], d, (v.x & 0xffffu) | (v.y << 16));
      g_st((uint32_t*)a.dst[1], d, (v.x >> 16) | (v.y & 0xffff0000u));
    }
  }
}

// Four consecutive elements e..e+3 (e a multiple of 4, so one innermost row:
// dimensions are >= 16): with an innermost stride of 1 they are consecutive on
// the caller's side too, and every access is a 8- to 32-byte vector.
MCFFT_DEV void layout_gather_four(const LayoutArgs& a, long long e) {
  const long long s = layout_offset(a, e);
  if (!a.split) {
    u32x4 v;
    if (!a.planar) {
      v = g_ld((const u32x4*)((const uint32_t*)a.src[0] + s), 0);
    } else {
      const u32x2 re = g_ld((const u32x2*)((const uint16_t*)a.src[0] + s), 0);
      const u32x2 im = g_ld((const u32x2*)((const uint16_t*)a.src[1] + s), 0);
      v.x = (re.x & 0xffffu) | (im.x << 16);
      v.y = (re.x >> 16) | (im.x & 0xffff0000u);
      v.z = (re.y & 0xffffu) | (im.y << 16);
      v.w = (re.y >> 16) | (im.y & 0xffff0000u);
    }
    g_st((u32x4*)((uint32_t*)a.dst[0] + e), 0, v);
  } else {
    u32x4 v0, v1;  // elements e, e+1 and e+2, e+3 (x, y of each)
    if (!a.planar) {
      v0 = g_ld((const u32x4*)((const u32x2*)a.src[0] + s), 0);
      v1 = g_ld((const u32x4*)((const u32x2*)a.src[0] + s + 2), 0);
    } else {
      const u32x4 re = g_ld((const u32x4*)((const uint32_t*)a.src[0] + s), 0);
      const u32x4 im = g_ld((const u32x4*)((const uint32_t*)a.src[1] + s), 0);
      v0.x = (re.x & 0xffffu) | (im.x << 16);
      v0.y = (re.x >> 16) | (im.x & 0xffff0000u);
      v0.z = (re.y & 0xffffu) | (im.y << 16);
      v0.w = (re.y >> 16) | (im.y & 0xffff0000u);
      v1.x = (re.z & 0xffffu) | (im.z << 16);
      v1.y = (re.z >> 16) | (im.z & 0xffff0000u);
      v1.z = (re.w & 0xffffu) | (im.w << 16);
      v1.w = (re.w >> 16) | (im.w & 0xffff0000u);
    }
    g_st((u32x4*)((u32x2*)a.dst[0] + e), 0, v0);
    g_st((u32x4*)((u32x2*)a.dst[0] + e + 2), 0, v1);
  }
}

MCFFT_DEV void layout_scatter_four(const LayoutArgs& a, long long e) {
  const long long d = layout_offset(a, e);
  if (!a.split) {
    const u32x4 v = g_ld((const u32x4*)((const uint32_t*)a.src[0] + e), 0);
    if (!a.planar) {
      g_st((u32x4*)((uint32_t*)a.dst[0] + d), 0, v);
    } else {
      u32x2 re, im;
      re.x = (v.x & 0xffffu) | (v.y << 16);
      re.y = (v.z & 0xffffu) | (v.w << 16);
      im.x = (v.x >> 16) | (v.y & 0xffff0000u);
      im.y = (v.z >> 16) | (v.w & 0xffff0000u);
      g_st((u32x2*)((uint16_t*)a.dst[0] + d), 0, re);
      g_st((u32x2*)((uint16_t*)a.dst[1] + d), 0, im);
    }
  } else {
    const u32x4 v0 = g_ld((const u32x4*)((const u32x2*)a.src[0] + e), 0);
    const u32x4 v1 = g_ld((const u32x4*)((const u32x2*)a.src[0] + e + 2), 0);
    if (!a.planar) {
      g_st((u32x4*)((u32x2*)a.dst[0] + d), 0, v0);
      g_st((u32x4*)((u32x2*)a.dst[0] + d + 2), 0, v1);
    } else {
      u32x4 re, im;  // element k = (x, y): scalar re = x.lo | y.lo << 16, im = x.hi | y.hi << 16
      re.x = (v0.x & 0xffffu) | (v0.y << 16);
      re.y = (v0.z & 0xffffu) | (v0.w << 16);
      re.z = (v1.x & 0xffffu) | (v1.y << 16);
      re.w = (v1.z & 0xffffu) | (v1.w << 16);
      im.x = (v0.x >> 16) | (v0.y & 0xffff0000u);
      im.y = (v0.z >> 16) | (v0.w & 0xffff0000u);
      im.z = (v1.x >> 16) | (v1.y & 0xffff0000u);
      im.w = (v1.z >> 16) | (v1.w & 0xffff0000u);
      g_st((u32x4*)((uint32_t*)a.dst[0] + d), 0, re);
      g_st((u32x4*)((uint32_t*)a.dst[1] + d), 0, im);
    }
  }
}

constexpr int kLayoutThreads = 256;

template <bool GATHER>
MCFFT_DEV void layout_body(const LayoutArgs& a) {
  const long long step = (long long)grid_size() * kLayoutThreads;
  const long long first = (long long)block_id() * kLayoutThreads + thread_id();
  if (a.vec) {
    for (long long q = first; q < a.total / 4; q += step) {
      if constexpr (GATHER) layout_gather_four(a, 4 * q);
      else layout_scatter_four(a, 4 * q);
    }
    return;
  }
  for (long long e = first; e < a.total; e += step) {
    if constexpr (GATHER) layout_gather_one(a, e);
    else layout_scatter_one(a, e);
  }
}

// ---- host side: validation and argument set-up, shared by the runtime and
// the CPU emulation ----

// Checks one side's layout against the transform shape: every stride and the
// distance positive, every dimension a power of two.  Returns 0 or ST_*.
inline int layout_validate(int rank, const long long* dims, long long batch, const long long* strides, long long dist,
                           int planar) {
  if (rank < 1 || rank > 3 || batch < 1 || (planar != 0 && planar != 1)) return ST_INVALID_VALUE;
  for (int d = 0; d < rank; ++d) {
    if (strides[d] < 1) return ST_INVALID_VALUE;
    if (dims[d] < 1 || (dims[d] & (dims[d] - 1))) return ST_UNSUPPORTED;
  }
  if (dist < 1) return ST_INVALID_VALUE;
  return ST_OK;
}

// True when the side is the contiguous interleaved layout (no staging pass).
// interleaved and dense inside each transform (the contiguous layout with any
// distance >= the transform's size between transforms)
inline bool layout_is_spaced(int rank, const long long* dims, const long long* strides, long long dist, int planar) {
  if (planar) return false;
  long long s = 1;
  for (int d = rank - 1; d >= 0; --d) {
    if (strides[d] != s) return false;
    s *= dims[d];
  }
  return dist >= s;
}

inline bool layout_is_dense(int rank, const long long* dims, const long long* strides, long long dist, int planar) {
  if (planar) return false;
  long long s = 1;
  for (int d = rank - 1; d >= 0; --d) {
    if (strides[d] != s) return false;
    s *= dims[d];
  }
  return dist == s;
}

// Elements the side spans: 1 + the largest offset any element has.
inline long long layout_extent(int rank, const long long* dims, long long batch, const long long* strides,
                               long long dist) {
  long long m = (batch - 1) * dist;
  for (int d = 0; d < rank; ++d) m += (dims[d] - 1) * strides[d];
  return m + 1;
}

// Four elements per thread when the side's innermost stride is 1, every group
// of four starts on a multiple of 4 elements (other strides and the distance)
// and the caller's arrays are aligned to four elements (scalars when planar).
inline bool layout_vec_ok(const LayoutArgs& a, const void* p0, const void* p1) {
  if (a.strides[2] != 1 || a.strides[0] % 4 || a.strides[1] % 4 || a.dist % 4) return false;
  const size_t scalar = (size_t)(a.split ? 8 : 4) / (a.planar ? 2 : 1);
  const size_t align = 4 * scalar < 16 ? 4 * scalar : 16;
  return (size_t)p0 % align == 0 && (!p1 || (size_t)p1 % align == 0);
}

inline void layout_fill(LayoutArgs& a, int rank, const long long* dims, long long batch, const long long* strides,
                        long long dist, int planar, int fmt) {
  long long n = 1;
  for (int d = 0; d < 3; ++d) {
    const int k = d - (3 - rank);  // padded dimension d is the caller's dimension k
    a.lg[d] = 0;
    a.strides[d] = 0;
    if (k >= 0) {
      while ((1ll << a.lg[d]) < dims[k]) ++a.lg[d];
      a.strides[d] = strides[k];
      n *= dims[k];
    }
  }
  a.dist = dist;
  a.total = n * batch;
  a.planar = planar;
  a.split = fmt == FMT_SPLIT ? 1 : 0;
  a.vec = 0;
}

}  // namespace mcfft

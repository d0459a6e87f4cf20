// mcfft_runtime.hip -- the C ABI (include/mcfft.h): plan objects, FP64 table
// generation on the GPU, and pass launches on the caller's stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/mcfft.h"
#include "fft_pass.h"
#include "plan_core.h"

using namespace mcfft;

struct mcfftPlan_st {
  PlanCore core;
  int device = 0;
  f32x2* tw[2] = {nullptr, nullptr};      // [0] forward, [1] inverse
  uint16_t* dft[2] = {nullptr, nullptr};
  u32x4* aops[2] = {nullptr, nullptr};    // per-lane DFT-matrix MFMA operands
  std::vector<int> grid_cap;              // resident workgroups per pass (occupancy x CUs)
  void* ws = nullptr;
  size_t ws_bytes = 0;
};

namespace {

// FP64 evaluation of exp(sign * 2 pi i * e / M) * scale, rounded to FP32.
// sincospi takes the exact rational argument 2e/M (M a power of two).
__global__ void gen_twiddles(const TwSpec* spec, long long n, int sign, f32x2* out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const TwSpec s = spec[i];
  double sn, cs;
  sincospi(2.0 * (double)s.e / (double)s.M, &sn, &cs);
  f32x2 v;
  v.x = (float)(cs * (double)s.scale);
  v.y = (float)((double)sign * sn * (double)s.scale);
  out[i] = v;
}

// W16^e, e = 0..15, as FP16 hi, FP16 lo (x 2^11) and BF16 (layout in fft_pass.h)
__global__ void gen_dft_table(int sign, uint16_t* t) {
  const int e = threadIdx.x;
  if (e >= 16) return;
  double sn, cs;
  sincospi(2.0 * e / 16.0, &sn, &cs);
  const double re = cs, im = sign * sn;
  const _Float16 rh = (_Float16)re, ih = (_Float16)im;
  t[e] = __builtin_bit_cast(uint16_t, rh);
  t[16 + e] = __builtin_bit_cast(uint16_t, ih);
  t[32 + e] = __builtin_bit_cast(uint16_t, (_Float16)((re - (double)rh) * 2048.0));
  t[48 + e] = __builtin_bit_cast(uint16_t, (_Float16)((im - (double)ih) * 2048.0));
  t[64 + e] = __builtin_bit_cast(uint16_t, (__bf16)(float)re);
  t[80 + e] = __builtin_bit_cast(uint16_t, (__bf16)(float)im);
}

// per-lane MFMA operand words of the DFT matrix (64 lanes x 16 words)
__global__ void gen_aops(const uint16_t* dft, int cm, uint32_t* out) {
  const int lane = threadIdx.x;
  if (lane >= 64) return;
  aops_words(dft, cm, lane, [&](int k, uint32_t w) { out[lane * 16 + k] = w; });
}

std::mutex g_attr_mu;
std::vector<const void*> g_attr_done;  // kernels whose LDS limit was raised

mcfftStatus set_lds_attr(const KernelInfo* k) {
  std::lock_guard<std::mutex> g(g_attr_mu);
  for (const void* f : g_attr_done)
    if (f == k->fn) return MCFFT_SUCCESS;
  if (hipFuncSetAttribute(k->fn, hipFuncAttributeMaxDynamicSharedMemorySize, k->lds_bytes) != hipSuccess)
    return MCFFT_HIP_ERROR;
  g_attr_done.push_back(k->fn);
  return MCFFT_SUCCESS;
}

void free_plan(mcfftPlan p) {
  if (!p) return;
  for (int d = 0; d < 2; ++d) {
    if (p->tw[d]) (void)hipFree(p->tw[d]);
    if (p->dft[d]) (void)hipFree(p->dft[d]);
    if (p->aops[d]) (void)hipFree(p->aops[d]);
  }
  if (p->ws) (void)hipFree(p->ws);
  delete p;
}

}  // namespace

extern "C" {

mcfftStatus mcfft_plan_many(mcfftPlan* plan, int rank, const long long* dims, long long batch, mcfftFormat in_fmt,
                            mcfftFormat out_fmt, mcfftScaling scaling) {
  if (!plan) return MCFFT_INVALID_VALUE;
  *plan = nullptr;
  mcfftPlan p = new (std::nothrow) mcfftPlan_st();
  if (!p) return MCFFT_ALLOC_FAILED;
  int st = plan_build(p->core, rank, dims, batch, (int)in_fmt, (int)out_fmt, (int)scaling, 0);
  if (st != ST_OK) {
    delete p;
    return (mcfftStatus)st;
  }
  if (hipGetDevice(&p->device) != hipSuccess) {
    delete p;
    return MCFFT_HIP_ERROR;
  }
  int n_cu = 0;
  if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, p->device) != hipSuccess || n_cu < 1) {
    delete p;
    return MCFFT_HIP_ERROR;
  }
  for (const PassPlan& pp : p->core.passes) {
    mcfftStatus s = set_lds_attr(pp.k);
    int per_cu = 0;
    if (s == MCFFT_SUCCESS &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pp.k->fn, pp.k->nw * 64, pp.k->lds_bytes) != hipSuccess)
      s = MCFFT_HIP_ERROR;
    if (s != MCFFT_SUCCESS) {
      delete p;
      return s;
    }
    p->grid_cap.push_back((per_cu < 1 ? 1 : per_cu) * n_cu);
  }
  // tables: FP64 on the GPU, then rounded to FP32 / FP16 / BF16
  const long long ntw = (long long)p->core.tw.size();
  TwSpec* dspec = nullptr;
  mcfftStatus rs = MCFFT_SUCCESS;
  if (ntw > 0 && hipMalloc(&dspec, sizeof(TwSpec) * ntw) != hipSuccess) rs = MCFFT_ALLOC_FAILED;
  for (int d = 0; d < 2 && rs == MCFFT_SUCCESS; ++d) {
    if (hipMalloc(&p->tw[d], sizeof(f32x2) * (ntw > 0 ? ntw : 1)) != hipSuccess) rs = MCFFT_ALLOC_FAILED;
    if (rs == MCFFT_SUCCESS && hipMalloc(&p->dft[d], sizeof(uint16_t) * kDftTableEntries) != hipSuccess)
      rs = MCFFT_ALLOC_FAILED;
    if (rs == MCFFT_SUCCESS && hipMalloc(&p->aops[d], sizeof(u32x4) * 64 * 4) != hipSuccess) rs = MCFFT_ALLOC_FAILED;
  }
  if (rs == MCFFT_SUCCESS && ntw > 0 &&
      hipMemcpy(dspec, p->core.tw.data(), sizeof(TwSpec) * ntw, hipMemcpyHostToDevice) != hipSuccess)
    rs = MCFFT_HIP_ERROR;
  for (int d = 0; d < 2 && rs == MCFFT_SUCCESS; ++d) {
    const int sign = d == 0 ? -1 : 1;
    if (ntw > 0) gen_twiddles<<<(unsigned)((ntw + 255) / 256), 256>>>(dspec, ntw, sign, p->tw[d]);
    gen_dft_table<<<1, 64>>>(sign, p->dft[d]);
    gen_aops<<<1, 64>>>(p->dft[d], p->core.cm, (uint32_t*)p->aops[d]);
    if (hipGetLastError() != hipSuccess) rs = MCFFT_HIP_ERROR;
  }
  if (rs == MCFFT_SUCCESS && hipDeviceSynchronize() != hipSuccess) rs = MCFFT_HIP_ERROR;
  if (dspec) (void)hipFree(dspec);
  if (rs == MCFFT_SUCCESS && p->core.ws_elems > 0) {
    const int ws_fmt = (p->core.cm == CM_SPLIT) ? FMT_SPLIT : p->core.in_fmt;
    p->ws_bytes = (size_t)p->core.ws_elems * (size_t)elem_bytes(ws_fmt);
    if (hipMalloc(&p->ws, p->ws_bytes) != hipSuccess) rs = MCFFT_ALLOC_FAILED;
  }
  if (rs != MCFFT_SUCCESS) {
    free_plan(p);
    return rs;
  }
  *plan = p;
  return MCFFT_SUCCESS;
}

mcfftStatus mcfft_plan_1d(mcfftPlan* plan, long long n, long long batch, mcfftFormat in_fmt, mcfftFormat out_fmt) {
  return mcfft_plan_many(plan, 1, &n, batch, in_fmt, out_fmt, MCFFT_SCALE_BALANCED);
}
mcfftStatus mcfft_plan_2d(mcfftPlan* plan, long long nx, long long ny, long long batch, mcfftFormat in_fmt,
                          mcfftFormat out_fmt) {
  const long long d[2] = {nx, ny};
  return mcfft_plan_many(plan, 2, d, batch, in_fmt, out_fmt, MCFFT_SCALE_BALANCED);
}
mcfftStatus mcfft_plan_3d(mcfftPlan* plan, long long nx, long long ny, long long nz, long long batch,
                          mcfftFormat in_fmt, mcfftFormat out_fmt) {
  const long long d[3] = {nx, ny, nz};
  return mcfft_plan_many(plan, 3, d, batch, in_fmt, out_fmt, MCFFT_SCALE_BALANCED);
}

mcfftStatus mcfft_exec(mcfftPlan plan, const void* in, void* out, int direction, void* stream) {
  if (!plan) return MCFFT_INVALID_PLAN;
  return mcfft_exec_passes(plan, in, out, direction, stream, 0, (int)plan->core.passes.size());
}

mcfftStatus mcfft_get_pass_count(mcfftPlan plan, int* count) {
  if (!plan) return MCFFT_INVALID_PLAN;
  if (!count) return MCFFT_INVALID_VALUE;
  *count = (int)plan->core.passes.size();
  return MCFFT_SUCCESS;
}

mcfftStatus mcfft_exec_passes(mcfftPlan plan, const void* in, void* out, int direction, void* stream, int first,
                              int count) {
  if (!plan) return MCFFT_INVALID_PLAN;
  if (first < 0 || count < 0 || first + count > (int)plan->core.passes.size()) return MCFFT_INVALID_VALUE;
  if (!in || !out) return MCFFT_INVALID_VALUE;
  if (direction != MCFFT_FORWARD && direction != MCFFT_INVERSE) return MCFFT_INVALID_VALUE;
  const PlanCore& c = plan->core;
  if (((size_t)in % (size_t)elem_bytes(c.in_fmt)) || ((size_t)out % (size_t)elem_bytes(c.out_fmt)))
    return MCFFT_MISALIGNED;
  if (c.ws_elems > 0 && !plan->ws) return MCFFT_INVALID_PLAN;
  const int d = direction == MCFFT_FORWARD ? 0 : 1;
  const int ws_fmt = (c.cm == CM_SPLIT) ? FMT_SPLIT : c.in_fmt;
  hipStream_t s = (hipStream_t)stream;
  for (long long ch = 0; ch < c.n_chunks; ++ch) {
    const long long t0 = ch * c.chunk;
    const long long nt = (t0 + c.chunk <= c.batch) ? c.chunk : c.batch - t0;
    for (int pi = first; pi < first + count; ++pi) {
      const PassPlan& pp = c.passes[pi];
      PassArgs a;
      memset(&a, 0, sizeof(a));
      auto bufptr = [&](int b, int fmt) -> char* {
        if (b == BUF_IN) return (char*)in + (size_t)(t0 * c.n) * (size_t)elem_bytes(fmt);
        if (b == BUF_OUT) return (char*)out + (size_t)(t0 * c.n) * (size_t)elem_bytes(fmt);
        return (char*)plan->ws;
      };
      const int dst_fmt = (pp.dst == BUF_OUT) ? c.out_fmt : ws_fmt;
      a.in = bufptr(pp.src, pp.src_fmt);
      a.out = bufptr(pp.dst, dst_fmt);
      a.stride_in = pp.stride_in;
      a.stride_out = pp.stride_out;
      a.in_sg = pp.in_sg;
      a.in_st = pp.in_st;
      a.out_sg = pp.out_sg;
      a.out_st = pp.out_st;
      a.n_inst = (int)pp.n_inst;
      a.tiles_per_group = pp.tiles_per_group;
      a.in_fmt = pp.src_fmt;
      auto view_elems = [&](int b) -> long long {
        return b == BUF_WS ? c.ws_elems : (c.batch - t0) * c.n;
      };
      a.in_elems = view_elems(pp.src);
      a.out_elems = view_elems(pp.dst);
      a.stage_tw_entries = (int)((long long)c.tw.size() - pp.stage_tw_off);
      a.inverse = d;
      a.out_scale = pp.out_scale;
      a.stage_tw = plan->tw[d] + pp.stage_tw_off;
      a.aops = plan->aops[d];
      if (pp.fourstep) {
        a.fs_hi = plan->tw[d] + pp.fs_hi_off;
        a.fs_lo = plan->tw[d] + pp.fs_lo_off;
        a.fs_shift = pp.fs_shift;
        a.fs_hi_entries = (int)(pp.fs_lo_off - pp.fs_hi_off);
        a.fs_lo_entries = 1 << pp.fs_shift;
        a.fs_mask = pp.fs_mask;
        a.fs_nmask = (unsigned)pp.fs_nmask;
      }
      const long long groups = pp.groups_per_transform > 0 ? nt * pp.groups_per_transform : pp.groups;
      const long long tiles = groups * (long long)pp.tiles_per_group;
      if (tiles <= 0 || tiles > 0x7fffffffLL) return MCFFT_UNSUPPORTED;
      a.n_tiles = (int)tiles;
      const long long grid = tiles < plan->grid_cap[pi] ? tiles : plan->grid_cap[pi];
      void* args[] = {&a};
      if (hipLaunchKernel(pp.k->fn, dim3((unsigned)grid), dim3((unsigned)(pp.k->nw * 64)), args,
                          (size_t)pp.k->lds_bytes, s) != hipSuccess)
        return MCFFT_HIP_ERROR;
    }
  }
  return MCFFT_SUCCESS;
}

mcfftStatus mcfft_get_output_exponent(mcfftPlan plan, int* exponent) {
  if (!plan) return MCFFT_INVALID_PLAN;
  if (!exponent) return MCFFT_INVALID_VALUE;
  *exponent = plan->core.out_exponent;
  return MCFFT_SUCCESS;
}

mcfftStatus mcfft_get_workspace_size(mcfftPlan plan, size_t* bytes) {
  if (!plan) return MCFFT_INVALID_PLAN;
  if (!bytes) return MCFFT_INVALID_VALUE;
  *bytes = plan->ws_bytes;
  return MCFFT_SUCCESS;
}

mcfftStatus mcfft_plan_describe(mcfftPlan plan, char* buf, size_t len) {
  if (!plan) return MCFFT_INVALID_PLAN;
  if (!buf || len == 0) return MCFFT_INVALID_VALUE;
  const std::string s = plan->core.describe();
  snprintf(buf, len, "%s", s.c_str());
  return MCFFT_SUCCESS;
}

mcfftStatus mcfft_destroy(mcfftPlan plan) {
  if (!plan) return MCFFT_INVALID_PLAN;
  free_plan(plan);
  return MCFFT_SUCCESS;
}

const char* mcfft_status_string(mcfftStatus s) {
  switch (s) {
    case MCFFT_SUCCESS: return "success";
    case MCFFT_INVALID_VALUE: return "invalid value";
    case MCFFT_UNSUPPORTED: return "unsupported";
    case MCFFT_ALLOC_FAILED: return "allocation failed";
    case MCFFT_HIP_ERROR: return "HIP error";
    case MCFFT_MISALIGNED: return "misaligned buffer";
    case MCFFT_INVALID_PLAN: return "invalid plan";
    case MCFFT_CACHE_ERROR: return "tuning cache error";
  }
  return "unknown status";
}

int mcfft_version(void) { return 100; }

}  // extern "C"

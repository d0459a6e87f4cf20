// mcfft_runtime.hip -- the C ABI (include/mcfft.h): plan objects, FP64 table
// generation on the GPU, and pass launches on the caller's stream.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <cmath>
#include <fstream>
#include <algorithm>
#include <map>
#include <sstream>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/mcfft.h"
#include "fft_pass.h"
#include "fft_big.h"
#include "plan_core.h"
#include "ragged.h"
#include "real_pass.h"
#include "layout_pass.h"
#include "ragged_plan.h"
#include "exec_args.h"

using namespace mcfft;

struct mcfftPlan_st {
  PlanCore core;
  int device = 0;
  std::string arch;
  f32x2* tw[2] = {nullptr, nullptr};      // [0] forward, [1] inverse
  uint16_t* dft[2] = {nullptr, nullptr};
  u32x4* aops[2] = {nullptr, nullptr};    // per-lane DFT-matrix MFMA operands
  uint32_t* dft256[2] = {nullptr, nullptr};  // row-scaled DFT matrices of the 65536-point kernel
  std::vector<int> grid_cap;              // resident workgroups per pass (occupancy x CUs)
  // PassArgs::stagger_* of the 65536-point kernel: half its workgroups start
  // 15 us late (4096 x 65536 FP16: 0.585 -> 0.561 ms); MCFFT_BIG_STAGGER=ticks[,phases]
  int stagger_ticks = 1500, stagger_phases = 2;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  // exec bookkeeping (bfp.h): status word and last-launch counter, per-transform
  // output exponents, and for AUTO multi-pass plans the exponent slots of every
  // transform and the first pass's tile exponents of one chunk
  ExecState* state = nullptr;
  int* out_exp = nullptr;
  unsigned long long* slots = nullptr;
  signed char* tile_e = nullptr;
  // ragged plans
  bool ragged = false;
  RaggedPlanCore rcore;
  RaggedPass* rpasses[2] = {nullptr, nullptr};
  RaggedTask* rtasks = nullptr;
  unsigned* rcounters = nullptr;
  int rgrid = 0;
  // runs of same-shape items large enough to fill the GPU on their own go to a
  // batched plan each (bulk); the rest to the persistent queue (has_queue)
  struct Bulk {
    mcfftPlan_st* plan;
    long long in_off, out_off;  // elements
  };
  std::vector<Bulk> bulk;
  bool has_queue = false;
  // a gathered batch (a ragged plan's bulk run of same-shape items wherever they
  // lie): per-transform element offsets into the caller's input [0] and output
  // [1] buffers (device), and the element counts addressable from those buffers
  long long* gather[2] = {nullptr, nullptr};
  long long gather_ext[2] = {0, 0};
  std::vector<int> item_exp;    // per item, from whichever path runs it
  // real-input / real-output plans: `core` is the length-H complex plan
  int real_kind = 0;          // 0 complex, 1 R2C, 2 C2R
  long long real_h = 0;       // H = N / 2 (N: the last dimension)
  long long real_rows = 1;    // rows of the last dimension per transform
  long long real_outer[2] = {1, 1};  // outer dimensions (2D: NX, 1; 3D: NX, NY)
  void* real_tmp = nullptr;   // H complex per transform (C2C output / input)
  f32x2* real_tw[2] = {nullptr, nullptr};  // W_N^k, k = 0..H, per direction
  int real_tmp_fmt = 0;
  int real_in_fmt = 0;        // C2R: storage format of the caller's spectrum
  int real_grid = 0;
  // advanced layouts (mcfft_plan_many_layout): caller side [0] input, [1] output
  bool layout = false;
  int lay_rank = 0;
  long long lay_dims[3] = {1, 1, 1};
  long long lay_strides[2][3] = {};
  long long lay_dist[2] = {0, 0};
  int lay_planar[2] = {0, 0};
  bool lay_dense[2] = {true, true};
  void* lay_buf[2] = {nullptr, nullptr};  // contiguous staging per side (one shared buffer when both stage)
  size_t lay_bytes = 0;
  int lay_grid = 0;
};

namespace mcfft {
extern const void* const kR2CPost[3];
extern const void* const kC2RPre[3];
extern const void* const kLayoutGather;
extern const void* const kLayoutScatter;
}  // namespace mcfft

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

// the 65536-point kernel's row-scaled DFT matrices (fft_big.h): FP64, rounded
// once to the plan's format as the 16-point table is (r2 = 0 gives its values)
__global__ void gen_dft256(int sign, int cm, uint32_t* t) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= kDft256Words) return;
  const Dft256Word d = dft256_word(w);
  uint32_t word = 0;
  for (int j = 0; j < 2; ++j) {
    double sn, cs;
    sincospi(2.0 * dft256_exponent(d.r2, d.i, d.k + j) / 256.0, &sn, &cs);
    const double v = d.b == 0 ? cs : d.b == 1 ? sign * sn : -(sign * sn);
    const uint32_t bits = cm == CM_BF16 ? __builtin_bit_cast(uint16_t, (__bf16)(float)v)
                                        : __builtin_bit_cast(uint16_t, (_Float16)v);
    word |= bits << (16 * j);
  }
  t[w] = word;
}

// per-lane MFMA operand words of the DFT matrix (64 lanes x 16 words)
__global__ void gen_aops(const uint16_t* dft, int cm, uint32_t* out) {
  const int lane = threadIdx.x;
  if (lane >= 64) return;
  aops_words(dft, cm, lane, [&](int k, uint32_t w) { out[lane * 16 + k] = w; });
}

// zero the ragged queue head, status and transform counters (a kernel, so the
// exec stays capturable into a hipGraph without memset nodes)
__global__ void reset_counters(unsigned* c, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) c[i] = 0u;
}

// FP64 error norms: per-block partial sums of |d|^2, |ref|^2 and max |d|
__global__ void error_norms_kernel(const void* y, int fmt, double scale, const double* ref, long long n,
                                   double* part) {
  __shared__ double s_num[256], s_den[256], s_max[256];
  double num = 0, den = 0, mx = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    double yr, yi;
    if (fmt == FMT_SPLIT) {
      const uint2 e = ((const uint2*)y)[i];
      yr = (double)f16_val(e.x & 0xffffu) + (double)f16_val(e.y & 0xffffu) / 2048.0;
      yi = (double)f16_val(e.x >> 16) + (double)f16_val(e.y >> 16) / 2048.0;
    } else {
      const uint32_t e = ((const uint32_t*)y)[i];
      yr = fmt == FMT_BF16 ? (double)bf16_val(e & 0xffffu) : (double)f16_val(e & 0xffffu);
      yi = fmt == FMT_BF16 ? (double)bf16_val(e >> 16) : (double)f16_val(e >> 16);
    }
    const double dr = yr * scale - ref[2 * i], di = yi * scale - ref[2 * i + 1];
    num += dr * dr + di * di;
    den += ref[2 * i] * ref[2 * i] + ref[2 * i + 1] * ref[2 * i + 1];
    mx = fmax(mx, fmax(fabs(dr), fabs(di)));
    if (dr != dr || di != di) mx = dr;  // NaN propagates
  }
  s_num[threadIdx.x] = num;
  s_den[threadIdx.x] = den;
  s_max[threadIdx.x] = mx;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s_num[threadIdx.x] += s_num[threadIdx.x + w];
      s_den[threadIdx.x] += s_den[threadIdx.x + w];
      const double o = s_max[threadIdx.x + w];
      if (!(o <= s_max[threadIdx.x])) s_max[threadIdx.x] = o;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = s_num[0];
    part[3 * blockIdx.x + 1] = s_den[0];
    part[3 * blockIdx.x + 2] = s_max[0];
  }
}

// Error-norm reduction grid: kNormBlocks blocks of 256 threads, each writing one
// (sum |d|^2, sum |ref|^2, max |d|) triple of partials.
constexpr int kNormBlocks = 1024;
// largest element count the error-norm entry points accept (far beyond any
// buffer a GPU holds; the partials do not grow with n)
constexpr long long kErrorNormsMaxElems = 1LL << 40;

// Partials of mcfft_error_norms_noalloc: device memory owned by the library for
// the whole process (one copy per GPU), so that entry point allocates nothing.
// Every block of the reduction writes its own triple before the host reads
// them, so nothing needs clearing between calls.  Calls that overlap in time on
// one GPU share this array and race.
static __device__ __align__(256) double scratch[3 * kNormBlocks];

std::mutex g_scratch_mu;
std::map<int, double*> g_scratch_ptr;  // device -> address of `scratch` on that device

// Address of `scratch` on the current device (looked up once per device; the
// symbol's host-side name is not a device address).
mcfftStatus scratch_on_current_device(double** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return MCFFT_HIP_ERROR;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it == g_scratch_ptr.end()) {
    double* p = nullptr;
    if (hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MCFFT_HIP_ERROR;
    it = g_scratch_ptr.emplace(dev, p).first;
  }
  *out = it->second;
  return MCFFT_SUCCESS;
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

// ---------------- tuning cache ----------------
// Text format and build check: tune_cache_parse / tune_cache_format (plan_core).
// MCFFT_BUILD_ID identifies the kernel sources this library was built from
// (a hash the Makefile computes); a cache written by another build is ignored
// and replaced on the next tune.
#ifndef MCFFT_BUILD_ID
#define MCFFT_BUILD_ID "unversioned"
#endif
std::mutex g_cache_mu;
std::map<std::string, std::string> g_cache;
bool g_cache_loaded = false;

std::string cache_path(const std::string& arch) {
  const char* env = getenv("MCFFT_TUNE_CACHE");
  if (env && std::string(env) == "off") return "";
  if (env && *env) return env;
  const char* home = getenv("HOME");
  if (!home || !*home) return "";
  return std::string(home) + "/.cache/mcfft/tune-" + arch + ".txt";
}

void load_cache_locked(const std::string& arch) {
  if (g_cache_loaded) return;
  g_cache_loaded = true;
  const std::string path = cache_path(arch);
  if (path.empty()) return;
  std::ifstream f(path);
  if (!f) return;
  std::stringstream text;
  text << f.rdbuf();
  g_cache = tune_cache_parse(text.str(), MCFFT_BUILD_ID);
}

bool save_cache_locked(const std::string& arch) {
  const std::string path = cache_path(arch);
  if (path.empty()) return true;
  const size_t slash = path.rfind('/');
  if (slash != std::string::npos) {
    std::string dir = path.substr(0, slash), cur;
    std::stringstream ss(dir);
    std::string part;
    while (std::getline(ss, part, '/')) {
      cur += part + "/";
      (void)mkdir(cur.c_str(), 0755);
    }
  }
  std::ofstream f(path + ".tmp");
  if (!f) return false;
  f << tune_cache_format(g_cache, MCFFT_BUILD_ID);
  f.close();
  return rename((path + ".tmp").c_str(), path.c_str()) == 0;
}

std::string arch_of(int device) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return "unknown";
  std::string a = prop.gcnArchName;
  const size_t c = a.find(':');
  return c == std::string::npos ? a : a.substr(0, c);
}

// the cached tuning choice for a plan of this signature, if any
bool cached_choice(const std::string& arch, const PlanCore& core, TuneChoice& c) {
  std::lock_guard<std::mutex> g(g_cache_mu);
  load_cache_locked(arch);
  const auto it = g_cache.find(arch + "|" + plan_signature(core));
  return it != g_cache.end() && parse_tune_choice(it->second, c);
}

// run pass i on kernel names[i] where that kernel can run it (silently keeps
// the default where a name no longer matches the compiled kernels)
void apply_kernels(mcfftPlan p, const std::vector<std::string>& names) {
  if (names.size() != p->core.passes.size()) return;
  for (size_t i = 0; i < names.size(); ++i) {
    PassPlan& pp = p->core.passes[i];
    for (const KernelInfo* k : pass_candidates(pp, p->core.cm, p->core.scaling))
      if (names[i] == k->name) {
        PassPlan trial = pp;
        if (set_pass_kernel(trial, k)) pp = trial;
      }
  }
}

// MCFFT_FORCE_KERNEL=<kernel name>: every pass of a new plan that can run that
// kernel runs it (pins a variant for testing and benchmarking; overrides the
// tuning cache)
void apply_forced(mcfftPlan p) {
  const char* env = getenv("MCFFT_FORCE_KERNEL");
  if (!env || !*env) return;
  for (PassPlan& pp : p->core.passes)
    for (const KernelInfo* k : pass_candidates(pp, p->core.cm, p->core.scaling))
      if (std::string(env) == k->name) {
        PassPlan trial = pp;
        if (set_pass_kernel(trial, k)) pp = trial;
      }
}

void free_plan(mcfftPlan p) {
  if (!p) return;
  for (int d = 0; d < 2; ++d) {
    if (p->tw[d]) (void)hipFree(p->tw[d]);
    if (p->dft[d]) (void)hipFree(p->dft[d]);
    if (p->aops[d]) (void)hipFree(p->aops[d]);
    if (p->dft256[d]) (void)hipFree(p->dft256[d]);
  }
  if (p->ws) (void)hipFree(p->ws);
  if (p->state) (void)hipFree(p->state);
  if (p->out_exp) (void)hipFree(p->out_exp);
  if (p->slots) (void)hipFree(p->slots);
  if (p->tile_e) (void)hipFree(p->tile_e);
  for (int d = 0; d < 2; ++d)
    if (p->rpasses[d]) (void)hipFree(p->rpasses[d]);
  if (p->rtasks) (void)hipFree(p->rtasks);
  if (p->rcounters) (void)hipFree(p->rcounters);
  for (auto& b : p->bulk) free_plan(b.plan);
  if (p->real_tmp) (void)hipFree(p->real_tmp);
  for (int d = 0; d < 2; ++d)
    if (p->real_tw[d]) (void)hipFree(p->real_tw[d]);
  for (int d = 0; d < 2; ++d)
    if (p->gather[d]) (void)hipFree(p->gather[d]);
  if (p->lay_buf[0]) (void)hipFree(p->lay_buf[0]);
  if (p->lay_buf[1] && p->lay_buf[1] != p->lay_buf[0]) (void)hipFree(p->lay_buf[1]);
  delete p;
}

// FP64 twiddle tables, DFT tables and MFMA operand tables of a plan (both directions)
mcfftStatus make_tables(mcfftPlan p, const std::vector<TwSpec>& tw, int cm) {
  const long long ntw = (long long)tw.size();
  TwSpec* dspec = nullptr;
  mcfftStatus rs = MCFFT_SUCCESS;
  if (ntw > 0 && hipMalloc(&dspec, sizeof(TwSpec) * ntw) != hipSuccess) rs = MCFFT_ALLOC_FAILED;
  for (int d = 0; d < 2 && rs == MCFFT_SUCCESS; ++d) {
    if (hipMalloc(&p->tw[d], sizeof(f32x2) * (ntw > 0 ? ntw : 1)) != hipSuccess) rs = MCFFT_ALLOC_FAILED;
    if (rs == MCFFT_SUCCESS && hipMalloc(&p->dft[d], sizeof(uint16_t) * kDftTableEntries) != hipSuccess)
      rs = MCFFT_ALLOC_FAILED;
    if (rs == MCFFT_SUCCESS && hipMalloc(&p->aops[d], sizeof(u32x4) * 64 * 4) != hipSuccess) rs = MCFFT_ALLOC_FAILED;
    if (rs == MCFFT_SUCCESS && hipMalloc(&p->dft256[d], sizeof(uint32_t) * kDft256Words) != hipSuccess)
      rs = MCFFT_ALLOC_FAILED;
  }
  if (rs == MCFFT_SUCCESS && ntw > 0 &&
      hipMemcpy(dspec, tw.data(), sizeof(TwSpec) * ntw, hipMemcpyHostToDevice) != hipSuccess)
    rs = MCFFT_HIP_ERROR;
  for (int d = 0; d < 2 && rs == MCFFT_SUCCESS; ++d) {
    const int sign = d == 0 ? -1 : 1;
    if (ntw > 0) gen_twiddles<<<(unsigned)((ntw + 255) / 256), 256>>>(dspec, ntw, sign, p->tw[d]);
    gen_dft_table<<<1, 64>>>(sign, p->dft[d]);
    gen_aops<<<1, 64>>>(p->dft[d], cm, (uint32_t*)p->aops[d]);
    gen_dft256<<<(kDft256Words + 255) / 256, 256>>>(sign, cm, p->dft256[d]);
    if (hipGetLastError() != hipSuccess) rs = MCFFT_HIP_ERROR;
  }
  if (rs == MCFFT_SUCCESS && hipDeviceSynchronize() != hipSuccess) rs = MCFFT_HIP_ERROR;
  if (dspec) (void)hipFree(dspec);
  return rs;
}

// Exec bookkeeping buffers of a regular plan (bfp.h), all zeroed: the status
// word, per-transform output exponents, and for AUTO multi-pass plans two
// exponent slots per transform plus one byte per first-pass instance of a chunk
// (at least as many as the first pass's tiles, whatever tile width it runs).
mcfftStatus alloc_exec_state(mcfftPlan p) {
  const PlanCore& c = p->core;
  const size_t n_slots = bfp_slot_count(c), tile_bytes = bfp_tile_bytes(c);
  if (hipMalloc(&p->state, sizeof(ExecState)) != hipSuccess ||
      hipMalloc(&p->out_exp, sizeof(int) * (size_t)c.batch) != hipSuccess ||
      (n_slots && hipMalloc(&p->slots, sizeof(unsigned long long) * n_slots) != hipSuccess) ||
      (tile_bytes && hipMalloc(&p->tile_e, tile_bytes) != hipSuccess))
    return MCFFT_ALLOC_FAILED;
  if (hipMemset(p->state, 0, sizeof(ExecState)) != hipSuccess ||
      hipMemset(p->out_exp, 0, sizeof(int) * (size_t)c.batch) != hipSuccess ||
      (n_slots && hipMemset(p->slots, 0, sizeof(unsigned long long) * n_slots) != hipSuccess) ||
      hipDeviceSynchronize() != hipSuccess)
    return MCFFT_HIP_ERROR;
  return MCFFT_SUCCESS;
}

// A regular plan: structure `opts` (null: the tuning cache's choice for this
// signature, else the planner's defaults), the cache's kernel choice when its
// structure matches, device tables, workspace and exec bookkeeping.
mcfftStatus create_plan(mcfftPlan* plan, int rank, const long long* dims, long long batch, mcfftFormat in_fmt,
                        mcfftFormat out_fmt, mcfftScaling scaling, const PlanOpts* opts) {
  mcfftPlan p = new (std::nothrow) mcfftPlan_st();
  if (!p) return MCFFT_ALLOC_FAILED;
  // MCFFT_WS_BYTES overrides the four-step workspace bound (bytes)
  long long ws_limit = 0;
  if (const char* env = getenv("MCFFT_WS_BYTES")) ws_limit = atoll(env);
  int st = plan_build(p->core, rank, dims, batch, (int)in_fmt, (int)out_fmt, (int)scaling, ws_limit,
                      opts ? *opts : PlanOpts());
  if (st == ST_OK && hipGetDevice(&p->device) != hipSuccess) st = ST_HIP_ERROR;
  int n_cu = 0;
  if (st == ST_OK &&
      (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, p->device) != hipSuccess || n_cu < 1))
    st = ST_HIP_ERROR;
  if (st != ST_OK) {
    delete p;
    return (mcfftStatus)st;
  }
  p->arch = arch_of(p->device);
  TuneChoice tc;
  if (!opts && cached_choice(p->arch, p->core, tc)) {
    const PlanOpts have = p->core.opts;
    if (tc.opts.outer_radix != have.outer_radix || tc.opts.fourstep_lg2 != have.fourstep_lg2) {
      PlanCore alt;
      if (plan_build(alt, rank, dims, batch, (int)in_fmt, (int)out_fmt, (int)scaling, ws_limit, tc.opts) == ST_OK)
        p->core = std::move(alt);
    }
    apply_kernels(p, tc.kernels);
  }
  apply_forced(p);
  if (const char* env = getenv("MCFFT_BIG_STAGGER")) {  // ticks of the 100 MHz clock[,phases]; 0: none
    const char* c = strchr(env, ',');
    p->stagger_ticks = std::min(std::max(atoi(env), 0), 100000);  // at most 1 ms
    p->stagger_phases = std::min(std::max(c ? atoi(c + 1) : 2, 1), 16);
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
  mcfftStatus rs = make_tables(p, p->core.tw, p->core.cm);
  if (rs == MCFFT_SUCCESS && p->core.ws_elems > 0) {
    const int ws_fmt = (p->core.cm == CM_SPLIT) ? FMT_SPLIT : p->core.in_fmt;
    p->ws_bytes = (size_t)p->core.ws_elems * (size_t)elem_bytes(ws_fmt);
    if (hipMalloc(&p->ws, p->ws_bytes) != hipSuccess) rs = MCFFT_ALLOC_FAILED;
  }
  if (rs == MCFFT_SUCCESS) rs = alloc_exec_state(p);
  if (rs != MCFFT_SUCCESS) {
    free_plan(p);
    return rs;
  }
  *plan = p;
  return MCFFT_SUCCESS;
}

}  // namespace

extern "C" {

mcfftStatus mcfft_plan_many(mcfftPlan* plan, int rank, const long long* dims, long long batch, mcfftFormat in_fmt,
                            mcfftFormat out_fmt, mcfftScaling scaling) {
  if (!plan) return MCFFT_INVALID_VALUE;
  *plan = nullptr;
  return create_plan(plan, rank, dims, batch, in_fmt, out_fmt, scaling, nullptr);
}

// Real-input forward / real-output inverse plans (see include/mcfft.h).
static mcfftStatus plan_real(mcfftPlan* plan, int kind, int rank, const long long* dims, long long batch,
                             mcfftFormat in_fmt, mcfftFormat out_fmt, mcfftScaling scaling) {
  if (!plan) return MCFFT_INVALID_VALUE;
  *plan = nullptr;
  if (rank < 1 || rank > 3 || !dims || batch < 1) return MCFFT_INVALID_VALUE;
  long long rows = 1;
  for (int d = 0; d < rank; ++d) {
    if (dims[d] < 1) return MCFFT_INVALID_VALUE;
    if (d < rank - 1) rows *= dims[d];
  }
  const long long n = dims[rank - 1];
  if (n < 32 || (n & (n - 1))) return MCFFT_UNSUPPORTED;
  const long long h = n / 2;
  // the passes address a transform's rows x (H + 1) elements with 32-bit offsets
  if (rows * (h + 1) * 8 > (1LL << 32)) return MCFFT_UNSUPPORTED;
  const int cm = compute_mode((int)in_fmt, (int)out_fmt);
  if (cm < 0) return MCFFT_UNSUPPORTED;
  // C2R keeps its pre-pass output in the complex plan's input format
  const int tmp_fmt = kind == 1 ? (int)out_fmt : (cm == CM_SPLIT ? FMT_SPLIT : (int)in_fmt);
  long long cdims[3] = {dims[0], rank > 1 ? dims[1] : 1, rank > 2 ? dims[2] : 1};
  cdims[rank - 1] = h;
  mcfftPlan p = nullptr;
  mcfftStatus st =
      mcfft_plan_many(&p, rank, cdims, batch, kind == 1 ? in_fmt : (mcfftFormat)tmp_fmt, out_fmt, scaling);
  if (st != MCFFT_SUCCESS) return st;
  p->real_kind = kind;
  p->real_h = h;
  p->real_rows = rows;
  p->real_outer[0] = rank > 1 ? dims[0] : 1;
  p->real_outer[1] = rank > 2 ? dims[1] : 1;
  p->real_tmp_fmt = tmp_fmt;
  p->real_in_fmt = (int)in_fmt;
  int n_cu = 0;
  if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, p->device) != hipSuccess) st = MCFFT_HIP_ERROR;
  p->real_grid = 8 * (n_cu > 0 ? n_cu : 1);
  if (st == MCFFT_SUCCESS &&
      hipMalloc(&p->real_tmp, (size_t)(rows * h * batch) * (size_t)elem_bytes(tmp_fmt)) != hipSuccess)
    st = MCFFT_ALLOC_FAILED;
  // W_N^k, k = 0..H, generated in FP64 (forward and inverse signs)
  TwSpec* dspec = nullptr;
  if (st == MCFFT_SUCCESS) {
    std::vector<TwSpec> spec((size_t)h + 1);
    for (long long k = 0; k <= h; ++k) spec[(size_t)k] = TwSpec{k, n, 1.0f};
    if (hipMalloc(&dspec, sizeof(TwSpec) * spec.size()) != hipSuccess ||
        hipMemcpy(dspec, spec.data(), sizeof(TwSpec) * spec.size(), hipMemcpyHostToDevice) != hipSuccess)
      st = MCFFT_HIP_ERROR;
    for (int d = 0; d < 2 && st == MCFFT_SUCCESS; ++d) {
      if (hipMalloc(&p->real_tw[d], sizeof(f32x2) * spec.size()) != hipSuccess) {
        st = MCFFT_ALLOC_FAILED;
        break;
      }
      gen_twiddles<<<(unsigned)((spec.size() + 255) / 256), 256>>>(dspec, (long long)spec.size(), d == 0 ? -1 : 1,
                                                                   p->real_tw[d]);
      if (hipGetLastError() != hipSuccess) st = MCFFT_HIP_ERROR;
    }
    if (st == MCFFT_SUCCESS && hipDeviceSynchronize() != hipSuccess) st = MCFFT_HIP_ERROR;
  }
  if (dspec) (void)hipFree(dspec);
  if (st != MCFFT_SUCCESS) {
    free_plan(p);
    return st;
  }
  *plan = p;
  return MCFFT_SUCCESS;
}

mcfftStatus mcfft_plan_r2c_1d(mcfftPlan* plan, long long n, long long batch, mcfftFormat in_fmt, mcfftFormat out_fmt,
                              mcfftScaling scaling) {
  return plan_real(plan, 1, 1, &n, batch, in_fmt, out_fmt, scaling);
}

mcfftStatus mcfft_plan_c2r_1d(mcfftPlan* plan, long long n, long long batch, mcfftFormat in_fmt, mcfftFormat out_fmt,
                              mcfftScaling scaling) {
  return plan_real(plan, 2, 1, &n, batch, in_fmt, out_fmt, scaling);
}

mcfftStatus mcfft_plan_r2c(mcfftPlan* plan, int rank, const long long* dims, long long batch, mcfftFormat in_fmt,
                           mcfftFormat out_fmt, mcfftScaling scaling) {
  return plan_real(plan, 1, rank, dims, batch, in_fmt, out_fmt, scaling);
}

mcfftStatus mcfft_plan_c2r(mcfftPlan* plan, int rank, const long long* dims, long long batch, mcfftFormat in_fmt,
                           mcfftFormat out_fmt, mcfftScaling scaling) {
  return plan_real(plan, 2, rank, dims, batch, in_fmt, out_fmt, scaling);
}

mcfftStatus mcfft_plan_many_layout(mcfftPlan* plan, int rank, const long long* dims, long long batch,
                                   const mcfftLayout* in_layout, const mcfftLayout* out_layout, mcfftFormat in_fmt,
                                   mcfftFormat out_fmt, mcfftScaling scaling) {
  if (!plan || !dims) return MCFFT_INVALID_VALUE;
  *plan = nullptr;
  if (rank < 1 || rank > 3 || batch < 1) return MCFFT_INVALID_VALUE;
  const mcfftLayout* lay[2] = {in_layout, out_layout};
  for (int sd = 0; sd < 2; ++sd)
    if (lay[sd]) {
      const int v = layout_validate(rank, dims, batch, lay[sd]->strides, lay[sd]->dist, lay[sd]->planar);
      if (v != ST_OK) return (mcfftStatus)v;
    }
  // Sides that only space whole transforms apart (interleaved, dense inside a
  // transform) need no staging: the plan runs as a gathered batch whose offset
  // tables are t * dist -- when its rows do not share tiles (2D/3D, or 1D plans
  // of more than one pass or with a one-row tile kernel, 2048 points and up;
  // shorter 1D rows keep staging).
  long long dist[2] = {0, 0};
  bool spaced = true;
  for (int sd = 0; sd < 2; ++sd) {
    long long s = 1;
    for (int d = rank - 1; d >= 0; --d) s *= dims[d];
    dist[sd] = s;
    if (lay[sd]) {
      spaced = spaced && layout_is_spaced(rank, dims, lay[sd]->strides, lay[sd]->dist, lay[sd]->planar);
      dist[sd] = lay[sd]->dist;
    }
  }
  if (spaced && rank == 1) {
    PlanCore probe;
    PlanOpts po;
    po.gathered = true;
    spaced = plan_build(probe, rank, dims, batch, (int)in_fmt, (int)out_fmt, (int)scaling, 0, po) == ST_OK &&
             (probe.passes.size() > 1 || probe.passes[0].k->Q == 1);
  }
  mcfftPlan p = nullptr;
  mcfftStatus st;
  if (spaced) {
    PlanOpts o;
    o.gathered = true;
    st = create_plan(&p, rank, dims, batch, in_fmt, out_fmt, scaling, &o);
    for (int sd = 0; sd < 2 && st == MCFFT_SUCCESS; ++sd) {
      std::vector<long long> off((size_t)batch);
      for (long long t = 0; t < batch; ++t) off[(size_t)t] = t * dist[sd];
      if (hipMalloc(&p->gather[sd], sizeof(long long) * (size_t)batch) != hipSuccess)
        st = MCFFT_ALLOC_FAILED;
      else if (hipMemcpy(p->gather[sd], off.data(), sizeof(long long) * (size_t)batch, hipMemcpyHostToDevice) !=
               hipSuccess)
        st = MCFFT_HIP_ERROR;
      p->gather_ext[sd] = (batch - 1) * dist[sd] + p->core.n;
    }
    if (st != MCFFT_SUCCESS && p) {
      free_plan(p);
      return st;
    }
  } else {
    st = mcfft_plan_many(&p, rank, dims, batch, in_fmt, out_fmt, scaling);
  }
  if (st != MCFFT_SUCCESS) return st;
  p->layout = true;
  p->lay_rank = rank;
  long long n = 1;
  for (int d = 0; d < rank; ++d) {
    p->lay_dims[d] = dims[d];
    n *= dims[d];
  }
  for (int sd = 0; sd < 2; ++sd) {
    if (lay[sd]) {
      for (int d = 0; d < rank; ++d) p->lay_strides[sd][d] = lay[sd]->strides[d];
      p->lay_dist[sd] = lay[sd]->dist;
      p->lay_planar[sd] = lay[sd]->planar;
    } else {
      long long s = 1;
      for (int d = rank - 1; d >= 0; --d) {
        p->lay_strides[sd][d] = s;
        s *= dims[d];
      }
      p->lay_dist[sd] = n;
    }
    // a gathered plan reads and writes the caller's spaced sides directly
    p->lay_dense[sd] =
        spaced || layout_is_dense(rank, dims, p->lay_strides[sd], p->lay_dist[sd], p->lay_planar[sd]);
  }
  // staging: one buffer serves both sides when both stage and the formats match
  // (the transform then runs in place on it)
  const size_t bytes[2] = {(size_t)(n * batch) * (size_t)elem_bytes(in_fmt),
                           (size_t)(n * batch) * (size_t)elem_bytes(out_fmt)};
  const bool share = !p->lay_dense[0] && !p->lay_dense[1] && in_fmt == out_fmt;
  for (int sd = 0; sd < 2 && st == MCFFT_SUCCESS; ++sd) {
    if (p->lay_dense[sd]) continue;
    if (sd == 1 && share) {
      p->lay_buf[1] = p->lay_buf[0];
      continue;
    }
    if (hipMalloc(&p->lay_buf[sd], bytes[sd]) != hipSuccess) st = MCFFT_ALLOC_FAILED;
    else p->lay_bytes += bytes[sd];
  }
  int n_cu = 0;
  if (st == MCFFT_SUCCESS &&
      (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, p->device) != hipSuccess || n_cu < 1))
    st = MCFFT_HIP_ERROR;
  p->lay_grid = 32 * (n_cu > 0 ? n_cu : 1);
  if (st != MCFFT_SUCCESS) {
    free_plan(p);
    return st;
  }
  *plan = p;
  return MCFFT_SUCCESS;
}

mcfftStatus mcfft_plan_1d(mcfftPlan* plan, long long n, long long batch, mcfftFormat in_fmt, mcfftFormat out_fmt) {
  return mcfft_plan_many(plan, 1, &n, batch, in_fmt, out_fmt, MCFFT_SCALE_AUTO);
}
mcfftStatus mcfft_plan_2d(mcfftPlan* plan, long long nx, long long ny, long long batch, mcfftFormat in_fmt,
                          mcfftFormat out_fmt) {
  const long long d[2] = {nx, ny};
  return mcfft_plan_many(plan, 2, d, batch, in_fmt, out_fmt, MCFFT_SCALE_AUTO);
}
mcfftStatus mcfft_plan_3d(mcfftPlan* plan, long long nx, long long ny, long long nz, long long batch,
                          mcfftFormat in_fmt, mcfftFormat out_fmt) {
  const long long d[3] = {nx, ny, nz};
  return mcfft_plan_many(plan, 3, d, batch, in_fmt, out_fmt, MCFFT_SCALE_AUTO);
}

mcfftStatus mcfft_plan_ragged(mcfftPlan* plan, int n_items, const mcfftRaggedItem* items, mcfftFormat in_fmt,
                              mcfftFormat out_fmt, mcfftScaling scaling) {
  if (!plan || !items || n_items < 1) return MCFFT_INVALID_VALUE;
  *plan = nullptr;
  const int cm = compute_mode((int)in_fmt, (int)out_fmt);
  if (cm < 0) return (int)in_fmt < 0 || (int)in_fmt > 2 || (int)out_fmt < 0 || (int)out_fmt > 2 ? MCFFT_INVALID_VALUE
                                                                                             : MCFFT_UNSUPPORTED;
  if ((int)scaling < 0 || (int)scaling > 3) return MCFFT_INVALID_VALUE;
  if (scaling == MCFFT_SCALE_AUTO) return MCFFT_UNSUPPORTED;  // ragged items are statically scaled
  mcfftPlan p = new (std::nothrow) mcfftPlan_st();
  if (!p) return MCFFT_ALLOC_FAILED;
  p->ragged = true;
  p->core.in_fmt = (int)in_fmt;
  p->core.out_fmt = (int)out_fmt;
  p->core.cm = cm;
  p->item_exp.assign((size_t)n_items, 0);
  // Bulk runs: consecutive items of one shape lying back to back in both buffers,
  // at least MCFFT_RAGGED_BULK elements together (default 2^23, enough to fill the
  // GPU; 0 = none), run as one batched plan each.  The other items (and any run
  // whose batched plan fails) go to the persistent queue.
  long long bulk_min = 1LL << 23;
  if (const char* e = getenv("MCFFT_RAGGED_BULK")) bulk_min = atoll(e);
  std::vector<RaggedItemSpec> spec;
  std::vector<int> rem_idx;
  auto same_shape = [&](const mcfftRaggedItem& a, const mcfftRaggedItem& b) {
    if (a.rank != b.rank) return false;
    for (int d = 0; d < a.rank; ++d)
      if (a.dims[d] != b.dims[d]) return false;
    return true;
  };
  for (int i = 0; i < n_items;) {
    const mcfftRaggedItem& a = items[i];
    long long n = 1;
    bool ok = a.rank >= 1 && a.rank <= 3 && a.in_offset >= 0 && a.out_offset >= 0;
    for (int d = 0; ok && d < a.rank; ++d) {
      ok = a.dims[d] >= 1;
      n *= a.dims[d];
    }
    int j = i + 1;
    if (ok)
      while (j < n_items && same_shape(items[j], a) && items[j].in_offset == items[j - 1].in_offset + n &&
             items[j].out_offset == items[j - 1].out_offset + n)
        ++j;
    if (ok && bulk_min > 0 && (long long)(j - i) * n >= bulk_min) {
      mcfftPlan sub = nullptr;
      const long long dims[3] = {a.dims[0], a.dims[1], a.dims[2]};
      int e = 0;
      if (mcfft_plan_many(&sub, a.rank, dims, j - i, in_fmt, out_fmt, scaling) == MCFFT_SUCCESS) {
        p->bulk.push_back({sub, a.in_offset, a.out_offset});
        if (mcfft_get_output_exponent(sub, &e) != MCFFT_SUCCESS) e = 0;
        for (int k = i; k < j; ++k) p->item_exp[(size_t)k] = e;
        i = j;
        continue;
      }
    }
    for (int k = i; k < j; ++k) {
      RaggedItemSpec it;
      it.rank = items[k].rank;
      for (int d = 0; d < 3; ++d) it.dims[d] = items[k].dims[d];
      it.in_off = items[k].in_offset;
      it.out_off = items[k].out_offset;
      spec.push_back(it);
      rem_idx.push_back(k);
    }
    i = j;
  }
  // Gathered bulk runs: the remaining items of one shape wherever they lie, at
  // least bulk_min elements or kRaggedGatherItems items together (enough
  // one-transform tiles to fill the GPU several times), run as one batched plan
  // whose transforms are located through offset tables -- 2D/3D items and 1D
  // items whose row pass has a one-row tile kernel (2048 points and up; shorter
  // rows share tiles, so they stay in the queue unless they lie back to back).
  // Items of a shape the queue's pass variants cannot run (1D above 2^18 points,
  // outer axes narrower than a variant's column tile) run this way whatever
  // their size.
  constexpr long long kRaggedGatherItems = 1024;
  if (!spec.empty()) {
    long long ext[2] = {0, 0};
    for (int i = 0; i < n_items; ++i) {
      long long n = 1;
      for (int d = 0; d < items[i].rank && d < 3; ++d) n *= items[i].dims[d];
      ext[0] = std::max(ext[0], items[i].in_offset + n);
      ext[1] = std::max(ext[1], items[i].out_offset + n);
    }
    std::map<std::vector<long long>, std::vector<size_t>> by_shape;
    for (size_t k = 0; k < spec.size(); ++k)
      by_shape[{spec[k].rank, spec[k].dims[0], spec[k].rank > 1 ? spec[k].dims[1] : 1,
                spec[k].rank > 2 ? spec[k].dims[2] : 1}].push_back(k);
    std::vector<bool> taken(spec.size(), false);
    for (const auto& kv : by_shape) {
      const RaggedItemSpec& a = spec[kv.second.front()];
      const long long cnt = (long long)kv.second.size();
      const long long n = kv.first[1] * kv.first[2] * kv.first[3];
      const bool queue_ok = ragged_shape_supported(a.rank, a.dims, (int)in_fmt, (int)out_fmt, (int)scaling) == ST_OK;
      const bool enough = bulk_min > 0 && (cnt * n >= bulk_min || cnt >= kRaggedGatherItems);
      if (queue_ok && ((a.rank == 1 && a.dims[0] < 2048) || !enough)) continue;
      PlanOpts o;
      o.gathered = true;
      const long long dims[3] = {a.dims[0], a.dims[1], a.dims[2]};
      mcfftPlan sub = nullptr;
      if (create_plan(&sub, a.rank, dims, cnt, in_fmt, out_fmt, scaling, &o) != MCFFT_SUCCESS) continue;
      if (queue_ok && sub->core.passes.size() == 1 && sub->core.passes[0].kind == KIND_ROW &&
          sub->core.passes[0].k->Q != 1) {  // rows would share tiles: the queue runs them
        free_plan(sub);
        continue;
      }
      std::vector<long long> off[2];
      for (size_t k : kv.second) {
        off[0].push_back(spec[k].in_off);
        off[1].push_back(spec[k].out_off);
      }
      bool ok = true;
      for (int d = 0; d < 2 && ok; ++d)
        ok = hipMalloc(&sub->gather[d], sizeof(long long) * (size_t)cnt) == hipSuccess &&
             hipMemcpy(sub->gather[d], off[d].data(), sizeof(long long) * (size_t)cnt, hipMemcpyHostToDevice) ==
                 hipSuccess;
      sub->gather_ext[0] = ext[0];
      sub->gather_ext[1] = ext[1];
      int e = 0;
      if (!ok || mcfft_get_output_exponent(sub, &e) != MCFFT_SUCCESS) {
        free_plan(sub);
        continue;
      }
      p->bulk.push_back({sub, 0, 0});
      for (size_t k : kv.second) {
        taken[k] = true;
        p->item_exp[(size_t)rem_idx[k]] = e;
      }
    }
    std::vector<RaggedItemSpec> spec2;
    std::vector<int> rem2;
    for (size_t k = 0; k < spec.size(); ++k)
      if (!taken[k]) {
        spec2.push_back(spec[k]);
        rem2.push_back(rem_idx[k]);
      }
    spec.swap(spec2);
    rem_idx.swap(rem2);
  }
  if (hipGetDevice(&p->device) != hipSuccess) {
    free_plan(p);
    return MCFFT_HIP_ERROR;
  }
  if (spec.empty()) {
    *plan = p;
    return MCFFT_SUCCESS;
  }
  p->has_queue = true;
  int st = ragged_build(p->rcore, (int)spec.size(), spec.data(), (int)in_fmt, (int)out_fmt, (int)scaling);
  if (st != ST_OK) {
    free_plan(p);
    return (mcfftStatus)st;
  }
  RaggedPlanCore& r = p->rcore;
  for (size_t k = 0; k < rem_idx.size(); ++k) p->item_exp[(size_t)rem_idx[k]] = r.exponent[k];
  const KernelInfo* k = &kRaggedKernels[r.cm];
  int n_cu = 0, per_cu = 0;
  mcfftStatus rs = MCFFT_SUCCESS;
  if (hipGetDevice(&p->device) != hipSuccess ||
      hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, p->device) != hipSuccess)
    rs = MCFFT_HIP_ERROR;
  if (rs == MCFFT_SUCCESS) rs = set_lds_attr(k);
  if (rs == MCFFT_SUCCESS &&
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k->fn, k->nw * 64, k->lds_bytes) != hipSuccess)
    rs = MCFFT_HIP_ERROR;
  p->rgrid = (per_cu < 1 ? 1 : per_cu) * (n_cu < 1 ? 1 : n_cu);
  if (rs == MCFFT_SUCCESS) rs = make_tables(p, r.tw, r.cm);
  const int ws_fmt = (r.cm == CM_SPLIT) ? FMT_SPLIT : r.in_fmt;
  if (rs == MCFFT_SUCCESS && r.ws_elems > 0) {
    p->ws_bytes = (size_t)r.ws_elems * (size_t)elem_bytes(ws_fmt);
    if (hipMalloc(&p->ws, p->ws_bytes) != hipSuccess) rs = MCFFT_ALLOC_FAILED;
  }
  // device pass descriptors per direction, tasks, counters
  const size_t np = r.passes.size(), nt = r.task_pass.size();
  for (int d = 0; d < 2 && rs == MCFFT_SUCCESS; ++d) {
    std::vector<RaggedPass> dp(np);
    for (size_t i = 0; i < np; ++i) {
      const RaggedPassHost& h = r.passes[i];
      RaggedPass& q = dp[i];
      memset(&q, 0, sizeof(q));
      PassArgs& a = q.args;
      const int src_esz = elem_bytes(h.pp.src_fmt);
      const int dst_esz = elem_bytes(h.pp.dst == BUF_WS ? ws_fmt : r.out_fmt);
      a.in = (const void*)(size_t)(h.in_off * src_esz);
      a.out = (void*)(size_t)(h.out_off * dst_esz);
      a.in_sg = h.pp.in_sg;
      a.in_st = h.pp.in_st;
      a.out_sg = h.pp.out_sg;
      a.out_st = h.pp.out_st;
      a.stride_in = h.pp.stride_in;
      a.stride_out = h.pp.stride_out;
      a.n_inst = (int)h.pp.n_inst;
      a.tiles_per_group = h.pp.tiles_per_group;
      a.n_tiles = (int)(h.pp.groups * h.pp.tiles_per_group);
      a.in_fmt = h.pp.src_fmt;
      a.inverse = d;
      a.out_scale = h.pp.out_scale;
      for (size_t si = 0; si < 6; ++si) a.st_scale[si] = si < h.pp.shifts.size() ? pow2(-h.pp.shifts[si]) : 1.f;
      a.stage_tw = p->tw[d] + h.pp.stage_tw_off;
      a.stage_tw_entries = (int)((long long)r.tw.size() - h.pp.stage_tw_off);
      a.aops = p->aops[d];
      a.in_elems = h.in_elems;
      a.out_elems = h.out_elems;
      if (h.pp.fourstep) {
        a.fs_tw = p->tw[d] + h.pp.fs_tw_off;
        a.fs_tw_entries = (int)((long long)r.tw.size() - h.pp.fs_tw_off);
        a.fs_cols = h.pp.fs_cols;
      }
      q.in_buf = h.pp.src;
      q.out_buf = h.pp.dst;
      q.variant = h.variant;
      q.counter = h.counter;
      q.wait_for = h.wait_for;
      q.signal = h.signal;
    }
    if (hipMalloc(&p->rpasses[d], sizeof(RaggedPass) * np) != hipSuccess) rs = MCFFT_ALLOC_FAILED;
    else if (hipMemcpy(p->rpasses[d], dp.data(), sizeof(RaggedPass) * np, hipMemcpyHostToDevice) != hipSuccess)
      rs = MCFFT_HIP_ERROR;
  }
  if (rs == MCFFT_SUCCESS) {
    std::vector<RaggedTask> tasks(nt);
    for (size_t i = 0; i < nt; ++i) tasks[i] = RaggedTask{r.task_pass[i], r.task_tile[i]};
    if (hipMalloc(&p->rtasks, sizeof(RaggedTask) * (nt ? nt : 1)) != hipSuccess ||
        hipMalloc(&p->rcounters, sizeof(unsigned) * (kRaggedCounterBase + r.n_counters + 1)) != hipSuccess)
      rs = MCFFT_ALLOC_FAILED;
    else if (hipMemcpy(p->rtasks, tasks.data(), sizeof(RaggedTask) * nt, hipMemcpyHostToDevice) != hipSuccess ||
             hipMemset(p->rcounters, 0, sizeof(unsigned) * (kRaggedCounterBase + r.n_counters + 1)) != hipSuccess)
      rs = MCFFT_HIP_ERROR;
  }
  if (rs != MCFFT_SUCCESS) {
    free_plan(p);
    return rs;
  }
  *plan = p;
  return MCFFT_SUCCESS;
}

static mcfftStatus exec_ragged(mcfftPlan plan, const void* in, void* out, int direction, hipStream_t s) {
  const RaggedPlanCore& r = plan->rcore;
  const int ein = elem_bytes(plan->core.in_fmt), eout = elem_bytes(plan->core.out_fmt);
  if (((size_t)in % (size_t)ein) || ((size_t)out % (size_t)eout)) return MCFFT_MISALIGNED;
  for (const auto& b : plan->bulk) {
    const mcfftStatus bs = mcfft_exec(b.plan, (const char*)in + (size_t)b.in_off * (size_t)ein,
                                      (char*)out + (size_t)b.out_off * (size_t)eout, direction, s);
    if (bs != MCFFT_SUCCESS) return bs;
  }
  if (!plan->has_queue) return MCFFT_SUCCESS;
  const int d = direction == MCFFT_FORWARD ? 0 : 1;
  const int nc = kRaggedCounterBase + r.n_counters;
  reset_counters<<<1, 256, 0, s>>>(plan->rcounters, nc);
  if (hipGetLastError() != hipSuccess) return MCFFT_HIP_ERROR;
  RaggedArgs ra;
  ra.in_base = (const char*)in;
  ra.out_base = (char*)out;
  ra.ws_base = (char*)plan->ws;
  ra.passes = plan->rpasses[d];
  ra.tasks = plan->rtasks;
  ra.counters = plan->rcounters;
  ra.spin_limit = 1u << 24;
  ra.aops = plan->aops[d];
  const KernelInfo* k = &kRaggedKernels[r.cm];
  // MCFFT_RAGGED_SINGLE=1: one launch with in-kernel pass dependencies;
  // default: one launch per pass phase (kernel boundaries order the phases)
  const char* env = getenv("MCFFT_RAGGED_SINGLE");
  const bool single = env && env[0] == '1';
  const int n_phase = (int)r.phase_begin.size() - 1;
  for (int ph = 0; ph < (single ? 1 : n_phase); ++ph) {
    ra.task_begin = single ? 0 : r.phase_begin[ph];
    ra.task_end = single ? r.phase_begin.back() : r.phase_begin[ph + 1];
    ra.queue = single ? 3 : ph;
    ra.use_deps = single ? 1 : 0;
    const int n = ra.task_end - ra.task_begin;
    if (n <= 0) continue;
    const int grid = plan->rgrid < n ? plan->rgrid : n;
    void* args[] = {&ra};
    if (hipLaunchKernel(k->fn, dim3((unsigned)grid), dim3((unsigned)(k->nw * 64)), args, (size_t)k->lds_bytes, s) !=
        hipSuccess)
      return MCFFT_HIP_ERROR;
  }
  return MCFFT_SUCCESS;
}

static mcfftStatus exec_complex(mcfftPlan plan, const void* in, void* out, int direction, void* stream, int first,
                                int count, void* const* start_events = nullptr, void* const* stop_events = nullptr,
                                const int* in_exp = nullptr);

// R2C: complex FFT of length H into the plan's buffer, then the post-pass;
// C2R: the pre-pass into the buffer, then the inverse complex FFT.
static mcfftStatus exec_real(mcfftPlan plan, const void* in, void* out, int direction, hipStream_t s,
                             const int* in_exp) {
  if (!in || !out) return MCFFT_INVALID_VALUE;
  if ((plan->real_kind == 1) != (direction == MCFFT_FORWARD)) return MCFFT_INVALID_VALUE;
  const int cm = plan->core.cm;
  const int np = (int)plan->core.passes.size();
  RealArgs ra;
  ra.h = plan->real_h;
  ra.batch = plan->core.batch;
  ra.rows = plan->real_rows;
  ra.ox = plan->real_outer[0];
  ra.oy = plan->real_outer[1];
  ra.tw = plan->real_tw[direction == MCFFT_FORWARD ? 0 : 1];
  const long long total = (plan->real_kind == 1 ? (ra.h / 2 + 1) : ra.h) * ra.rows * ra.batch;
  const long long need = (total + 1023) / 1024;
  const unsigned grid = (unsigned)(need < plan->real_grid ? need : plan->real_grid);
  void* args[] = {&ra};
  if (plan->real_kind == 1) {
    mcfftStatus st = exec_complex(plan, in, plan->real_tmp, direction, s, 0, np, nullptr, nullptr, in_exp);
    if (st != MCFFT_SUCCESS) return st;
    ra.in = plan->real_tmp;
    ra.out = out;
    ra.in_fmt = plan->core.out_fmt;
    if (hipLaunchKernel(kR2CPost[cm], dim3(grid), dim3(1024), args, 0, s) != hipSuccess) return MCFFT_HIP_ERROR;
    return MCFFT_SUCCESS;
  }
  ra.in = in;
  ra.out = plan->real_tmp;
  ra.in_fmt = plan->real_in_fmt;
  if (hipLaunchKernel(kC2RPre[cm], dim3(grid), dim3(1024), args, 0, s) != hipSuccess) return MCFFT_HIP_ERROR;
  return exec_complex(plan, plan->real_tmp, out, direction, s, 0, np, nullptr, nullptr, in_exp);
}

// Gather a non-contiguous input into the staging buffer, run the transform,
// scatter the result to a non-contiguous output.
static mcfftStatus exec_layout(mcfftPlan plan, const void* in_re, const void* in_im, void* out_re, void* out_im,
                               int direction, hipStream_t s, const int* in_exp = nullptr) {
  if (direction != MCFFT_FORWARD && direction != MCFFT_INVERSE) return MCFFT_INVALID_VALUE;
  if (!in_re || !out_re) return MCFFT_INVALID_VALUE;
  if ((in_im != nullptr) != (plan->lay_planar[0] != 0) || (out_im != nullptr) != (plan->lay_planar[1] != 0))
    return MCFFT_INVALID_VALUE;
  const int fmt[2] = {plan->core.in_fmt, plan->core.out_fmt};
  const void* ptr[2][2] = {{in_re, in_im}, {out_re, out_im}};
  for (int sd = 0; sd < 2; ++sd) {
    const size_t al = (size_t)elem_bytes(fmt[sd]) / (plan->lay_planar[sd] ? 2 : 1);
    for (int c = 0; c < 2; ++c)
      if ((size_t)ptr[sd][c] % al) return MCFFT_MISALIGNED;
  }
  const int np = (int)plan->core.passes.size();
  if (plan->gather[0]) {
    // no staging: the sides may alias only exactly
    const char* a0 = (const char*)in_re;
    const char* b0 = (const char*)out_re;
    const char* a1 = a0 + plan->gather_ext[0] * elem_bytes(fmt[0]);
    const char* b1 = b0 + plan->gather_ext[1] * elem_bytes(fmt[1]);
    const bool overlap = a0 < b1 && b0 < a1;
    if (overlap && !(a0 == b0 && plan->lay_dist[0] == plan->lay_dist[1] && fmt[0] == fmt[1]))
      return MCFFT_INVALID_VALUE;
  }
  auto launch = [&](int sd, const void* k, LayoutArgs& la) -> mcfftStatus {
    layout_fill(la, plan->lay_rank, plan->lay_dims, plan->core.batch, plan->lay_strides[sd], plan->lay_dist[sd],
                plan->lay_planar[sd], fmt[sd]);
    la.vec = layout_vec_ok(la, ptr[sd][0], ptr[sd][1]) ? 1 : 0;
    const long long need = ((la.vec ? la.total / 4 : la.total) + kLayoutThreads - 1) / kLayoutThreads;
    const unsigned grid = (unsigned)(need < plan->lay_grid ? need : plan->lay_grid);
    void* args[] = {&la};
    return hipLaunchKernel(k, dim3(grid), dim3(kLayoutThreads), args, 0, s) == hipSuccess ? MCFFT_SUCCESS
                                                                                        : MCFFT_HIP_ERROR;
  };
  const void* src = in_re;
  void* dst = out_re;
  if (!plan->lay_dense[0]) {
    LayoutArgs la{};
    la.src[0] = in_re;
    la.src[1] = in_im;
    la.dst[0] = plan->lay_buf[0];
    const mcfftStatus st = launch(0, kLayoutGather, la);
    if (st != MCFFT_SUCCESS) return st;
    src = plan->lay_buf[0];
  }
  if (!plan->lay_dense[1]) dst = plan->lay_buf[1];
  const mcfftStatus st = exec_complex(plan, src, dst, direction, s, 0, np, nullptr, nullptr, in_exp);
  if (st != MCFFT_SUCCESS || plan->lay_dense[1]) return st;
  LayoutArgs la{};
  la.src[0] = plan->lay_buf[1];
  la.dst[0] = out_re;
  la.dst[1] = out_im;
  return launch(1, kLayoutScatter, la);
}

mcfftStatus mcfft_exec_layout(mcfftPlan plan, const void* in_re, const void* in_im, void* out_re, void* out_im,
                              int direction, void* stream) {
  if (!plan || !plan->layout) return MCFFT_INVALID_PLAN;
  return exec_layout(plan, in_re, in_im, out_re, out_im, direction, (hipStream_t)stream);
}

mcfftStatus mcfft_exec_exp(mcfftPlan plan, const void* in, const int* in_exp, void* out, int direction,
                           void* stream) {
  if (!plan) return MCFFT_INVALID_PLAN;
  if (plan->layout) return exec_layout(plan, in, nullptr, out, nullptr, direction, (hipStream_t)stream, in_exp);
  if (plan->ragged) {
    if (in_exp) return MCFFT_UNSUPPORTED;
    if (!in || !out) return MCFFT_INVALID_VALUE;
    if (direction != MCFFT_FORWARD && direction != MCFFT_INVERSE) return MCFFT_INVALID_VALUE;
    return exec_ragged(plan, in, out, direction, (hipStream_t)stream);
  }
  if (plan->real_kind) return exec_real(plan, in, out, direction, (hipStream_t)stream, in_exp);
  return exec_complex(plan, in, out, direction, stream, 0, (int)plan->core.passes.size(), nullptr, nullptr, in_exp);
}

mcfftStatus mcfft_exec(mcfftPlan plan, const void* in, void* out, int direction, void* stream) {
  return mcfft_exec_exp(plan, in, nullptr, out, direction, stream);
}

// Every compiled kernel variant of each pass of `p`, timed over `reps`
// launches on the given buffers; the fastest is kept per pass.
static mcfftStatus tune_passes(mcfftPlan p, const void* in, void* out, int direction, void* stream, int reps,
                               int n_cu, hipEvent_t e0, hipEvent_t e1, bool verbose) {
  hipStream_t s = (hipStream_t)stream;
  mcfftStatus rs = MCFFT_SUCCESS;
  for (size_t pi = 0; pi < p->core.passes.size() && rs == MCFFT_SUCCESS; ++pi) {
    PassPlan best = p->core.passes[pi];
    int best_cap = p->grid_cap[pi];
    float best_ms = 1e30f;
    for (const KernelInfo* k : pass_candidates(p->core.passes[pi], p->core.cm, p->core.scaling)) {
      PassPlan trial = p->core.passes[pi];
      if (!set_pass_kernel(trial, k)) continue;
      int per_cu = 0;
      if (set_lds_attr(k) != MCFFT_SUCCESS ||
          hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k->fn, k->nw * 64, k->lds_bytes) != hipSuccess ||
          per_cu < 1)
        continue;
      const PassPlan saved = p->core.passes[pi];
      const int saved_cap = p->grid_cap[pi];
      p->core.passes[pi] = trial;
      p->grid_cap[pi] = per_cu * n_cu;
      rs = mcfft_exec_passes(p, in, out, direction, stream, (int)pi, 1);  // warm-up
      float ms = 0.f;
      if (rs == MCFFT_SUCCESS && hipEventRecord(e0, s) == hipSuccess) {
        for (int r = 0; r < reps && rs == MCFFT_SUCCESS; ++r)
          rs = mcfft_exec_passes(p, in, out, direction, stream, (int)pi, 1);
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
          rs = MCFFT_HIP_ERROR;
      }
      if (rs == MCFFT_SUCCESS && verbose)
        fprintf(stderr, "mcfft tune: %s pass %zu %s %.4f ms/launch\n", opts_string(p->core.opts).c_str(), pi,
                k->name, ms / (float)reps);
      if (rs == MCFFT_SUCCESS && ms < best_ms) {
        best_ms = ms;
        best = trial;
        best_cap = per_cu * n_cu;
      }
      p->core.passes[pi] = saved;
      p->grid_cap[pi] = saved_cap;
    }
    p->core.passes[pi] = best;
    p->grid_cap[pi] = best_cap;
  }
  return rs;
}

// mean time of one whole exec of `p` (ms) over `reps` execs after one warm-up
static mcfftStatus time_exec(mcfftPlan p, const void* in, void* out, int direction, void* stream, int reps,
                             hipEvent_t e0, hipEvent_t e1, float* ms) {
  hipStream_t s = (hipStream_t)stream;
  const int np = (int)p->core.passes.size();
  mcfftStatus rs = mcfft_exec_passes(p, in, out, direction, stream, 0, np);
  if (rs == MCFFT_SUCCESS && hipEventRecord(e0, s) != hipSuccess) rs = MCFFT_HIP_ERROR;
  for (int r = 0; r < reps && rs == MCFFT_SUCCESS; ++r) rs = mcfft_exec_passes(p, in, out, direction, stream, 0, np);
  if (rs == MCFFT_SUCCESS && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                              hipEventElapsedTime(ms, e0, e1) != hipSuccess))
    rs = MCFFT_HIP_ERROR;
  *ms /= (float)reps;
  return rs;
}

mcfftStatus mcfft_plan_tune(mcfftPlan plan, const void* in, void* out, int direction, void* stream, int reps) {
  if (!plan || plan->ragged || plan->real_kind) return MCFFT_INVALID_PLAN;
  if (plan->gather[0]) return MCFFT_UNSUPPORTED;  // its buffers are laid out for its offset tables
  if (!in || !out || reps < 1) return MCFFT_INVALID_VALUE;
  int n_cu = 0;
  if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, plan->device) != hipSuccess)
    return MCFFT_HIP_ERROR;
  const char* venv = getenv("MCFFT_TUNE_VERBOSE");  // print every variant's time to stderr
  const bool verbose = venv && venv[0] == '1';
  const PlanCore& c0 = plan->core;
  // structures to try (a layout plan keeps its own: its staging is sized for it)
  std::vector<PlanOpts> structs;
  if (!plan->layout) {
    long long ws_limit = 0;
    if (const char* env = getenv("MCFFT_WS_BYTES")) ws_limit = atoll(env);
    structs = plan_structures(c0.rank, c0.dims, c0.batch, c0.in_fmt, c0.out_fmt, c0.scaling, ws_limit);
  }
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return MCFFT_HIP_ERROR;
  mcfftStatus rs = tune_passes(plan, in, out, direction, stream, reps, n_cu, e0, e1, verbose);
  float best_ms = 0.f;
  if (rs == MCFFT_SUCCESS) rs = time_exec(plan, in, out, direction, stream, reps, e0, e1, &best_ms);
  if (rs == MCFFT_SUCCESS && verbose) fprintf(stderr, "mcfft tune: %s exec %.4f ms\n", opts_string(plan->core.opts).c_str(), best_ms);
  mcfftPlan best = plan;
  for (const PlanOpts& o : structs) {
    if (rs != MCFFT_SUCCESS) break;
    if (o.outer_radix == c0.opts.outer_radix && o.fourstep_lg2 == c0.opts.fourstep_lg2) continue;
    mcfftPlan cand = nullptr;
    if (create_plan(&cand, c0.rank, c0.dims, c0.batch, (mcfftFormat)c0.in_fmt, (mcfftFormat)c0.out_fmt,
                    (mcfftScaling)c0.scaling, &o) != MCFFT_SUCCESS)
      continue;  // e.g. the structure's workspace does not fit: not a candidate
    float ms = 0.f;
    rs = tune_passes(cand, in, out, direction, stream, reps, n_cu, e0, e1, verbose);
    if (rs == MCFFT_SUCCESS) rs = time_exec(cand, in, out, direction, stream, reps, e0, e1, &ms);
    if (rs == MCFFT_SUCCESS && verbose) fprintf(stderr, "mcfft tune: %s exec %.4f ms\n", opts_string(o).c_str(), ms);
    if (rs == MCFFT_SUCCESS && ms < best_ms) {
      best_ms = ms;
      if (best != plan) free_plan(best);
      best = cand;
    } else {
      free_plan(cand);
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (best != plan) {
    if (rs == MCFFT_SUCCESS) std::swap(*plan, *best);  // the plan takes the faster structure's resources
    free_plan(best);
  }
  if (rs != MCFFT_SUCCESS) return rs;
  TuneChoice choice;
  choice.opts = plan->core.opts;
  for (const PassPlan& pp : plan->core.passes) choice.kernels.push_back(pp.k->name);
  std::lock_guard<std::mutex> g(g_cache_mu);
  load_cache_locked(plan->arch);
  g_cache[plan->arch + "|" + plan_signature(plan->core)] = format_tune_choice(choice);
  return save_cache_locked(plan->arch) ? MCFFT_SUCCESS : MCFFT_CACHE_ERROR;
}

mcfftStatus mcfft_build_id(const char** id) {
  if (!id) return MCFFT_INVALID_VALUE;
  *id = MCFFT_BUILD_ID;
  return MCFFT_SUCCESS;
}

// The error-norm reduction into `part` (3 * kNormBlocks doubles of device
// memory), then the combine of the partials on the host (blocking).
static mcfftStatus error_norms_into(const void* y, mcfftFormat fmt, int exponent, const double* ref, long long n,
                                    double* rel_l2, double* max_abs, hipStream_t s, double* part) {
  error_norms_kernel<<<kNormBlocks, 256, 0, s>>>(y, (int)fmt, ldexp(1.0, exponent), ref, n, part);
  double h[3 * kNormBlocks];
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(h, part, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return MCFFT_HIP_ERROR;
  double num = 0, den = 0, mx = 0;
  for (int b = 0; b < kNormBlocks; ++b) {
    num += h[3 * b];
    den += h[3 * b + 1];
    if (!(h[3 * b + 2] <= mx)) mx = h[3 * b + 2];
  }
  *rel_l2 = sqrt(num / (den > 0 ? den : 1.0));
  *max_abs = mx;
  return MCFFT_SUCCESS;
}

static bool error_norms_args_ok(const void* y, mcfftFormat fmt, const double* ref, long long n, const double* rel_l2,
                                const double* max_abs) {
  return y && ref && n >= 1 && n <= kErrorNormsMaxElems && rel_l2 && max_abs && (int)fmt >= 0 && (int)fmt <= 2;
}

mcfftStatus mcfft_error_norms(const void* y, mcfftFormat fmt, int exponent, const double* ref, long long n,
                              double* rel_l2, double* max_abs, void* stream) {
  if (!error_norms_args_ok(y, fmt, ref, n, rel_l2, max_abs)) return MCFFT_INVALID_VALUE;
  double* part = nullptr;
  if (hipMalloc(&part, sizeof(double) * 3 * kNormBlocks) != hipSuccess) return MCFFT_ALLOC_FAILED;
  const mcfftStatus rs = error_norms_into(y, fmt, exponent, ref, n, rel_l2, max_abs, (hipStream_t)stream, part);
  (void)hipFree(part);
  return rs;
}

mcfftStatus mcfft_error_norms_noalloc(const void* y, mcfftFormat fmt, int exponent, const double* ref, long long n,
                                      double* rel_l2, double* max_abs, void* stream) {
  if (!error_norms_args_ok(y, fmt, ref, n, rel_l2, max_abs)) return MCFFT_INVALID_VALUE;
  double* part = nullptr;
  const mcfftStatus st = scratch_on_current_device(&part);
  if (st != MCFFT_SUCCESS) return st;
  return error_norms_into(y, fmt, exponent, ref, n, rel_l2, max_abs, (hipStream_t)stream, part);
}

mcfftStatus mcfft_get_item_exponents(mcfftPlan plan, int* exponents) {
  if (!plan) return MCFFT_INVALID_PLAN;
  if (!exponents) return MCFFT_INVALID_VALUE;
  if (!plan->ragged) return mcfft_get_output_exponent(plan, exponents);
  for (size_t i = 0; i < plan->item_exp.size(); ++i) exponents[i] = plan->item_exp[i];
  return MCFFT_SUCCESS;
}

// read and clear the sticky status word of a regular plan (device synchronised)
static mcfftStatus take_status(mcfftPlan plan, unsigned* flags) {
  if (!plan->state) return MCFFT_SUCCESS;
  ExecState st;
  if (hipMemcpy(&st, plan->state, sizeof(st), hipMemcpyDeviceToHost) != hipSuccess) return MCFFT_HIP_ERROR;
  *flags |= st.status;
  if (st.status && hipMemset(&plan->state->status, 0, sizeof(unsigned)) != hipSuccess) return MCFFT_HIP_ERROR;
  return MCFFT_SUCCESS;
}

mcfftStatus mcfft_get_status_flags(mcfftPlan plan, unsigned* flags) {
  if (!plan) return MCFFT_INVALID_PLAN;
  if (!flags) return MCFFT_INVALID_VALUE;
  *flags = 0;
  if (hipDeviceSynchronize() != hipSuccess) return MCFFT_HIP_ERROR;
  if (!plan->ragged) return take_status(plan, flags);
  if (plan->rcounters && hipMemcpy(flags, plan->rcounters + 4, sizeof(unsigned), hipMemcpyDeviceToHost) != hipSuccess)
    return MCFFT_HIP_ERROR;
  for (const auto& b : plan->bulk) {
    const mcfftStatus st = take_status(b.plan, flags);
    if (st != MCFFT_SUCCESS) return st;
  }
  return MCFFT_SUCCESS;
}

mcfftStatus mcfft_get_transform_exponents(mcfftPlan plan, int* exponents) {
  if (!plan) return MCFFT_INVALID_PLAN;
  if (!exponents) return MCFFT_INVALID_VALUE;
  if (plan->ragged) return mcfft_get_item_exponents(plan, exponents);
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(exponents, plan->out_exp, sizeof(int) * (size_t)plan->core.batch, hipMemcpyDeviceToHost) != hipSuccess)
    return MCFFT_HIP_ERROR;
  return MCFFT_SUCCESS;
}

mcfftStatus mcfft_get_exponent_buffer(mcfftPlan plan, const int** exponents) {
  if (!plan || plan->ragged) return MCFFT_INVALID_PLAN;
  if (!exponents) return MCFFT_INVALID_VALUE;
  *exponents = plan->out_exp;
  return MCFFT_SUCCESS;
}

mcfftStatus mcfft_get_pass_count(mcfftPlan plan, int* count) {
  if (!plan || plan->ragged) return MCFFT_INVALID_PLAN;
  if (!count) return MCFFT_INVALID_VALUE;
  *count = (int)plan->core.passes.size();
  return MCFFT_SUCCESS;
}

mcfftStatus mcfft_exec_passes(mcfftPlan plan, const void* in, void* out, int direction, void* stream, int first,
                              int count) {
  if (!plan || plan->ragged || plan->real_kind) return MCFFT_INVALID_PLAN;
  return exec_complex(plan, in, out, direction, stream, first, count);
}

mcfftStatus mcfft_exec_timed(mcfftPlan plan, const void* in, void* out, int direction, void* stream,
                             void* const* start_events, void* const* stop_events) {
  if (!plan || plan->ragged || plan->real_kind) return MCFFT_INVALID_PLAN;
  if (!start_events || !stop_events) return MCFFT_INVALID_VALUE;
  if (plan->core.n_chunks != 1) return MCFFT_UNSUPPORTED;
  return exec_complex(plan, in, out, direction, stream, 0, (int)plan->core.passes.size(), start_events,
                      stop_events);
}

// start_events / stop_events (one per pass, single-chunk plans): each pass is
// launched with hipExtLaunchKernel, which records the events at the start and
// the end of that kernel's execution.
static mcfftStatus exec_complex(mcfftPlan plan, const void* in, void* out, int direction, void* stream, int first,
                                int count, void* const* start_events, void* const* stop_events, const int* in_exp) {
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
      const bool gathered = plan->gather[0] != nullptr;  // transforms located by offset tables
      auto bufptr = [&](int b, int fmt) -> char* {
        const size_t skip = gathered ? 0 : (size_t)(t0 * c.n) * (size_t)elem_bytes(fmt);
        if (b == BUF_IN) return (char*)in + skip;
        if (b == BUF_OUT) return (char*)out + skip;
        return (char*)plan->ws;
      };
      auto view_elems = [&](int b) -> long long {
        if (b == BUF_WS) return c.ws_elems;
        return gathered ? plan->gather_ext[b == BUF_IN ? 0 : 1] : (c.batch - t0) * c.n;
      };
      const int dst_fmt = (pp.dst == BUF_OUT) ? c.out_fmt : ws_fmt;
      PassArgs a;
      fill_pass_args(a, c, pp, bufptr(pp.src, pp.src_fmt), bufptr(pp.dst, dst_fmt), view_elems(pp.src),
                     view_elems(pp.dst), d, plan->tw[d], plan->aops[d], plan->dft256[d]);
      if (gathered) {
        const long long* const tab[2] = {plan->gather[0], plan->gather[1]};
        gather_args(a, pp, tab, t0);
      }
      a.stagger_ticks = plan->stagger_ticks;
      a.stagger_phases = plan->stagger_phases;
      fill_bfp(a, c, pi, t0, nt, plan->real_kind == 2, in_exp,
               BfpBuffers{plan->state, plan->slots, plan->tile_e, plan->out_exp});
      const long long groups = pp.groups_per_transform > 0 ? nt * pp.groups_per_transform : pp.groups;
      const long long tiles = groups * (long long)pp.tiles_per_group;
      if (tiles <= 0 || tiles > 0x7fffffffLL) return MCFFT_UNSUPPORTED;
      a.n_tiles = (int)tiles;
      const long long grid = tiles < plan->grid_cap[pi] ? tiles : plan->grid_cap[pi];
      void* args[] = {&a};
      if (start_events) {
        if (hipExtLaunchKernel(pp.k->fn, dim3((unsigned)grid), dim3((unsigned)(pp.k->nw * 64)), args,
                               (size_t)pp.k->lds_bytes, s, (hipEvent_t)start_events[pi],
                               (hipEvent_t)stop_events[pi], 0) != hipSuccess)
          return MCFFT_HIP_ERROR;
      } else if (hipLaunchKernel(pp.k->fn, dim3((unsigned)grid), dim3((unsigned)(pp.k->nw * 64)), args,
                                 (size_t)pp.k->lds_bytes, s) != hipSuccess) {
        return MCFFT_HIP_ERROR;
      }
    }
  }
  return MCFFT_SUCCESS;
}

mcfftStatus mcfft_get_output_exponent(mcfftPlan plan, int* exponent) {
  if (!plan) return MCFFT_INVALID_PLAN;
  if (!exponent) return MCFFT_INVALID_VALUE;
  if (!plan->ragged && plan->core.scaling == SCALE_AUTO) return MCFFT_UNSUPPORTED;  // per transform, found on the device
  *exponent = plan->core.out_exponent + (plan->real_kind == 2 ? 1 : 0);
  return MCFFT_SUCCESS;
}

mcfftStatus mcfft_get_workspace_size(mcfftPlan plan, size_t* bytes) {
  if (!plan) return MCFFT_INVALID_PLAN;
  if (!bytes) return MCFFT_INVALID_VALUE;
  size_t b = plan->ws_bytes;
  if (plan->real_kind)
    b += (size_t)(plan->real_rows * plan->real_h * plan->core.batch) * (size_t)elem_bytes(plan->real_tmp_fmt);
  b += plan->lay_bytes;
  for (const auto& bk : plan->bulk) {
    size_t sb = 0;
    if (mcfft_get_workspace_size(bk.plan, &sb) == MCFFT_SUCCESS) b += sb;
  }
  *bytes = b;
  return MCFFT_SUCCESS;
}

mcfftStatus mcfft_plan_describe(mcfftPlan plan, char* buf, size_t len) {
  if (!plan) return MCFFT_INVALID_PLAN;
  if (!buf || len == 0) return MCFFT_INVALID_VALUE;
  std::string s;
  if (plan->ragged) {
    s = plan->has_queue ? plan->rcore.describe() : std::string("ragged items=0 (every item in a bulk plan)\n");
    for (const auto& b : plan->bulk) {
      const std::string d = b.plan->core.describe();
      s += (b.plan->gather[0] ? std::string("  bulk gathered: ")
                              : "  bulk in=" + std::to_string(b.in_off) + " out=" + std::to_string(b.out_off) + ": ") +
           d.substr(0, d.find('\n')) + "\n";
    }
  } else {
    s = plan->core.describe();
  }
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
    case MCFFT_NONFINITE: return "non-finite result (overflow)";
  }
  return "unknown status";
}

int mcfft_version(void) { return 200; }

}  // extern "C"

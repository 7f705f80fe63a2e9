// extern "C" entry points of liblettuce_amd.so (declared in
// include/lettuce_amd.h) and the FP64 utility kernels: equilibrium
// initialisation, storage encode/decode, moments, global reductions,
// enstrophy and the halfway bounce-back link force.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>
#include <type_traits>

#include "lbm_step_impl.h"
#include "lbm_slab.h"
#include "lbm_tb2.h"

namespace lbm {
KernelFn pick_kernel_d2q9(int storage, int coll, bool mask, int mode);
KernelFn pick_kernel_d3q19(int storage, int coll, bool mask, int mode);
KernelFn pick_kernel_d3q27(int storage, int coll, bool mask, int mode);
int launch_learned(const KParams& p, const lbm_learned* l, int stencil, int storage, bool mask,
                   int mode, bool wide, hipStream_t stream, int64_t nsteps, const void* f_in,
                   void* f_out);
}  // namespace lbm

// Launch tuning, for measurements and tests.  Each knob starts from its
// environment variable, read once when the library is loaded, and
// lbm_set_tuning changes it at run time (value < 0 restores the start value):
//  two_step           LETTUCE_AMD_TB2           0 never, 1 the default
//                     selection (see two_step_kernel), 2 wherever it applies
//  two_step_planes    LETTUCE_AMD_TB2_PLANES    x chunk of the two-step kernel
//                                               (0: chosen per grid)
//  two_step_barriers  LETTUCE_AMD_TB2_BARRIERS  1 or 2 barriers per plane
//                                               (0: chosen per format)
//  wide_offsets       LETTUCE_AMD_WIDE          1 forces the 64-bit i0 plane
//                                               offsets (0: chosen per grid)
//  post_split         LETTUCE_AMD_POST_SPLIT    0 runs post-streaming with the
//                                               push kernel only (see lbm_run)
namespace {
struct Knob {
  const char* name;
  const char* env;
  int dflt;
  int start;
  int value;
};
Knob g_knobs[] = {
    {"two_step", "LETTUCE_AMD_TB2", 1, 0, 0},
    {"two_step_planes", "LETTUCE_AMD_TB2_PLANES", 0, 0, 0},
    {"two_step_barriers", "LETTUCE_AMD_TB2_BARRIERS", 0, 0, 0},
    {"wide_offsets", "LETTUCE_AMD_WIDE", 0, 0, 0},
    {"post_split", "LETTUCE_AMD_POST_SPLIT", 1, 0, 0},
};
enum KnobId { K_TWO_STEP, K_TB2_PLANES, K_TB2_BARRIERS, K_WIDE, K_POST_SPLIT, K_COUNT };
struct KnobInit {
  KnobInit() {
    for (Knob& k : g_knobs) {
      const char* v = getenv(k.env);
      k.start = k.value = v ? atoi(v) : k.dflt;
    }
  }
} g_knob_init;
inline int knob(KnobId id) { return g_knobs[id].value; }
}  // namespace

using namespace lbm;

static thread_local char g_err[512] = "";

static int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

namespace lbm {
// error reporting for the entry points defined in other translation units
int report_error(int code, const char* msg) { return fail(code, "%s", msg); }
}  // namespace lbm

static int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(LBM_E_HIP, "%s: %s", what, hipGetErrorString(e));
  return LBM_OK;
}

static int stencil_q(int stencil) {
  switch (stencil) {
    case LBM_D2Q9: return 9;
    case LBM_D3Q19: return 19;
    case LBM_D3Q27: return 27;
  }
  return -1;
}
static int stencil_d(int stencil) { return stencil == LBM_D2Q9 ? 2 : 3; }

static int storage_bytes(int storage) {
  switch (storage) {
    case LBM_F32: return 4;
    case LBM_F64: return 8;
    case LBM_F16_SHIFTED:
    case LBM_BF16_SHIFTED: return 2;
  }
  return -1;
}

// Shape / stride validation shared by every entry point.  The population
// array is [Q][n0][n1][n2] with unit stride along i2, rows of at least n2
// elements and no overlap between populations.
static int check_layout(int stencil, int storage, const int64_t n[3], const int64_t stride[4]) {
  const int Q = stencil_q(stencil);
  if (Q < 0) return fail(LBM_E_ARG, "unknown stencil %d", stencil);
  if (storage_bytes(storage) < 0) return fail(LBM_E_ARG, "unknown storage %d", storage);
  if (!n || !stride) return fail(LBM_E_ARG, "NULL extents or strides");
  if (n[0] < 1 || n[1] < 1 || n[2] < 1)
    return fail(LBM_E_ARG, "extents must be positive, got (%lld, %lld, %lld)",
                (long long)n[0], (long long)n[1], (long long)n[2]);
  if (stencil_d(stencil) == 2 && n[0] != 1)
    return fail(LBM_E_ARG, "2D lattices need n[0] == 1, got %lld", (long long)n[0]);
  if (stride[3] != 1) return fail(LBM_E_ARG, "the last axis must be contiguous (stride 1)");
  // the strides must address every (q, i0, i1, i2) at a distinct element:
  // sorted by stride, each axis must step over the whole extent of the
  // axes below it
  int64_t ext[4] = {Q, n[0], n[1], n[2]}, str[4] = {stride[0], stride[1], stride[2], 1};
  int ord[4] = {3, 2, 1, 0};
  for (int i = 1; i < 4; ++i)
    for (int j = i; j > 0 && str[ord[j]] < str[ord[j - 1]]; --j) {
      const int t = ord[j];
      ord[j] = ord[j - 1];
      ord[j - 1] = t;
    }
  int64_t span = 1;
  for (int k = 0; k < 4; ++k) {
    const int a = ord[k];
    if (ext[a] == 1) continue;
    if (str[a] < span) return fail(LBM_E_ARG, "overlapping population strides");
    span = str[a] * ext[a];
  }
  // Inside the kernels a cell's offset is its i0 plane's offset (64-bit,
  // added to the population's base) plus a 32-bit offset inside the plane,
  // which formats of up to 4 bytes address in bytes: the offsets inside a
  // plane stay below 2^31 elements and below 2^32 bytes.
  const int64_t plane_span = (n[1] - 1) * stride[2] + n[2];
  const int bytes = storage_bytes(storage);
  const int64_t plane_limit =
      bytes <= 4 ? std::min<int64_t>((int64_t)1 << 31, ((int64_t)1 << 32) / bytes)
                 : ((int64_t)1 << 31);
  if (plane_span > plane_limit)
    return fail(LBM_E_ARG, "an i0 plane spans %lld elements, above the %lld that 32-bit offsets inside a plane address",
                (long long)plane_span, (long long)plane_limit);
  if (span > ((int64_t)1 << 40)) return fail(LBM_E_ARG, "population array too large");
  return LBM_OK;
}

// Whether a layout needs the wide (64-bit i0 plane) offsets of the
// two-step and learned kernels: their narrow variants address a cell by one
// 32-bit offset, in bytes for formats of up to 4 bytes.
// The knob wide_offsets = 1 selects the wide variants for every grid (tests
// run them on small grids against the narrow ones).
static bool wide_offsets(int storage, const int64_t n[3], const int64_t stride[4]) {
  if (knob(K_WIDE) == 1) return true;
  const int64_t span = (n[0] - 1) * stride[1] + (n[1] - 1) * stride[2] + n[2];
  const int bytes = storage_bytes(storage);
  const int64_t limit = bytes <= 4 ? std::min<int64_t>((int64_t)1 << 31, ((int64_t)1 << 32) / bytes)
                                   : ((int64_t)1 << 31);
  return span > limit;
}

// masked kernels index the class map by the storage offset of a cell, which
// is its linear index only for the contiguous [Q][n0][n1][n2] layout
static bool contiguous_cells(const int64_t n[3], const int64_t stride[4]) {
  return stride[2] == n[2] && (n[0] == 1 || stride[1] == n[1] * n[2]);
}

static int block_size_for(int64_t n2) {
  if (n2 % 256 == 0) return 256;
  if (n2 % 192 == 0) return 192;
  if (n2 % 128 == 0) return 128;
  if (n2 <= 64) return 64;
  if (n2 <= 128) return 128;
  return 256;
}

// the two-step kernel for this step, if one applies: periodic D3Q19 or
// D3Q27 without boundaries, pull streaming, z a multiple of 64 and y of
// TB_Y.  Used by default for D3Q19 in every format and for D3Q27 in FP32,
// with one barrier per plane.  Measured on MI355X, TGV 256^3: D3Q19 BGK FP32
// 38.3k (single steps) -> 63.4k MLUPS, FP16-shifted 106k; D3Q27 KBC FP32
// 26.7k -> 40.0k, but D3Q27 KBC BF16 45.9k -> 39.7k (its arithmetic, not
// its memory traffic, bounds it, and the two-step kernel recomputes the
// halo rows), so 16-bit D3Q27 stays on single steps.
// The knob two_step = 0 turns it off, = 2 uses it wherever it applies;
// two_step_planes sets the x chunk, two_step_barriers = 1 or 2 the variant
// with one barrier per plane or two with shallower LDS rings.  For D3Q19,
// loading the next plane during the current one measured slower (its extra
// VGPRs leave one block per CU; D3Q27 has one anyway and does load ahead),
// and so did the plain tile order (XCD order: 54.7k -> 57.8k MLUPS).
static TbFn two_step_kernel(const lbm_step_args* a) {
  const int mode = knob(K_TWO_STEP);
  if (mode == 0) return nullptr;
  if (mode == 1 && !(a->stencil == LBM_D3Q19 || (a->stencil == LBM_D3Q27 && a->storage == LBM_F32)))
    return nullptr;
  if (a->cell_class || a->streaming != LBM_PRE_STREAMING) return nullptr;
  if (a->n[2] % 64 != 0 || a->n[1] % TB_Y != 0) return nullptr;
  if (a->n[2] * a->n[1] * a->n[0] > (int64_t)0x7fffffff) return nullptr;
  // one barrier per plane for FP32, two (with the shallower LDS rings) for
  // the 16-bit formats: measured on D3Q19 BGK 256^3 with the wave
  // priorities, FP16 99.3k -> 102.5k MLUPS and BF16 99.6k -> 102.7k with
  // two, FP32 63.5k -> 63.1k
  const int barriers =
      knob(K_TB2_BARRIERS) > 0 ? knob(K_TB2_BARRIERS) : (storage_bytes(a->storage) == 2 ? 2 : 1);
  return pick_tb2_kernel(a->stencil, codec_id(a->storage, a->sr_seed), a->collision,
                         barriers == 2 ? 2 : 1,
                         wide_offsets(a->storage, a->n, a->stride));
}

// the two-step kernel for a 2D grid, if one applies: D2Q9 (with or without
// a cell-class map), pull streaming, a collision other than the learned
// one, 32- or 16-bit storage, rows of a multiple of 64 and at most 1024
// cells (contiguous, where a class map indexes them), element offsets
// within 32 bits (bytes), and the LDS ring within the device's per-block
// limit.  Two rows per iteration where
// a block of 2 n2 threads and its ring fit (measured on the 2048 x 512
// channel, BGK FP32: 61.2k MLUPS with one row, 77.2k with two), else one.
// *lds: the block's dynamic LDS, *nr: rows per iteration.
static Tb2dFn two_step_2d(const lbm_step_args* a, int64_t* lds, int* nr) {
  if (knob(K_TWO_STEP) == 0) return nullptr;
  if (a->stencil != LBM_D2Q9 || a->streaming != LBM_PRE_STREAMING) return nullptr;
  if (a->collision == LBM_COLL_LEARNED || a->storage == LBM_F64) return nullptr;
  if (a->n[0] != 1 || a->n[2] % 64 != 0 || a->n[2] > 1024) return nullptr;
  if (a->stride[3] != 1 || (a->cell_class && a->stride[2] != a->n[2])) return nullptr;
  const int bytes = storage_bytes(a->storage);
  const int64_t last = a->stride[0] * 8 + a->stride[2] * (a->n[1] - 1) + a->n[2] - 1;
  if ((last + 1) * bytes >= ((int64_t)1 << 32)) return nullptr;
  int dev = 0, max_lds = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    return nullptr;
  for (int rows = 2; rows >= 1; --rows) {
    *nr = rows;
    *lds = tb2d_lds_bytes((int)a->n[2], bytes, rows);
    if (rows * a->n[2] <= 1024 && *lds + tb2d_static_lds_bytes() <= max_lds)
      return pick_tb2d_kernel(a->stencil, codec_id(a->storage, a->sr_seed), a->collision, rows,
                              a->cell_class != nullptr);
  }
  return nullptr;
}

static bool has_two_step(const lbm_step_args* a) {
  int64_t lds = 0;
  int nr = 0;
  return two_step_kernel(a) != nullptr || two_step_2d(a, &lds, &nr) != nullptr;
}

// pairs of steps fused into one pass by the two-step kernel.  After nsteps
// the result must sit in bufs[nsteps & 1] as with single steps, so an odd
// number of pairs runs one pair as two single steps.
static int64_t two_step_pairs(bool two_step, int64_t nsteps) {
  int64_t pairs = two_step ? nsteps / 2 : 0;
  if (pairs % 2 != 0) --pairs;
  return pairs;
}

// Validates the arguments and prepares the single-step kernel (one cell per
// thread, block_tile indexing) of a step with a->streaming.
static int prepare_step(const lbm_step_args* a, KParams& p, KernelFn& fn, dim3& grid, dim3& block) {
  if (!a) return fail(LBM_E_ARG, "NULL args");
  if (a->abi_version != LBM_ABI_VERSION)
    return fail(LBM_E_VERSION, "abi_version %d != library %d", a->abi_version, LBM_ABI_VERSION);
  int rc = check_layout(a->stencil, a->storage, a->n, a->stride);
  if (rc) return rc;
  if (!a->f_in || !a->f_out) return fail(LBM_E_ARG, "NULL population pointer");
  if (a->f_in == a->f_out) return fail(LBM_E_ARG, "f_in and f_out must be distinct buffers");
  if (a->streaming < 0 || a->streaming > 3) return fail(LBM_E_ARG, "unknown streaming %d", a->streaming);
  if (a->collision < 0 || a->collision > LBM_COLL_BGK_GUO)
    return fail(LBM_E_ARG, "unknown collision %d", a->collision);
  const bool mask = a->cell_class != nullptr;
  if (mask && (!a->classes || a->n_classes < 1 || a->n_classes > 256))
    return fail(LBM_E_ARG, "cell_class given without a class table of 1..256 entries");
  if (mask && !contiguous_cells(a->n, a->stride))
    return fail(LBM_E_UNSUPPORTED, "boundaries need each population stored as one contiguous block");
  // the class map and the boundary-value field are indexed by 32-bit cell
  // offsets
  if (mask && a->n[0] * a->n[1] * a->n[2] > (int64_t)0x7fffffff)
    return fail(LBM_E_ARG, "grids with boundaries are limited to 2^31 - 1 cells");
  if (a->collision == LBM_COLL_BGK && !(a->tau > 0)) return fail(LBM_E_ARG, "BGK needs tau > 0");
  if (a->collision == LBM_COLL_BGK_GUO) {
    if (!(a->tau > 0) || !(a->force_tau > 0))
      return fail(LBM_E_ARG, "Guo forcing needs tau > 0 and force_tau > 0");
    for (int d = 0; d < 3; ++d)
      if (!std::isfinite(a->accel[d])) return fail(LBM_E_ARG, "acceleration must be finite");
    if (a->stencil == LBM_D2Q9 && a->accel[2] != 0)
      return fail(LBM_E_ARG, "a 2D acceleration has no z component");
  }
  if (a->collision == LBM_COLL_TRT && !(a->tau > 0 && a->tau_minus > 0))
    return fail(LBM_E_ARG, "TRT needs tau > 0 and tau_minus > 0");
  if (a->collision == LBM_COLL_KBC && !(a->tau > 0)) return fail(LBM_E_ARG, "KBC needs tau > 0");
  if (a->sr_seed != 0 && a->storage != LBM_BF16_SHIFTED)
    return fail(LBM_E_UNSUPPORTED, "stochastic rounding (sr_seed != 0) applies to BF16_SHIFTED storage only");

  fn = nullptr;
  if (a->collision == LBM_COLL_LEARNED) {
    if (a->stencil != LBM_D2Q9 && a->stencil != LBM_D3Q27)
      return fail(LBM_E_UNSUPPORTED, "the learned collision is defined for D2Q9 and D3Q27");
    if (a->storage == LBM_F64)
      return fail(LBM_E_UNSUPPORTED, "the learned collision computes in FP32; use F32 or shifted storage");
    if (a->n[2] % 64 != 0)
      return fail(LBM_E_UNSUPPORTED, "the learned collision needs the contiguous extent to be a multiple of 64");
    if (!a->learned || !a->learned->packed) return fail(LBM_E_ARG, "learned parameters missing");
    if (((uintptr_t)a->learned->packed) % 16 != 0) return fail(LBM_E_ARG, "packed parameters must be 16-byte aligned");
    if (!(a->tau > 0.5)) return fail(LBM_E_ARG, "the learned collision needs tau > 1/2");
    if (!(a->learned->rate_lo <= a->learned->rate_hi))
      return fail(LBM_E_ARG, "rate_lo must not exceed rate_hi");
  } else {
    const int codec = codec_id(a->storage, a->sr_seed);
    switch (a->stencil) {
      case LBM_D2Q9: fn = pick_kernel_d2q9(codec, a->collision, mask, a->streaming); break;
      case LBM_D3Q19: fn = pick_kernel_d3q19(codec, a->collision, mask, a->streaming); break;
      case LBM_D3Q27: fn = pick_kernel_d3q27(codec, a->collision, mask, a->streaming); break;
      default: fn = nullptr;
    }
    if (!fn)
      return fail(LBM_E_UNSUPPORTED, "no native kernel for stencil %d, storage %d, collision %d",
                  a->stencil, a->storage, a->collision);
  }

  p.fin = a->f_in;
  p.fout = a->f_out;
  p.cls = a->cell_class;
  p.ct = a->classes;
  p.nclasses = mask ? (int)a->n_classes : 0;
  p.bcf = a->bc_field;
  p.n0 = (int)a->n[0];
  p.n1 = (int)a->n[1];
  p.n2 = (int)a->n[2];
  p.s0 = a->stride[1];
  p.s1 = (int)a->stride[2];
  p.sq = a->stride[0];
  p.ncells = a->n[0] * a->n[1] * a->n[2];
  const int bs = block_size_for(a->n[2]);
  p.nb2 = (int)((a->n[2] + bs - 1) / bs);
  const int64_t nb = (int64_t)p.nb2 * a->n[1] * a->n[0];
  if (nb > 0x7fffffff) return fail(LBM_E_ARG, "too many blocks");
  p.nblocks = (int)nb;
  const bool learned = a->collision == LBM_COLL_LEARNED;
  const bool guo = a->collision == LBM_COLL_BGK_GUO;
  p.tau_inv = (a->collision == LBM_COLL_BGK || learned || guo) ? 1.0 / a->tau : 0.0;
  for (int d = 0; d < 3; ++d) p.accel[d] = guo ? a->accel[d] : 0.0;
  p.guo = guo ? 1.0 - 1.0 / (2.0 * a->force_tau) : 0.0;
  p.trt_p = a->collision == LBM_COLL_TRT ? 1.0 / (2.0 * a->tau) : 0.0;
  p.trt_m = a->collision == LBM_COLL_TRT ? 1.0 / (2.0 * a->tau_minus) : 0.0;
  p.beta = (a->collision == LBM_COLL_KBC || learned) ? 1.0 / (2.0 * a->tau) : 0.0;
  p.beta_inv = (a->collision == LBM_COLL_KBC || learned) ? 2.0 * a->tau : 0.0;
  p.rho0 = (a->storage == LBM_F16_SHIFTED || a->storage == LBM_BF16_SHIFTED) ? a->rho0_shift : 0.0;
  p.sr_seed = a->sr_seed;
  p.sr_step = a->sr_step;
  grid = dim3((unsigned)nb);
  block = dim3(bs);
  return LBM_OK;
}

// `count` single-step launches of fn, alternating the buffers; `step` is
// the index of the next step (stochastic rounding keys) and advances by one
// per launch
static void launch_singles(KernelFn fn, dim3 grid, dim3 block, KParams p, const void* const (&bufs)[2],
                           int64_t& launch, int64_t count, uint32_t& step, hipStream_t s) {
  for (int64_t k = 0; k < count; ++k, ++launch) {
    p.fin = bufs[launch & 1];
    p.fout = (void*)bufs[(launch + 1) & 1];
    p.sr_step = step++;
    hipLaunchKernelGGL(fn, grid, block, 0, s, p);
  }
}

// `nsteps` pull steps of a (a->streaming == LBM_PRE_STREAMING) from
// bufs[launch & 1]: pairs fused by the two-step kernel where it applies,
// the rest as single steps.  The number of launches has the parity of
// nsteps, so the result sits where single steps would leave it.
static void run_pull_steps(const lbm_step_args* a, const KParams& p, KernelFn fn, dim3 grid,
                           dim3 block, const void* const (&bufs)[2], int64_t& launch,
                           int64_t nsteps, uint32_t& step, hipStream_t s) {
  TbFn tb = two_step_kernel(a);
  int64_t lds = 0;
  int nr = 1;
  Tb2dFn t2 = tb ? nullptr : two_step_2d(a, &lds, &nr);
  const int64_t pairs = two_step_pairs(tb || t2, nsteps);
  if (pairs > 0 && t2) {
    Tb2dParams tp;
    tp.k = p;
    // rows per block: the grid's rows over the blocks the device holds at
    // once, at least 8 (each block computes S1 on 4 rows beyond its chunk)
    const int n1 = (int)a->n[1], n2 = (int)a->n[2];
    // the ring may exceed the default 64 KB of dynamic LDS per block
    (void)hipFuncSetAttribute((const void*)t2, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    tp.cx = knob(K_TB2_PLANES);
    if (tp.cx < 1) {
      tp.cx = 8;
      int per_cu = 0, dev = 0, cus = 0;
      if (hipGetDevice(&dev) == hipSuccess &&
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
          hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)t2, nr * n2,
                                                       (size_t)lds) == hipSuccess &&
          per_cu > 0) {
        const int64_t resident = (int64_t)per_cu * cus;
        tp.cx = (int)std::max<int64_t>(8, (n1 + resident - 1) / resident);
      }
    }
    const unsigned nb = (unsigned)((n1 + tp.cx - 1) / tp.cx);
    for (int64_t k = 0; k < pairs; ++k, ++launch) {
      tp.k.fin = bufs[launch & 1];
      tp.k.fout = (void*)bufs[(launch + 1) & 1];
      tp.k.sr_step = step;  // the launch's two steps: step and step + 1
      step += 2;
      hipLaunchKernelGGL(t2, dim3(nb), dim3(nr * n2), (size_t)lds, s, tp);
    }
  } else if (pairs > 0) {
    TbParams tp;
    tp.k = p;
    tp.nzt = (int)(a->n[2] / 64);
    tp.nyt = (int)(a->n[1] / TB_Y);
    // x chunk: 64 planes, or 128 where 64-plane tiles would make between one
    // and two rounds of the blocks the device holds at once, so that they
    // make one (or 48, below).  Measured on MI355X: D3Q19 256^3 (512 tiles, 512 resident
    // blocks) and 1024^3 (32768 tiles) are fastest at 64 planes (larger
    // chunks: -4 % to -17 %); D3Q27 KBC FP32 256^3, one block per CU, runs
    // 30.7k MLUPS at 64 planes and 31.3k at 128.
    tp.cx = knob(K_TB2_PLANES);
    if (tp.cx < 1) {
      tp.cx = 64;
      int per_cu = 0, dev = 0, cus = 0;
      if (hipGetDevice(&dev) == hipSuccess &&
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
          hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)tb, TB_THREADS, 0) ==
              hipSuccess &&
          per_cu > 0) {
        const int64_t resident = (int64_t)per_cu * cus;
        const int64_t tiles64 = (int64_t)tp.nzt * tp.nyt * ((a->n[0] + 63) / 64);
        // one block per CU (D3Q27): otherwise 48 planes where the rounds of
        // resident blocks times the planes a block sweeps (cx + 2) come out
        // lower than for 64.  D3Q27 KBC FP32 384^3: 64 planes make 6.75
        // rounds, 48 make 9; 35.6k -> 36.2k-37.5k MLUPS (the model ranks the
        // seven chunk lengths measured there, 32 to 192, in measured order).
        auto cost = [&](int64_t cx) {
          const int64_t tiles = (int64_t)tp.nzt * tp.nyt * ((a->n[0] + cx - 1) / cx);
          return (tiles + resident - 1) / resident * (cx + 2);
        };
        if (tiles64 > resident && tiles64 <= 2 * resident) tp.cx = 128;
        else if (per_cu == 1 && cost(48) < cost(64)) tp.cx = 48;
      }
    }
    const int64_t nxt = (a->n[0] + tp.cx - 1) / tp.cx;
    const int64_t ntiles = (int64_t)tp.nzt * tp.nyt * nxt;
    tp.ntiles = (int)ntiles;
    // block b runs on XCD b % 8 and takes tile (b % 8) * per_xcd + b / 8
    tp.per_xcd = (int)((ntiles + 7) / 8);
    const int64_t nb = 8 * (int64_t)tp.per_xcd;
    for (int64_t k = 0; k < pairs; ++k, ++launch) {
      tp.k.fin = bufs[launch & 1];
      tp.k.fout = (void*)bufs[(launch + 1) & 1];
      tp.k.sr_step = step;  // the launch's two steps: step and step + 1
      step += 2;
      hipLaunchKernelGGL(tb, dim3((unsigned)nb), dim3(TB_THREADS), 0, s, tp);
    }
  }
  launch_singles(fn, grid, block, p, bufs, launch, nsteps - 2 * pairs, step, s);
}

// Whether lbm_run splits a post-streaming run (lettuce's default order,
// collide then stream) into collide-only, pull and double-streaming launches:
// for n >= 2 steps without the learned collision, unless the knob
// post_split is 0.  The two orders are one composition of the same
// operators, (S C)^n = (S C S) (C S)^(n-2) C, so every cell takes the same
// values; the pull kernels run faster than the push kernel and can fuse
// steps in pairs.
static bool post_split(const lbm_step_args* a, int64_t nsteps) {
  return a->streaming == LBM_POST_STREAMING && a->collision != LBM_COLL_LEARNED && nsteps >= 2 &&
         knob(K_POST_SPLIT) != 0;
}

extern "C" {

int lbm_abi_version(void) { return LBM_ABI_VERSION; }

const char* lbm_last_error(void) { return g_err; }

int64_t lbm_struct_size(int which) {
  switch (which) {
    case 0: return (int64_t)sizeof(lbm_step_args);
    case 1: return (int64_t)sizeof(lbm_cell_class);
    case 2: return (int64_t)sizeof(lbm_learned);
  }
  return -1;
}

int lbm_set_tuning(const char* name, int32_t value) {
  if (!name) return fail(LBM_E_ARG, "NULL tuning name");
  for (Knob& k : g_knobs)
    if (!strcmp(k.name, name)) {
      k.value = value < 0 ? k.start : value;
      return LBM_OK;
    }
  return fail(LBM_E_ARG, "unknown tuning knob '%s'", name);
}

int lbm_step(const lbm_step_args* a, void* stream) {
  return lbm_run(a, 1, stream);
}

int lbm_run_plan(const lbm_step_args* a, int64_t nsteps, int64_t* plan) {
  if (nsteps < 0) return fail(LBM_E_ARG, "negative step count");
  if (!plan) return fail(LBM_E_ARG, "NULL plan");
  KParams p;
  KernelFn fn;
  dim3 grid, block;
  int rc = prepare_step(a, p, fn, grid, block);
  if (rc) return rc;
  if (a->collision == LBM_COLL_LEARNED) {
    plan[0] = 0;
    plan[1] = nsteps;
  } else if (post_split(a, nsteps)) {
    lbm_step_args k = *a;
    k.streaming = LBM_PRE_STREAMING;
    plan[0] = two_step_pairs(has_two_step(&k), nsteps - 2);
    plan[1] = nsteps - 2 * plan[0];
  } else {
    plan[0] = two_step_pairs(has_two_step(a), nsteps);
    plan[1] = nsteps - 2 * plan[0];
  }
  return LBM_OK;
}

int lbm_run(const lbm_step_args* a, int64_t nsteps, void* stream) {
  if (nsteps < 0) return fail(LBM_E_ARG, "negative step count");
  KParams p;
  KernelFn fn;
  dim3 grid, block;
  int rc = prepare_step(a, p, fn, grid, block);
  if (rc) return rc;
  if (nsteps == 0) return LBM_OK;
  g_err[0] = 0;
  hipStream_t s = (hipStream_t)stream;
  if (a->collision == LBM_COLL_LEARNED) {
    rc = launch_learned(p, a->learned, a->stencil, codec_id(a->storage, a->sr_seed),
                        a->cell_class != nullptr, a->streaming,
                        wide_offsets(a->storage, a->n, a->stride), s, nsteps, a->f_in, a->f_out);
    if (rc) return fail(rc, "no learned kernel for storage %d", a->storage);
    return check_launch("lbm_run (learned)");
  }
  const void* bufs[2] = {a->f_in, a->f_out};
  int64_t launch = 0;  // launches so far; launch k reads bufs[k & 1]
  uint32_t step = a->sr_step;  // index of the next step (stochastic rounding)
  if (post_split(a, nsteps)) {
    // (S C)^n = (S C S) (C S)^(n-2) C: collide only, n - 2 pull steps (two
    // fused per launch where the two-step kernel applies), then stream,
    // collide and stream.  n launches as with the push kernel, or fewer.
    lbm_step_args c = *a, k = *a, d = *a;
    c.streaming = LBM_NO_STREAMING;
    k.streaming = LBM_PRE_STREAMING;
    d.streaming = LBM_DOUBLE_STREAMING;
    KParams pc, pk, pd;
    KernelFn fc, fk, fd;
    dim3 gc, gk, gd, bc, bk, bd;
    if ((rc = prepare_step(&c, pc, fc, gc, bc)) || (rc = prepare_step(&k, pk, fk, gk, bk)) ||
        (rc = prepare_step(&d, pd, fd, gd, bd)))
      return rc;
    launch_singles(fc, gc, bc, pc, bufs, launch, 1, step, s);
    run_pull_steps(&k, pk, fk, gk, bk, bufs, launch, nsteps - 2, step, s);
    launch_singles(fd, gd, bd, pd, bufs, launch, 1, step, s);
    return check_launch("lbm_run (post-streaming)");
  }
  if (a->streaming == LBM_PRE_STREAMING) run_pull_steps(a, p, fn, grid, block, bufs, launch, nsteps, step, s);
  else launch_singles(fn, grid, block, p, bufs, launch, nsteps, step, s);
  return check_launch("lbm_run");
}

int64_t lbm_halo_planes(int32_t stencil) {
  const int k = halo_planes(stencil);
  if (k < 0) fail(LBM_E_ARG, "x-slabs are defined for D3Q19 and D3Q27, not stencil %d", stencil);
  return k;
}

}  // extern "C"

// the halo planes of an x-slab must be addressable with 32-bit offsets
static int check_slab_extent(int stencil, const int64_t n[3]) {
  const int k = halo_planes(stencil);
  if (k < 0) return fail(LBM_E_UNSUPPORTED, "x-slabs are defined for D3Q19 and D3Q27");
  if ((int64_t)k * n[1] * n[2] > (int64_t)0x7fffffff)
    return fail(LBM_E_ARG, "halo planes too large for 32-bit offsets");
  return LBM_OK;
}

static bool distinct4(const void* a, const void* b, const void* c, const void* d) {
  return a != b && a != c && a != d && b != c && b != d && c != d;
}

extern "C" {

int lbm_halo_pack(int32_t stencil, int32_t storage, const int64_t n[3], const int64_t stride[4],
                  const void* f, void* send_lo, void* send_hi, void* stream) {
  int rc = check_layout(stencil, storage, n, stride);
  if (rc) return rc;
  if (!f || !send_lo || !send_hi) return fail(LBM_E_ARG, "NULL population or halo pointer");
  if (send_lo == send_hi) return fail(LBM_E_ARG, "send_lo and send_hi must be distinct buffers");
  rc = check_slab_extent(stencil, n);
  if (rc) return rc;
  FaceFn fn = pick_pack_kernel(stencil, storage);
  if (!fn) return fail(LBM_E_UNSUPPORTED, "no pack kernel for storage %d", storage);
  KParams p;
  p.fin = f;
  p.n0 = (int)n[0];
  p.n1 = (int)n[1];
  p.n2 = (int)n[2];
  p.s0 = stride[1];
  p.s1 = (int)stride[2];
  p.sq = stride[0];
  const int bs = block_size_for(n[2]);
  p.nb2 = (int)((n[2] + bs - 1) / bs);
  const int64_t nb = (int64_t)p.nb2 * n[1] * (n[0] == 1 ? 1 : 2);
  if (nb > 0x7fffffff) return fail(LBM_E_ARG, "too many blocks");
  SlabHalo h{nullptr, nullptr, send_lo, send_hi};
  g_err[0] = 0;
  hipLaunchKernelGGL(fn, dim3((unsigned)nb), dim3(bs), 0, (hipStream_t)stream, p, h);
  return check_launch("lbm_halo_pack");
}

int lbm_slab_step(const lbm_step_args* a, int32_t part, const void* recv_lo, const void* recv_hi,
                  void* send_lo, void* send_hi, void* stream) {
  KParams p;
  KernelFn fn;
  dim3 grid, block;
  int rc = prepare_step(a, p, fn, grid, block);
  if (rc) return rc;
  rc = check_slab_extent(a->stencil, a->n);
  if (rc) return rc;
  if (a->cell_class) return fail(LBM_E_UNSUPPORTED, "x-slabs run without boundaries (cell_class must be NULL)");
  if (a->streaming != LBM_PRE_STREAMING)
    return fail(LBM_E_UNSUPPORTED, "x-slabs use pull streaming (LBM_PRE_STREAMING)");
  if (a->collision == LBM_COLL_LEARNED)
    return fail(LBM_E_UNSUPPORTED, "x-slabs have no learned-collision kernel");
  if (a->sr_seed != 0) return fail(LBM_E_UNSUPPORTED, "x-slabs round to nearest (sr_seed must be 0)");
  if (part < LBM_SLAB_ALL || part > LBM_SLAB_FACES) return fail(LBM_E_ARG, "unknown slab part %d", part);
  const bool faces = part != LBM_SLAB_INTERIOR;
  if (faces) {
    if (!recv_lo || !recv_hi || !send_lo || !send_hi) return fail(LBM_E_ARG, "NULL halo pointer");
    if (!distinct4(recv_lo, recv_hi, send_lo, send_hi))
      return fail(LBM_E_ARG, "the four halo buffers must be distinct");
  }
  FaceFn ff = faces ? pick_face_kernel(a->stencil, a->storage, a->collision) : nullptr;
  if (faces && !ff)
    return fail(LBM_E_UNSUPPORTED, "no face kernel for stencil %d, storage %d, collision %d",
                a->stencil, a->storage, a->collision);
  g_err[0] = 0;
  const int64_t n0 = a->n[0];
  if (part != LBM_SLAB_FACES && n0 > 2) {
    KParams pi = p;
    pi.x0 = 1;
    const int64_t nb = (int64_t)p.nb2 * a->n[1] * (n0 - 2);
    pi.nblocks = (int)nb;
    hipLaunchKernelGGL(fn, dim3((unsigned)nb), block, 0, (hipStream_t)stream, pi);
  }
  if (faces) {
    const int64_t nb = (int64_t)p.nb2 * a->n[1] * (n0 == 1 ? 1 : 2);
    SlabHalo h{recv_lo, recv_hi, send_lo, send_hi};
    hipLaunchKernelGGL(ff, dim3((unsigned)nb), block, 0, (hipStream_t)stream, p, h);
  }
  return check_launch("lbm_slab_step");
}

}  // extern "C"

// ---------------------------------------------------------------------------
// utility kernels (grid-stride over cells; arrays are [Q][cells] views given
// by the population stride and the spatial strides)
// ---------------------------------------------------------------------------
namespace {

struct Grid {
  int64_t n0, n1, n2, s0, s1, sq, cells;
};

__device__ __forceinline__ int64_t cell_offset(const Grid& g, int64_t c) {
  const int64_t i2 = c % g.n2;
  const int64_t r = c / g.n2;
  const int64_t i1 = r % g.n1;
  const int64_t i0 = r / g.n1;
  return i0 * g.s0 + i1 * g.s1 + i2;
}

template <class S, class ST>
__device__ __forceinline__ void load_cell(const Grid& g, const void* f, int64_t off, double rho0,
                                          typename ST::C (&v)[S::Q]) {
  using T = typename ST::T;
  using C = typename ST::C;
#pragma unroll
  for (int q = 0; q < S::Q; ++q) {
    const C sh = ST::shifted ? (C)(S::w[q] * rho0) : C(0);
    v[q] = ST::ld((const T*)f + q * g.sq + off, 0, sh);
  }
}

template <class S, class ST, class IN>
__global__ void k_equilibrium(Grid g, const IN* rho, const IN* u, double rho0, void* f) {
  using C = typename ST::C;
  using T = typename ST::T;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < g.cells;
       c += (int64_t)gridDim.x * blockDim.x) {
    C r = (C)rho[c];
    C uu[3] = {(C)u[c], (C)u[g.cells + c], S::D == 3 ? (C)u[2 * g.cells + c] : C(0)};
    C fe[S::Q];
    equilibrium<S, C>(r, uu, fe);
    const int64_t off = cell_offset(g, c);
#pragma unroll
    for (int q = 0; q < S::Q; ++q) {
      const C sh = ST::shifted ? (C)(S::w[q] * rho0) : C(0);
      ST::st((T*)f + q * g.sq + off, 0, fe[q], sh);
    }
  }
}

template <class S, class ST, class OUT>
__global__ void k_decode(Grid g, const void* f, double rho0, OUT* out) {
  using C = typename ST::C;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < g.cells;
       c += (int64_t)gridDim.x * blockDim.x) {
    C v[S::Q];
    load_cell<S, ST>(g, f, cell_offset(g, c), rho0, v);
#pragma unroll
    for (int q = 0; q < S::Q; ++q) out[q * g.cells + c] = (OUT)v[q];
  }
}

template <class S, class ST, class IN>
__global__ void k_encode(Grid g, const IN* in, double rho0, void* f) {
  using C = typename ST::C;
  using T = typename ST::T;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < g.cells;
       c += (int64_t)gridDim.x * blockDim.x) {
    const int64_t off = cell_offset(g, c);
#pragma unroll
    for (int q = 0; q < S::Q; ++q) {
      const C sh = ST::shifted ? (C)(S::w[q] * rho0) : C(0);
      ST::st((T*)f + q * g.sq + off, 0, (C)in[q * g.cells + c], sh);
    }
  }
}

// FP64 moments of one cell from its (decoded) populations
template <class S, class C>
__device__ __forceinline__ void moments_f64(const C (&v)[S::Q], double& rho, double (&j)[3]) {
  rho = 0;
  j[0] = j[1] = j[2] = 0;
#pragma unroll
  for (int q = 0; q < S::Q; ++q) {
    const double x = (double)v[q];
    rho += x;
    j[0] += S::e[q][0] * x;
    j[1] += S::e[q][1] * x;
    j[2] += S::e[q][2] * x;
  }
}

template <class S, class ST>
__global__ void k_moments(Grid g, const void* f, double rho0, double* rho, double* u) {
  using C = typename ST::C;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < g.cells;
       c += (int64_t)gridDim.x * blockDim.x) {
    C v[S::Q];
    load_cell<S, ST>(g, f, cell_offset(g, c), rho0, v);
    double r, j[3];
    moments_f64<S, C>(v, r, j);
    rho[c] = r;
#pragma unroll
    for (int d = 0; d < S::D; ++d) u[d * g.cells + c] = j[d] / r;
  }
}

constexpr int RED_BLOCK = 256;
constexpr int RED_MAX_BLOCKS = 2048;
constexpr int RED_N = 5;

// block-wide sum of RED_N doubles; result valid in thread 0
__device__ void block_sum(double (&v)[RED_N]) {
  __shared__ double sh[RED_N][RED_BLOCK / 64];
#pragma unroll
  for (int k = 0; k < RED_N; ++k) {
    double x = v[k];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if ((threadIdx.x & 63) == 0) sh[k][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < RED_N; ++k) {
      double s = 0;
      for (int w = 0; w < RED_BLOCK / 64; ++w) s += sh[k][w];
      v[k] = s;
    }
  }
}

template <class S, class ST>
__global__ void __launch_bounds__(RED_BLOCK) k_reduce(Grid g, const void* f, double rho0,
                                                      const uint8_t* mask, double* part) {
  using C = typename ST::C;
  double acc[RED_N] = {0, 0, 0, 0, 0};
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < g.cells;
       c += (int64_t)gridDim.x * blockDim.x) {
    if (mask && mask[c]) continue;
    C v[S::Q];
    load_cell<S, ST>(g, f, cell_offset(g, c), rho0, v);
    double r, j[3];
    moments_f64<S, C>(v, r, j);
    acc[0] += r;
    acc[1] += j[0];
    acc[2] += j[1];
    acc[3] += j[2];
    const double ri = 1.0 / r;
    acc[4] += 0.5 * (j[0] * j[0] + j[1] * j[1] + j[2] * j[2]) * ri * ri;
  }
  block_sum(acc);
  if (threadIdx.x == 0)
    for (int k = 0; k < RED_N; ++k) part[blockIdx.x * RED_N + k] = acc[k];
}

// fixed-order final sum of per-block partials (deterministic)
__global__ void k_finish(const double* part, int nparts, double* out, int nout) {
  double acc[RED_N] = {0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < nparts; b += RED_BLOCK)
    for (int k = 0; k < nout; ++k) acc[k] += part[b * RED_N + k];
  block_sum(acc);
  if (threadIdx.x == 0)
    for (int k = 0; k < nout; ++k) out[k] = acc[k];
}

// Failure scan: part[b] = { NaN populations, cells with non-finite u,
// max |u|^2 over the cells with finite u } of block b
template <class S, class ST>
__global__ void __launch_bounds__(RED_BLOCK) k_scan(Grid g, const void* f, double rho0,
                                                    double* part) {
  using C = typename ST::C;
  double acc[RED_N] = {0, 0, 0, 0, 0};
  double umax = 0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < g.cells;
       c += (int64_t)gridDim.x * blockDim.x) {
    C v[S::Q];
    load_cell<S, ST>(g, f, cell_offset(g, c), rho0, v);
#pragma unroll
    for (int q = 0; q < S::Q; ++q)
      if (v[q] != v[q]) acc[0] += 1.0;
    double r, j[3];
    moments_f64<S, C>(v, r, j);
    const double uu = (j[0] * j[0] + j[1] * j[1] + j[2] * j[2]) / (r * r);
    if (isfinite(uu)) umax = fmax(umax, uu);
    else acc[1] += 1.0;
  }
  acc[2] = umax;
  __shared__ double shmax[RED_BLOCK / 64];
  for (int o = 32; o > 0; o >>= 1) umax = fmax(umax, __shfl_xor(umax, o, 64));
  if ((threadIdx.x & 63) == 0) shmax[threadIdx.x >> 6] = umax;
  block_sum(acc);
  if (threadIdx.x == 0) {
    double m = 0;
    for (int w = 0; w < RED_BLOCK / 64; ++w) m = fmax(m, shmax[w]);
    part[blockIdx.x * RED_N + 0] = acc[0];
    part[blockIdx.x * RED_N + 1] = acc[1];
    part[blockIdx.x * RED_N + 2] = m;
  }
}

// counts summed, maxima maxed, in a fixed order (deterministic)
__global__ void k_finish_scan(const double* part, int nparts, double* out) {
  double acc[RED_N] = {0, 0, 0, 0, 0};
  double m = 0;
  for (int b = threadIdx.x; b < nparts; b += RED_BLOCK) {
    acc[0] += part[b * RED_N + 0];
    acc[1] += part[b * RED_N + 1];
    m = fmax(m, part[b * RED_N + 2]);
  }
  __shared__ double shmax[RED_BLOCK / 64];
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) shmax[threadIdx.x >> 6] = m;
  block_sum(acc);
  if (threadIdx.x == 0) {
    for (int w = 0; w < RED_BLOCK / 64; ++w) m = fmax(m, shmax[w]);
    out[0] = acc[0];
    out[1] = acc[1];
    out[2] = m;
  }
}

// 6th-order periodic central difference along axis `a` (weights of
// torch_gradient order 6, lettuce/util/utility.py:52-56)
__device__ __forceinline__ double d6(const double* u, int64_t i0, int64_t i1, int64_t i2,
                                     const int64_t (&n)[3], int a) {
  const double w[3] = {3.0 / 4.0, -3.0 / 20.0, 1.0 / 60.0};
  double s = 0;
  for (int k = 1; k <= 3; ++k) {
    int64_t p[3] = {i0, i1, i2}, m[3] = {i0, i1, i2};
    p[a] = (p[a] + k) % n[a];
    m[a] = (m[a] - k + n[a]) % n[a];
    const double up = u[(p[0] * n[1] + p[1]) * n[2] + p[2]];
    const double um = u[(m[0] * n[1] + m[1]) * n[2] + m[2]];
    s += w[k - 1] * (up - um);
  }
  return s;
}

__global__ void __launch_bounds__(RED_BLOCK) k_enstrophy(int64_t n0, int64_t n1, int64_t n2, int dim,
                                                         const double* u, double* part) {
  const int64_t n[3] = {n0, n1, n2};
  const int64_t cells = n0 * n1 * n2;
  double acc[RED_N] = {0, 0, 0, 0, 0};
  // lattice axes (x, y, z) -> storage axes: 2D (i1, i2), 3D (i0, i1, i2)
  const int ax = dim == 2 ? 1 : 0, ay = dim == 2 ? 2 : 1, az = 2;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < cells;
       c += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i2 = c % n2, r = c / n2, i1 = r % n1, i0 = r / n1;
    const double* ux = u;
    const double* uy = u + cells;
    const double w3 = d6(ux, i0, i1, i2, n, ay) - d6(uy, i0, i1, i2, n, ax);
    double s = w3 * w3;
    if (dim == 3) {
      const double* uz = u + 2 * cells;
      const double w1 = d6(uz, i0, i1, i2, n, ay) - d6(uy, i0, i1, i2, n, az);
      const double w2 = d6(ux, i0, i1, i2, n, az) - d6(uz, i0, i1, i2, n, ax);
      s += w1 * w1 + w2 * w2;
    }
    acc[0] += s;
  }
  block_sum(acc);
  if (threadIdx.x == 0)
    for (int k = 0; k < RED_N; ++k) part[blockIdx.x * RED_N + k] = acc[k];
}

template <class S, class ST>
__global__ void __launch_bounds__(RED_BLOCK) k_link_force(Grid g, const void* f, double rho0,
                                                          const int32_t* links, int64_t nl,
                                                          double* part) {
  using T = typename ST::T;
  using C = typename ST::C;
  double acc[RED_N] = {0, 0, 0, 0, 0};
  for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < nl;
       l += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = links[2 * l];
    const int d = links[2 * l + 1];
    const int64_t off = cell_offset(g, c);
    double v = 0;
#pragma unroll
    for (int q = 0; q < S::Q; ++q)
      if (q == d) {
        const C sh = ST::shifted ? (C)(S::w[q] * rho0) : C(0);
        v = (double)ST::ld((const T*)f + q * g.sq + off, 0, sh);
        acc[0] += 2.0 * v * S::e[q][0];
        acc[1] += 2.0 * v * S::e[q][1];
        acc[2] += 2.0 * v * S::e[q][2];
      }
  }
  block_sum(acc);
  if (threadIdx.x == 0)
    for (int k = 0; k < RED_N; ++k) part[blockIdx.x * RED_N + k] = acc[k];
}

Grid make_grid(const int64_t n[3], const int64_t stride[4]) {
  Grid g;
  g.n0 = n[0];
  g.n1 = n[1];
  g.n2 = n[2];
  g.sq = stride[0];
  g.s0 = stride[1];
  g.s1 = stride[2];
  g.cells = n[0] * n[1] * n[2];
  return g;
}

unsigned grid_blocks(int64_t cells, int block, int maxb) {
  int64_t b = (cells + block - 1) / block;
  if (b > maxb) b = maxb;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// dispatch helpers over (stencil, storage)
template <template <class, class> class F, class... A>
int with_lattice(int stencil, int storage, A... args) {
#define LBM_CASE_ST(S)                                                 \
  switch (storage) {                                                   \
    case LBM_F32: return F<S, StF32>::run(args...);                    \
    case LBM_F64: return F<S, StF64>::run(args...);                    \
    case LBM_F16_SHIFTED: return F<S, StF16>::run(args...);            \
    case LBM_BF16_SHIFTED: return F<S, StBF16>::run(args...);          \
  }                                                                    \
  break;
  switch (stencil) {
    case LBM_D2Q9: LBM_CASE_ST(D2Q9)
    case LBM_D3Q19: LBM_CASE_ST(D3Q19)
    case LBM_D3Q27: LBM_CASE_ST(D3Q27)
  }
#undef LBM_CASE_ST
  return fail(LBM_E_ARG, "unknown stencil/storage %d/%d", stencil, storage);
}

template <class S, class ST>
struct EqRun {
  static int run(Grid g, const void* rho, const void* u, double rho0, void* f, hipStream_t s) {
    const unsigned nb = grid_blocks(g.cells, 256, 65536);
    if (std::is_same<ST, StF64>::value)
      hipLaunchKernelGGL((k_equilibrium<S, ST, double>), dim3(nb), dim3(256), 0, s, g,
                         (const double*)rho, (const double*)u, rho0, f);
    else
      hipLaunchKernelGGL((k_equilibrium<S, ST, float>), dim3(nb), dim3(256), 0, s, g,
                         (const float*)rho, (const float*)u, rho0, f);
    return check_launch("lbm_equilibrium");
  }
};

template <class S, class ST>
struct DecodeRun {
  static int run(Grid g, const void* f, double rho0, void* out, int f64, hipStream_t s) {
    const unsigned nb = grid_blocks(g.cells, 256, 65536);
    if (f64)
      hipLaunchKernelGGL((k_decode<S, ST, double>), dim3(nb), dim3(256), 0, s, g, f, rho0, (double*)out);
    else
      hipLaunchKernelGGL((k_decode<S, ST, float>), dim3(nb), dim3(256), 0, s, g, f, rho0, (float*)out);
    return check_launch("lbm_decode");
  }
};

template <class S, class ST>
struct EncodeRun {
  static int run(Grid g, const void* in, int f64, double rho0, void* f, hipStream_t s) {
    const unsigned nb = grid_blocks(g.cells, 256, 65536);
    if (f64)
      hipLaunchKernelGGL((k_encode<S, ST, double>), dim3(nb), dim3(256), 0, s, g, (const double*)in, rho0, f);
    else
      hipLaunchKernelGGL((k_encode<S, ST, float>), dim3(nb), dim3(256), 0, s, g, (const float*)in, rho0, f);
    return check_launch("lbm_encode");
  }
};

template <class S, class ST>
struct MomRun {
  static int run(Grid g, const void* f, double rho0, double* rho, double* u, hipStream_t s) {
    const unsigned nb = grid_blocks(g.cells, 256, 65536);
    hipLaunchKernelGGL((k_moments<S, ST>), dim3(nb), dim3(256), 0, s, g, f, rho0, rho, u);
    return check_launch("lbm_moments_f64");
  }
};

template <class S, class ST>
struct RedRun {
  static int run(Grid g, const void* f, double rho0, const uint8_t* mask, double* out, double* ws,
                 hipStream_t s) {
    const unsigned nb = grid_blocks(g.cells, RED_BLOCK, RED_MAX_BLOCKS);
    hipLaunchKernelGGL((k_reduce<S, ST>), dim3(nb), dim3(RED_BLOCK), 0, s, g, f, rho0, mask, ws);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(RED_BLOCK), 0, s, (const double*)ws, (int)nb, out, RED_N);
    return check_launch("lbm_reduce_f64");
  }
};

template <class S, class ST>
struct ScanRun {
  static int run(Grid g, const void* f, double rho0, double* out, double* ws, hipStream_t s) {
    const unsigned nb = grid_blocks(g.cells, RED_BLOCK, RED_MAX_BLOCKS);
    hipLaunchKernelGGL((k_scan<S, ST>), dim3(nb), dim3(RED_BLOCK), 0, s, g, f, rho0, ws);
    hipLaunchKernelGGL(k_finish_scan, dim3(1), dim3(RED_BLOCK), 0, s, (const double*)ws, (int)nb, out);
    return check_launch("lbm_scan_f64");
  }
};

template <class S, class ST>
struct ForceRun {
  static int run(Grid g, const void* f, double rho0, const int32_t* links, int64_t nl, double* out,
                 double* ws, hipStream_t s) {
    const unsigned nb = grid_blocks(nl, RED_BLOCK, RED_MAX_BLOCKS);
    hipLaunchKernelGGL((k_link_force<S, ST>), dim3(nb), dim3(RED_BLOCK), 0, s, g, f, rho0, links, nl, ws);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(RED_BLOCK), 0, s, (const double*)ws, (int)nb, out, 3);
    return check_launch("lbm_link_force_f64");
  }
};

double shift_for(int storage, double rho0) {
  return (storage == LBM_F16_SHIFTED || storage == LBM_BF16_SHIFTED) ? rho0 : 0.0;
}

}  // namespace

// Per-block partial sums of lbm_reduce_f64_static: a device array of the
// library in place of the caller's workspace, RED_MAX_BLOCKS partials of
// RED_N doubles (81,920 bytes) on every device the library runs on, held
// until the process ends.  Host code must not use this name as a device
// address: each device holds its own copy, whose address
// static_scratch_of_current_device returns.
static __device__ __align__(256) double scratch[RED_MAX_BLOCKS * RED_N];
static_assert(sizeof(scratch) == LBM_REDUCE_STATIC_BYTES, "header size of the static scratch");

namespace {
constexpr int SCRATCH_MAX_DEVICES = 256;
std::atomic<double*> g_scratch[SCRATCH_MAX_DEVICES];
std::mutex g_scratch_mutex;

// The address of `scratch` on the device current on this thread, looked up
// with hipGetSymbolAddress on the first call per device and kept.
int static_scratch_of_current_device(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(LBM_E_HIP, "hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES)
    return fail(LBM_E_UNSUPPORTED, "device %d: the static scratch tracks devices 0..%d", dev,
                SCRATCH_MAX_DEVICES - 1);
  double* p = g_scratch[dev].load(std::memory_order_acquire);
  if (!p) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    p = g_scratch[dev].load(std::memory_order_relaxed);
    if (!p) {
      void* addr = nullptr;
      e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !addr)
        return fail(LBM_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                    hipGetErrorString(e));
      p = (double*)addr;
      g_scratch[dev].store(p, std::memory_order_release);
    }
  }
  *out = p;
  return LBM_OK;
}
}  // namespace

extern "C" {

int lbm_equilibrium(int32_t stencil, int32_t storage, const int64_t n[3], const int64_t stride[4],
                    const void* rho, const void* u, double rho0_shift, void* f, void* stream) {
  int rc = check_layout(stencil, storage, n, stride);
  if (rc) return rc;
  if (!rho || !u || !f) return fail(LBM_E_ARG, "NULL pointer");
  return with_lattice<EqRun>(stencil, storage, make_grid(n, stride), rho, u,
                             shift_for(storage, rho0_shift), f, (hipStream_t)stream);
}

int lbm_decode(int32_t stencil, int32_t storage, const int64_t n[3], const int64_t stride[4],
               const void* f, double rho0_shift, void* out, int32_t out_f64, void* stream) {
  int rc = check_layout(stencil, storage, n, stride);
  if (rc) return rc;
  if (!f || !out) return fail(LBM_E_ARG, "NULL pointer");
  return with_lattice<DecodeRun>(stencil, storage, make_grid(n, stride), f,
                                 shift_for(storage, rho0_shift), out, (int)out_f64, (hipStream_t)stream);
}

int lbm_encode(int32_t stencil, int32_t storage, const int64_t n[3], const int64_t stride[4],
               const void* in, int32_t in_f64, double rho0_shift, void* f, void* stream) {
  int rc = check_layout(stencil, storage, n, stride);
  if (rc) return rc;
  if (!f || !in) return fail(LBM_E_ARG, "NULL pointer");
  return with_lattice<EncodeRun>(stencil, storage, make_grid(n, stride), in, (int)in_f64,
                                 shift_for(storage, rho0_shift), f, (hipStream_t)stream);
}

int lbm_moments_f64(int32_t stencil, int32_t storage, const int64_t n[3], const int64_t stride[4],
                    const void* f, double rho0_shift, double* rho, double* u, void* stream) {
  int rc = check_layout(stencil, storage, n, stride);
  if (rc) return rc;
  if (!f || !rho || !u) return fail(LBM_E_ARG, "NULL pointer");
  return with_lattice<MomRun>(stencil, storage, make_grid(n, stride), f,
                              shift_for(storage, rho0_shift), rho, u, (hipStream_t)stream);
}

int64_t lbm_reduce_workspace_bytes(const int64_t n[3]) {
  (void)n;
  return (int64_t)RED_MAX_BLOCKS * RED_N * (int64_t)sizeof(double);
}

int lbm_reduce_f64(int32_t stencil, int32_t storage, const int64_t n[3], const int64_t stride[4],
                   const void* f, double rho0_shift, const uint8_t* mask, double* out, void* workspace,
                   void* stream) {
  int rc = check_layout(stencil, storage, n, stride);
  if (rc) return rc;
  if (!f || !out || !workspace) return fail(LBM_E_ARG, "NULL pointer");
  return with_lattice<RedRun>(stencil, storage, make_grid(n, stride), f,
                              shift_for(storage, rho0_shift), mask, out, (double*)workspace,
                              (hipStream_t)stream);
}

int lbm_reduce_f64_static(int32_t stencil, int32_t storage, const int64_t n[3],
                          const int64_t stride[4], const void* f, double rho0_shift,
                          const uint8_t* mask, double* out, void* stream) {
  int rc = check_layout(stencil, storage, n, stride);
  if (rc) return rc;
  if (!f || !out) return fail(LBM_E_ARG, "NULL pointer");
  // the partials lbm_reduce_f64 would need; every block writes its own
  // before the final sum reads them, so nothing must start at zero
  const int64_t need = lbm_reduce_workspace_bytes(n);
  if (need > (int64_t)sizeof(scratch))
    return fail(LBM_E_ARG, "the reduction needs %lld bytes of partial sums, above the %lld of the static scratch",
                (long long)need, (long long)sizeof(scratch));
  double* ws = nullptr;
  rc = static_scratch_of_current_device(&ws);
  if (rc) return rc;
  return with_lattice<RedRun>(stencil, storage, make_grid(n, stride), f,
                              shift_for(storage, rho0_shift), mask, out, ws, (hipStream_t)stream);
}

int lbm_scan_f64(int32_t stencil, int32_t storage, const int64_t n[3], const int64_t stride[4],
                 const void* f, double rho0_shift, double* out, void* workspace, void* stream) {
  int rc = check_layout(stencil, storage, n, stride);
  if (rc) return rc;
  if (!f || !out || !workspace) return fail(LBM_E_ARG, "NULL pointer");
  return with_lattice<ScanRun>(stencil, storage, make_grid(n, stride), f,
                               shift_for(storage, rho0_shift), out, (double*)workspace,
                               (hipStream_t)stream);
}

int lbm_enstrophy_f64(int32_t stencil, const int64_t n[3], const double* u, double* out,
                      void* workspace, void* stream) {
  if (stencil_q(stencil) < 0) return fail(LBM_E_ARG, "unknown stencil %d", stencil);
  if (!n || !u || !out || !workspace) return fail(LBM_E_ARG, "NULL pointer");
  const int dim = stencil_d(stencil);
  if (dim == 2 && n[0] != 1) return fail(LBM_E_ARG, "2D lattices need n[0] == 1");
  if (n[0] < 1 || n[1] < 1 || n[2] < 1) return fail(LBM_E_ARG, "extents must be positive");
  const int64_t cells = n[0] * n[1] * n[2];
  const unsigned nb = grid_blocks(cells, RED_BLOCK, RED_MAX_BLOCKS);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_enstrophy, dim3(nb), dim3(RED_BLOCK), 0, s, n[0], n[1], n[2], dim, u,
                     (double*)workspace);
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(RED_BLOCK), 0, s, (const double*)workspace, (int)nb, out, 1);
  return check_launch("lbm_enstrophy_f64");
}

int lbm_link_force_f64(int32_t stencil, int32_t storage, const int64_t n[3], const int64_t stride[4],
                       const void* f, double rho0_shift, const int32_t* links, int64_t n_links,
                       double* out, void* workspace, void* stream) {
  int rc = check_layout(stencil, storage, n, stride);
  if (rc) return rc;
  if (!f || !out || !workspace || (n_links > 0 && !links)) return fail(LBM_E_ARG, "NULL pointer");
  if (n_links < 0) return fail(LBM_E_ARG, "negative link count");
  return with_lattice<ForceRun>(stencil, storage, make_grid(n, stride), f,
                                shift_for(storage, rho0_shift), links, n_links, out, (double*)workspace,
                                (hipStream_t)stream);
}

}  // extern "C"

namespace {
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
template <bool NT>
__global__ void __launch_bounds__(256) k_stream_copy(const u32x4* __restrict__ src,
                                                     u32x4* __restrict__ dst, int64_t n16) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n16) return;
  const u32x4 v = src[i];
  if (NT) __builtin_nontemporal_store(v, dst + i);
  else dst[i] = v;
}
}  // namespace

extern "C" int lbm_stream_copy(const void* src, void* dst, int64_t bytes, int32_t nontemporal,
                               void* stream) {
  if (!src || !dst) return fail(LBM_E_ARG, "NULL pointer");
  if (bytes < 0 || bytes % 16 != 0) return fail(LBM_E_ARG, "bytes must be a non-negative multiple of 16");
  if (((uintptr_t)src | (uintptr_t)dst) % 16 != 0) return fail(LBM_E_ARG, "pointers must be 16-byte aligned");
  const int64_t n16 = bytes / 16;
  if (n16 == 0) return LBM_OK;
  const int64_t blocks = (n16 + 255) / 256;
  if (blocks > 0x7fffffff) return fail(LBM_E_ARG, "copy too large");
  g_err[0] = 0;
  if (nontemporal)
    hipLaunchKernelGGL(k_stream_copy<true>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       (const u32x4*)src, (u32x4*)dst, n16);
  else
    hipLaunchKernelGGL(k_stream_copy<false>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       (const u32x4*)src, (u32x4*)dst, n16);
  return check_launch("lbm_stream_copy");
}

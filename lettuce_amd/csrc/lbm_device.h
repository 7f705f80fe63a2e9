// Device-side building blocks of the lattice-Boltzmann step for gfx950:
// lattice tables, population storage codecs, equilibrium and collision
// operators.  Everything here is computed per cell in registers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "../../include/lettuce_amd.h"

namespace lbm {

// ---------------------------------------------------------------------------
// Lattices.  Direction order, weights and opposites are lettuce's
// (lettuce/ext/_stencil/d2q9.py:8-10, d3q19.py:8-13, d3q27.py:8-12) so that
// population q here is population q of lettuce's flow.f.
// e[q] holds the lattice components (x, y, z); go(q, a) is the grid offset of
// direction q along storage axis a (i0, i1, i2).  2D lattices live on axes
// (i1, i2) = (x, y) with i0 == 0.
// ---------------------------------------------------------------------------
struct D2Q9 {
  static constexpr int Q = 9, D = 2, ID = LBM_D2Q9;
  static constexpr int e[9][3] = {{0, 0, 0},  {1, 0, 0},  {0, 1, 0},
                                  {-1, 0, 0}, {0, -1, 0}, {1, 1, 0},
                                  {-1, 1, 0}, {-1, -1, 0}, {1, -1, 0}};
  static constexpr int opp[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
  static constexpr double w[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,
                                  1.0 / 9.0,  1.0 / 9.0,  1.0 / 36.0,
                                  1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
};

struct D3Q19 {
  static constexpr int Q = 19, D = 3, ID = LBM_D3Q19;
  static constexpr int e[19][3] = {
      {0, 0, 0},  {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},  {0, -1, 0},
      {0, 0, 1},  {0, 0, -1},  {0, 1, 1},  {0, -1, -1}, {0, 1, -1},
      {0, -1, 1}, {1, 0, 1},   {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1},
      {1, 1, 0},  {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}};
  static constexpr int opp[19] = {0,  2,  1,  4,  3,  6,  5,  8,  7, 10,
                                  9, 12, 11, 14, 13, 16, 15, 18, 17};
  static constexpr double w[19] = {
      1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
      1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
      1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
      1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
};

struct D3Q27 {
  static constexpr int Q = 27, D = 3, ID = LBM_D3Q27;
  static constexpr int e[27][3] = {
      {0, 0, 0},   {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},
      {0, 0, 1},   {0, 0, -1},  {0, 1, 1},   {0, -1, -1}, {0, 1, -1},
      {0, -1, 1},  {1, 0, 1},   {-1, 0, -1}, {1, 0, -1},  {-1, 0, 1},
      {1, 1, 0},   {-1, -1, 0}, {1, -1, 0},  {-1, 1, 0},  {1, 1, 1},
      {-1, -1, -1}, {1, 1, -1}, {-1, -1, 1}, {1, -1, 1},  {-1, 1, -1},
      {1, -1, -1}, {-1, 1, 1}};
  static constexpr int opp[27] = {0,  2,  1,  4,  3,  6,  5,  8,  7,
                                  10, 9,  12, 11, 14, 13, 16, 15, 18,
                                  17, 20, 19, 22, 21, 24, 23, 26, 25};
  static constexpr double w[27] = {
      8.0 / 27.0,   2.0 / 27.0,   2.0 / 27.0,   2.0 / 27.0,   2.0 / 27.0,
      2.0 / 27.0,   2.0 / 27.0,   1.0 / 54.0,   1.0 / 54.0,   1.0 / 54.0,
      1.0 / 54.0,   1.0 / 54.0,   1.0 / 54.0,   1.0 / 54.0,   1.0 / 54.0,
      1.0 / 54.0,   1.0 / 54.0,   1.0 / 54.0,   1.0 / 54.0,   1.0 / 216.0,
      1.0 / 216.0,  1.0 / 216.0,  1.0 / 216.0,  1.0 / 216.0,  1.0 / 216.0,
      1.0 / 216.0,  1.0 / 216.0};
};

// grid offset of direction q along storage axis a
template <class S>
__host__ __device__ constexpr int go(int q, int a) {
  return S::D == 2 ? (a == 0 ? 0 : S::e[q][a - 1]) : S::e[q][a];
}

// ---------------------------------------------------------------------------
// Storage codecs.  C is the arithmetic type.  Shifted formats store
// f_q - w_q * rho0 so that the stored values are O(u) deviations.
// ---------------------------------------------------------------------------
// FP32 -> BF16 with round to nearest even: one v_cvt_pk_bf16_f32 on gfx950
// (in place of five integer operations for the carry and the NaN check)
__device__ __forceinline__ uint16_t f32_to_bf16_rne(float x) {
  return __builtin_bit_cast(uint16_t, (__bf16)x);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
  return __uint_as_float(((uint32_t)h) << 16);
}

// Element i of a population array p, where p (the population's base, plus
// the offset of an i0 plane in the kernels' wide variants) is uniform
// across the wave and i is a 32-bit element offset from it.  For formats of
// up to 4 bytes the host keeps i * sizeof(T) below 2^32 (a plane's extent,
// or in the narrow variants the whole array's), so p stays in scalar
// registers and each access takes one 32-bit vector offset (the global_load
// saddr form) instead of a 64-bit address computed per lane.  FP64 keeps
// 64-bit addresses.  The accesses go through global (address space 1)
// pointers.
template <class T>
using gptr = __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gptr<T> elem_ptr(T* p, uint32_t i) {
  if constexpr (sizeof(T) <= 4)
    return (gptr<T>)((gptr<char>)p + i * (uint32_t)sizeof(T));
  else
    return (gptr<T>)p + i;
}
template <class T>
__device__ __forceinline__ T elem(const T* p, uint32_t i) {
  return *elem_ptr(const_cast<T*>(p), i);
}
// p kept whole in scalar registers: in a loop whose in-plane offsets do not
// change (the two-step kernel's sweep over planes), the compiler would
// otherwise hoist base + offset out of the loop as one 64-bit vector
// register pair per population and add the plane offset per access
template <class T>
__device__ __forceinline__ T* scalar_base(T* p) {
  uint64_t v = (uint64_t)p;
  asm volatile("" : "+s"(v));
  return (T*)v;
}

// Stochastic rounding of shifted BF16 (opt-in per run): the 16 bits that
// the conversion to BF16 drops are replaced by 16 random bits added before
// truncation, so a value rounds away from zero with a probability equal to
// the dropped fraction and the rounding is unbiased on average.  Rounding to
// nearest, applied step after step to smooth fields, biases the kinetic
// energy.  The bits are a function of the run's seed, the step, the stored
// element's offset and the population, so every kernel that stores the same
// element at the same step draws the same bits.  Hash: lowbias32
// (C. Wellons, "Prospecting for hash functions", 2018).
__host__ __device__ constexpr uint32_t lowbias32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
// uniform per launch: the key of step `step` of a run seeded with `seed`
__host__ __device__ constexpr uint32_t sr_step_key(uint32_t seed, uint32_t step) {
  return lowbias32(seed ^ lowbias32(step + 0x9e3779b9u));
}
// the element's 32 random bits at a step
__device__ __forceinline__ uint32_t sr_cell(uint32_t key, uint32_t elem) {
  return lowbias32(elem ^ key);
}
// population q's 16 bits from the element's hash: bits 16-31 of the product
// of its low 24 bits with an odd 24-bit constant per population (distinct
// for distinct q).  24-bit operands make the product one full-rate
// v_mul_u32_u24; a 32-bit product is a quarter-rate instruction, and with
// one per stored value it was the largest cost of the rounding.
__device__ __forceinline__ uint32_t sr_bits(uint32_t h, int q) {
  const uint32_t c = ((0x9e3779b1u + 0x3c6ef372u * (uint32_t)q) & 0xffffffu) | 1u;
  return ((h & 0xffffffu) * c) >> 16;
}

// dec: stored value -> population, enc: population -> stored value;
// enc_sr: the same with stochastic rounding from 16 random bits r (formats
// other than BF16 ignore r); ld / st: one element, stores bypassing the
// caches (streamed output)
struct StF32 {
  using T = float;
  using C = float;
  static constexpr bool shifted = false;
  static constexpr bool stochastic = false;
  __device__ static C val(T x) { return x; }  // without the shift
  __device__ static C dec(T x, C) { return x; }
  __device__ static T enc(C v, C) { return v; }
  __device__ static T enc_sr(C v, C, uint32_t) { return v; }
  // two populations at once (the 16-bit formats convert a pair in one
  // instruction)
  __device__ static void enc2(C a, C b, C, C, T (&e)[2]) {
    e[0] = a;
    e[1] = b;
  }
  __device__ static C ld(const T* p, uint32_t i, C) { return elem(p, i); }
  __device__ static void st(T* p, uint32_t i, C v, C) {
    __builtin_nontemporal_store(v, elem_ptr(p, i));
  }
};
struct StF64 {
  using T = double;
  using C = double;
  static constexpr bool shifted = false;
  static constexpr bool stochastic = false;
  __device__ static C val(T x) { return x; }  // without the shift
  __device__ static C dec(T x, C) { return x; }
  __device__ static T enc(C v, C) { return v; }
  __device__ static T enc_sr(C v, C, uint32_t) { return v; }
  __device__ static C ld(const T* p, uint32_t i, C) { return p[i]; }
  __device__ static void st(T* p, uint32_t i, C v, C) {
    __builtin_nontemporal_store(v, p + i);
  }
};
struct StF16 {
  using T = _Float16;
  using C = float;
  static constexpr bool shifted = true;
  static constexpr bool stochastic = false;
  __device__ static C val(T x) { return (float)x; }  // without the shift
  __device__ static C dec(T x, C sh) { return (float)x + sh; }
  __device__ static T enc(C v, C sh) { return (_Float16)(v - sh); }
  __device__ static T enc_sr(C v, C sh, uint32_t) { return enc(v, sh); }
  __device__ static void enc2(C a, C b, C sha, C shb, T (&e)[2]) {
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    const h2 v = {(_Float16)(a - sha), (_Float16)(b - shb)};
    uint32_t u = __builtin_bit_cast(uint32_t, v);
    asm volatile("" : "+v"(u));
    e[0] = __builtin_bit_cast(_Float16, (uint16_t)u);
    e[1] = __builtin_bit_cast(_Float16, (uint16_t)(u >> 16));
  }
  __device__ static C ld(const T* p, uint32_t i, C sh) {
    return (float)elem(p, i) + sh;
  }
  __device__ static void st(T* p, uint32_t i, C v, C sh) {
    __builtin_nontemporal_store((_Float16)(v - sh), elem_ptr(p, i));
  }
};
struct StBF16 {
  using T = uint16_t;
  using C = float;
  static constexpr bool shifted = true;
  static constexpr bool stochastic = false;
  __device__ static C val(T x) { return bf16_to_f32(x); }  // without the shift
  __device__ static C dec(T x, C sh) { return bf16_to_f32(x) + sh; }
  __device__ static T enc(C v, C sh) { return f32_to_bf16_rne(v - sh); }
  // r < 2^16; infinities and NaN round to nearest (adding to their bits
  // could turn one into the other), by a select, not a branch
  __device__ static T enc_sr(C v, C sh, uint32_t r) {
    const float x = v - sh;
    const uint32_t u = __float_as_uint(x);
    const uint16_t near = f32_to_bf16_rne(x);
    const uint16_t sto = (uint16_t)((u + r) >> 16);
    // class mask: signalling and quiet NaN, -inf, +inf
    return __builtin_amdgcn_classf(x, 0x207) ? near : sto;
  }
  __device__ static void enc2(C a, C b, C sha, C shb, T (&e)[2]) {
    typedef __bf16 b2 __attribute__((ext_vector_type(2)));
    const b2 v = {(__bf16)(a - sha), (__bf16)(b - shb)};
    // opaque: split into two halves, the pair would be converted one by one
    uint32_t u = __builtin_bit_cast(uint32_t, v);
    asm volatile("" : "+v"(u));
    e[0] = (uint16_t)u;
    e[1] = (uint16_t)(u >> 16);
  }
  __device__ static C ld(const T* p, uint32_t i, C sh) {
    return bf16_to_f32(elem(p, i)) + sh;
  }
  __device__ static void st(T* p, uint32_t i, C v, C sh) {
    __builtin_nontemporal_store(f32_to_bf16_rne(v - sh), elem_ptr(p, i));
  }
};
// BF16 storage whose kernels store with stochastic rounding (enc_sr with
// bits keyed to the run's seed, the step and the element).  A codec of its
// own, so the kernels that round to nearest carry none of its code.
struct StBF16SR : StBF16 {
  static constexpr bool stochastic = true;
};
// Kernel-table id of StBF16SR (the public storage ids are 0-3): what the
// dispatchers take for BF16_SHIFTED storage with a non-zero sr_seed.
constexpr int kCodecBF16SR = 16;
__host__ __device__ inline int codec_id(int storage, uint32_t sr_seed) {
  return storage == LBM_BF16_SHIFTED && sr_seed != 0 ? kCodecBF16SR : storage;
}

// ---------------------------------------------------------------------------
// Moments and the quadratic equilibrium
// (lettuce/ext/_equilibrium/quadratic_equilibrium.py:11-24 with cs^2 = 1/3:
//  feq = w rho ((2 e.u - u.u) / (2 cs^2) + (e.u / cs^2)^2 / 2 + 1)).
// ---------------------------------------------------------------------------
// 1 / x.  FP32: v_rcp_f32 (1 ulp) refined by one Newton step,
// r + r (1 - x r), which leaves it within rounding of the correctly
// rounded quotient at 3 instructions instead of the ~10 of a division (a
// bare v_rcp_f32 biased FP32 kinetic energies by up to 6e-5 over 1000
// steps).  FP64: a correctly rounded division.
__device__ __forceinline__ float recip(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  return __builtin_fmaf(r, __builtin_fmaf(-x, r, 1.0f), r);
}
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }
// 1 / x within 1 ulp (FP32: a bare v_rcp_f32; FP64: a division), for KBC's
// entropic weights 1 / feq.  gamma is a quotient of two sums that carry the
// same weights, so their 1-ulp errors shift it by about that much, against
// the per-term rounding the sums add anyway; the Newton step cost 2 VALU
// instructions per population.  Measured: D3Q27 KBC BF16 256^3 46.7k ->
// 48.7k MLUPS, the 2048 x 512 channel with KBC (FP32) 66.5k -> 69.4k.
__device__ __forceinline__ float recip_1ulp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double recip_1ulp(double x) { return 1.0 / x; }

template <class S, class C>
__device__ __forceinline__ void moments(const C (&f)[S::Q], C& rho, C (&u)[3]) {
  C r = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
  for (int q = 0; q < S::Q; ++q) {
    r += f[q];
    if (S::e[q][0] != 0) jx += S::e[q][0] > 0 ? f[q] : -f[q];
    if (S::e[q][1] != 0) jy += S::e[q][1] > 0 ? f[q] : -f[q];
    if (S::D == 3 && S::e[q][2] != 0) jz += S::e[q][2] > 0 ? f[q] : -f[q];
  }
  rho = r;
  const C ri = recip(r);
  u[0] = jx * ri;
  u[1] = jy * ri;
  u[2] = S::D == 3 ? jz * ri : C(0);
}

template <class S, class C>
__device__ __forceinline__ C eu_of(int q, const C (&u)[3]) {
  C s = 0;
  if (S::e[q][0] != 0) s += S::e[q][0] > 0 ? u[0] : -u[0];
  if (S::e[q][1] != 0) s += S::e[q][1] > 0 ? u[1] : -u[1];
  if (S::D == 3 && S::e[q][2] != 0) s += S::e[q][2] > 0 ? u[2] : -u[2];
  return s;
}

template <class S, class C>
__device__ __forceinline__ C feq_q(int q, C rho, const C (&u)[3], C uu) {
  const C eu = eu_of<S, C>(q, u);
  return C(S::w[q]) * rho * (C(1.5) * (C(2) * eu - uu) + C(4.5) * eu * eu + C(1));
}

// all Q values at once: opposite populations share the even part
// w rho (c + 4.5 (e.u)^2), c = 1 - 1.5 u.u, and differ in the sign of the
// odd part 3 w rho (e.u)
template <class S, class C>
__device__ __forceinline__ void equilibrium(C rho, const C (&u)[3], C (&feq)[S::Q]) {
  const C uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const C c = C(1) - C(1.5) * uu;
#pragma unroll
  for (int q = 0; q < S::Q; ++q) {
    const int o = S::opp[q];
    if (o < q) continue;
    const C wr = C(S::w[q]) * rho;
    if (o == q) {
      feq[q] = wr * c;
      continue;
    }
    const C eu = eu_of<S, C>(q, u);
    const C even = wr * (c + C(4.5) * eu * eu);
    const C odd = C(3) * wr * eu;
    feq[q] = even + odd;
    feq[o] = even - odd;
  }
}

// ---------------------------------------------------------------------------
// Collision operators (the cell's own populations, in place)
// ---------------------------------------------------------------------------
template <class C>
struct CollParams {
  C tau_inv;      // BGK: 1 / tau
  C trt_p;        // TRT: 1 / (2 tau_plus)
  C trt_m;        // TRT: 1 / (2 tau_minus)
  C beta;         // KBC: 1 / (2 tau)
  C beta_inv;     // KBC: 2 tau
  C accel[3];     // BGK_GUO: acceleration (lattice units)
  C guo;          // BGK_GUO: 1 - 1 / (2 force_tau)
};

// BGK: f - (f - feq) / tau  (lettuce/ext/_collision/bgk_collision.py:17-22),
// evaluated as (1 - w) f + w feq with w = 1 / tau and, for each pair of
// opposite populations (e.u of one is -e.u of the other),
//   w feq = w_q rho w (c + 4.5 (e.u)^2) +- 3 w_q rho w (e.u),  c = 1 - 1.5 u.u
// so the even and odd parts of the equilibrium are formed once per pair.
template <class S, class C>
__device__ __forceinline__ void collide_bgk(C (&f)[S::Q], const CollParams<C>& cp) {
  C rho, u[3];
  moments<S, C>(f, rho, u);
  const C uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const C om = cp.tau_inv, keep = C(1) - om;
  const C a = om * rho;
  const C c = C(1) - C(1.5) * uu;
#pragma unroll
  for (int q = 0; q < S::Q; ++q) {
    const int o = S::opp[q];
    if (o < q) continue;  // the pair was relaxed at its first index
    const C wa = C(S::w[q]) * a;
    if (o == q) {
      f[q] = keep * f[q] + wa * c;
      continue;
    }
    const C eu = eu_of<S, C>(q, u);
    const C even = wa * (c + C(4.5) * eu * eu);
    const C odd = C(3) * wa * eu;
    const C fq = f[q], fo = f[o];
    f[q] = keep * fq + (even + odd);
    f[o] = keep * fo + (even - odd);
  }
}

// BGK with Guo forcing (lettuce/ext/_collision/bgk_collision.py:17-22,
// lettuce/ext/_force/guo.py:15-36): u* = j / rho + a / (2 rho),
// f - (f - feq(rho, u*)) / tau + guo w_q (3 (e_q - u*) + 9 (e_q.u*) e_q) . a
template <class S, class C>
__device__ __forceinline__ void collide_bgk_guo(C (&f)[S::Q], const CollParams<C>& cp) {
  C rho, u[3];
  moments<S, C>(f, rho, u);
  const C half_ri = C(0.5) / rho;
#pragma unroll
  for (int d = 0; d < 3; ++d) u[d] += cp.accel[d] * half_ri;
  const C uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const C ua = u[0] * cp.accel[0] + u[1] * cp.accel[1] + u[2] * cp.accel[2];
#pragma unroll
  for (int q = 0; q < S::Q; ++q) {
    const C fe = feq_q<S, C>(q, rho, u, uu);
    const C eu = eu_of<S, C>(q, u);
    const C ea = eu_of<S, C>(q, cp.accel);
    const C si = cp.guo * C(S::w[q]) * (C(3) * (ea - ua) + C(9) * eu * ea);
    f[q] = f[q] - cp.tau_inv * (f[q] - fe) + si;
  }
}

// TRT (lettuce/ext/_collision/trt_collision.py:16-27):
// f - [(f+fb)-(feq+feqb)]/(2 tau+) - [(f-fb)-(feq-feqb)]/(2 tau-), fb = f[opp]
template <class S, class C>
__device__ __forceinline__ void collide_trt(C (&f)[S::Q], const CollParams<C>& cp) {
  C rho, u[3], fe[S::Q];
  moments<S, C>(f, rho, u);
  equilibrium<S, C>(rho, u, fe);
  C g[S::Q];
#pragma unroll
  for (int q = 0; q < S::Q; ++q) {
    const int o = S::opp[q];
    const C dp = (f[q] + f[o]) - (fe[q] + fe[o]);
    const C dm = (f[q] - f[o]) - (fe[q] - fe[o]);
    g[q] = f[q] - dp * cp.trt_p - dm * cp.trt_m;
  }
#pragma unroll
  for (int q = 0; q < S::Q; ++q) f[q] = g[q];
}

// KBC shear part Delta s from the raw second-order moments of f - feq.
// lettuce normalises the moments by rho and multiplies back by m000
// (kbc_collision.py:22-39, :41-94); in raw moments the s-part is:
//  2D: s0 = -T, s1 = s3 = (T+N)/4, s2 = s4 = (T-N)/4, s5 = s7 = Pxy/4,
//      s6 = s8 = -Pxy/4                                          (:76-94)
//  3D: s0 = -T, s1,2 = (2Nxz - Nyz + T)/6, s3,4 = (2Nyz - Nxz + T)/6,
//      s5,6 = (-Nxz - Nyz + T)/6, s7,8 = Pyz/4, s9,10 = -Pyz/4,
//      s11,12 = Pxz/4, s13,14 = -Pxz/4, s15,16 = Pxy/4, s17,18 = -Pxy/4,
//      corners s19..26 = 0                                       (:41-74)
template <class S, class C>
__device__ __forceinline__ void kbc_shear(const C (&df)[S::Q], C (&ds)[S::Q]) {
  if constexpr (S::D == 2) {
    C xx = 0, yy = 0, xy = 0;
#pragma unroll
    for (int q = 1; q < S::Q; ++q) {
      const int ex = S::e[q][0], ey = S::e[q][1];
      if (ex != 0) xx += df[q];
      if (ey != 0) yy += df[q];
      if (ex * ey != 0) xy += (ex * ey > 0) ? df[q] : -df[q];
    }
    const C T = xx + yy, N = xx - yy;
    ds[0] = -T;
    ds[1] = C(0.25) * (T + N);
    ds[2] = C(0.25) * (T - N);
    ds[3] = ds[1];
    ds[4] = ds[2];
    ds[5] = C(0.25) * xy;
    ds[6] = -ds[5];
    ds[7] = ds[5];
    ds[8] = -ds[5];
  } else {
    C xx = 0, yy = 0, zz = 0, xy = 0, xz = 0, yz = 0;
#pragma unroll
    for (int q = 1; q < S::Q; ++q) {
      const int ex = S::e[q][0], ey = S::e[q][1], ez = S::e[q][2];
      if (ex != 0) xx += df[q];
      if (ey != 0) yy += df[q];
      if (ez != 0) zz += df[q];
      if (ex * ey != 0) xy += (ex * ey > 0) ? df[q] : -df[q];
      if (ex * ez != 0) xz += (ex * ez > 0) ? df[q] : -df[q];
      if (ey * ez != 0) yz += (ey * ez > 0) ? df[q] : -df[q];
    }
    const C T = xx + yy + zz, Nxz = xx - zz, Nyz = yy - zz;
    const C sixth = C(1.0 / 6.0);
    ds[0] = -T;
    ds[1] = sixth * (C(2) * Nxz - Nyz + T);
    ds[2] = ds[1];
    ds[3] = sixth * (C(2) * Nyz - Nxz + T);
    ds[4] = ds[3];
    ds[5] = sixth * (-Nxz - Nyz + T);
    ds[6] = ds[5];
    ds[7] = C(0.25) * yz;
    ds[8] = ds[7];
    ds[9] = -ds[7];
    ds[10] = ds[9];
    ds[11] = C(0.25) * xz;
    ds[12] = ds[11];
    ds[13] = -ds[11];
    ds[14] = ds[13];
    ds[15] = C(0.25) * xy;
    ds[16] = ds[15];
    ds[17] = -ds[15];
    ds[18] = ds[17];
#pragma unroll
    for (int q = 19; q < S::Q; ++q) ds[q] = C(0);
  }
}

// KBC (lettuce/ext/_collision/kbc_collision.py:96-160):
// dh = f - feq - ds, gamma = 1/beta - (2 - 1/beta) <ds dh/feq> / <dh dh/feq>
// (gamma < 1e-15 or NaN -> 2), f' = f - beta (2 ds + gamma dh)
template <class S, class C>
__device__ __forceinline__ C kbc_gamma(const C (&ds)[S::Q], const C (&dh)[S::Q],
                                       const C (&fe)[S::Q], const CollParams<C>& cp) {
  C sum_s = 0, sum_h = 0;
#pragma unroll
  for (int q = 0; q < S::Q; ++q) {
    const C r = dh[q] * recip_1ulp(fe[q]);
    sum_s += ds[q] * r;
    sum_h += dh[q] * r;
  }
  C gamma = cp.beta_inv - (C(2) - cp.beta_inv) * sum_s / sum_h;
  if (!(gamma >= C(1e-15))) gamma = C(2);  // also catches NaN
  return gamma;
}

template <class S, class C>
__device__ __forceinline__ void collide_kbc(C (&f)[S::Q], const CollParams<C>& cp) {
  C rho, u[3], fe[S::Q], df[S::Q], ds[S::Q], dh[S::Q];
  moments<S, C>(f, rho, u);
  equilibrium<S, C>(rho, u, fe);
#pragma unroll
  for (int q = 0; q < S::Q; ++q) df[q] = f[q] - fe[q];
  kbc_shear<S, C>(df, ds);
#pragma unroll
  for (int q = 0; q < S::Q; ++q) dh[q] = df[q] - ds[q];
  const C gamma = kbc_gamma<S, C>(ds, dh, fe, cp);
#pragma unroll
  for (int q = 0; q < S::Q; ++q) f[q] = f[q] - cp.beta * (C(2) * ds[q] + gamma * dh[q]);
}

}  // namespace lbm

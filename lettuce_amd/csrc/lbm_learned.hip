// Learned collision for D2Q9, fused into the stream + collide step.
//
// Per cell: KBC split of df = f - feq into the shear part ds and the
// higher-order part dh (raw moments xxy, xyy, xxyy only); a 6 -> 32 -> 32 -> 3
// SiLU MLP maps the six non-equilibrium moments to the relaxation rates of
// xxy, xyy, xxyy:
//   f' = f - ds / tau - sum_g r_g m_g(dh) phi_g.
// Cells whose raw rates leave [stable_lo, stable_hi] (or are NaN) take the
// KBC update (lettuce/ext/_collision/kbc_collision.py:96-160).
//
// The MLP runs on v_mfma_f32_32x32x16_bf16.  A wave owns 64 cells, one per
// lane, and treats them as two 32-cell blocks (lanes 0-31 = block A, lanes
// 32-63 = block B).  Each layer is H^T = W X^T with the cells as the N
// dimension, so the output of a layer (D operand, 16 rows per lane) is
// already the B operand of the next layer once the weights' K index is
// permuted to match (krow below); no lane shuffles are needed.  Block A
// feeds its features in K rows 0-7 and block B in rows 8-15 against a first
// layer duplicated in both halves, and the last layer places its three
// outputs in rows 0-2 and 4-6, so every lane ends up holding the rates of
// its own cell.
#include <string.h>

#include "lbm_step_impl.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace lbm {

// 32x32 output layout: item r of lane l holds row drow(l, r), column l % 32
__host__ __device__ constexpr int drow(int l, int r) {
  return 8 * (r / 4) + 4 * (l / 32) + r % 4;
}
// K index 8 * (l / 32) + t of K-chunk c, as a hidden unit: the row that lane
// l holds in item 8 c + t of the previous layer's output
__host__ __device__ constexpr int krow(int c, int l, int t) { return drow(l, 8 * c + t); }

struct LearnedPacked {
  uint16_t a1[64][8];     // layer 1 A operand (bf16), W1 in both K halves
  uint16_t a2[2][64][8];  // layer 2, K chunks 0 and 1 (bf16)
  uint16_t a3[2][64][8];  // layer 3, outputs in rows 0-2 and 4-6 (bf16)
  float c1[64][16];       // biases in the output layout
  float c2[64][16];
  float c3[64][16];
};
static_assert(sizeof(LearnedPacked) == LBM_LEARNED_PACKED_BYTES, "packed layout size");

static uint16_t bf16_bits_rne(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float silu(float v) {
  return v * __builtin_amdgcn_rcpf(1.0f + __expf(-v));
}

__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) {
  return *reinterpret_cast<const bf16x8*>(p);
}
__device__ __forceinline__ f32x16 ld16(const float* p) {
  return *reinterpret_cast<const f32x16*>(p);
}

// SiLU of an accumulator, split into the two bf16 K chunks of the next layer
__device__ __forceinline__ void activate(const f32x16& d, bf16x8& lo, bf16x8& hi) {
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    lo[t] = (__bf16)silu(d[t]);
    hi[t] = (__bf16)silu(d[8 + t]);
  }
}

struct LearnedArgs {
  const LearnedPacked* packed;
  float in_scale, rate_lo, rate_hi, stable_lo, stable_hi;
  int ntiles;  // 64-cell tiles
  int nt2;     // tiles along i2
};

// D2Q9 dual vectors of the raw moments xxy, xyy, xxyy
__device__ constexpr float PHI[3][9] = {
    {0.f, 0.f, -0.5f, 0.f, 0.5f, 0.25f, 0.25f, -0.25f, -0.25f},
    {0.f, -0.5f, 0.f, 0.5f, 0.f, 0.25f, -0.25f, -0.25f, 0.25f},
    {1.f, -0.5f, -0.5f, -0.5f, -0.5f, 0.25f, 0.25f, 0.25f, 0.25f}};

template <class ST, bool MASK, int MODE>
__global__ void __launch_bounds__(256) learned_kernel(const KParams p, const LearnedArgs la) {
  using S = D2Q9;
  using C = float;
  constexpr bool PULL = (MODE & LBM_PRE_STREAMING) != 0;
  constexpr bool PUSH = (MODE & LBM_POST_STREAMING) != 0;
  const int lane = threadIdx.x & 63;
  const int nwaves = gridDim.x * (blockDim.x >> 6);
  const LearnedPacked* P = la.packed;
  const bf16x8 A1 = ld8(P->a1[lane]);
  const bf16x8 A20 = ld8(P->a2[0][lane]), A21 = ld8(P->a2[1][lane]);
  const bf16x8 A30 = ld8(P->a3[0][lane]), A31 = ld8(P->a3[1][lane]);
  const CollParams<C> cp = coll_params<C>(p);
  const C tau_inv = (C)p.tau_inv;
  const bf16x8 zero8 = {};

  // the tile loop is uniform across the wave: every MFMA runs with all lanes
  for (int tile = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); tile < la.ntiles;
       tile += nwaves) {
    const int b2 = tile % la.nt2;
    const int r = tile / la.nt2;
    const int i1 = r % p.n1, i0 = r / p.n1;
    const int i2 = b2 * 64 + lane;
    int a0[3], a1[3], a2[3];
    axis_offsets(i0, p.n0, p.s0, a0);
    axis_offsets(i1, p.n1, p.s1, a1);
    axis_offsets(i2, p.n2, 1, a2);
    const int cell = a0[1] + a1[1] + a2[1];
    uint32_t nsm = 0, flags = 0, op = LBM_OP_COLLIDE;
    if (MASK) {
      const lbm_cell_class& K = p.ct[p.cls[cell]];
      nsm = K.no_stream;
      flags = K.flags;
      op = K.op;
    }
    C f[9];
    if (PULL) gather_pull<S, ST, MASK>(p, cell, a0, a1, a2, nsm, flags, f);
    else gather_own<S, ST>(p, cell, f);

    // non-equilibrium part and the MLP inputs
    C rho, u[3], fe[9], df[9], ds[9], dh[9];
    moments<S, C>(f, rho, u);
    equilibrium<S, C>(rho, u, fe);
#pragma unroll
    for (int q = 0; q < 9; ++q) df[q] = f[q] - fe[q];
    kbc_shear<S, C>(df, ds);
#pragma unroll
    for (int q = 0; q < 9; ++q) dh[q] = df[q] - ds[q];
    const C xx = df[1] + df[3] + df[5] + df[6] + df[7] + df[8];
    const C yy = df[2] + df[4] + df[5] + df[6] + df[7] + df[8];
    const C xy = df[5] - df[6] + df[7] - df[8];
    const C sc = la.in_scale / rho;
    const C feat[6] = {(xx + yy) * sc, (xx - yy) * sc, xy * sc,
                       (df[5] + df[6] - df[7] - df[8]) * sc,
                       (df[5] - df[6] - df[7] + df[8]) * sc,
                       (df[5] + df[6] + df[7] + df[8]) * sc};
    bf16x8 x8 = zero8;
#pragma unroll
    for (int k = 0; k < 6; ++k) x8[k] = (__bf16)feat[k];
    const bf16x8 bA = lane < 32 ? x8 : zero8;
    const bf16x8 bB = lane < 32 ? zero8 : x8;

    // MLP, two 32-cell blocks per layer
    bf16x8 hA0, hA1, hB0, hB1;
    {
      const f32x16 c1 = ld16(P->c1[lane]);
      f32x16 dA = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A1, bA, c1, 0, 0, 0);
      f32x16 dB = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A1, bB, c1, 0, 0, 0);
      activate(dA, hA0, hA1);
      activate(dB, hB0, hB1);
    }
    {
      const f32x16 c2 = ld16(P->c2[lane]);
      f32x16 dA = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A20, hA0, c2, 0, 0, 0);
      dA = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A21, hA1, dA, 0, 0, 0);
      f32x16 dB = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A20, hB0, c2, 0, 0, 0);
      dB = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A21, hB1, dB, 0, 0, 0);
      activate(dA, hA0, hA1);
      activate(dB, hB0, hB1);
    }
    C o[3];
    {
      const f32x16 c3 = ld16(P->c3[lane]);
      f32x16 dA = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A30, hA0, c3, 0, 0, 0);
      dA = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A31, hA1, dA, 0, 0, 0);
      f32x16 dB = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A30, hB0, c3, 0, 0, 0);
      dB = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A31, hB1, dB, 0, 0, 0);
#pragma unroll
      for (int g = 0; g < 3; ++g) o[g] = lane < 32 ? dA[g] : dB[g];
    }

    const bool out_of_grid = i2 >= p.n2;  // never true: n2 % 64 == 0 is required
    if (!MASK || op == LBM_OP_COLLIDE) {
      C rate[3];
      bool fallback = false;
#pragma unroll
      for (int g = 0; g < 3; ++g) {
        const C raw = tau_inv + o[g];
        if (!(raw >= la.stable_lo && raw <= la.stable_hi)) fallback = true;
        rate[g] = fminf(fmaxf(raw, la.rate_lo), la.rate_hi);
      }
      if (fallback) {
        collide_kbc<S, C>(f, cp);
      } else {
        const C mg[3] = {dh[5] + dh[6] - dh[7] - dh[8], dh[5] - dh[6] - dh[7] + dh[8],
                         dh[5] + dh[6] + dh[7] + dh[8]};
#pragma unroll
        for (int q = 0; q < 9; ++q) {
          C upd = rate[0] * mg[0] * PHI[0][q];
          upd += rate[1] * mg[1] * PHI[1][q];
          upd += rate[2] * mg[2] * PHI[2][q];
          f[q] = f[q] - ds[q] * tau_inv - upd;
        }
      }
    } else {
      apply_boundary_op<S, ST, PULL>(p, op, flags, cell, i0, i1, i2, f);
    }
    if (!out_of_grid) store_cell<S, ST, MASK, PUSH>(p, cell, a0, a1, a2, nsm, flags, f);
  }
}

using LearnedFn = void (*)(const KParams, const LearnedArgs);

template <class ST, bool MASK>
LearnedFn learned_mode(int mode) {
  switch (mode) {
    case 0: return learned_kernel<ST, MASK, 0>;
    case 1: return learned_kernel<ST, MASK, 1>;
    case 2: return learned_kernel<ST, MASK, 2>;
    case 3: return learned_kernel<ST, MASK, 3>;
  }
  return nullptr;
}

template <class ST>
LearnedFn learned_mask(bool mask, int mode) {
  return mask ? learned_mode<ST, true>(mode) : learned_mode<ST, false>(mode);
}

LearnedFn pick_learned_kernel(int storage, bool mask, int mode) {
  switch (storage) {
    case LBM_F32: return learned_mask<StF32>(mask, mode);
    case LBM_F16_SHIFTED: return learned_mask<StF16>(mask, mode);
    case LBM_BF16_SHIFTED: return learned_mask<StBF16>(mask, mode);
  }
  return nullptr;  // FP64 storage: the learned collision computes in FP32
}

int launch_learned(const KParams& p, const lbm_learned* l, int storage, bool mask, int mode,
                   hipStream_t stream, int64_t nsteps, const void* f_in, void* f_out) {
  LearnedFn fn = pick_learned_kernel(storage, mask, mode);
  if (!fn) return LBM_E_UNSUPPORTED;
  LearnedArgs la;
  la.packed = (const LearnedPacked*)l->packed;
  la.in_scale = l->in_scale;
  la.rate_lo = l->rate_lo;
  la.rate_hi = l->rate_hi;
  la.stable_lo = l->stable_lo;
  la.stable_hi = l->stable_hi;
  la.nt2 = p.n2 / 64;
  la.ntiles = p.n0 * p.n1 * la.nt2;
  const int waves_per_block = 4;
  int blocks = (la.ntiles + waves_per_block - 1) / waves_per_block;
  if (blocks > 4096) blocks = 4096;
  KParams q = p;
  const void* bufs[2] = {f_in, f_out};
  for (int64_t k = 0; k < nsteps; ++k) {
    q.fin = bufs[k & 1];
    q.fout = (void*)bufs[(k + 1) & 1];
    hipLaunchKernelGGL(fn, dim3(blocks), dim3(64 * waves_per_block), 0, stream, q, la);
  }
  return LBM_OK;
}

}  // namespace lbm

extern "C" int lbm_learned_pack(const float* w1, const float* b1, const float* w2,
                                const float* b2, const float* w3, const float* b3, void* out) {
  using namespace lbm;
  if (!w1 || !b1 || !w2 || !b2 || !w3 || !b3 || !out) return LBM_E_ARG;
  constexpr int H = LBM_LEARNED_HIDDEN;
  LearnedPacked* P = (LearnedPacked*)out;
  memset(P, 0, sizeof(LearnedPacked));
  // layer 3 weights and bias with the three outputs in rows 0-2 and 4-6
  float w3p[32][H], b3p[32];
  for (int i = 0; i < 32; ++i) {
    const int src = i < 3 ? i : (i >= 4 && i < 7 ? i - 4 : -1);
    b3p[i] = src >= 0 ? b3[src] : 0.f;
    for (int k = 0; k < H; ++k) w3p[i][k] = src >= 0 ? w3[src * H + k] : 0.f;
  }
  for (int l = 0; l < 64; ++l) {
    const int i = l % 32;
    for (int t = 0; t < 8; ++t) {
      P->a1[l][t] = bf16_bits_rne(t < 6 ? w1[i * 6 + t] : 0.f);
      for (int c = 0; c < 2; ++c) {
        P->a2[c][l][t] = bf16_bits_rne(w2[i * H + krow(c, l, t)]);
        P->a3[c][l][t] = bf16_bits_rne(w3p[i][krow(c, l, t)]);
      }
    }
    for (int r = 0; r < 16; ++r) {
      P->c1[l][r] = b1[drow(l, r)];
      P->c2[l][r] = b2[drow(l, r)];
      P->c3[l][r] = b3p[drow(l, r)];
    }
  }
  return LBM_OK;
}

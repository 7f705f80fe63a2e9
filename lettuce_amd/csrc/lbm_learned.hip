// Learned collision for D2Q9 and D3Q27, fused into the stream + collide step.
//
// Per cell: KBC split of df = f - feq into the shear part ds and the
// higher-order part dh; a 6 -> 32 -> 32 -> 3 SiLU MLP maps six
// non-equilibrium moments to three relaxation rates r_g of groups of
// higher-order raw moments of dh:
//   f' = f - ds / tau - sum_g r_g sum_{k in g} m_k(dh) phi_k,
// with phi_k the dual vector of raw moment m_k.
//   D2Q9   inputs T, N, Pxy, xxy, xyy, xxyy of df; groups {xxy}, {xyy},
//          {xxyy} (dh holds only these three moments).
//   D3Q27  inputs T, Nxz, Nyz, Pxy, Pxz, Pyz of df (the second-order
//          moments KBC's shear part carries); groups: the raw moments
//          x^a y^b z^c (a, b, c <= 2) of order a + b + c = 3, of order 4,
//          and of order 5 and 6 (dh holds no moment of order <= 2).
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
#include <stdlib.h>
#include <string.h>

#include <algorithm>

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

__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) {
  return *reinterpret_cast<const bf16x8*>(p);
}
__device__ __forceinline__ f32x16 ld16(const float* p) {
  return *reinterpret_cast<const f32x16*>(p);
}

// The MLP runs in the log2 domain: a layer's accumulator holds
// s = -log2(e) z for the pre-activation z (the factor is folded into the
// features' scale and into the biases), and
//   s / (1 + 2^s) = s * rcp(1 + 2^s) = -log2(e) silu(z)
// is the next layer's input in the same domain.  Since log2(e) ln(2) = 1 the
// weights of layers 2 and 3 need no factor; the last layer's outputs are
// multiplied by -ln(2).  This saves the multiplication of exp(-z) =
// 2^(-log2(e) z) per activation; the bf16 roundings of the inputs and
// activations happen on the scaled values.
__host__ __device__ constexpr double kLog2e = 1.4426950408889634;
__host__ __device__ constexpr double kLn2 = 0.6931471805599453;

// SiLU of an accumulator in the log2 domain, split into the two bf16 K
// chunks of the next layer.  The 16 exponentials go out first, then the 16
// reciprocals, then the products: each transcendental's result is used 16
// instructions later, so no wait states sit between a transcendental and
// its use.
__device__ __forceinline__ void activate(const f32x16& d, bf16x8& lo, bf16x8& hi) {
  float r[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) r[t] = __builtin_amdgcn_exp2f(d[t]);
#pragma unroll
  for (int t = 0; t < 16; ++t) r[t] = __builtin_amdgcn_rcpf(1.0f + r[t]);
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    lo[t] = (__bf16)(d[t] * r[t]);
    hi[t] = (__bf16)(d[8 + t] * r[8 + t]);
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

// lattice-specific parts: the six MLP inputs and the rate-weighted update
template <class S>
struct LearnedModel;

template <>
struct LearnedModel<D2Q9> {
  static constexpr bool kMomentSpace = false;
  __device__ static void features(const float (&df)[9], float sc, float (&x)[6]) {
    const float xx = df[1] + df[3] + df[5] + df[6] + df[7] + df[8];
    const float yy = df[2] + df[4] + df[5] + df[6] + df[7] + df[8];
    const float xy = df[5] - df[6] + df[7] - df[8];
    x[0] = (xx + yy) * sc;
    x[1] = (xx - yy) * sc;
    x[2] = xy * sc;
    x[3] = (df[5] + df[6] - df[7] - df[8]) * sc;
    x[4] = (df[5] - df[6] - df[7] + df[8]) * sc;
    x[5] = (df[5] + df[6] + df[7] + df[8]) * sc;
  }
  __device__ static void update(float (&f)[9], const float (&ds)[9], const float (&dh)[9],
                                const float (&rate)[3], float tau_inv) {
    const float mg[3] = {dh[5] + dh[6] - dh[7] - dh[8], dh[5] - dh[6] - dh[7] + dh[8],
                         dh[5] + dh[6] + dh[7] + dh[8]};
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      float upd = rate[0] * mg[0] * PHI[0][q];
      upd += rate[1] * mg[1] * PHI[1][q];
      upd += rate[2] * mg[2] * PHI[2][q];
      f[q] = f[q] - ds[q] * tau_inv - upd;
    }
  }
};

// D3Q27 works in the basis of raw moments x^a y^b z^c (a, b, c in 0..2), a
// tensor product of the 1D moments (1, x, x^2) at x in {-1, 0, 1}, so
// moments and dual vectors are three 1D transforms:
//   to moments   (v-, v0, v+) -> (v- + v0 + v+, v+ - v-, v+ + v-)
//   back         (m0, m1, m2) -> ((m2 - m1) / 2, m0 - m2, (m1 + m2) / 2)
// KBC's shear part ds carries exactly the second-order moments of df and
// dh = df - ds the moments of order >= 3 (df has none of order 0 and 1), so
//   f' = f - ds / tau - Phi(r . M dh) = f - Phi(w . M df),
//   w = 1/tau on order 2, r_0 / r_1 / r_2 on orders 3 / 4 / 5-6, 0 below,
// and the MLP inputs are second-order entries of the same M df.  M df is
// M f (whose entries 000 and 100 / 010 / 001 are rho and j) minus the
// equilibrium's raw moments in closed form: one forward and one backward
// transform per cell, no per-population equilibrium, ds or dh.
template <>
struct LearnedModel<D3Q27> {
  using S = D3Q27;
  static constexpr bool kMomentSpace = true;
  // t[a][b][c] <- sum_q e_x^a e_y^b e_z^c v_q
  __device__ static void to_moments(const float (&v)[27], float (&t)[3][3][3]) {
#pragma unroll
    for (int q = 0; q < 27; ++q) t[S::e[q][0] + 1][S::e[q][1] + 1][S::e[q][2] + 1] = v[q];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float a = t[0][j][k], b = t[1][j][k], c = t[2][j][k];
        t[0][j][k] = a + b + c;
        t[1][j][k] = c - a;
        t[2][j][k] = c + a;
      }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float a = t[i][0][k], b = t[i][1][k], c = t[i][2][k];
        t[i][0][k] = a + b + c;
        t[i][1][k] = c - a;
        t[i][2][k] = c + a;
      }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const float a = t[i][j][0], b = t[i][j][1], c = t[i][j][2];
        t[i][j][0] = a + b + c;
        t[i][j][1] = c - a;
        t[i][j][2] = c + a;
      }
  }
  // 1D moments sum_x W(x) x^p of the weights W(0) = 2/3, W(+-1) = 1/6, whose
  // products are the D3Q27 weights; P3 = the raw moment abc of w
  __host__ __device__ static constexpr float mu1(int p) {
    return p == 0 ? 1.f : (p % 2 ? 0.f : 1.f / 3.f);
  }
  __host__ __device__ static constexpr float P3(int a, int b, int c) {
    return mu1(a) * mu1(b) * mu1(c);
  }
  // t <- t + M (w rho0): the raw moments of the shift that shifted formats
  // leave out of the stored values, rho0 P3(abc) where a, b, c are all even
  __device__ static void add_shift(float (&t)[3][3][3], double rho0) {
#pragma unroll
    for (int a = 0; a < 3; a += 2)
#pragma unroll
      for (int b = 0; b < 3; b += 2)
#pragma unroll
        for (int c = 0; c < 3; c += 2) t[a][b][c] += (float)(rho0 * P3(a, b, c));
  }
  // t <- t - M feq(rho, u) for the quadratic equilibrium
  // feq = w rho (1 + 3 e.u + 4.5 (e.u)^2 - 1.5 u.u): each raw moment is
  // rho (c0 P3(abc) + 3 sum_d u_d P3(abc + d) + 4.5 sum_dd' u_d u_d' P3(abc + d + d'));
  // orders 0 and 1 are left as they are (relax gives them weight 0)
  __device__ static void subtract_equilibrium(float (&t)[3][3][3], float rho,
                                              const float (&u)[3]) {
    const float uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const float c0 = rho * (1.f - 1.5f * uu);
    const float lx = 3.f * rho * u[0], ly = 3.f * rho * u[1], lz = 3.f * rho * u[2];
    const float h = 4.5f * rho;
    const float qxx = h * u[0] * u[0], qyy = h * u[1] * u[1], qzz = h * u[2] * u[2];
    const float qxy = 2.f * h * u[0] * u[1], qxz = 2.f * h * u[0] * u[2],
                qyz = 2.f * h * u[1] * u[2];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          if (a + b + c <= 1) continue;  // weight 0 in relax
          float m = 0.f;
          if (P3(a, b, c) != 0.f) m += P3(a, b, c) * c0;
          if (P3(a + 1, b, c) != 0.f) m += P3(a + 1, b, c) * lx;
          if (P3(a, b + 1, c) != 0.f) m += P3(a, b + 1, c) * ly;
          if (P3(a, b, c + 1) != 0.f) m += P3(a, b, c + 1) * lz;
          if (P3(a + 2, b, c) != 0.f) m += P3(a + 2, b, c) * qxx;
          if (P3(a, b + 2, c) != 0.f) m += P3(a, b + 2, c) * qyy;
          if (P3(a, b, c + 2) != 0.f) m += P3(a, b, c + 2) * qzz;
          if (P3(a + 1, b + 1, c) != 0.f) m += P3(a + 1, b + 1, c) * qxy;
          if (P3(a + 1, b, c + 1) != 0.f) m += P3(a + 1, b, c + 1) * qxz;
          if (P3(a, b + 1, c + 1) != 0.f) m += P3(a, b + 1, c + 1) * qyz;
          t[a][b][c] -= m;
        }
  }
  // T, Nxz, Nyz, Pxy, Pxz, Pyz of df from its raw moments
  __device__ static void features_of(const float (&t)[3][3][3], float sc, float (&x)[6]) {
    const float xx = t[2][0][0], yy = t[0][2][0], zz = t[0][0][2];
    x[0] = (xx + yy + zz) * sc;
    x[1] = (xx - zz) * sc;
    x[2] = (yy - zz) * sc;
    x[3] = t[1][1][0] * sc;
    x[4] = t[1][0][1] * sc;
    x[5] = t[0][1][1] * sc;
  }
  // f <- f - Phi(w . t).  The 1D backward transform is B' D with
  // D = diag(1, 1/2, 1/2) and B' (m0, m1, m2) -> (m2 - m1, m0 - 2 m2, m1 + m2);
  // the three D factors are powers of two and fold exactly into the weights
  // as 2^-(number of non-zero exponents), which leaves 3 operations per 1D
  // transform instead of 5.  Moments of order 0 and 1 have weight 0.
  __device__ static void relax(float (&f)[27], float (&t)[3][3][3], const float (&rate)[3],
                               float tau_inv) {
    // weights of (order, non-zero exponents) (2,1) (2,2) (3,2) (3,3) (4,2) (4,3) (5|6,3)
    const float w21 = 0.5f * tau_inv, w22 = 0.25f * tau_inv;
    const float w32 = 0.25f * rate[0], w33 = 0.125f * rate[0];
    const float w42 = 0.25f * rate[1], w43 = 0.125f * rate[1];
    const float w53 = 0.125f * rate[2];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const int o = a + b + c, nz = (a > 0) + (b > 0) + (c > 0);
          if (o <= 1) {
            t[a][b][c] = 0.f;
            continue;
          }
          t[a][b][c] *= o == 2 ? (nz == 1 ? w21 : w22)
                        : o == 3 ? (nz == 2 ? w32 : w33)
                        : o == 4 ? (nz == 2 ? w42 : w43)
                                 : w53;
        }
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float m0 = t[0][j][k], m1 = t[1][j][k], m2 = t[2][j][k];
        t[0][j][k] = m2 - m1;
        t[1][j][k] = __builtin_fmaf(-2.f, m2, m0);
        t[2][j][k] = m1 + m2;
      }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float m0 = t[i][0][k], m1 = t[i][1][k], m2 = t[i][2][k];
        t[i][0][k] = m2 - m1;
        t[i][1][k] = __builtin_fmaf(-2.f, m2, m0);
        t[i][2][k] = m1 + m2;
      }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const float m0 = t[i][j][0], m1 = t[i][j][1], m2 = t[i][j][2];
        t[i][j][0] = m2 - m1;
        t[i][j][1] = __builtin_fmaf(-2.f, m2, m0);
        t[i][j][2] = m1 + m2;
      }
#pragma unroll
    for (int q = 0; q < 27; ++q) f[q] -= t[S::e[q][0] + 1][S::e[q][1] + 1][S::e[q][2] + 1];
  }
};

// WIDE: 64-bit i0 plane offsets, for arrays whose offsets (in bytes, for
// formats of up to 4 bytes) do not fit 32 bits
template <class S, class ST, bool MASK, int MODE, bool WIDE>
__global__ void __launch_bounds__(256, 3) learned_kernel(const KParams p, const LearnedArgs la) {
  using C = float;
  constexpr int Q = S::Q;
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
  // the biases (16 per lane and layer) in LDS as [layer][quarter][lane], so
  // that each ds_read_b128 of a wave reads 1 KB without bank conflicts
  __shared__ float4 bias[3][4][64];
  for (int i = threadIdx.x; i < 3 * 4 * 64; i += blockDim.x) {
    const int L = i / 256, c = (i / 64) % 4, l = i % 64;
    const float* src = (L == 0 ? P->c1[l] : (L == 1 ? P->c2[l] : P->c3[l])) + 4 * c;
    bias[L][c][l] = make_float4(src[0], src[1], src[2], src[3]);
  }
  __shared__ lbm_cell_class ctl[MASK ? 256 : 1];  // the cell-class table
  if (MASK) stage_classes(p, ctl);
  __syncthreads();
  auto ld_bias = [&](int L) {
    f32x16 v;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const float4 b = bias[L][c][lane];
      v[4 * c] = b.x;
      v[4 * c + 1] = b.y;
      v[4 * c + 2] = b.z;
      v[4 * c + 3] = b.w;
    }
    return v;
  };

  // D2Q9 pull steps issue the next tile's pull loads before the current
  // tile's MLP (raw, c0: reloaded in place once the current tile has
  // decoded them).  The 2048 x 512 channel, a small grid, is bound by
  // those loads' latency rather than by arithmetic.  D3Q27 with the same
  // prefetch measured slower (learned BF16 256^3: 33.3k against 37.3k MLUPS
  // at two waves per SIMD; spills at three).
  constexpr bool PF = S::Q == 9 && PULL;
  using X = std::conditional_t<WIDE, long long, int>;
  // the offsets of a tile: i0 plane, i1 row, and the lane's cell along i2
  auto coords = [&](int tile, int& i0, int& i1, int& i2, int& b2, X (&a0)[3], int (&a1)[3],
                    int (&a2)[3]) {
    b2 = tile % la.nt2;
    const int r = tile / la.nt2;
    i1 = r % p.n1;
    i0 = r / p.n1;
    i2 = b2 * 64 + lane;
    axis_offsets(i0, p.n0, (X)p.s0, a0);
    axis_offsets(i1, p.n1, p.s1, a1);
    axis_offsets(i2, p.n2, 1, a2);
    return (int)(a0[1] + a1[1] + a2[1]);
  };
  typename ST::T raw[Q];
  uint32_t c0 = 0;
  const int first = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) +
                                                   (threadIdx.x >> 6));
  if constexpr (PF) {
    if (first < la.ntiles) {
      int j0, j1, j2, c2;
      X n0_[3];
      int n1_[3], n2_[3];
      const int ncell = coords(first, j0, j1, j2, c2, n0_, n1_, n2_);
      gather_issue<S, ST, MASK>(p, ncell, n0_, n1_, n2_, raw, c0);
    }
  }

  // the tile loop is uniform across the wave: every MFMA runs with all lanes
  // (readfirstlane tells the compiler so: the tile's i0 plane offset then
  // lives in scalar registers)
  for (int tile = first; tile < la.ntiles; tile += nwaves) {
    int i0, i1, i2, b2;
    X a0[3];
    int a1[3], a2[3];
    const int cell = coords(tile, i0, i1, i2, b2, a0, a1, a2);  // used with boundaries only
    using LM = LearnedModel<S>;
    // D3Q27 with a shifted format works on the stored values g = f - w rho0
    // throughout: the update f' = f - Phi(...) is g' = g - Phi(...), and the
    // shift enters only the moments (add_shift).  The KBC fallback and the
    // boundary ops decode and re-encode.
    constexpr bool RAW = LM::kMomentSpace && ST::shifted;
    auto shift_all = [&](C (&v)[Q], C sign) {
#pragma unroll
      for (int q = 0; q < Q; ++q) v[q] += sign * shift_of<ST>(S::w[q], p);
    };
    uint32_t nsm, flags, op;
    C f[Q];
    if constexpr (PF) {
      gather_finish<S, ST, MASK, RAW>(p, cell, a0, a1, a2, ctl, raw, c0, nsm, flags, op, f);
      const int next = tile + nwaves;
      if (next < la.ntiles) {
        int j0, j1, j2, c2;
        X n0_[3];
        int n1_[3], n2_[3];
        const int ncell = coords(next, j0, j1, j2, c2, n0_, n1_, n2_);
        gather_issue<S, ST, MASK>(p, ncell, n0_, n1_, n2_, raw, c0);
      }
    } else {
      gather_batched<S, ST, MASK, PULL, RAW, true>(p, cell, a0, a1, a2, ctl, nsm, flags, op, f);
    }

    // non-equilibrium part and the MLP inputs
    C feat[6];
    // D3Q27: the raw moments of df (mt); D2Q9: the KBC split (ds, dh)
    C mt[LM::kMomentSpace ? 3 : 1][3][3], ds[LM::kMomentSpace ? 1 : Q], dh[LM::kMomentSpace ? 1 : Q];
    if constexpr (LM::kMomentSpace) {
      // rho and j are moments 000 and 100 / 010 / 001 of f
      LM::to_moments(f, mt);
      if constexpr (RAW) LM::add_shift(mt, p.rho0);
      const C rho = mt[0][0][0], ri = recip(rho);
      const C u[3] = {mt[1][0][0] * ri, mt[0][1][0] * ri, mt[0][0][1] * ri};
      LM::subtract_equilibrium(mt, rho, u);
      LM::features_of(mt, la.in_scale * ri, feat);
    } else {
      C rho, u[3], df[Q], fe[Q];
      moments<S, C>(f, rho, u);
      equilibrium<S, C>(rho, u, fe);
#pragma unroll
      for (int q = 0; q < Q; ++q) df[q] = f[q] - fe[q];
      kbc_shear<S, C>(df, ds);
#pragma unroll
      for (int q = 0; q < Q; ++q) dh[q] = df[q] - ds[q];
      LM::features(df, la.in_scale / rho, feat);
    }
    bf16x8 x8 = zero8;
#pragma unroll
    for (int k = 0; k < 6; ++k) x8[k] = (__bf16)feat[k];
    const bf16x8 bA = lane < 32 ? x8 : zero8;
    const bf16x8 bB = lane < 32 ? zero8 : x8;

    // MLP, two 32-cell blocks per layer
    bf16x8 hA0, hA1, hB0, hB1;
    {
      const f32x16 c1 = ld_bias(0);
      f32x16 dA = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A1, bA, c1, 0, 0, 0);
      f32x16 dB = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A1, bB, c1, 0, 0, 0);
      activate(dA, hA0, hA1);
      activate(dB, hB0, hB1);
    }
    {
      const f32x16 c2 = ld_bias(1);
      f32x16 dA = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A20, hA0, c2, 0, 0, 0);
      dA = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A21, hA1, dA, 0, 0, 0);
      f32x16 dB = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A20, hB0, c2, 0, 0, 0);
      dB = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A21, hB1, dB, 0, 0, 0);
      activate(dA, hA0, hA1);
      activate(dB, hB0, hB1);
    }
    C o[3];
    {
      const f32x16 c3 = ld_bias(2);
      f32x16 dA = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A30, hA0, c3, 0, 0, 0);
      dA = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A31, hA1, dA, 0, 0, 0);
      f32x16 dB = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A30, hB0, c3, 0, 0, 0);
      dB = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A31, hB1, dB, 0, 0, 0);
#pragma unroll
      for (int g = 0; g < 3; ++g) o[g] = (lane < 32 ? dA[g] : dB[g]) * (float)-kLn2;
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
        if constexpr (RAW) shift_all(f, C(1));
        collide_kbc<S, C>(f, cp);
        if constexpr (RAW) shift_all(f, C(-1));
      } else if constexpr (LM::kMomentSpace) {
        LM::relax(f, mt, rate, tau_inv);
      } else {
        LM::update(f, ds, dh, rate, tau_inv);
      }
    } else {
      if constexpr (RAW) shift_all(f, C(1));
      apply_boundary_op<S, ST, PULL>(p, op, flags, cell, i0, i1, i2, f);
      if constexpr (RAW) shift_all(f, C(-1));
    }
    if (out_of_grid) continue;
    if constexpr (!PUSH) {
      // the cell's own slots: a scalar base, the tile's first cell of
      // population q, stepped by the population stride, plus the lane's
      // byte offset, which is the same for every population and tile, so
      // no address takes a vector instruction.  The asm keeps the base in
      // one scalar register pair, stepped between the stores: the gather's
      // per-population bases already fill the scalar registers, and 27 more
      // would spill.  16-bit formats convert two populations per
      // instruction and store the second from the register's high half.
      using T = typename ST::T;
      uint64_t sb = (uint64_t)p.fout +
                    (uint64_t)(a0[1] + (long long)(a1[1] + b2 * 64)) * sizeof(T);
      const uint64_t step = (uint64_t)p.sq * sizeof(T);
      // opaque here, so that its zero extension is not hoisted out of the
      // tile loop as a 64-bit value, which would take a 64-bit vector add
      // per store instead of the 32-bit vector offset of the saddr form
      uint32_t lane_off = (uint32_t)lane * (uint32_t)sizeof(T);
      asm volatile("" : "+v"(lane_off));
      auto slot = [&]() { return (gptr<T>)((gptr<char>)sb + lane_off); };
      auto next = [&]() {
        sb += step;
        asm volatile("" : "+s"(sb));
      };
      if constexpr (ST::stochastic) {
        // stochastic rounding keyed to the step and this cell's offset
        const uint32_t hs = sr_cell(sr_step_key(p.sr_seed, p.sr_step),
                                    (uint32_t)(a0[1] + a1[1] + a2[1]));
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          __builtin_nontemporal_store(
              ST::enc_sr(f[q], RAW ? C(0) : shift_of<ST>(S::w[q], p), sr_bits(hs, q)), slot());
          next();
        }
      } else {
#pragma unroll
        for (int q = 0; q < Q; q += 2) {
          const C sa = RAW ? C(0) : shift_of<ST>(S::w[q], p);
          if (q + 1 < Q) {
            const C sb1 = RAW ? C(0) : shift_of<ST>(S::w[q + 1], p);
            T e[2];
            ST::enc2(f[q], f[q + 1], sa, sb1, e);
            __builtin_nontemporal_store(e[0], slot());
            next();
            __builtin_nontemporal_store(e[1], slot());
            next();
          } else {
            __builtin_nontemporal_store(ST::enc(f[q], sa), slot());
          }
        }
      }
    } else {
      store_cell<S, ST, MASK, PUSH, RAW>(p, cell, a0, a1, a2, nsm, flags, f);
    }
  }
}

using LearnedFn = void (*)(const KParams, const LearnedArgs);

template <class S, class ST, bool MASK, bool WIDE>
LearnedFn learned_mode(int mode) {
  switch (mode) {
    case 0: return learned_kernel<S, ST, MASK, 0, WIDE>;
    case 1: return learned_kernel<S, ST, MASK, 1, WIDE>;
    case 2: return learned_kernel<S, ST, MASK, 2, WIDE>;
    case 3: return learned_kernel<S, ST, MASK, 3, WIDE>;
  }
  return nullptr;
}

template <class S, class ST, bool WIDE>
LearnedFn learned_mask(bool mask, int mode) {
  return mask ? learned_mode<S, ST, true, WIDE>(mode) : learned_mode<S, ST, false, WIDE>(mode);
}

template <class S, bool WIDE>
LearnedFn learned_storage(int storage, bool mask, int mode) {
  switch (storage) {
    case LBM_F32: return learned_mask<S, StF32, WIDE>(mask, mode);
    case LBM_F16_SHIFTED: return learned_mask<S, StF16, WIDE>(mask, mode);
    case LBM_BF16_SHIFTED: return learned_mask<S, StBF16, WIDE>(mask, mode);
    case kCodecBF16SR: return learned_mask<S, StBF16SR, WIDE>(mask, mode);
  }
  return nullptr;  // FP64 storage: the learned collision computes in FP32
}

template <class S>
LearnedFn learned_wide(int storage, bool mask, int mode, bool wide) {
  return wide ? learned_storage<S, true>(storage, mask, mode)
              : learned_storage<S, false>(storage, mask, mode);
}

LearnedFn pick_learned_kernel(int stencil, int storage, bool mask, int mode, bool wide) {
  if (stencil == LBM_D2Q9) return learned_wide<D2Q9>(storage, mask, mode, wide);
  if (stencil == LBM_D3Q27) return learned_wide<D3Q27>(storage, mask, mode, wide);
  return nullptr;
}

int launch_learned(const KParams& p, const lbm_learned* l, int stencil, int storage, bool mask,
                   int mode, bool wide, hipStream_t stream, int64_t nsteps, const void* f_in,
                   void* f_out) {
  LearnedFn fn = pick_learned_kernel(stencil, storage, mask, mode, wide);
  if (!fn) return LBM_E_UNSUPPORTED;
  LearnedArgs la;
  la.packed = (const LearnedPacked*)l->packed;
  la.in_scale = (float)(-kLog2e * (double)l->in_scale);  // the log2 domain
  la.rate_lo = l->rate_lo;
  la.rate_hi = l->rate_hi;
  la.stable_lo = l->stable_lo;
  la.stable_hi = l->stable_hi;
  la.nt2 = p.n2 / 64;
  la.ntiles = p.n0 * p.n1 * la.nt2;
  // Blocks: about three 64-cell tiles per wave (at most 4096 blocks), but
  // at least two rounds of blocks at the kernel's occupancy.  Measured on
  // MI355X: the 2048 x 512 D2Q9 channel (16384 tiles) runs 16 % faster with
  // ~2.7 tiles per wave than with one (the per-block set-up of weights,
  // biases and class flags is amortised), D3Q27 256^3 (262144 tiles) 3 %
  // faster with 4096 blocks than with one round.
  const int waves_per_block = 4;
  const int64_t all = (la.ntiles + waves_per_block - 1) / waves_per_block;
  int64_t blocks = std::min<int64_t>(4096, (la.ntiles + 3 * waves_per_block - 1) /
                                               (3 * waves_per_block));
  int per_cu = 0, dev = 0, cus = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)fn, 64 * waves_per_block,
                                                   0) == hipSuccess)
    blocks = std::max<int64_t>(blocks, std::min<int64_t>(all, 2 * (int64_t)per_cu * cus));
  blocks = std::max<int64_t>(1, std::min(blocks, all));
  KParams q = p;
  const void* bufs[2] = {f_in, f_out};
  for (int64_t k = 0; k < nsteps; ++k) {
    q.fin = bufs[k & 1];
    q.fout = (void*)bufs[(k + 1) & 1];
    q.sr_step = p.sr_step + (uint32_t)k;
    hipLaunchKernelGGL(fn, dim3((unsigned)blocks), dim3(64 * waves_per_block), 0, stream, q, la);
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
    // the biases in the log2 domain
    for (int r = 0; r < 16; ++r) {
      P->c1[l][r] = (float)(-kLog2e * (double)b1[drow(l, r)]);
      P->c2[l][r] = (float)(-kLog2e * (double)b2[drow(l, r)]);
      P->c3[l][r] = (float)(-kLog2e * (double)b3p[drow(l, r)]);
    }
  }
  return LBM_OK;
}

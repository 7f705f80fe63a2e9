// Training kernels of the learned collision (D2Q9 and D3Q27): the forward
// step with the MLP evaluated in the arithmetic type (FP32, or FP64 for
// gradient checks) and its hand-written adjoint.
//
// D2Q9 forward per cell (g = pulled populations, s = in_scale):
//   rho = sum g, u = j / rho, feq, df = g - feq
//   ds = KBC shear part of df (T = xx + yy, N = xx - yy, Pxy), dh = df - ds
//   x  = s / rho (T, N, Pxy, xxy, xyy, xxyy)(df)
//   z1 = W1 x + b1, a1 = silu(z1), z2 = W2 a1 + b2, a2 = silu(z2),
//   o  = W3 a2 + b3, r = clamp(1/tau + o, rate_lo, rate_hi)
//   f' = g - ds / tau - sum_k r_k m_k(dh) phi_k
// (the inference kernel's formula without bf16 rounding and without the
// KBC fallback; the clamp has zero derivative outside [rate_lo, rate_hi]).
//
// D3Q27 forward per cell, in the raw-moment basis t = M g (moments
// x^a y^b z^c, a, b, c <= 2, the inference kernel's formulation):
//   rho = t_000, u = (t_100, t_010, t_001) / rho
//   t  <- t - M feq(rho, u) on orders >= 2 (closed form), so t = M df there
//   x  = s / rho (t200 + t020 + t002, t200 - t002, t020 - t002, t110, t101, t011)
//   the MLP and the clamped rates r as for D2Q9
//   f' = g - Phi(w . t), Phi = M^-1, w = 1/tau on order 2, r_0 / r_1 / r_2
//        on orders 3 / 4 / 5-6, 0 below
//
// Adjoint of one step f_{n+1} = C(S f_n), S the periodic pull:
//   mu(x)         = (dC/dg)^T lambda_{n+1}(x)        per cell, analytic
//   lambda_n,q(y) = mu_q(y + e_q)                    transpose of the pull
// The backward kernel reads lambda_{n+1} either directly or as the gather
// mu_{n+1,q}(x + e_q) of the previous backward step, and writes mu_n and
// the per-cell record (x, a1, a2, dz1, dz2, do) from which the weight
// gradients are sums of outer products over the cells.
#include <string.h>

#include "lbm_step_impl.h"

namespace lbm {

constexpr int HID = LBM_LEARNED_HIDDEN;
constexpr int NPARAM = HID * 6 + HID + HID * HID + HID + 3 * HID + 3;  // 1379
constexpr int NREC = 6 + HID + HID + HID + HID + 3;                    // 137

// offsets of the parameter vector [W1 (32x6), b1, W2 (32x32), b2, W3 (3x32), b3]
constexpr int P_W1 = 0, P_B1 = P_W1 + HID * 6, P_W2 = P_B1 + HID, P_B2 = P_W2 + HID * HID,
              P_W3 = P_B2 + HID, P_B3 = P_W3 + 3 * HID;
// offsets of the per-cell record [x (6), a1, a2, dz1, dz2, do (3)]
constexpr int R_X = 0, R_A1 = 6, R_A2 = R_A1 + HID, R_DZ1 = R_A2 + HID, R_DZ2 = R_DZ1 + HID,
              R_DO = R_DZ2 + HID;
static_assert(P_B3 + 3 == NPARAM, "parameter count");
static_assert(R_DO + 3 == NREC, "record size");

struct TrainArgs {
  const void* params;  // NPARAM values of the arithmetic type
  double in_scale, rate_lo, rate_hi;
};

// D2Q9 raw moment rows T, N, Pxy, xxy, xyy, xxyy (as 9-vectors)
__device__ constexpr int MROW[6][9] = {
    {0, 1, 1, 1, 1, 2, 2, 2, 2},    {0, 1, -1, 1, -1, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 1, -1, 1, -1},  {0, 0, 0, 0, 0, 1, 1, -1, -1},
    {0, 0, 0, 0, 0, 1, -1, -1, 1},  {0, 0, 0, 0, 0, 1, 1, 1, 1}};
__device__ constexpr double PHI_D[3][9] = {
    {0., 0., -0.5, 0., 0.5, 0.25, 0.25, -0.25, -0.25},
    {0., -0.5, 0., 0.5, 0., 0.25, -0.25, -0.25, 0.25},
    {1., -0.5, -0.5, -0.5, -0.5, 0.25, 0.25, 0.25, 0.25}};

template <class C>
__device__ __forceinline__ C sigmoid(C z) {
  return C(1) / (C(1) + exp(-z));
}

// A scheduling barrier after each row of the MLP's matrix-vector products:
// the weights come from LDS, and left alone the scheduler issues their
// loads far ahead of their uses, which needed more than 256 VGPRs in the
// backward kernel (spills to AGPRs and scratch, one wave per SIMD).
__device__ __forceinline__ void row_fence() { __builtin_amdgcn_sched_barrier(0); }

// z1 = W1 x + b1, a1 = silu(z1), z2 = W2 a1 + b2, a2 = silu(z2),
// raw = 1/tau + W3 a2 + b3, r = clamp(raw); with rec set, x, a1 and a2 go to
// the record (stride rstride)
template <class C>
__device__ __forceinline__ void mlp_forward(const C* __restrict__ P, const C (&x)[6], C tau_inv,
                                            C lo, C hi, C (&z1)[HID], C (&z2)[HID], C (&raw)[3],
                                            C (&r)[3], C* rec, int64_t rstride) {
  C a[HID];
#pragma unroll
  for (int i = 0; i < HID; ++i) {
    C z = P[P_B1 + i];
#pragma unroll
    for (int k = 0; k < 6; ++k) z += P[P_W1 + i * 6 + k] * x[k];
    z1[i] = z;
    a[i] = z * sigmoid(z);
    row_fence();
  }
  if (rec) {
#pragma unroll
    for (int k = 0; k < 6; ++k) rec[(R_X + k) * rstride] = x[k];
#pragma unroll
    for (int i = 0; i < HID; ++i) rec[(R_A1 + i) * rstride] = a[i];
  }
#pragma unroll
  for (int i = 0; i < HID; ++i) {
    C z = P[P_B2 + i];
#pragma unroll
    for (int j = 0; j < HID; ++j) z += P[P_W2 + i * HID + j] * a[j];
    z2[i] = z;
    row_fence();
  }
#pragma unroll
  for (int i = 0; i < HID; ++i) a[i] = z2[i] * sigmoid(z2[i]);
  if (rec) {
#pragma unroll
    for (int i = 0; i < HID; ++i) rec[(R_A2 + i) * rstride] = a[i];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    C o = P[P_B3 + k];
#pragma unroll
    for (int j = 0; j < HID; ++j) o += P[P_W3 + k * HID + j] * a[j];
    raw[k] = tau_inv + o;
    r[k] = raw[k] < lo ? lo : (raw[k] > hi ? hi : raw[k]);
    row_fence();
  }
}

// silu'(z) = sigma(z) (1 + z (1 - sigma(z)))
template <class C>
__device__ __forceinline__ C dsilu(C z) {
  const C sg = sigmoid(z);
  return sg * (C(1) + z * (C(1) - sg));
}

// dL/dx from dL/do (go); z2 is overwritten by dL/dz2; with rec set, dz1,
// dz2 and do go to the record
template <class C>
__device__ __forceinline__ void mlp_backward(const C* __restrict__ P, const C (&go)[3],
                                             const C (&z1)[HID], C (&z2)[HID], C (&gx)[6],
                                             C* rec, int64_t rstride) {
#pragma unroll
  for (int i = 0; i < HID; ++i) {
    C v = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) v += P[P_W3 + k * HID + i] * go[k];
    z2[i] = v * dsilu(z2[i]);
    row_fence();
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) gx[k] = 0;
#pragma unroll
  for (int j = 0; j < HID; ++j) {
    C v = 0;
#pragma unroll
    for (int i = 0; i < HID; ++i) v += P[P_W2 + i * HID + j] * z2[i];
    const C gz1 = v * dsilu(z1[j]);
    if (rec) rec[(R_DZ1 + j) * rstride] = gz1;
#pragma unroll
    for (int k = 0; k < 6; ++k) gx[k] += P[P_W1 + j * 6 + k] * gz1;
    row_fence();
  }
  if (rec) {
#pragma unroll
    for (int i = 0; i < HID; ++i) rec[(R_DZ2 + i) * rstride] = z2[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) rec[(R_DO + k) * rstride] = go[k];
  }
}

// ---------------------------------------------------------------------------
// D3Q27 raw moments.  t[a][b][c] is the moment x^a y^b z^c (a, b, c <= 2);
// population q sits at t[e_x + 1][e_y + 1][e_z + 1] before the transform.
// The basis is the tensor product of the 1D moments (1, x, x^2) at
// x in {-1, 0, 1}, so M, M^-1 and their transposes are one 3-point map
// along each of the three axes.
// ---------------------------------------------------------------------------
template <int AXIS, class C, class F>
__device__ __forceinline__ void map_axis(C (&t)[3][3][3], const F& A) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      if constexpr (AXIS == 0) A(t[0][i][j], t[1][i][j], t[2][i][j]);
      else if constexpr (AXIS == 1) A(t[i][0][j], t[i][1][j], t[i][2][j]);
      else A(t[i][j][0], t[i][j][1], t[i][j][2]);
    }
}
template <class C, class F>
__device__ __forceinline__ void map_all(C (&t)[3][3][3], const F& A) {
  map_axis<0>(t, A);
  map_axis<1>(t, A);
  map_axis<2>(t, A);
}

// t <- M v
template <class C>
__device__ __forceinline__ void m27_moments(const C (&v)[27], C (&t)[3][3][3]) {
  using S = D3Q27;
#pragma unroll
  for (int q = 0; q < 27; ++q) t[S::e[q][0] + 1][S::e[q][1] + 1][S::e[q][2] + 1] = v[q];
  map_all(t, [](C& a, C& b, C& c) {  // (v-, v0, v+) -> (v- + v0 + v+, v+ - v-, v+ + v-)
    const C m0 = a + b + c, m1 = c - a, m2 = c + a;
    a = m0, b = m1, c = m2;
  });
}

// v <- v + sign * (M^-1 t); t is overwritten
template <class C>
__device__ __forceinline__ void m27_add_inverse(C (&t)[3][3][3], C sign, C (&v)[27]) {
  using S = D3Q27;
  map_all(t, [](C& a, C& b, C& c) {  // (m0, m1, m2) -> ((m2 - m1)/2, m0 - m2, (m1 + m2)/2)
    const C vm = C(0.5) * (c - b), v0 = a - c, vp = C(0.5) * (b + c);
    a = vm, b = v0, c = vp;
  });
#pragma unroll
  for (int q = 0; q < 27; ++q) v[q] += sign * t[S::e[q][0] + 1][S::e[q][1] + 1][S::e[q][2] + 1];
}

// nu <- (M^-1)^T lam
template <class C>
__device__ __forceinline__ void m27_inverse_transposed(const C (&lam)[27], C (&nu)[3][3][3]) {
  using S = D3Q27;
#pragma unroll
  for (int q = 0; q < 27; ++q) nu[S::e[q][0] + 1][S::e[q][1] + 1][S::e[q][2] + 1] = lam[q];
  map_all(nu, [](C& a, C& b, C& c) {  // (l-, l0, l+) -> (l0, (l+ - l-)/2, (l- + l+)/2 - l0)
    const C n0 = b, n1 = C(0.5) * (c - a), n2 = C(0.5) * (a + c) - b;
    a = n0, b = n1, c = n2;
  });
}

// v <- v + M^T g; g is overwritten
template <class C>
__device__ __forceinline__ void m27_add_transposed(C (&g)[3][3][3], C (&v)[27]) {
  using S = D3Q27;
  map_all(g, [](C& a, C& b, C& c) {  // (g0, g1, g2) -> (g0 - g1 + g2, g0, g0 + g1 + g2)
    const C vm = a - b + c, v0 = a, vp = a + b + c;
    a = vm, b = v0, c = vp;
  });
#pragma unroll
  for (int q = 0; q < 27; ++q) v[q] += g[S::e[q][0] + 1][S::e[q][1] + 1][S::e[q][2] + 1];
}

// 1D moments sum_x W(x) x^p of W(0) = 2/3, W(+-1) = 1/6, whose products are
// the D3Q27 weights, and their products P3(a, b, c) = raw moment abc of w
__host__ __device__ constexpr double mu1d(int p) {
  return p == 0 ? 1.0 : (p % 2 ? 0.0 : 1.0 / 3.0);
}
__host__ __device__ constexpr double p3(int a, int b, int c) { return mu1d(a) * mu1d(b) * mu1d(c); }

// raw moment abc of feq(rho, u) / rho for the quadratic equilibrium
// feq = w rho (1 + 3 e.u + 4.5 (e.u)^2 - 1.5 u.u), and its derivatives in u
template <class C>
__device__ __forceinline__ C meq_unit(int a, int b, int c, const C (&u)[3], C (&du)[3]) {
  const C uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const int ex[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  C m = (C)p3(a, b, c) * (C(1) - C(1.5) * uu);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const C pd = (C)p3(a + ex[d][0], b + ex[d][1], c + ex[d][2]);
    m += C(3) * u[d] * pd;
    du[d] = C(-3) * u[d] * (C)p3(a, b, c) + C(3) * pd;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      const C pde = (C)p3(a + ex[d][0] + ex[e][0], b + ex[d][1] + ex[e][1], c + ex[d][2] + ex[e][2]);
      m += C(4.5) * u[d] * u[e] * pde;
      du[d] += C(9) * u[e] * pde;
    }
  }
  return m;
}

// relaxation weight of moment abc: 1/tau on order 2, r_0 / r_1 / r_2 on
// orders 3 / 4 / 5-6, 0 on orders 0 and 1; g = the rate group (-1: none)
__device__ __forceinline__ int m27_group(int a, int b, int c) {
  const int o = a + b + c;
  return o <= 2 ? -1 : (o == 3 ? 0 : (o == 4 ? 1 : 2));
}

template <class C>
struct CellFwd27 {
  C rho, ri, u[3], t[3][3][3], x[6], z1[HID], z2[HID], raw[3], r[3];
};

// One D3Q27 cell's collision of the pulled populations g into out
template <class C>
__device__ __forceinline__ void cell_forward27(const C (&g)[27], const C* __restrict__ P,
                                               C tau_inv, C s, C lo, C hi, CellFwd27<C>& F,
                                               C (&out)[27], C* rec, int64_t rstride) {
  m27_moments(g, F.t);
  F.rho = F.t[0][0][0];
  F.ri = C(1) / F.rho;
  F.u[0] = F.t[1][0][0] * F.ri;
  F.u[1] = F.t[0][1][0] * F.ri;
  F.u[2] = F.t[0][0][1] * F.ri;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (a + b + c <= 1) continue;
        C du[3];
        F.t[a][b][c] -= F.rho * meq_unit(a, b, c, F.u, du);
      }
  const C xx = F.t[2][0][0], yy = F.t[0][2][0], zz = F.t[0][0][2];
  const C sc = s * F.ri;
  F.x[0] = (xx + yy + zz) * sc;
  F.x[1] = (xx - zz) * sc;
  F.x[2] = (yy - zz) * sc;
  F.x[3] = F.t[1][1][0] * sc;
  F.x[4] = F.t[1][0][1] * sc;
  F.x[5] = F.t[0][1][1] * sc;
  mlp_forward<C>(P, F.x, tau_inv, lo, hi, F.z1, F.z2, F.raw, F.r, rec, rstride);
  C y[3][3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int o = a + b + c, grp = m27_group(a, b, c);
        y[a][b][c] = o <= 1 ? C(0) : (grp < 0 ? tau_inv : F.r[grp]) * F.t[a][b][c];
      }
#pragma unroll
  for (int q = 0; q < 27; ++q) out[q] = g[q];
  m27_add_inverse(y, C(-1), out);
}

// mu = (d out / d g)^T lam for one D3Q27 cell; with rec set, also the record
template <class C>
__device__ __forceinline__ void cell_backward27(const C (&g)[27], const C (&lam)[27],
                                                const C* __restrict__ P, C tau_inv, C s, C lo,
                                                C hi, C (&mu)[27], C* rec, int64_t rstride) {
  CellFwd27<C> F;
  {
    C fout[27];
    cell_forward27<C>(g, P, tau_inv, s, lo, hi, F, fout, rec, rstride);
  }
  // out = g - Phi y, y = w . t: dL/dy = -Phi^T lam
  C nu[3][3][3];
  m27_inverse_transposed(lam, nu);
  C gr[3] = {0, 0, 0};
  C gt[3][3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int o = a + b + c, grp = m27_group(a, b, c);
        const C gy = -nu[a][b][c];
        gt[a][b][c] = o <= 1 ? C(0) : (grp < 0 ? tau_inv : F.r[grp]) * gy;
        if (grp >= 0) gr[grp] += F.t[a][b][c] * gy;
      }
  C go[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) go[k] = (F.raw[k] >= lo && F.raw[k] <= hi) ? gr[k] : C(0);
  C gx[6];
  mlp_backward<C>(P, go, F.z1, F.z2, gx, rec, rstride);
  // x = s / rho (T, Nxz, Nyz, Pxy, Pxz, Pyz) of t
  const C sc = s * F.ri;
  C grho = 0;
#pragma unroll
  for (int k = 0; k < 6; ++k) grho -= gx[k] * F.x[k] * F.ri;
  gt[2][0][0] += sc * (gx[0] + gx[1]);
  gt[0][2][0] += sc * (gx[0] + gx[2]);
  gt[0][0][2] += sc * (gx[0] - gx[1] - gx[2]);
  gt[1][1][0] += sc * gx[3];
  gt[1][0][1] += sc * gx[4];
  gt[0][1][1] += sc * gx[5];
  // t = M g - rho meq_unit(u) on orders >= 2
  C gu[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (a + b + c <= 1) continue;
        C du[3];
        const C m = meq_unit(a, b, c, F.u, du);
        grho -= gt[a][b][c] * m;
#pragma unroll
        for (int d = 0; d < 3; ++d) gu[d] -= gt[a][b][c] * F.rho * du[d];
      }
  // u = j / rho, rho = (M g)_000, j = (M g)_100, _010, _001
  gt[0][0][0] = grho;
  gt[1][0][0] = gu[0] * F.ri;
  gt[0][1][0] = gu[1] * F.ri;
  gt[0][0][1] = gu[2] * F.ri;
  gt[0][0][0] -= (gu[0] * F.u[0] + gu[1] * F.u[1] + gu[2] * F.u[2]) * F.ri;
#pragma unroll
  for (int q = 0; q < 27; ++q) mu[q] = lam[q];
  m27_add_transposed(gt, mu);
}

// what the adjoint needs of the forward pass of one cell
template <class C>
struct CellFwd {
  C rho, ri, u[3], fe[9], x[6], z1[HID], z2[HID], raw[3], r[3], mg[3];
};

// One cell's collision of the pulled populations g into out.  With rec set,
// the inputs x and activations a1, a2 go to the record (stride rstride).
template <class C>
__device__ __forceinline__ void cell_forward(const C (&g)[9], const C* __restrict__ P,
                                             C tau_inv, C s, C lo, C hi, CellFwd<C>& F,
                                             C (&out)[9], C* rec, int64_t rstride) {
  using S = D2Q9;
  moments<S, C>(g, F.rho, F.u);
  F.ri = C(1) / F.rho;
  equilibrium<S, C>(F.rho, F.u, F.fe);
  C df[9], ds[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) df[q] = g[q] - F.fe[q];
  kbc_shear<S, C>(df, ds);
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    C m = 0;
#pragma unroll
    for (int q = 0; q < 9; ++q)
      if (MROW[k][q]) m += MROW[k][q] * df[q];
    F.x[k] = s * F.ri * m;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    C m = 0;
#pragma unroll
    for (int q = 0; q < 9; ++q)
      if (MROW[3 + k][q]) m += MROW[3 + k][q] * (df[q] - ds[q]);
    F.mg[k] = m;
  }
  mlp_forward<C>(P, F.x, tau_inv, lo, hi, F.z1, F.z2, F.raw, F.r, rec, rstride);
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    C upd = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) upd += F.r[k] * F.mg[k] * (C)PHI_D[k][q];
    out[q] = g[q] - ds[q] * tau_inv - upd;
  }
}

// mu = (d out / d g)^T lam for one cell; with rec set, also the record
template <class C>
__device__ __forceinline__ void cell_backward(const C (&g)[9], const C (&lam)[9],
                                              const C* __restrict__ P, C tau_inv, C s, C lo,
                                              C hi, C (&mu)[9], C* rec, int64_t rstride) {
  using S = D2Q9;
  CellFwd<C> F;
  {
    C fout[9];
    cell_forward<C>(g, P, tau_inv, s, lo, hi, F, fout, rec, rstride);
  }
  // out = g - ds tau_inv - sum_k r_k mg_k phi_k
  C gmg[3], go[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    C lp = 0;
#pragma unroll
    for (int q = 0; q < 9; ++q) lp += lam[q] * (C)PHI_D[k][q];
    gmg[k] = -F.r[k] * lp;
    const bool inside = F.raw[k] >= lo && F.raw[k] <= hi;
    go[k] = inside ? -F.mg[k] * lp : C(0);
  }
  // mg = rows of dh, dh = df - ds
  C gds[9], gdf[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    C v = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if (MROW[3 + k][q]) v += MROW[3 + k][q] * gmg[k];
    gds[q] = -lam[q] * tau_inv - v;
    gdf[q] = v;
  }
  // ds from (T, N, Pxy): ds0 = -T, ds1 = ds3 = (T+N)/4, ds2 = ds4 = (T-N)/4,
  // ds5 = ds7 = Pxy/4, ds6 = ds8 = -Pxy/4
  const C gT = -gds[0] + C(0.25) * (gds[1] + gds[3] + gds[2] + gds[4]);
  const C gN = C(0.25) * (gds[1] + gds[3] - gds[2] - gds[4]);
  const C gP = C(0.25) * (gds[5] + gds[7] - gds[6] - gds[8]);
  // T = xx + yy, N = xx - yy; xx, yy, Pxy are rows of df
  const C gxx = gT + gN, gyy = gT - gN;
#pragma unroll
  for (int q = 1; q < 9; ++q) {
    const int ex = S::e[q][0], ey = S::e[q][1];
    if (ex != 0) gdf[q] += gxx;
    if (ey != 0) gdf[q] += gyy;
    if (ex * ey != 0) gdf[q] += (ex * ey > 0) ? gP : -gP;
  }
  // MLP backward; z2 is overwritten by dL/dz2
  C gx[6];
  mlp_backward<C>(P, go, F.z1, F.z2, gx, rec, rstride);
  // x_k = s / rho * M_k df
  C grho = 0;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    grho -= gx[k] * F.x[k] * F.ri;
#pragma unroll
    for (int q = 0; q < 9; ++q)
      if (MROW[k][q]) gdf[q] += s * F.ri * gx[k] * MROW[k][q];
  }
  // df = g - feq(rho, u), feq = w rho (1 + 3 eu - 3/2 uu + 9/2 eu^2)
  C gu[2] = {0, 0};
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    const C gfe = -gdf[q];
    grho += gfe * F.fe[q] * F.ri;
    const C eu = S::e[q][0] * F.u[0] + S::e[q][1] * F.u[1];
    const C c = gfe * (C)S::w[q] * F.rho;
#pragma unroll
    for (int d = 0; d < 2; ++d)
      gu[d] += c * (C(3) * S::e[q][d] - C(3) * F.u[d] + C(9) * eu * S::e[q][d]);
  }
  // u = j / rho
  C gj[2];
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    gj[d] = gu[d] * F.ri;
    grho -= gu[d] * F.u[d] * F.ri;
  }
#pragma unroll
  for (int q = 0; q < 9; ++q)
    mu[q] = lam[q] + gdf[q] + grho + gj[0] * S::e[q][0] + gj[1] * S::e[q][1];
}

// cell index c -> (i0, i1, i2) and the periodic neighbour tables
__device__ __forceinline__ void cell_tables(const KParams& p, int64_t c, int (&a0)[3], int (&a1)[3],
                                            int (&a2)[3]) {
  const int i2 = (int)(c % p.n2);
  const int64_t r = c / p.n2;
  const int i1 = (int)(r % p.n1), i0 = (int)(r / p.n1);
  axis_offsets(i0, p.n0, (int)p.s0, a0);  // fewer than 2^31 populations
  axis_offsets(i1, p.n1, p.s1, a1);
  axis_offsets(i2, p.n2, 1, a2);
}

// the stencil's cell collision and its adjoint
template <class S, class C>
__device__ __forceinline__ void collide_cell(const C (&g)[S::Q], const C* __restrict__ P,
                                             C tau_inv, C s, C lo, C hi, C (&out)[S::Q]) {
  if constexpr (S::Q == 9) {
    CellFwd<C> F;
    cell_forward<C>(g, P, tau_inv, s, lo, hi, F, out, nullptr, 0);
  } else {
    CellFwd27<C> F;
    cell_forward27<C>(g, P, tau_inv, s, lo, hi, F, out, nullptr, 0);
  }
}
template <class S, class C>
__device__ __forceinline__ void adjoint_cell(const C (&g)[S::Q], const C (&lam)[S::Q],
                                             const C* __restrict__ P, C tau_inv, C s, C lo, C hi,
                                             C (&mu)[S::Q], C* rec, int64_t rstride) {
  if constexpr (S::Q == 9) cell_backward<C>(g, lam, P, tau_inv, s, lo, hi, mu, rec, rstride);
  else cell_backward27<C>(g, lam, P, tau_inv, s, lo, hi, mu, rec, rstride);
}

// The parameters, read by every cell, are staged in LDS once per block.
// Read from global memory they are wave-uniform scalar loads; the 1379 of
// them do not fit the scalar registers, and the compiler kept them in
// vector-register lanes, read back with one v_readlane per use (15,616 of
// the backward kernel's 23,831 instructions for D2Q9 FP32).
template <class C>
__device__ __forceinline__ void stage_params(const TrainArgs& t, C* sp) {
  const C* P = (const C*)t.params;
  for (int i = threadIdx.x; i < NPARAM; i += blockDim.x) sp[i] = P[i];
  __syncthreads();
}

template <class S, class C>
__global__ void __launch_bounds__(128) train_forward_kernel(const KParams p, const TrainArgs t) {
  __shared__ C P[NPARAM];
  stage_params<C>(t, P);
  const C tau_inv = (C)p.tau_inv, s = (C)t.in_scale, lo = (C)t.rate_lo, hi = (C)t.rate_hi;
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= p.ncells) return;
  int a0[3], a1[3], a2[3];
  cell_tables(p, c, a0, a1, a2);
  const C* fin = (const C*)p.fin;
  C g[S::Q];
#pragma unroll
  for (int q = 0; q < S::Q; ++q) g[q] = fin[q * p.sq + nb_offset<S>(q, -1, a0, a1, a2)];
  C out[S::Q];
  collide_cell<S, C>(g, P, tau_inv, s, lo, hi, out);
  C* fout = (C*)p.fout;
  const int cell = a0[1] + a1[1] + a2[1];
#pragma unroll
  for (int q = 0; q < S::Q; ++q) fout[q * p.sq + cell] = out[q];
}

// lam: the adjoint of the step's output, either directly (gathered == 0) or
// as the previous backward step's mu, read at x + e_q (gathered == 1)
// Two waves per SIMD: the kernel then fits 256 VGPRs with a few spills to
// scratch (D2Q9 FP32: 212 bytes per lane); with one wave per SIMD and no
// spills it ran slower (C4 training: 22.5 against 27.8 ms per iteration).
template <class S, class C>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2)))
train_backward_kernel(const KParams p, const TrainArgs t,
                                                             const C* __restrict__ lam_in,
                                                             int gathered, C* __restrict__ mu_out,
                                                             C* __restrict__ rec) {
  __shared__ C P[NPARAM];
  stage_params<C>(t, P);
  const C tau_inv = (C)p.tau_inv, s = (C)t.in_scale, lo = (C)t.rate_lo, hi = (C)t.rate_hi;
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= p.ncells) return;
  int a0[3], a1[3], a2[3];
  cell_tables(p, c, a0, a1, a2);
  const C* fin = (const C*)p.fin;
  C g[S::Q], lam[S::Q], mu[S::Q];
#pragma unroll
  for (int q = 0; q < S::Q; ++q) {
    g[q] = fin[q * p.sq + nb_offset<S>(q, -1, a0, a1, a2)];
    lam[q] = lam_in[q * p.sq + (gathered ? nb_offset<S>(q, +1, a0, a1, a2) : a0[1] + a1[1] + a2[1])];
  }
  adjoint_cell<S, C>(g, lam, P, tau_inv, s, lo, hi, mu, rec ? rec + c : nullptr, p.ncells);
  const int cell = a0[1] + a1[1] + a2[1];
#pragma unroll
  for (int q = 0; q < S::Q; ++q) mu_out[q * p.sq + cell] = mu[q];
}

// lam_q(x) = mu_q(x + e_q): the transpose of the pull applied to mu
template <class S, class C>
__global__ void __launch_bounds__(256) unpull_kernel(const KParams p, const C* __restrict__ mu,
                                                     C* __restrict__ lam) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= p.ncells) return;
  int a0[3], a1[3], a2[3];
  cell_tables(p, c, a0, a1, a2);
  const int cell = a0[1] + a1[1] + a2[1];
#pragma unroll
  for (int q = 0; q < S::Q; ++q) lam[q * p.sq + cell] = mu[q * p.sq + nb_offset<S>(q, +1, a0, a1, a2)];
}

// ---------------------------------------------------------------------------
// Parameter gradients from the records of the backward kernel: sums over the
// cells of products of record rows, in the parameter layout
//   dW1[i][j] = sum_c dz1_i x_j    db1[i] = sum_c dz1_i
//   dW2[i][j] = sum_c dz2_i a1_j   db2[i] = sum_c dz2_i
//   dW3[i][j] = sum_c do_i a2_j    db3[i] = sum_c do_i
// (1379 sums, against 67 x 70 for the full product of the record's two
// halves).  Stage 1: block b sums cells [b * RG_CHUNK, (b + 1) * RG_CHUNK)
// in tiles of RG_TILE cells staged in LDS, with a row of ones (for the
// biases) and a row of zeros after the record's rows; LDS rows are padded
// by one element, so that the rows the lanes of a wave read at one cell
// fall in distinct banks.  Thread t holds dW2 rows 2 (t / 16) + u, columns
// 2 (t % 16) + v (u, v < 2), and items t and t + 256 of the other 355 sums
// (dW1, dW3, db1, db2, db3), each a product of two staged rows.  Stage 2
// adds the blocks' partial sums to the gradient in block order, so the
// result does not depend on scheduling.  (A GEMM over K = cells gives the
// 67 x 70 product 25 output tiles of 16 x 16 for the whole device.)
constexpr int RG_TILE = 32;
constexpr int RG_CHUNK = 2048;
constexpr int RG_ONE = NREC, RG_ZERO = NREC + 1, RG_ROWS = NREC + 2;
constexpr int RG_REST = NPARAM - HID * HID;  // 355

// the two staged rows whose product item r (< RG_REST) sums, and its index
// in the parameter vector
__device__ __forceinline__ void rest_item(int r, int& row, int& col, int& out) {
  if (r < HID * 6) {  // dW1
    row = R_DZ1 + r / 6;
    col = R_X + r % 6;
    out = P_W1 + r;
  } else if (r < HID * 6 + 3 * HID) {  // dW3
    const int q = r - HID * 6;
    row = R_DO + q / HID;
    col = R_A2 + q % HID;
    out = P_W3 + q;
  } else if (r < HID * 6 + 3 * HID + HID) {  // db1
    const int q = r - HID * 6 - 3 * HID;
    row = R_DZ1 + q;
    col = RG_ONE;
    out = P_B1 + q;
  } else if (r < HID * 6 + 3 * HID + 2 * HID) {  // db2
    const int q = r - HID * 6 - 4 * HID;
    row = R_DZ2 + q;
    col = RG_ONE;
    out = P_B2 + q;
  } else if (r < RG_REST) {  // db3
    const int q = r - HID * 6 - 5 * HID;
    row = R_DO + q;
    col = RG_ONE;
    out = P_B3 + q;
  } else {  // no item
    row = RG_ZERO;
    col = RG_ZERO;
    out = -1;
  }
}

template <class C>
__global__ void __launch_bounds__(256) grad_partial_kernel(const C* __restrict__ rec, int64_t cells,
                                                           C* __restrict__ part) {
  __shared__ C s[RG_ROWS][RG_TILE + 1];
  const int t = threadIdx.x;
  const int r2 = R_DZ2 + 2 * (t >> 4), c2 = R_A1 + 2 * (t & 15);
  int ra, ca, oa, rb, cb, ob;
  rest_item(t, ra, ca, oa);
  rest_item(t + 256, rb, cb, ob);
  C g2[2][2] = {{C(0), C(0)}, {C(0), C(0)}}, ga = C(0), gb = C(0);
  const int64_t c0 = (int64_t)blockIdx.x * RG_CHUNK;
  const int64_t c1 = c0 + RG_CHUNK < cells ? c0 + RG_CHUNK : cells;
  // the next tile's values are loaded into registers while the current
  // tile is summed (element e = t + 256 i of a tile: row e / RG_TILE, cell
  // e % RG_TILE)
  constexpr int PER = (NREC * RG_TILE + 255) / 256;
  C nxt[PER];
  auto fetch = [&](int64_t base) {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int e = t + 256 * i;
      const int r = e / RG_TILE, k = e % RG_TILE;
      const int64_t c = base + k;
      nxt[i] = (e < NREC * RG_TILE && c < c1) ? rec[(int64_t)r * cells + c] : C(0);
    }
  };
  fetch(c0);
  for (int64_t base = c0; base < c1; base += RG_TILE) {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int e = t + 256 * i;
      if (e < NREC * RG_TILE) s[e / RG_TILE][e % RG_TILE] = nxt[i];
    }
    if (t < RG_TILE) {
      s[RG_ONE][t] = base + t < c1 ? C(1) : C(0);
      s[RG_ZERO][t] = C(0);
    }
    __syncthreads();
    if (base + RG_TILE < c1) fetch(base + RG_TILE);
#pragma unroll 8
    for (int k = 0; k < RG_TILE; ++k) {
      const C d0 = s[r2][k], d1 = s[r2 + 1][k], a0 = s[c2][k], a1 = s[c2 + 1][k];
      g2[0][0] += d0 * a0;
      g2[0][1] += d0 * a1;
      g2[1][0] += d1 * a0;
      g2[1][1] += d1 * a1;
      ga += s[ra][k] * s[ca][k];
      gb += s[rb][k] * s[cb][k];
    }
    __syncthreads();
  }
  C* dst = part + (int64_t)blockIdx.x * NPARAM;
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v)
      dst[P_W2 + (r2 - R_DZ2 + u) * HID + (c2 - R_A1 + v)] = g2[u][v];
  if (oa >= 0) dst[oa] = ga;
  if (ob >= 0) dst[ob] = gb;
}

// grad[e] += sum over the nparts partial sums: thread group g (of 16) of a
// block adds parts g, g + 16, g + 32, ..., and the groups' sums are
// combined in group order
constexpr int RG_GROUPS = 16;
template <class C>
__global__ void __launch_bounds__(64 * RG_GROUPS) grad_sum_kernel(const C* __restrict__ part,
                                                                  int nparts, C* __restrict__ grad) {
  __shared__ C red[RG_GROUPS][64];
  const int el = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int e = blockIdx.x * 64 + el;
  C s = C(0);
  if (e < NPARAM)
    for (int b = g; b < nparts; b += RG_GROUPS) s += part[(int64_t)b * NPARAM + e];
  red[g][el] = s;
  __syncthreads();
  if (g == 0 && e < NPARAM) {
    C sum = red[0][el];
#pragma unroll
    for (int i = 1; i < RG_GROUPS; ++i) sum += red[i][el];
    grad[e] += sum;
  }
}

int64_t grad_parts(int64_t cells) { return (cells + RG_CHUNK - 1) / RG_CHUNK; }

}  // namespace lbm

using namespace lbm;

static_assert(NPARAM == LBM_LEARNED_PARAMS, "parameter count");
static_assert(NREC == LBM_LEARNED_RECORD, "record size");

namespace {
// D2Q9: periodic fields with the batch on i0, contiguous [9][n0][n1][n2];
// D3Q27: one periodic field, contiguous [27][n0][n1][n2]
int train_params(int32_t stencil, int32_t storage, const int64_t n[3], double tau, KParams& p) {
  if (stencil != LBM_D2Q9 && stencil != LBM_D3Q27)
    return report_error(LBM_E_UNSUPPORTED, "training kernels exist for D2Q9 and D3Q27");
  if (storage != LBM_F32 && storage != LBM_F64)
    return report_error(LBM_E_UNSUPPORTED, "training kernels take F32 or F64 populations");
  if (!n || n[0] < 1 || n[1] < 1 || n[2] < 1) return report_error(LBM_E_ARG, "extents must be positive");
  if (!(tau > 0.5)) return report_error(LBM_E_ARG, "tau must exceed 1/2");
  const int64_t cells = n[0] * n[1] * n[2];
  const int q = stencil == LBM_D2Q9 ? 9 : 27;
  if (q * cells > (int64_t)0x7fffffff) return report_error(LBM_E_ARG, "more than 2^31 populations");
  memset(&p, 0, sizeof(p));
  p.n0 = (int)n[0];
  p.n1 = (int)n[1];
  p.n2 = (int)n[2];
  p.s1 = p.n2;
  p.s0 = (long long)p.n1 * p.n2;
  p.sq = cells;
  p.ncells = cells;
  p.tau_inv = 1.0 / tau;
  return LBM_OK;
}

// one thread per cell
unsigned train_blocks(int64_t cells, int bs) { return (unsigned)((cells + bs - 1) / bs); }

int launched() {
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? LBM_OK : report_error(LBM_E_HIP, hipGetErrorString(e));
}

template <class C>
void launch_forward(int32_t stencil, const KParams& p, const TrainArgs& t, hipStream_t s) {
  const unsigned nb = train_blocks(p.ncells, 128);
  if (stencil == LBM_D2Q9)
    hipLaunchKernelGGL((train_forward_kernel<D2Q9, C>), dim3(nb), dim3(128), 0, s, p, t);
  else
    hipLaunchKernelGGL((train_forward_kernel<D3Q27, C>), dim3(nb), dim3(128), 0, s, p, t);
}

template <class C>
void launch_backward(int32_t stencil, const KParams& p, const TrainArgs& t, const void* lam,
                     int gathered, void* mu, void* rec, hipStream_t s) {
  const unsigned nb = train_blocks(p.ncells, 128);
  if (stencil == LBM_D2Q9)
    hipLaunchKernelGGL((train_backward_kernel<D2Q9, C>), dim3(nb), dim3(128), 0, s, p, t,
                       (const C*)lam, gathered, (C*)mu, (C*)rec);
  else
    hipLaunchKernelGGL((train_backward_kernel<D3Q27, C>), dim3(nb), dim3(128), 0, s, p, t,
                       (const C*)lam, gathered, (C*)mu, (C*)rec);
}

template <class C>
void launch_unpull(int32_t stencil, const KParams& p, const void* mu, void* lam, hipStream_t s) {
  const unsigned nb = train_blocks(p.ncells, 256);
  if (stencil == LBM_D2Q9)
    hipLaunchKernelGGL((unpull_kernel<D2Q9, C>), dim3(nb), dim3(256), 0, s, p, (const C*)mu, (C*)lam);
  else
    hipLaunchKernelGGL((unpull_kernel<D3Q27, C>), dim3(nb), dim3(256), 0, s, p, (const C*)mu, (C*)lam);
}
}  // namespace

extern "C" {

int lbm_learned_train_forward(int32_t stencil, int32_t storage, const int64_t n[3], double tau,
                              const void* params, double in_scale, double rate_lo,
                              double rate_hi, const void* f_in, void* f_out, void* stream) {
  KParams p;
  const int rc = train_params(stencil, storage, n, tau, p);
  if (rc) return rc;
  if (!params || !f_in || !f_out) return report_error(LBM_E_ARG, "null buffer");
  if (f_in == f_out) return report_error(LBM_E_ARG, "f_in and f_out must differ");
  if (!(rate_lo <= rate_hi)) return report_error(LBM_E_ARG, "rate_lo must not exceed rate_hi");
  p.fin = f_in;
  p.fout = f_out;
  const TrainArgs t{params, in_scale, rate_lo, rate_hi};
  if (storage == LBM_F64) launch_forward<double>(stencil, p, t, (hipStream_t)stream);
  else launch_forward<float>(stencil, p, t, (hipStream_t)stream);
  return launched();
}

int lbm_learned_train_backward(int32_t stencil, int32_t storage, const int64_t n[3], double tau,
                               const void* params, double in_scale, double rate_lo,
                               double rate_hi, const void* f_in, const void* lam,
                               int32_t lam_gathered, void* mu, void* record, void* stream) {
  KParams p;
  const int rc = train_params(stencil, storage, n, tau, p);
  if (rc) return rc;
  if (!params || !f_in || !lam || !mu) return report_error(LBM_E_ARG, "null buffer");
  if (lam == mu || f_in == mu) return report_error(LBM_E_ARG, "mu must not alias f_in or lam");
  if (!(rate_lo <= rate_hi)) return report_error(LBM_E_ARG, "rate_lo must not exceed rate_hi");
  p.fin = f_in;
  const TrainArgs t{params, in_scale, rate_lo, rate_hi};
  const int g = lam_gathered != 0;
  if (storage == LBM_F64) launch_backward<double>(stencil, p, t, lam, g, mu, record, (hipStream_t)stream);
  else launch_backward<float>(stencil, p, t, lam, g, mu, record, (hipStream_t)stream);
  return launched();
}

int lbm_unpull(int32_t stencil, int32_t storage, const int64_t n[3], const void* mu, void* lam,
               void* stream) {
  KParams p;
  const int rc = train_params(stencil, storage, n, 1.0, p);
  if (rc) return rc;
  if (!mu || !lam) return report_error(LBM_E_ARG, "null buffer");
  if (mu == lam) return report_error(LBM_E_ARG, "mu and lam must differ");
  if (storage == LBM_F64) launch_unpull<double>(stencil, p, mu, lam, (hipStream_t)stream);
  else launch_unpull<float>(stencil, p, mu, lam, (hipStream_t)stream);
  return launched();
}

int64_t lbm_learned_grad_workspace(int32_t storage, int64_t cells) {
  if (storage != LBM_F32 && storage != LBM_F64) return -1;
  if (cells < 1) return -1;
  return grad_parts(cells) * NPARAM * (storage == LBM_F64 ? 8 : 4);
}

int lbm_learned_grad_reduce(int32_t storage, int64_t cells, const void* record, void* workspace,
                            void* grad, void* stream) {
  if (storage != LBM_F32 && storage != LBM_F64)
    return report_error(LBM_E_UNSUPPORTED, "the gradient reduction takes F32 or F64 records");
  if (cells < 1) return report_error(LBM_E_ARG, "cells must be positive");
  if (!record || !workspace || !grad) return report_error(LBM_E_ARG, "null buffer");
  const int64_t parts = grad_parts(cells);
  if (parts > 0x7fffffff) return report_error(LBM_E_ARG, "too many cells");
  const hipStream_t s = (hipStream_t)stream;
  const unsigned sum_blocks = (NPARAM + 63) / 64;
  if (storage == LBM_F64) {
    hipLaunchKernelGGL(grad_partial_kernel<double>, dim3((unsigned)parts), dim3(256), 0, s,
                       (const double*)record, cells, (double*)workspace);
    hipLaunchKernelGGL(grad_sum_kernel<double>, dim3(sum_blocks), dim3(64 * RG_GROUPS), 0, s,
                       (const double*)workspace, (int)parts, (double*)grad);
  } else {
    hipLaunchKernelGGL(grad_partial_kernel<float>, dim3((unsigned)parts), dim3(256), 0, s,
                       (const float*)record, cells, (float*)workspace);
    hipLaunchKernelGGL(grad_sum_kernel<float>, dim3(sum_blocks), dim3(64 * RG_GROUPS), 0, s,
                       (const float*)workspace, (int)parts, (float*)grad);
  }
  return launched();
}

}  // extern "C"

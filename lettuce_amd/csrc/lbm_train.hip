// Training kernels of the D2Q9 learned collision: the forward step with
// the MLP evaluated in the arithmetic type (FP32, or FP64 for gradient
// checks) and its hand-written adjoint.
//
// Forward per cell (g = pulled populations, s = in_scale):
//   rho = sum g, u = j / rho, feq, df = g - feq
//   ds = KBC shear part of df (T = xx + yy, N = xx - yy, Pxy), dh = df - ds
//   x  = s / rho (T, N, Pxy, xxy, xyy, xxyy)(df)
//   z1 = W1 x + b1, a1 = silu(z1), z2 = W2 a1 + b2, a2 = silu(z2),
//   o  = W3 a2 + b3, r = clamp(1/tau + o, rate_lo, rate_hi)
//   f' = g - ds / tau - sum_k r_k m_k(dh) phi_k
// (the inference kernel's formula without bf16 rounding and without the
// KBC fallback; the clamp has zero derivative outside [rate_lo, rate_hi]).
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
  C a[HID];
#pragma unroll
  for (int i = 0; i < HID; ++i) {
    C z = P[P_B1 + i];
#pragma unroll
    for (int k = 0; k < 6; ++k) z += P[P_W1 + i * 6 + k] * F.x[k];
    F.z1[i] = z;
    a[i] = z * sigmoid(z);
  }
  if (rec) {
#pragma unroll
    for (int k = 0; k < 6; ++k) rec[(R_X + k) * rstride] = F.x[k];
#pragma unroll
    for (int i = 0; i < HID; ++i) rec[(R_A1 + i) * rstride] = a[i];
  }
#pragma unroll
  for (int i = 0; i < HID; ++i) {
    C z = P[P_B2 + i];
#pragma unroll
    for (int j = 0; j < HID; ++j) z += P[P_W2 + i * HID + j] * a[j];
    F.z2[i] = z;
  }
#pragma unroll
  for (int i = 0; i < HID; ++i) a[i] = F.z2[i] * sigmoid(F.z2[i]);
  if (rec) {
#pragma unroll
    for (int i = 0; i < HID; ++i) rec[(R_A2 + i) * rstride] = a[i];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    C o = P[P_B3 + k];
#pragma unroll
    for (int j = 0; j < HID; ++j) o += P[P_W3 + k * HID + j] * a[j];
    F.raw[k] = tau_inv + o;
    F.r[k] = F.raw[k] < lo ? lo : (F.raw[k] > hi ? hi : F.raw[k]);
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    C upd = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) upd += F.r[k] * F.mg[k] * (C)PHI_D[k][q];
    out[q] = g[q] - ds[q] * tau_inv - upd;
  }
}

// silu'(z) = sigma(z) (1 + z (1 - sigma(z)))
template <class C>
__device__ __forceinline__ C dsilu(C z) {
  const C sg = sigmoid(z);
  return sg * (C(1) + z * (C(1) - sg));
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
#pragma unroll
  for (int i = 0; i < HID; ++i) {
    C v = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) v += P[P_W3 + k * HID + i] * go[k];
    F.z2[i] = v * dsilu(F.z2[i]);
  }
  C gx[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int j = 0; j < HID; ++j) {
    C v = 0;
#pragma unroll
    for (int i = 0; i < HID; ++i) v += P[P_W2 + i * HID + j] * F.z2[i];
    const C gz1 = v * dsilu(F.z1[j]);
    if (rec) rec[(R_DZ1 + j) * rstride] = gz1;
#pragma unroll
    for (int k = 0; k < 6; ++k) gx[k] += P[P_W1 + j * 6 + k] * gz1;
  }
  if (rec) {
#pragma unroll
    for (int i = 0; i < HID; ++i) rec[(R_DZ2 + i) * rstride] = F.z2[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) rec[(R_DO + k) * rstride] = go[k];
  }
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

template <class C>
__global__ void __launch_bounds__(128) train_forward_kernel(const KParams p, const TrainArgs t) {
  using S = D2Q9;
  const C* P = (const C*)t.params;
  const C tau_inv = (C)p.tau_inv, s = (C)t.in_scale, lo = (C)t.rate_lo, hi = (C)t.rate_hi;
  {
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= p.ncells) return;
    int a0[3], a1[3], a2[3];
    cell_tables(p, c, a0, a1, a2);
    const C* fin = (const C*)p.fin;
    C g[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) g[q] = fin[q * p.sq + nb_offset<S>(q, -1, a0, a1, a2)];
    CellFwd<C> F;
    C out[9];
    cell_forward<C>(g, P, tau_inv, s, lo, hi, F, out, nullptr, 0);
    C* fout = (C*)p.fout;
    const int cell = a0[1] + a1[1] + a2[1];
#pragma unroll
    for (int q = 0; q < 9; ++q) fout[q * p.sq + cell] = out[q];
  }
}

// lam: the adjoint of the step's output, either directly (gathered == 0) or
// as the previous backward step's mu, read at x + e_q (gathered == 1)
template <class C>
__global__ void __launch_bounds__(128) train_backward_kernel(const KParams p, const TrainArgs t,
                                                             const C* __restrict__ lam_in,
                                                             int gathered, C* __restrict__ mu_out,
                                                             C* __restrict__ rec) {
  using S = D2Q9;
  const C* P = (const C*)t.params;
  const C tau_inv = (C)p.tau_inv, s = (C)t.in_scale, lo = (C)t.rate_lo, hi = (C)t.rate_hi;
  {
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= p.ncells) return;
    int a0[3], a1[3], a2[3];
    cell_tables(p, c, a0, a1, a2);
    const C* fin = (const C*)p.fin;
    C g[9], lam[9], mu[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      g[q] = fin[q * p.sq + nb_offset<S>(q, -1, a0, a1, a2)];
      lam[q] = lam_in[q * p.sq + (gathered ? nb_offset<S>(q, +1, a0, a1, a2) : a0[1] + a1[1] + a2[1])];
    }
    cell_backward<C>(g, lam, P, tau_inv, s, lo, hi, mu, rec ? rec + c : nullptr, p.ncells);
    const int cell = a0[1] + a1[1] + a2[1];
#pragma unroll
    for (int q = 0; q < 9; ++q) mu_out[q * p.sq + cell] = mu[q];
  }
}

// lam_q(x) = mu_q(x + e_q): the transpose of the pull applied to mu
template <class C>
__global__ void __launch_bounds__(256) unpull_kernel(const KParams p, const C* __restrict__ mu,
                                                     C* __restrict__ lam) {
  using S = D2Q9;
  {
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= p.ncells) return;
    int a0[3], a1[3], a2[3];
    cell_tables(p, c, a0, a1, a2);
    const int cell = a0[1] + a1[1] + a2[1];
#pragma unroll
    for (int q = 0; q < 9; ++q) lam[q * p.sq + cell] = mu[q * p.sq + nb_offset<S>(q, +1, a0, a1, a2)];
  }
}

}  // namespace lbm

using namespace lbm;

static_assert(NPARAM == LBM_LEARNED_PARAMS, "parameter count");
static_assert(NREC == LBM_LEARNED_RECORD, "record size");

namespace {
// periodic D2Q9 with the batch on i0, contiguous [9][n0][n1][n2]
int train_params(int32_t storage, const int64_t n[3], double tau, KParams& p) {
  if (storage != LBM_F32 && storage != LBM_F64)
    return report_error(LBM_E_UNSUPPORTED, "training kernels take F32 or F64 populations");
  if (!n || n[0] < 1 || n[1] < 1 || n[2] < 1) return report_error(LBM_E_ARG, "extents must be positive");
  if (!(tau > 0.5)) return report_error(LBM_E_ARG, "tau must exceed 1/2");
  const int64_t cells = n[0] * n[1] * n[2];
  if (9 * cells > (int64_t)0x7fffffff) return report_error(LBM_E_ARG, "more than 2^31 populations");
  memset(&p, 0, sizeof(p));
  p.n0 = (int)n[0];
  p.n1 = (int)n[1];
  p.n2 = (int)n[2];
  p.s1 = p.n2;
  p.s0 = p.n1 * p.n2;
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
}  // namespace

extern "C" {

int lbm_learned_train_forward(int32_t storage, const int64_t n[3], double tau,
                              const void* params, double in_scale, double rate_lo,
                              double rate_hi, const void* f_in, void* f_out, void* stream) {
  KParams p;
  const int rc = train_params(storage, n, tau, p);
  if (rc) return rc;
  if (!params || !f_in || !f_out) return report_error(LBM_E_ARG, "null buffer");
  if (f_in == f_out) return report_error(LBM_E_ARG, "f_in and f_out must differ");
  if (!(rate_lo <= rate_hi)) return report_error(LBM_E_ARG, "rate_lo must not exceed rate_hi");
  p.fin = f_in;
  p.fout = f_out;
  const TrainArgs t{params, in_scale, rate_lo, rate_hi};
  const unsigned nb = train_blocks(p.ncells, 128);
  if (storage == LBM_F64)
    hipLaunchKernelGGL(train_forward_kernel<double>, dim3(nb), dim3(128), 0, (hipStream_t)stream, p, t);
  else
    hipLaunchKernelGGL(train_forward_kernel<float>, dim3(nb), dim3(128), 0, (hipStream_t)stream, p, t);
  return launched();
}

int lbm_learned_train_backward(int32_t storage, const int64_t n[3], double tau,
                               const void* params, double in_scale, double rate_lo,
                               double rate_hi, const void* f_in, const void* lam,
                               int32_t lam_gathered, void* mu, void* record, void* stream) {
  KParams p;
  const int rc = train_params(storage, n, tau, p);
  if (rc) return rc;
  if (!params || !f_in || !lam || !mu) return report_error(LBM_E_ARG, "null buffer");
  if (lam == mu || f_in == mu) return report_error(LBM_E_ARG, "mu must not alias f_in or lam");
  if (!(rate_lo <= rate_hi)) return report_error(LBM_E_ARG, "rate_lo must not exceed rate_hi");
  p.fin = f_in;
  const TrainArgs t{params, in_scale, rate_lo, rate_hi};
  const unsigned nb = train_blocks(p.ncells, 128);
  if (storage == LBM_F64)
    hipLaunchKernelGGL(train_backward_kernel<double>, dim3(nb), dim3(128), 0, (hipStream_t)stream, p,
                       t, (const double*)lam, (int)(lam_gathered != 0), (double*)mu, (double*)record);
  else
    hipLaunchKernelGGL(train_backward_kernel<float>, dim3(nb), dim3(128), 0, (hipStream_t)stream, p,
                       t, (const float*)lam, (int)(lam_gathered != 0), (float*)mu, (float*)record);
  return launched();
}

int lbm_unpull(int32_t storage, const int64_t n[3], const void* mu, void* lam, void* stream) {
  KParams p;
  const int rc = train_params(storage, n, 1.0, p);
  if (rc) return rc;
  if (!mu || !lam) return report_error(LBM_E_ARG, "null buffer");
  if (mu == lam) return report_error(LBM_E_ARG, "mu and lam must differ");
  const unsigned nb = train_blocks(p.ncells, 256);
  if (storage == LBM_F64)
    hipLaunchKernelGGL(unpull_kernel<double>, dim3(nb), dim3(256), 0, (hipStream_t)stream, p,
                       (const double*)mu, (double*)lam);
  else
    hipLaunchKernelGGL(unpull_kernel<float>, dim3(nb), dim3(256), 0, (hipStream_t)stream, p,
                       (const float*)mu, (float*)lam);
  return launched();
}

}  // extern "C"

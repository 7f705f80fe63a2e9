// Kernels of an x-slab: one device's share of a lattice split along i0.
//
// A pull step of the slab needs, besides its own planes 0 .. n0-1, the
// populations that enter through its two i0 faces: those with e_x = +1 from
// the lower neighbour's last plane and those with e_x = -1 from the upper
// neighbour's first plane.  They travel as packed halo buffers
// [k][n1][n2] (k = 5 for D3Q19, 9 for D3Q27: only the populations that cross
// the face, about a quarter / a third of a plane of f).
//
// The planes 1 .. n0-2 need no halo and run as step_kernel over an i0 range
// (KParams::x0), so they can overlap the exchange.  face_kernel runs the two
// face planes: it pulls from the received halos where the source plane lies
// outside the slab and, after the collision, writes the populations that
// leave the slab into the send buffers of the next step, so no separate pack
// pass runs per step.  pack_kernel fills the send buffers once from a state
// that was not produced by a slab step (the initial condition, a restart).
//
// A slab whose recv buffers hold its own send buffers of the previous step
// (one slab, wrapped onto itself) steps exactly as the periodic lattice.
#include "lbm_slab.h"

namespace lbm {

// index of population q among the populations with the same e_x
template <class S>
__host__ __device__ constexpr int cross_index(int q) {
  int k = 0;
  for (int r = 0; r < q; ++r) k += go<S>(r, 0) == go<S>(q, 0) ? 1 : 0;
  return k;
}

template <class S>
constexpr int cross_count() {
  int k = 0;
  for (int q = 0; q < S::Q; ++q) k += go<S>(q, 0) == 1 ? 1 : 0;
  return k;
}

static_assert(cross_count<D3Q19>() == 5 && cross_count<D3Q27>() == 9, "crossing populations");
static_assert(cross_index<D3Q19>(1) == 0 && cross_index<D3Q19>(2) == 0, "first crossing index");

// block b of a face launch: face plane i0 (0, or n0 - 1 for the second
// face), row i1 and run b2 of the row
__device__ __forceinline__ void face_tile(const KParams& p, int b, int& i0, int& i1, int& b2) {
  b2 = b % p.nb2;
  const int r = b / p.nb2;
  i1 = r % p.n1;
  i0 = (r / p.n1) == 0 ? 0 : p.n0 - 1;
}

template <class S, class ST, int COLL>
__global__ void __launch_bounds__(256) face_kernel(const KParams p, const SlabHalo h) {
  using C = typename ST::C;
  using T = typename ST::T;
  int i0, i1, b2;
  face_tile(p, blockIdx.x, i0, i1, b2);
  const int i2 = b2 * blockDim.x + threadIdx.x;
  if (i2 >= p.n2) return;

  int a1[3], a2[3], hr[3];
  axis_offsets(i1, p.n1, p.s1, a1);  // rows of f
  axis_offsets(i1, p.n1, p.n2, hr);  // rows of a halo plane
  axis_offsets(i2, p.n2, 1, a2);
  const long long own_plane = i0 * p.s0;  // offset of plane i0 in f
  const uint32_t plane = (uint32_t)p.n1 * (uint32_t)p.n2;
  const T* fin = (const T*)p.fin;
  const T* rlo = (const T*)h.recv_lo;
  const T* rhi = (const T*)h.recv_hi;

  C f[S::Q];
#pragma unroll
  for (int q = 0; q < S::Q; ++q) {
    const C sh = shift_of<ST>(S::w[q], p);
    const int ex = go<S>(q, 0);
    const int xs = i0 - ex;  // source plane of the pull
    const int ry = 1 - go<S>(q, 1), rz = 1 - go<S>(q, 2);
    if (ex == 1 && xs < 0)
      f[q] = ST::ld(rlo + cross_index<S>(q) * plane, (uint32_t)(hr[ry] + a2[rz]), sh);
    else if (ex == -1 && xs >= p.n0)
      f[q] = ST::ld(rhi + cross_index<S>(q) * plane, (uint32_t)(hr[ry] + a2[rz]), sh);
    else
      f[q] = ST::ld(fin + q * p.sq + xs * p.s0, (uint32_t)(a1[ry] + a2[rz]), sh);
  }

  const CollParams<C> cp = coll_params<C>(p);
  collide<S, C, COLL>(f, cp);

  T* fout = (T*)p.fout;
  T* slo = (T*)h.send_lo;
  T* shi = (T*)h.send_hi;
  const uint32_t own = (uint32_t)(hr[1] + a2[1]);
#pragma unroll
  for (int q = 0; q < S::Q; ++q) {
    const C sh = shift_of<ST>(S::w[q], p);
    ST::st(fout + q * p.sq + own_plane, (uint32_t)(a1[1] + a2[1]), f[q], sh);
    // the same encoded value leaves through the face it points at
    if (go<S>(q, 0) == -1 && i0 == 0) ST::st(slo + cross_index<S>(q) * plane, own, f[q], sh);
    if (go<S>(q, 0) == 1 && i0 == p.n0 - 1) ST::st(shi + cross_index<S>(q) * plane, own, f[q], sh);
  }
}

// raw copy of the leaving populations of planes 0 and n0 - 1 (U: an
// unsigned integer of the storage element's width)
template <class S, class U>
__global__ void __launch_bounds__(256) pack_kernel(const KParams p, const SlabHalo h) {
  int i0, i1, b2;
  face_tile(p, blockIdx.x, i0, i1, b2);
  const int i2 = b2 * blockDim.x + threadIdx.x;
  if (i2 >= p.n2) return;
  const long long cell = i0 * p.s0 + i1 * p.s1 + i2;
  const uint32_t plane = (uint32_t)p.n1 * (uint32_t)p.n2;
  const uint32_t own = (uint32_t)(i1 * p.n2 + i2);
  const U* fin = (const U*)p.fin;
  U* slo = (U*)h.send_lo;
  U* shi = (U*)h.send_hi;
#pragma unroll
  for (int q = 0; q < S::Q; ++q) {
    if (go<S>(q, 0) == -1 && i0 == 0) slo[cross_index<S>(q) * plane + own] = fin[q * p.sq + cell];
    if (go<S>(q, 0) == 1 && i0 == p.n0 - 1) shi[cross_index<S>(q) * plane + own] = fin[q * p.sq + cell];
  }
}

template <class S, class ST>
FaceFn pick_face_coll(int coll) {
  switch (coll) {
    case LBM_COLL_NONE: return face_kernel<S, ST, LBM_COLL_NONE>;
    case LBM_COLL_BGK: return face_kernel<S, ST, LBM_COLL_BGK>;
    case LBM_COLL_TRT: return face_kernel<S, ST, LBM_COLL_TRT>;
    case LBM_COLL_BGK_GUO: return face_kernel<S, ST, LBM_COLL_BGK_GUO>;
    case LBM_COLL_KBC:
      if constexpr (S::ID == LBM_D3Q19) return nullptr;  // lettuce: D2Q9/D3Q27 only
      else return face_kernel<S, ST, LBM_COLL_KBC>;
  }
  return nullptr;
}

template <class S>
FaceFn pick_face(int storage, int coll) {
  switch (storage) {
    case LBM_F32: return pick_face_coll<S, StF32>(coll);
    case LBM_F64: return pick_face_coll<S, StF64>(coll);
    case LBM_F16_SHIFTED: return pick_face_coll<S, StF16>(coll);
    case LBM_BF16_SHIFTED: return pick_face_coll<S, StBF16>(coll);
  }
  return nullptr;
}

template <class S>
FaceFn pick_pack(int storage) {
  switch (storage) {
    case LBM_F32: return pack_kernel<S, uint32_t>;
    case LBM_F64: return pack_kernel<S, uint64_t>;
    case LBM_F16_SHIFTED:
    case LBM_BF16_SHIFTED: return pack_kernel<S, uint16_t>;
  }
  return nullptr;
}

int halo_planes(int stencil) {
  if (stencil == LBM_D3Q19) return cross_count<D3Q19>();
  if (stencil == LBM_D3Q27) return cross_count<D3Q27>();
  return -1;
}

FaceFn pick_face_kernel(int stencil, int storage, int coll) {
  if (stencil == LBM_D3Q19) return pick_face<D3Q19>(storage, coll);
  if (stencil == LBM_D3Q27) return pick_face<D3Q27>(storage, coll);
  return nullptr;
}

FaceFn pick_pack_kernel(int stencil, int storage) {
  if (stencil == LBM_D3Q19) return pick_pack<D3Q19>(storage);
  if (stencil == LBM_D3Q27) return pick_pack<D3Q27>(storage);
  return nullptr;
}

}  // namespace lbm

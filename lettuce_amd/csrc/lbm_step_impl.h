// The fused stream + collide kernel.  One thread owns one cell; a block is a
// run of cells along the contiguous axis i2, so every population load and
// store of a wavefront is one coalesced 256-byte (FP32) transaction.
//
// Streaming follows lettuce's semantics exactly:
//  - pre-streaming (pull), lettuce/_simulation.py:241-256:
//      f_q(x) <- f_q(x - e_q) periodic, or f_q(x) if no_streaming_mask[q](x)
//    plus the halfway bounce-back rule of
//    examples/.../boundary/halfway_bounce_back_boundary.py:170-190:
//      f_q(x) <- f*_opp(q)(x) when x is fluid and x - e_q is solid;
//  - post-streaming (push) writes the same values from the source side, so
//    each destination slot has exactly one writer (no races).
// The per-cell operator (collision or boundary) is chosen by the cell class,
// i.e. by lettuce's no_collision_mask (lettuce/_simulation.py:258-305).
#pragma once

#include "lbm_device.h"

namespace lbm {

struct KParams {
  const void* fin;
  void* fout;
  const uint8_t* cls;
  const lbm_cell_class* ct;
  const double* bcf;
  int n0, n1, n2;     // extents
  long long s0;       // cell stride of i0 (elements)
  int s1;             // cell stride of i1 (elements); i2 stride is 1
  long long sq;       // population stride (elements)
  long long ncells;   // n0*n1*n2 (bc_field plane stride)
  int nb2;            // blocks along i2
  int nblocks;        // total blocks
  double tau_inv, trt_p, trt_m, beta, beta_inv;
  double rho0;
  double accel[3], guo;  // BGK_GUO
  int x0 = 0;             // first i0 plane of the launch (slab interiors)
  int nclasses = 0;       // entries of ct (0 without boundaries)
  uint32_t sr_seed = 0;   // stochastic rounding of BF16 storage (0: to nearest)
  uint32_t sr_step = 0;   // step index of the launch (its first step)
};

// One stored element: to nearest, or with stochastic rounding when the
// format supports it and the run asks for it (key: sr_step_key of the
// step, elem: the element's offset, used for the random bits only).
template <class ST>
__device__ __forceinline__ void store_elem(const KParams& p, typename ST::T* base, uint32_t i,
                                           typename ST::C v, typename ST::C sh, uint32_t key,
                                           uint32_t elem, int q) {
  if constexpr (ST::stochastic)
    __builtin_nontemporal_store(ST::enc_sr(v, sh, sr_bits(sr_cell(key, elem), q)),
                                elem_ptr(base, i));
  else
    ST::st(base, i, v, sh);
}

// Which row (i0, i1) and which run b2 of the row a block takes: rows in
// storage order, i1 fastest.  (XCD-aware orders, in which each of the 8
// XCDs sweeps its own contiguous share of the rows, measured slower on
// MI355X for D3Q19 256^3 FP32: 5.57 and 4.76 TB/s against 5.69 TB/s.)
__device__ __forceinline__ void block_tile(const KParams& p, int b, int& i0, int& i1, int& b2) {
  b2 = b % p.nb2;
  const int r = b / p.nb2;
  i1 = r % p.n1;
  i0 = p.x0 + r / p.n1;
}

template <class C>
__device__ __forceinline__ CollParams<C> coll_params(const KParams& p) {
  CollParams<C> cp;
  cp.tau_inv = (C)p.tau_inv;
  cp.trt_p = (C)p.trt_p;
  cp.trt_m = (C)p.trt_m;
  cp.beta = (C)p.beta;
  cp.beta_inv = (C)p.beta_inv;
  cp.accel[0] = (C)p.accel[0];
  cp.accel[1] = (C)p.accel[1];
  cp.accel[2] = (C)p.accel[2];
  cp.guo = (C)p.guo;
  return cp;
}

// periodic neighbour offsets along one axis: a[0] = i-1, a[1] = i, a[2] = i+1
// (wrapped), each multiplied by the axis stride.  Offsets along i0 are
// 64-bit: a cell's offset is its i0 plane's offset (uniform across a block,
// added to the population's base pointer in scalar registers) plus a 32-bit
// offset inside the plane, so only a plane must stay below 2^31 elements.
template <class I>
__device__ __forceinline__ void axis_offsets(int i, int n, I s, I (&a)[3]) {
  a[0] = (I)(i == 0 ? n - 1 : i - 1) * s;
  a[1] = (I)i * s;
  a[2] = (I)(i == n - 1 ? 0 : i + 1) * s;
}

// offset of x + sign * e_q inside its i0 plane, and of its i0 plane
template <class S>
__device__ __forceinline__ int nb_in_plane(int q, int sign, const int (&a1)[3],
                                           const int (&a2)[3]) {
  return a1[1 + sign * go<S>(q, 1)] + a2[1 + sign * go<S>(q, 2)];
}
template <class S, class X>
__device__ __forceinline__ X nb_plane(int q, int sign, const X (&a0)[3]) {
  return a0[1 + sign * go<S>(q, 0)];
}
// A kernel keeps its i0 offsets (X) 64-bit where the array needs it: the
// plane offset then goes into the population's base pointer.  With 32-bit
// i0 offsets (arrays whose offsets fit 32 bits) it goes into the 32-bit
// offset, which leaves the per-population bases unchanged across cells.
template <class T, class X>
__device__ __forceinline__ T* at_plane(T* base, X plane) {
  if constexpr (std::is_same_v<X, int>) return base;
  else return base + plane;
}
template <class X>
__device__ __forceinline__ uint32_t in_plane(X plane, int in) {
  if constexpr (std::is_same_v<X, int>) return (uint32_t)(plane + in);
  else return (uint32_t)in;
}
// full offset of x + sign * e_q; the masked kernels' cell offsets (class map
// lookups) are below 2^31, which the host checks for grids with boundaries
template <class S>
__device__ __forceinline__ int nb_offset(int q, int sign, const long long (&a0)[3],
                                         const int (&a1)[3], const int (&a2)[3]) {
  return (int)(nb_plane<S>(q, sign, a0) + nb_in_plane<S>(q, sign, a1, a2));
}
// the same with 32-bit i0 offsets (grids of fewer than 2^31 populations)
template <class S>
__device__ __forceinline__ int nb_offset(int q, int sign, const int (&a0)[3], const int (&a1)[3],
                                         const int (&a2)[3]) {
  return a0[1 + sign * go<S>(q, 0)] + nb_in_plane<S>(q, sign, a1, a2);
}

template <class ST>
__device__ __forceinline__ typename ST::C shift_of(double w, const KParams& p) {
  using C = typename ST::C;
  return ST::shifted ? (C)(w * p.rho0) : C(0);
}

__device__ __forceinline__ bool cls_solid(const KParams& p, int cell) {
  return (p.ct[p.cls[cell]].flags & LBM_CLASS_SOLID) != 0;
}

// Post-streaming populations of the cell at offset `cell` whose neighbour
// tables are a0/a1/a2, with the cell's own class fields (nsm, flags).
template <class S, class ST, bool MASK>
__device__ __forceinline__ void gather_pull(const KParams& p, int cell,
                                            const long long (&a0)[3], const int (&a1)[3],
                                            const int (&a2)[3], uint32_t nsm,
                                            uint32_t flags, typename ST::C (&f)[S::Q]) {
  using T = typename ST::T;
  const T* fin = (const T*)p.fin;
#pragma unroll
  for (int q = 0; q < S::Q; ++q) {
    const typename ST::C sh = shift_of<ST>(S::w[q], p);
    if (!MASK) {
      const long long plane = nb_plane<S>(q, -1, a0);
      f[q] = ST::ld(at_plane(fin + q * p.sq, plane), in_plane(plane, nb_in_plane<S>(q, -1, a1, a2)),
                    sh);
    } else {
      const int src = nb_offset<S>(q, -1, a0, a1, a2);
      bool hw = false;
      if ((flags & (LBM_CLASS_NEAR_SOLID | LBM_CLASS_SOLID)) == LBM_CLASS_NEAR_SOLID && q != 0)
        hw = cls_solid(p, src);
      if (hw) {
        const int o = S::opp[q];
        f[q] = ST::ld(fin + o * p.sq, cell, shift_of<ST>(S::w[o], p));
      } else {
        const int idx = ((nsm >> q) & 1u) ? cell : src;
        f[q] = ST::ld(fin + q * p.sq, idx, sh);
      }
    }
  }
}

// The first p.nclasses entries of the cell-class table, copied into LDS by
// the whole block before its first gather_batched (entries past them are
// never read for a valid class map; the array covers every uint8 id).
__device__ __forceinline__ void stage_classes(const KParams& p, lbm_cell_class* ctl) {
  const uint32_t* src = (const uint32_t*)p.ct;
  uint32_t* dst = (uint32_t*)ctl;
  const int words = p.nclasses * (int)(sizeof(lbm_cell_class) / 4);
  for (int i = threadIdx.x; i < words; i += blockDim.x) dst[i] = src[i];
}

// gather_pull / gather_own with the loads of a cell grouped into rounds.
// With boundaries (MASK) and pull streaming there are two schedules:
//  SPEC = false: first the class ids of the cell and of all its pull
//    sources (uint8 reads of the class map, mostly L2 hits; the class
//    fields come from the LDS table ctl), then every population;
//  SPEC = true: every lane loads its plain pull sources together with its
//    own class id; only lanes next to a solid (halfway bounce-back) or with
//    no-stream bits then load the class ids of their sources and their own
//    populations and replace the affected values.  A wave away from
//    boundaries makes one round trip instead of two, a wave at a boundary
//    two population round trips.  Measured on the 2048 x 512 D2Q9 channel:
//    the learned kernel (long compute after the loads) gains 3-4 %, the
//    step kernel (whose blocks wait for their wall wavefronts) loses 3-5 %.
// Left alone, the scheduler interleaves loads with their uses to save
// registers, which serialises the round trips of these latency-bound
// kernels; the scheduling barriers keep each round together.
// Returns the cell's no-stream bits, flags and op in nsm, flags, op.
// RAW: f receives the stored values without the shift of shifted formats.
template <class S, class ST, bool MASK, bool PULL, bool RAW = false, bool SPEC = false, class X>
__device__ __forceinline__ void gather_batched(const KParams& p, int cell, const X (&a0)[3],
                                               const int (&a1)[3], const int (&a2)[3],
                                               const lbm_cell_class* ctl, uint32_t& nsm,
                                               uint32_t& flags, uint32_t& op,
                                               typename ST::C (&f)[S::Q]) {
  using T = typename ST::T;
  const T* fin = (const T*)p.fin;
  // pull sources: plane and in-plane offsets (masked: the full offset)
  X splane[S::Q];
  int sinp[S::Q], src[S::Q];
#pragma unroll
  for (int q = 0; q < S::Q; ++q) {
    splane[q] = PULL ? nb_plane<S>(q, -1, a0) : a0[1];
    sinp[q] = PULL ? nb_in_plane<S>(q, -1, a1, a2) : a1[1] + a2[1];
    src[q] = MASK ? (int)(splane[q] + sinp[q]) : 0;
  }
  constexpr bool TWO_ROUNDS = MASK && PULL && !SPEC;
  uint32_t c0 = 0, cid0[S::Q];
  if (MASK) c0 = p.cls[cell];
  if constexpr (TWO_ROUNDS) {
#pragma unroll
    for (int q = 1; q < S::Q; ++q) cid0[q] = p.cls[src[q]];
    __builtin_amdgcn_sched_barrier(0);
  }
  nsm = 0;
  flags = 0;
  op = LBM_OP_COLLIDE;
  if (MASK) {
    nsm = ctl[c0].no_stream;
    flags = ctl[c0].flags;
    op = ctl[c0].op;
  }
  T raw[S::Q];
  if constexpr (TWO_ROUNDS) {
    const bool near =
        (flags & (LBM_CLASS_NEAR_SOLID | LBM_CLASS_SOLID)) == LBM_CLASS_NEAR_SOLID;
#pragma unroll
    for (int q = 0; q < S::Q; ++q) {
      int pop = q, idx = src[q];
      if (near && q != 0 && (ctl[cid0[q]].flags & LBM_CLASS_SOLID)) {
        pop = S::opp[q];  // the cell's own opposite population
        idx = cell;
      } else if ((nsm >> q) & 1u) {
        idx = cell;
      }
      raw[q] = fin[pop * p.sq + idx];
    }
  } else {
#pragma unroll
    for (int q = 0; q < S::Q; ++q)
      raw[q] = elem(at_plane(fin + q * p.sq, splane[q]), in_plane(splane[q], sinp[q]));
  }
  __builtin_amdgcn_sched_barrier(0);
  if constexpr (MASK && PULL && SPEC) {
    // halfway bounce-back: a near-solid fluid cell whose pull source is
    // solid takes its own opposite population; a no-stream bit keeps the
    // cell's own value (source offsets are in the grid, so the first round
    // read valid memory for these lanes too)
    const bool near =
        (flags & (LBM_CLASS_NEAR_SOLID | LBM_CLASS_SOLID)) == LBM_CLASS_NEAR_SOLID;
    if (near || nsm != 0) {
      // D2Q9 / D3Q19: the own populations in the same round as the class
      // ids; D3Q27 loads only those it needs, a round later (27 more
      // registers would cost the learned kernel occupancy)
      constexpr bool BATCH = S::Q <= 19;
      uint32_t cid[S::Q];
      T own[S::Q];
#pragma unroll
      for (int q = 1; q < S::Q; ++q) {
        cid[q] = p.cls[src[q]];
        if (BATCH) own[q] = fin[q * p.sq + cell];
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int q = 1; q < S::Q; ++q) {
        if (near && (ctl[cid[q]].flags & LBM_CLASS_SOLID)) {
          const int o = S::opp[q];
          raw[q] = BATCH ? own[o] : fin[o * p.sq + cell];
        } else if ((nsm >> q) & 1u) {
          raw[q] = BATCH ? own[q] : fin[q * p.sq + cell];
        }
      }
    }
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int q = 0; q < S::Q; ++q)
    f[q] = RAW ? ST::val(raw[q]) : ST::dec(raw[q], shift_of<ST>(S::w[q], p));
}

// The pull gather of gather_batched's SPEC schedule in two halves, for a
// kernel that issues the next cell's loads before it computes the current
// one.  gather_issue loads every plain pull source and the cell's own class
// id into raw and c0; gather_finish applies the class (lanes next to a
// solid or with no-stream bits load their class ids and own populations, a
// round later) and decodes into f.
template <class S, class ST, bool MASK, class X>
__device__ __forceinline__ void gather_issue(const KParams& p, int cell, const X (&a0)[3],
                                             const int (&a1)[3], const int (&a2)[3],
                                             typename ST::T (&raw)[S::Q], uint32_t& c0) {
  using T = typename ST::T;
  const T* fin = (const T*)p.fin;
#pragma unroll
  for (int q = 0; q < S::Q; ++q) {
    const X sp = nb_plane<S>(q, -1, a0);
    raw[q] = elem(at_plane(fin + q * p.sq, sp), in_plane(sp, nb_in_plane<S>(q, -1, a1, a2)));
  }
  c0 = MASK ? (uint32_t)p.cls[cell] : 0u;
}

template <class S, class ST, bool MASK, bool RAW, class X>
__device__ __forceinline__ void gather_finish(const KParams& p, int cell, const X (&a0)[3],
                                              const int (&a1)[3], const int (&a2)[3],
                                              const lbm_cell_class* ctl,
                                              typename ST::T (&raw)[S::Q], uint32_t c0,
                                              uint32_t& nsm, uint32_t& flags, uint32_t& op,
                                              typename ST::C (&f)[S::Q]) {
  using T = typename ST::T;
  const T* fin = (const T*)p.fin;
  nsm = 0;
  flags = 0;
  op = LBM_OP_COLLIDE;
  if constexpr (MASK) {
    nsm = ctl[c0].no_stream;
    flags = ctl[c0].flags;
    op = ctl[c0].op;
    const bool near =
        (flags & (LBM_CLASS_NEAR_SOLID | LBM_CLASS_SOLID)) == LBM_CLASS_NEAR_SOLID;
    if (near || nsm != 0) {
      uint32_t cid[S::Q];
      T own[S::Q];
#pragma unroll
      for (int q = 1; q < S::Q; ++q) {
        cid[q] = p.cls[nb_offset<S>(q, -1, a0, a1, a2)];
        own[q] = fin[q * p.sq + cell];
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int q = 1; q < S::Q; ++q) {
        if (near && (ctl[cid[q]].flags & LBM_CLASS_SOLID)) raw[q] = own[S::opp[q]];
        else if ((nsm >> q) & 1u) raw[q] = own[q];
      }
    }
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int q = 0; q < S::Q; ++q)
    f[q] = RAW ? ST::val(raw[q]) : ST::dec(raw[q], shift_of<ST>(S::w[q], p));
}

// the populations of the cell at offset plane + in (its i0 plane, and inside it)
template <class S, class ST>
__device__ __forceinline__ void gather_own(const KParams& p, long long plane, int in,
                                           typename ST::C (&f)[S::Q]) {
  using T = typename ST::T;
  const T* fin = (const T*)p.fin + plane;
#pragma unroll
  for (int q = 0; q < S::Q; ++q) f[q] = ST::ld(fin + q * p.sq, in, shift_of<ST>(S::w[q], p));
}

template <class S, class C, int COLL>
__device__ __forceinline__ void collide(C (&f)[S::Q], const CollParams<C>& cp) {
  if constexpr (COLL == LBM_COLL_BGK) collide_bgk<S, C>(f, cp);
  else if constexpr (COLL == LBM_COLL_TRT) collide_trt<S, C>(f, cp);
  else if constexpr (COLL == LBM_COLL_KBC) collide_kbc<S, C>(f, cp);
  else if constexpr (COLL == LBM_COLL_BGK_GUO) collide_bgk_guo<S, C>(f, cp);
  // LBM_COLL_NONE: f unchanged
}

// Populations of the neighbour x - dir of an outlet cell as the streaming
// step leaves them (its moments are those of its post-collision state when
// it collides with a momentum-conserving operator).
template <class S, class ST, bool PULL>
__device__ __forceinline__ void outlet_neighbour(const KParams& p, const lbm_cell_class& K,
                                                 int i0, int i1, int i2,
                                                 typename ST::C (&g)[S::Q]) {
  const int j0 = i0 - K.dir[0], j1 = i1 - K.dir[1], j2 = i2 - K.dir[2];
  const int w0 = j0 < 0 ? j0 + p.n0 : (j0 >= p.n0 ? j0 - p.n0 : j0);
  const int w1 = j1 < 0 ? j1 + p.n1 : (j1 >= p.n1 ? j1 - p.n1 : j1);
  const int w2 = j2 < 0 ? j2 + p.n2 : (j2 >= p.n2 ? j2 - p.n2 : j2);
  long long c0[3];
  int c1[3], c2[3];
  axis_offsets(w0, p.n0, p.s0, c0);
  axis_offsets(w1, p.n1, p.s1, c1);
  axis_offsets(w2, p.n2, 1, c2);
  const int nb = (int)(c0[1] + c1[1] + c2[1]);
  if (PULL) {
    const lbm_cell_class& KN = p.ct[p.cls[nb]];
    gather_pull<S, ST, true>(p, nb, c0, c1, c2, KN.no_stream, KN.flags, g);
  } else {
    gather_own<S, ST>(p, 0, nb, g);
  }
}

// The operator of a non-collision cell (one of LBM_OP_BOUNCE_BACK,
// LBM_OP_EQUILIBRIUM, LBM_OP_OUTLET, LBM_OP_ANTI_BOUNCE_BACK, LBM_OP_KEEP)
// on its populations f.  neighbour(K, g) fills g with the populations of
// the neighbour x - K.dir as the streaming step leaves them (OUTLET,
// ANTI_BOUNCE_BACK).
template <class S, class ST, class NB>
__device__ __forceinline__ void apply_boundary_op_nb(const KParams& p, uint32_t op, uint32_t flags,
                                                     int cell, typename ST::C (&f)[S::Q],
                                                     NB&& neighbour) {
  using C = typename ST::C;
  if (op == LBM_OP_BOUNCE_BACK) {
    // fullway bounce-back: f_q <- f_opp(q) (bounce_back_boundary.py:17-18)
    C g[S::Q];
#pragma unroll
    for (int q = 0; q < S::Q; ++q) g[q] = f[S::opp[q]];
#pragma unroll
    for (int q = 0; q < S::Q; ++q) f[q] = g[q];
  } else if (op == LBM_OP_EQUILIBRIUM) {
    // f <- feq(rho_bc, u_bc) (equilibrium_boundary_pu.py:79-84)
    const lbm_cell_class& K = p.ct[p.cls[cell]];
    C rho, u[3];
    if (flags & LBM_CLASS_BC_FIELD) {
      rho = (C)p.bcf[cell];
      u[0] = (C)p.bcf[p.ncells + cell];
      u[1] = (C)p.bcf[2 * p.ncells + cell];
      u[2] = (C)p.bcf[3 * p.ncells + cell];
    } else {
      rho = (C)K.rho;
      u[0] = (C)K.u[0];
      u[1] = (C)K.u[1];
      u[2] = (C)K.u[2];
    }
    equilibrium<S, C>(rho, u, f);
  } else if (op == LBM_OP_OUTLET) {
    // f <- feq(rho_out, u(x - dir)) (equilibrium_outlet_p.py:63-73)
    const lbm_cell_class& K = p.ct[p.cls[cell]];
    C g[S::Q];
    neighbour(K, g);
    C rho_n, u[3];
    moments<S, C>(g, rho_n, u);
    equilibrium<S, C>((C)K.rho, u, f);
  } else if (op == LBM_OP_ANTI_BOUNCE_BACK) {
    // anti-bounce-back outlet (anti_bounce_back_outlet.py:79-92): for the
    // populations v leaving through the outlet (e_v . dir = 1),
    // f_opp(v) <- -f_v + w_v rho (2 + (e_v.u_w)^2 / cs^4 - |u_w|^2 / cs^2)
    // with u_w = u + (u - u(x - dir)) / 2 and rho, u of the cell itself
    const lbm_cell_class& K = p.ct[p.cls[cell]];
    C g[S::Q];
    neighbour(K, g);
    C rho, u[3], rho_n, un[3];
    moments<S, C>(f, rho, u);
    moments<S, C>(g, rho_n, un);
    C uw[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) uw[d] = u[d] + C(0.5) * (u[d] - un[d]);
    const C uwuw = uw[0] * uw[0] + uw[1] * uw[1] + uw[2] * uw[2];
    const int d0 = K.dir[0], d1 = K.dir[1], d2 = K.dir[2];
#pragma unroll
    for (int q = 0; q < S::Q; ++q) {
      if (go<S>(q, 0) * d0 + go<S>(q, 1) * d1 + go<S>(q, 2) * d2 == 1) {
        const C eu = eu_of<S, C>(q, uw);
        f[S::opp[q]] = -f[q] + C(S::w[q]) * rho * (C(2) + C(9) * eu * eu - C(3) * uwuw);
      }
    }
  }
  // LBM_OP_KEEP: unchanged
}

template <class S, class ST, bool PULL>
__device__ __forceinline__ void apply_boundary_op(const KParams& p, uint32_t op, uint32_t flags,
                                                  int cell, int i0, int i1, int i2,
                                                  typename ST::C (&f)[S::Q]) {
  apply_boundary_op_nb<S, ST>(p, op, flags, cell, f,
                              [&](const lbm_cell_class& K, typename ST::C(&g)[S::Q]) {
                                outlet_neighbour<S, ST, PULL>(p, K, i0, i1, i2, g);
                              });
}

// Write the cell's post-collision populations: in place (no post-streaming)
// or pushed to the neighbours with the streaming rules of the destination.
// RAW: f holds values to store as they are (no shift subtracted).
template <class S, class ST, bool MASK, bool PUSH, bool RAW = false, class X>
__device__ __forceinline__ void store_cell(const KParams& p, int cell, const X (&a0)[3],
                                           const int (&a1)[3], const int (&a2)[3], uint32_t nsm,
                                           uint32_t flags, const typename ST::C (&f)[S::Q]) {
  using C = typename ST::C;
  using T = typename ST::T;
  T* fout = (T*)p.fout;
  const uint32_t key = ST::stochastic ? sr_step_key(p.sr_seed, p.sr_step) : 0u;
  if (!PUSH) {
    const uint32_t own = (uint32_t)(a0[1] + a1[1] + a2[1]);
#pragma unroll
    for (int q = 0; q < S::Q; ++q)
      store_elem<ST>(p, at_plane(fout + q * p.sq, a0[1]), in_plane(a0[1], a1[1] + a2[1]), f[q],
                     RAW ? C(0) : shift_of<ST>(S::w[q], p), key, own, q);
  } else if (!MASK) {
#pragma unroll
    for (int q = 0; q < S::Q; ++q)
      store_elem<ST>(p, at_plane(fout + q * p.sq, nb_plane<S>(q, +1, a0)),
                     in_plane(nb_plane<S>(q, +1, a0), nb_in_plane<S>(q, +1, a1, a2)), f[q],
                     RAW ? C(0) : shift_of<ST>(S::w[q], p), key,
                     (uint32_t)(nb_plane<S>(q, +1, a0) + nb_in_plane<S>(q, +1, a1, a2)), q);
  } else {
    // Push with the pull rule evaluated at the destination d = x + e_q:
    //   d fluid and x solid (halfway bounce-back)  -> d fills itself
    //   no_streaming_mask[q](d)                     -> d keeps its own value
    //   otherwise                                   -> d receives f_q(x)
    const bool self_solid = (flags & LBM_CLASS_SOLID) != 0;
#pragma unroll
    for (int q = 0; q < S::Q; ++q) {
      const C sh = RAW ? C(0) : shift_of<ST>(S::w[q], p);
      // (a) to the neighbour
      const int dst = nb_offset<S>(q, +1, a0, a1, a2);
      const lbm_cell_class& KD = p.ct[p.cls[dst]];
      const bool d_solid = (KD.flags & LBM_CLASS_SOLID) != 0;
      const bool d_keeps = ((KD.no_stream >> q) & 1u) != 0;
      const bool d_bounce = q != 0 && self_solid && !d_solid;
      if (!d_keeps && !d_bounce) store_elem<ST>(p, fout + q * p.sq, dst, f[q], sh, key, dst, q);
      // (b) the cell's own slot q
      bool own_bounce = false;
      if (q != 0 && !self_solid && (flags & LBM_CLASS_NEAR_SOLID))
        own_bounce = cls_solid(p, nb_offset<S>(q, -1, a0, a1, a2));
      if (own_bounce) store_elem<ST>(p, fout + q * p.sq, cell, f[S::opp[q]], sh, key, cell, q);
      else if ((nsm >> q) & 1u) store_elem<ST>(p, fout + q * p.sq, cell, f[q], sh, key, cell, q);
    }
  }
}

template <class S, class ST, int COLL, bool MASK, int MODE>
__global__ void __launch_bounds__(256) step_kernel(const KParams p) {
  using C = typename ST::C;
  using T = typename ST::T;
  constexpr bool PULL = (MODE & LBM_PRE_STREAMING) != 0;
  constexpr bool PUSH = (MODE & LBM_POST_STREAMING) != 0;

  // with boundaries, the cell-class table in LDS
  __shared__ lbm_cell_class ctl[MASK ? 256 : 1];
  if (MASK) {
    stage_classes(p, ctl);
    __syncthreads();
  }

  int i0, i1, b2;
  block_tile(p, blockIdx.x, i0, i1, b2);
  const int i2 = b2 * blockDim.x + threadIdx.x;
  if (i2 >= p.n2) return;

  long long a0[3];
  int a1[3], a2[3];
  axis_offsets(i0, p.n0, p.s0, a0);
  axis_offsets(i1, p.n1, p.s1, a1);
  axis_offsets(i2, p.n2, 1, a2);
  const int cell = (int)(a0[1] + a1[1] + a2[1]);  // used with boundaries only

  uint32_t nsm = 0, flags = 0, op = LBM_OP_COLLIDE;
  C f[S::Q];
  if (MASK) {
    gather_batched<S, ST, MASK, PULL>(p, cell, a0, a1, a2, ctl, nsm, flags, op, f);
  } else if (PULL) {
    gather_pull<S, ST, MASK>(p, cell, a0, a1, a2, nsm, flags, f);
  } else {
    gather_own<S, ST>(p, a0[1], a1[1] + a2[1], f);
  }

  const CollParams<C> cp = coll_params<C>(p);
  if (!MASK || op == LBM_OP_COLLIDE) collide<S, C, COLL>(f, cp);
  else apply_boundary_op<S, ST, PULL>(p, op, flags, cell, i0, i1, i2, f);
  store_cell<S, ST, MASK, PUSH>(p, cell, a0, a1, a2, nsm, flags, f);
}


// Shift a value across the lanes of the wavefront with DPP: DIR = +1 gives
// lane i the value of lane i - 1 (wave_shr:1), DIR = -1 the value of lane
// i + 1 (wave_shl:1).  The lane at the wave's edge keeps its own value.
template <int DIR>
__device__ __forceinline__ int wave_shift_i32(int v) {
  return __builtin_amdgcn_update_dpp(v, v, DIR > 0 ? 0x138 : 0x130, 0xf, 0xf, false);
}
template <int DIR, class T>
__device__ __forceinline__ T wave_shift(T v) {
  if constexpr (sizeof(T) == 8) {
    long long b = __builtin_bit_cast(long long, v);
    const int lo = wave_shift_i32<DIR>((int)(b & 0xffffffffll));
    const int hi = wave_shift_i32<DIR>((int)(b >> 32));
    return __builtin_bit_cast(T, ((long long)hi << 32) | (unsigned int)lo);
  } else if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, wave_shift_i32<DIR>(__builtin_bit_cast(int, v)));
  } else {
    const unsigned short s = __builtin_bit_cast(unsigned short, v);
    return __builtin_bit_cast(T, (unsigned short)wave_shift_i32<DIR>((int)s));
  }
}

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------
using KernelFn = void (*)(const KParams);

template <class S, class ST, int COLL, bool MASK>
KernelFn pick_mode(int mode) {
  switch (mode) {
    case 0: return step_kernel<S, ST, COLL, MASK, 0>;
    case 1: return step_kernel<S, ST, COLL, MASK, 1>;
    case 2: return step_kernel<S, ST, COLL, MASK, 2>;
    case 3: return step_kernel<S, ST, COLL, MASK, 3>;
  }
  return nullptr;
}

template <class S, class ST, int COLL>
KernelFn pick_mask(bool mask, int mode) {
  return mask ? pick_mode<S, ST, COLL, true>(mode) : pick_mode<S, ST, COLL, false>(mode);
}

template <class S, class ST>
KernelFn pick_coll(int coll, bool mask, int mode) {
  switch (coll) {
    case LBM_COLL_NONE: return pick_mask<S, ST, LBM_COLL_NONE>(mask, mode);
    case LBM_COLL_BGK: return pick_mask<S, ST, LBM_COLL_BGK>(mask, mode);
    case LBM_COLL_TRT: return pick_mask<S, ST, LBM_COLL_TRT>(mask, mode);
    case LBM_COLL_BGK_GUO: return pick_mask<S, ST, LBM_COLL_BGK_GUO>(mask, mode);
    case LBM_COLL_KBC:
      if constexpr (S::ID == LBM_D3Q19) return nullptr;  // lettuce: D2Q9/D3Q27 only
      else return pick_mask<S, ST, LBM_COLL_KBC>(mask, mode);
  }
  return nullptr;
}

template <class S>
KernelFn pick_kernel(int storage, int coll, bool mask, int mode) {
  switch (storage) {
    case LBM_F32: return pick_coll<S, StF32>(coll, mask, mode);
    case LBM_F64: return pick_coll<S, StF64>(coll, mask, mode);
    case LBM_F16_SHIFTED: return pick_coll<S, StF16>(coll, mask, mode);
    case LBM_BF16_SHIFTED: return pick_coll<S, StBF16>(coll, mask, mode);
    case kCodecBF16SR: return pick_coll<S, StBF16SR>(coll, mask, mode);
  }
  return nullptr;
}

// sets the message returned by lbm_last_error and returns code
int report_error(int code, const char* msg);

}  // namespace lbm

// Two fused pull steps per pass over memory for 2D lattices with boundaries
// (D2Q9 with a cell-class map): f_out = K(K(f_in)) with K the masked pull
// step of step_kernel (halfway bounce-back, no-stream bits, fullway
// bounce-back, equilibrium boundaries, equilibrium and anti-bounce-back
// outlets).  The intermediate populations stay in LDS, so a pair of steps
// reads every population about (cx + 4) / cx times and writes it once,
// instead of reading and writing it twice.
//
// A block owns whole rows of the grid (blockDim = n2, the contiguous
// extent, one thread per cell of a row) and sweeps a chunk of cx rows along
// i1.  Per iteration it
//  A: computes the first step S1 of row x0 - 2 + it from the input (every
//     boundary rule reads the input, as a single step does) and writes it,
//     rounded to the storage format, with the row's class ids and SOLID
//     flags, into an LDS ring of R = 6 rows;
//  B: computes the second step S2 of row x0 - 4 + it from the ring rows
//     it - 4 .. it and writes it to the output.
// S2 reaches two rows: one for the pull, one more for the neighbour an
// outlet cell gathers (x - dir, pulled).  Rows wrap in the ring as the grid
// wraps periodically, and a row is complete in the block, so the periodic
// wrap along i2 is the ring's own.  The loads of S1 are software-pipelined:
// the populations of row it + 1 and the class ids of row it + 2 are in
// flight while the block computes S2 of the current iteration.
// S1 values are rounded as a stored step rounds them (stochastic rounding
// with the first step's bits), so the result equals two step_kernel
// launches up to the rounding of the collision arithmetic (bit for bit
// without collision).
#include "lbm_tb2.h"

namespace lbm {

namespace {

constexpr int kTb2dHalo = 2;
constexpr int kTb2dRing = 2 * kTb2dHalo + 2;

// the class ids of a cell and of its pull sources x - e_q (the first load
// round of a masked pull step); cell offsets are in-plane (2D: the grid)
template <class S>
struct PullIds {
  uint32_t c0;
  uint32_t src[S::Q];
};

template <class S>
__device__ __forceinline__ void load_pull_ids(const KParams& p, int cell, const int (&a1)[3],
                                              const int (&a2)[3], PullIds<S>& ids) {
  ids.c0 = p.cls[cell];
#pragma unroll
  for (int q = 1; q < S::Q; ++q) ids.src[q] = p.cls[nb_in_plane<S>(q, -1, a1, a2)];
  ids.src[0] = ids.c0;
}

// the second round: the stored value the pull step takes for each
// population (halfway bounce-back: the cell's own opposite population where
// the fluid cell's source is solid; a no-stream bit: the cell's own value)
template <class S, class ST>
__device__ __forceinline__ void load_pull_pops(const KParams& p, int cell, const int (&a1)[3],
                                               const int (&a2)[3], const PullIds<S>& ids,
                                               const lbm_cell_class* ctl,
                                               typename ST::T (&raw)[S::Q]) {
  using T = typename ST::T;
  const T* fin = (const T*)p.fin;
  const uint32_t flags = ctl[ids.c0].flags, nsm = ctl[ids.c0].no_stream;
  const bool near = (flags & (LBM_CLASS_NEAR_SOLID | LBM_CLASS_SOLID)) == LBM_CLASS_NEAR_SOLID;
  // selects, not branches: every lane reads its sources' flags
#pragma unroll
  for (int q = 0; q < S::Q; ++q) {
    const bool bounce = near & (q != 0) & ((ctl[ids.src[q]].flags & LBM_CLASS_SOLID) != 0);
    const bool own = bounce | (((nsm >> q) & 1u) != 0);
    const long long off = (bounce ? S::opp[q] : q) * p.sq;
    raw[q] = fin[off + (own ? cell : nb_in_plane<S>(q, -1, a1, a2))];
  }
}

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// A block-wide barrier that orders LDS only: each wave's LDS operations
// complete before it, while its global loads stay in flight (a full
// __syncthreads waits for them too, which would empty the load pipeline
// every iteration)
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

}  // namespace

template <class S, class ST, int COLL>
__global__ void __launch_bounds__(1024) tb2d_kernel(const Tb2dParams P) {
  using C = typename ST::C;
  using T = typename ST::T;
  constexpr int Q = S::Q;
  constexpr int H = kTb2dHalo, R = kTb2dRing;
  extern __shared__ __align__(16) unsigned char tb2d_lds[];
  const KParams& p = P.k;
  const int n2 = p.n2;
  lbm_cell_class* ctl = (lbm_cell_class*)tb2d_lds;
  T* ring = (T*)(tb2d_lds + P.ring_offset);           // [R][Q][n2]
  uint8_t* cring = (uint8_t*)(ring + R * Q * n2);      // [R][n2] class ids
  uint8_t* sring = cring + R * n2;                     // [R][n2] SOLID flags
  stage_classes(p, ctl);

  const int y = threadIdx.x;
  const int x0 = blockIdx.x * P.cx;
  const int nit = min(P.cx, p.n1 - x0) + 2 * H;  // S1 rows x0 - H .. x0 + cx - 1 + H
  int a2[3];
  axis_offsets(y, n2, 1, a2);
  long long a0[3];
  axis_offsets(0, 1, 0LL, a0);  // the grid's single i0 plane
  const CollParams<C> cp = coll_params<C>(p);
  constexpr bool sr = ST::stochastic;
  const uint32_t key1 = sr ? sr_step_key(p.sr_seed, p.sr_step) : 0u;
  KParams p2 = p;  // the second step's stores
  p2.sr_step = p.sr_step + 1;

  // in-plane offsets of S1 row counter c (grid row x0 - H + c)
  auto rows = [&](int c, int (&a1)[3]) {
    axis_offsets(wrap(x0 - H + c, p.n1), p.n1, p.s1, a1);
    return a1[1] + a2[1];
  };
  // ring pull: the populations of ring cell (rr, yy) after the step's
  // streaming, with the rules of its own class; rows[k], cols[k]: the ring
  // rows rr - 1 + k and the columns yy - 1 + k, wrapped
  auto ring_pull_at = [&](const int (&rows3)[3], const int (&cols3)[3], C(&g)[Q]) {
    const int own = rows3[1] * n2 + cols3[1];
    const uint32_t c0 = cring[own];
    const uint32_t flags = ctl[c0].flags, nsm = ctl[c0].no_stream;
    const bool near = (flags & (LBM_CLASS_NEAR_SOLID | LBM_CLASS_SOLID)) == LBM_CLASS_NEAR_SOLID;
    const int own_pop = own + rows3[1] * (Q - 1) * n2;  // + q n2: the cell's own population q
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int src = rows3[1 - go<S>(q, 1)] * n2 + cols3[1 - go<S>(q, 2)];
      const bool bounce = near & (q != 0) & (sring[src] != 0);
      const bool keep = ((nsm >> q) & 1u) != 0;
      const int idx = bounce ? own_pop + S::opp[q] * n2
                             : (keep ? own_pop + q * n2
                                     : src + rows3[1 - go<S>(q, 1)] * (Q - 1) * n2 + q * n2);
      g[q] = ST::dec(ring[idx], shift_of<ST>(S::w[q], p));
    }
  };
  auto ring_pull = [&](int rr, int yy, C(&g)[Q]) {
    const int rows3[3] = {wrap(rr - 1, R), rr, wrap(rr + 1, R)};
    const int cols3[3] = {wrap(yy - 1, n2), yy, wrap(yy + 1, n2)};
    ring_pull_at(rows3, cols3, g);
  };
  const int cols_own[3] = {wrap(y - 1, n2), y, wrap(y + 1, n2)};

  __syncthreads();  // the class table
  // Load pipeline.  Entering phase A of row counter it: raw holds (or
  // receives) the populations of row it, ids the class ids of row it + 1,
  // c0 the class id of row it.  Phase A decodes raw, then reloads it with
  // row it + 1 (which needs ids, loaded an iteration earlier) and ids with
  // row it + 2, in place: no loop-carried copy waits for a load.  The loop
  // is split so that every iteration of the main loop stores (phase B):
  // the wait for row it + 1's ids then lets the previous stores run on.
  PullIds<S> ids;
  T raw[Q];
  uint32_t c0 = 0;
  {
    int a1[3];
    const int cell = rows(0, a1);
    load_pull_ids<S>(p, cell, a1, a2, ids);
    load_pull_pops<S, ST>(p, cell, a1, a2, ids, ctl, raw);
    c0 = ids.c0;
    int b1[3];
    load_pull_ids<S>(p, rows(1, b1), b1, a2, ids);
  }

  // ---- A: S1 of row counter it into ring slot it % R ----
  auto phase_a = [&](int it) {
    int a1[3];
    const int cell = rows(it, a1);
    const uint32_t c = c0;
    const uint32_t flags = ctl[c].flags, op = ctl[c].op;
    C f[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = ST::dec(raw[q], shift_of<ST>(S::w[q], p));
    // the next rows' loads go out before this row's arithmetic
    if (it + 1 < nit) {
      int b1[3];
      load_pull_pops<S, ST>(p, rows(it + 1, b1), b1, a2, ids, ctl, raw);
      c0 = ids.c0;
      if (it + 2 < nit) {
        int c1[3];
        load_pull_ids<S>(p, rows(it + 2, c1), c1, a2, ids);
      }
    }
    if (op == LBM_OP_COLLIDE) collide<S, C, COLL>(f, cp);
    else apply_boundary_op<S, ST, true>(p, op, flags, cell, 0, wrap(x0 - H + it, p.n1), y, f);
    const int slot = it % R;
    const uint32_t hs = sr ? sr_cell(key1, (uint32_t)cell) : 0u;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const C sh = shift_of<ST>(S::w[q], p);
      ring[(slot * Q + q) * n2 + y] = sr ? ST::enc_sr(f[q], sh, sr_bits(hs, q)) : ST::enc(f[q], sh);
    }
    cring[slot * n2 + y] = (uint8_t)c;
    sring[slot * n2 + y] = (flags & LBM_CLASS_SOLID) ? 1 : 0;
  };

  // ---- B: S2 of row counter c = it - H from ring rows c - H .. c + H ----
  auto phase_b = [&](int it) {
    const int c = it - H;
    const int rr = c % R;
    C f[Q];
    const int rows3[3] = {(c - 1) % R, rr, (c + 1) % R};
    ring_pull_at(rows3, cols_own, f);
    const uint32_t cb = cring[rr * n2 + y];
    const uint32_t flags = ctl[cb].flags, op = ctl[cb].op, nsm = ctl[cb].no_stream;
    int a1[3];
    const int cell = rows(c, a1);
    if (op == LBM_OP_COLLIDE) {
      collide<S, C, COLL>(f, cp);
    } else {
      apply_boundary_op_nb<S, ST>(p, op, flags, cell, f,
                                  [&](const lbm_cell_class& K, C(&g)[Q]) {
                                    ring_pull(wrap(rr - K.dir[1], R), wrap(y - K.dir[2], n2), g);
                                  });
    }
    store_cell<S, ST, true, false>(p2, cell, a0, a1, a2, nsm, flags, f);
  };

  for (int it = 0; it < 2 * H; ++it) {
    phase_a(it);
    lds_barrier();
  }
  for (int it = 2 * H; it < nit; ++it) {
    phase_a(it);
    lds_barrier();
    phase_b(it);
  }
}

namespace {

template <class S, class ST>
Tb2dFn pick_tb2d_coll(int coll) {
  switch (coll) {
    case LBM_COLL_NONE: return tb2d_kernel<S, ST, LBM_COLL_NONE>;
    case LBM_COLL_BGK: return tb2d_kernel<S, ST, LBM_COLL_BGK>;
    case LBM_COLL_TRT: return tb2d_kernel<S, ST, LBM_COLL_TRT>;
    case LBM_COLL_KBC: return tb2d_kernel<S, ST, LBM_COLL_KBC>;
    case LBM_COLL_BGK_GUO: return tb2d_kernel<S, ST, LBM_COLL_BGK_GUO>;
  }
  return nullptr;
}

}  // namespace

Tb2dFn pick_tb2d_kernel(int stencil, int codec, int coll) {
  if (stencil != LBM_D2Q9) return nullptr;
  switch (codec) {
    case LBM_F32: return pick_tb2d_coll<D2Q9, StF32>(coll);
    case LBM_F16_SHIFTED: return pick_tb2d_coll<D2Q9, StF16>(coll);
    case LBM_BF16_SHIFTED: return pick_tb2d_coll<D2Q9, StBF16>(coll);
    case kCodecBF16SR: return pick_tb2d_coll<D2Q9, StBF16SR>(coll);
  }
  return nullptr;
}

int64_t tb2d_lds_bytes(int nclasses, int n2, int storage_bytes, int64_t* ring_offset) {
  const int64_t table = ((int64_t)nclasses * (int64_t)sizeof(lbm_cell_class) + 15) / 16 * 16;
  if (ring_offset) *ring_offset = table;
  return table + (int64_t)kTb2dRing * D2Q9::Q * n2 * storage_bytes + 2 * (int64_t)kTb2dRing * n2;
}

int tb2d_halo() { return kTb2dHalo; }

}  // namespace lbm

// Two fused pull steps per pass over memory for 2D lattices (D2Q9), with a
// cell-class map or periodic without one: f_out = K(K(f_in)) with K the pull
// step of step_kernel (halfway bounce-back, no-stream bits, fullway
// bounce-back, equilibrium boundaries, equilibrium and anti-bounce-back
// outlets).  The intermediate populations stay in LDS, so a pair of steps
// reads every population about (cx + 4) / cx times and writes it once,
// instead of reading and writing it twice.
//
// A block owns whole rows of the grid (NR x n2 threads, n2 the contiguous
// extent: one thread per cell of NR rows) and sweeps a chunk of cx rows
// along i1, NR rows per iteration.  Per iteration, for each of its rows, it
//  A: computes the first step S1 of row x0 - 2 + rc from the input (every
//     boundary rule reads the input, as a single step does) and writes it,
//     rounded to the storage format, with the row's class ids and SOLID
//     flags, into an LDS ring of 4 + 2 NR rows;
//  B: computes the second step S2 of row x0 - 4 + rc from the ring rows
//     rc - 4 .. rc and writes it to the output.
// Two rows per iteration (where 2 n2 <= 1024 and the ring fits) give a
// block twice the wavefronts to hide the load and LDS latencies behind.
// S2 reaches two rows: one for the pull, one more for the neighbour an
// outlet cell gathers (x - dir, pulled).  Rows wrap in the ring as the grid
// wraps periodically, and a row is complete in the block, so the periodic
// wrap along i2 is the ring's own.  The loads of S1 are software-pipelined:
// the populations of row rc + NR and the class ids of row rc + 2 NR are in
// flight while the block computes S1 and S2 of the current iteration.
// Wavefronts without boundary cells pull plainly, without reading their
// sources' classes.
// S1 values are rounded as a stored step rounds them (stochastic rounding
// with the first step's bits), so the result equals two step_kernel
// launches up to the rounding of the collision arithmetic (bit for bit
// without collision).
#include "lbm_tb2.h"

namespace lbm {

namespace {

constexpr int kTb2dHalo = 2;
// S1 rows in the ring with nr rows per iteration and one barrier per
// iteration: phase B of iteration k reads rows k nr - 2 H .. k nr + nr - 1
// while phase A of iteration k + 1 (of faster wavefronts) writes the next
// nr rows
__host__ __device__ constexpr int tb2d_ring(int nr) { return 2 * kTb2dHalo + 2 * nr; }

// the class ids of a cell and of its pull sources x - e_q (the first load
// round of a masked pull step); cell offsets are in-plane (2D: the grid)
template <class S>
struct PullIds {
  uint32_t c0;
  uint32_t src[S::Q];
};

// (a grid without a class map: every cell is class 0, a plain collision)
template <class S, bool MASK>
__device__ __forceinline__ void load_pull_ids(const KParams& p, int cell, const int (&a1)[3],
                                              const int (&a2)[3], PullIds<S>& ids) {
  if constexpr (!MASK) {
#pragma unroll
    for (int q = 0; q < S::Q; ++q) ids.src[q] = 0;
    ids.c0 = 0;
    return;
  }
  ids.c0 = elem(p.cls, (uint32_t)cell);
#pragma unroll
  for (int q = 1; q < S::Q; ++q) ids.src[q] = elem(p.cls, (uint32_t)nb_in_plane<S>(q, -1, a1, a2));
  ids.src[0] = ids.c0;
}

// The classes' fields the kernel reads per cell, in LDS: flags, op and
// no-stream bits by class id
struct ClassLds {
  uint8_t flags[256];
  uint8_t op[256];
  uint32_t nsm[256];
};

// the second round: the stored value the pull step takes for each
// population (halfway bounce-back: the cell's own opposite population where
// the fluid cell's source is solid; a no-stream bit: the cell's own value).
// Element offsets are 32-bit (the host checks that the array fits).
template <class S, class ST>
__device__ __forceinline__ void load_pull_pops(const KParams& p, int cell, const int (&a1)[3],
                                               const int (&a2)[3], const PullIds<S>& ids,
                                               const ClassLds& cl, typename ST::T (&raw)[S::Q]) {
  using T = typename ST::T;
  const T* fin = (const T*)p.fin;
  const uint32_t flags = cl.flags[ids.c0], nsm = cl.nsm[ids.c0];
  const bool near = (flags & (LBM_CLASS_NEAR_SOLID | LBM_CLASS_SOLID)) == LBM_CLASS_NEAR_SOLID;
  const uint32_t sq = (uint32_t)p.sq;
  if (!__any(near || nsm != 0)) {
    // a wavefront away from boundaries: the plain pull
#pragma unroll
    for (int q = 0; q < S::Q; ++q)
      raw[q] = elem(fin, (uint32_t)q * sq + (uint32_t)nb_in_plane<S>(q, -1, a1, a2));
    return;
  }
  // selects, not branches: every lane reads its sources' flags
#pragma unroll
  for (int q = 0; q < S::Q; ++q) {
    const bool bounce = near & (q != 0) & ((cl.flags[ids.src[q]] & LBM_CLASS_SOLID) != 0);
    const bool own = bounce | (((nsm >> q) & 1u) != 0);
    const uint32_t pop = (uint32_t)(bounce ? S::opp[q] : q) * sq;
    raw[q] = elem(fin, pop + (uint32_t)(own ? cell : nb_in_plane<S>(q, -1, a1, a2)));
  }
}

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

}  // namespace

template <class S, class ST, int COLL, int NR, bool MASK>
__global__ void __launch_bounds__(1024) tb2d_kernel(const Tb2dParams P) {
  using C = typename ST::C;
  using T = typename ST::T;
  constexpr int Q = S::Q;
  constexpr int H = kTb2dHalo, R = tb2d_ring(NR);
  extern __shared__ __align__(16) unsigned char tb2d_lds[];
  __shared__ ClassLds cl;
  const KParams& p = P.k;
  const int n2 = p.n2;
  T* ring = (T*)tb2d_lds;                          // [R][Q][n2]
  uint8_t* cring = (uint8_t*)(ring + R * Q * n2);  // [R][n2] class ids
  uint8_t* sring = cring + R * n2;                 // [R][n2] SOLID flags
  if (MASK) {
    for (int i = threadIdx.x; i < p.nclasses; i += blockDim.x) {
      cl.flags[i] = p.ct[i].flags;
      cl.op[i] = p.ct[i].op;
      cl.nsm[i] = p.ct[i].no_stream;
    }
  } else if (threadIdx.x == 0) {
    cl.flags[0] = 0;
    cl.op[0] = LBM_OP_COLLIDE;
    cl.nsm[0] = 0;
  }

  // NR rows per iteration: thread row r (whole wavefronts, n2 % 64 == 0)
  const int r = NR == 1 ? 0 : (int)threadIdx.x / n2;
  const int y = (int)threadIdx.x - r * n2;
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
  // crow[k]: ring row rr - 1 + k times n2 (class and flag rings), prow[k]:
  // the same times Q (population ring); cols[k]: column yy - 1 + k, wrapped
  auto ring_pull_at = [&](const int (&crow)[3], const int (&prow)[3], const int (&cols3)[3],
                          C(&g)[Q]) {
    const int own = crow[1] + cols3[1];
    const uint32_t c0 = cring[own];
    const uint32_t flags = cl.flags[c0], nsm = cl.nsm[c0];
    const bool near = (flags & (LBM_CLASS_NEAR_SOLID | LBM_CLASS_SOLID)) == LBM_CLASS_NEAR_SOLID;
    const int own_pop = prow[1] + cols3[1];  // + q n2: the cell's own population q
    if (!__any(near || nsm != 0)) {
      // a wavefront away from boundaries: the plain pull
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int ex = 1 - go<S>(q, 1), ey = 1 - go<S>(q, 2);
        g[q] = ST::dec(ring[prow[ex] + q * n2 + cols3[ey]], shift_of<ST>(S::w[q], p));
      }
      return;
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int ex = 1 - go<S>(q, 1), ey = 1 - go<S>(q, 2);
      const bool bounce = near & (q != 0) & (sring[crow[ex] + cols3[ey]] != 0);
      const bool keep = ((nsm >> q) & 1u) != 0;
      const int idx = bounce ? own_pop + S::opp[q] * n2
                             : (keep ? own_pop + q * n2 : prow[ex] + q * n2 + cols3[ey]);
      g[q] = ST::dec(ring[idx], shift_of<ST>(S::w[q], p));
    }
  };
  auto ring_pull = [&](int rr, int yy, C(&g)[Q]) {
    const int r3[3] = {wrap(rr - 1, R), rr, wrap(rr + 1, R)};
    const int crow[3] = {r3[0] * n2, r3[1] * n2, r3[2] * n2};
    const int prow[3] = {crow[0] * Q, crow[1] * Q, crow[2] * Q};
    const int cols3[3] = {wrap(yy - 1, n2), yy, wrap(yy + 1, n2)};
    ring_pull_at(crow, prow, cols3, g);
  };
  const int cols_own[3] = {wrap(y - 1, n2), y, wrap(y + 1, n2)};

  __syncthreads();  // the class table
  // Load pipeline.  Iteration k handles S1 row counters k NR + r.  Entering
  // phase A for row counter rc: raw holds (or receives) the populations of
  // row rc, ids the class ids of row rc + NR, c0 the class id of row rc.
  // Phase A decodes raw, then reloads it with row rc + NR (which needs ids,
  // loaded an iteration earlier) and ids with row rc + 2 NR, in place: no
  // loop-carried copy waits for a load.  The loop is split so that every
  // iteration of the main loop stores (phase B): the wait for the next
  // rows' ids then lets the previous stores run on.
  PullIds<S> ids;
  T raw[Q];
  uint32_t c0 = 0;
  {
    int a1[3];
    const int cell = rows(r, a1);
    load_pull_ids<S, MASK>(p, cell, a1, a2, ids);
    load_pull_pops<S, ST>(p, cell, a1, a2, ids, cl, raw);
    c0 = ids.c0;
    if (r + NR < nit) {
      int b1[3];
      load_pull_ids<S, MASK>(p, rows(r + NR, b1), b1, a2, ids);
    }
  }

  // ---- A: S1 of row counter rc into ring slot rc % R ----
  auto phase_a = [&](int rc) {
    int a1[3];
    const int cell = rows(rc, a1);
    const uint32_t c = c0;
    const uint32_t flags = cl.flags[c], op = cl.op[c];
    C f[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = ST::dec(raw[q], shift_of<ST>(S::w[q], p));
    // the next rows' loads go out before this row's arithmetic
    if (rc + NR < nit) {
      int b1[3];
      load_pull_pops<S, ST>(p, rows(rc + NR, b1), b1, a2, ids, cl, raw);
      c0 = ids.c0;
      if (rc + 2 * NR < nit) {
        int c1[3];
        load_pull_ids<S, MASK>(p, rows(rc + 2 * NR, c1), c1, a2, ids);
      }
    }
    if (op == LBM_OP_COLLIDE) collide<S, C, COLL>(f, cp);
    else apply_boundary_op<S, ST, true>(p, op, flags, cell, 0, wrap(x0 - H + rc, p.n1), y, f);
    const int slot = rc % R;
    const uint32_t hs = sr ? sr_cell(key1, (uint32_t)cell) : 0u;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const C sh = shift_of<ST>(S::w[q], p);
      ring[(slot * Q + q) * n2 + y] = sr ? ST::enc_sr(f[q], sh, sr_bits(hs, q)) : ST::enc(f[q], sh);
    }
    cring[slot * n2 + y] = (uint8_t)c;
    sring[slot * n2 + y] = (flags & LBM_CLASS_SOLID) ? 1 : 0;
  };

  // ---- B: S2 of row counter c = rc - H from ring rows c - H .. c + H ----
  auto phase_b = [&](int rc) {
    const int c = rc - H;
    const int rr = c % R;
    C f[Q];
    const int crow[3] = {(c - 1) % R * n2, rr * n2, (c + 1) % R * n2};
    const int prow[3] = {crow[0] * Q, crow[1] * Q, crow[2] * Q};
    ring_pull_at(crow, prow, cols_own, f);
    const uint32_t cb = cring[crow[1] + y];
    const uint32_t flags = cl.flags[cb], op = cl.op[cb], nsm = cl.nsm[cb];
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

  // S2 of row counter rc - H needs S1 up to row rc, written by this
  // iteration or earlier; rows below 2 H have no S2 (halo)
  int k = 0;
  for (; k * NR < 2 * H; ++k) {
    phase_a(k * NR + r);
    __syncthreads();
  }
  for (; (k + 1) * NR <= nit; ++k) {
    phase_a(k * NR + r);
    __syncthreads();
    phase_b(k * NR + r);
  }
  if (k * NR < nit) {  // a last iteration with fewer than NR rows
    const bool valid = k * NR + r < nit;
    if (valid) phase_a(k * NR + r);
    __syncthreads();
    if (valid) phase_b(k * NR + r);
  }
}

namespace {

template <class S, class ST, int COLL>
Tb2dFn pick_tb2d_rows(int nr, bool mask) {
  if (mask) return nr == 2 ? tb2d_kernel<S, ST, COLL, 2, true> : tb2d_kernel<S, ST, COLL, 1, true>;
  return nr == 2 ? tb2d_kernel<S, ST, COLL, 2, false> : tb2d_kernel<S, ST, COLL, 1, false>;
}

template <class S, class ST>
Tb2dFn pick_tb2d_coll(int coll, int nr, bool mask) {
  switch (coll) {
    case LBM_COLL_NONE: return pick_tb2d_rows<S, ST, LBM_COLL_NONE>(nr, mask);
    case LBM_COLL_BGK: return pick_tb2d_rows<S, ST, LBM_COLL_BGK>(nr, mask);
    case LBM_COLL_TRT: return pick_tb2d_rows<S, ST, LBM_COLL_TRT>(nr, mask);
    case LBM_COLL_KBC: return pick_tb2d_rows<S, ST, LBM_COLL_KBC>(nr, mask);
    case LBM_COLL_BGK_GUO: return pick_tb2d_rows<S, ST, LBM_COLL_BGK_GUO>(nr, mask);
  }
  return nullptr;
}

}  // namespace

Tb2dFn pick_tb2d_kernel(int stencil, int codec, int coll, int nr, bool mask) {
  if (stencil != LBM_D2Q9 || (nr != 1 && nr != 2)) return nullptr;
  switch (codec) {
    case LBM_F32: return pick_tb2d_coll<D2Q9, StF32>(coll, nr, mask);
    case LBM_F16_SHIFTED: return pick_tb2d_coll<D2Q9, StF16>(coll, nr, mask);
    case LBM_BF16_SHIFTED: return pick_tb2d_coll<D2Q9, StBF16>(coll, nr, mask);
    case kCodecBF16SR: return pick_tb2d_coll<D2Q9, StBF16SR>(coll, nr, mask);
  }
  return nullptr;
}

int64_t tb2d_lds_bytes(int n2, int storage_bytes, int nr) {
  return (int64_t)tb2d_ring(nr) * n2 * (D2Q9::Q * storage_bytes + 2);
}

int64_t tb2d_static_lds_bytes() { return (int64_t)sizeof(ClassLds); }

int tb2d_halo() { return kTb2dHalo; }

}  // namespace lbm

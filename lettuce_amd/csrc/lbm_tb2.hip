// Two fused pull steps per pass over memory (temporal blocking) for periodic,
// boundary-free 3D lattices.  f_out = K(K(f_in)) with K the fused
// stream + collide step of step_kernel; the intermediate populations never
// leave the chip, so a pair of steps reads and writes every population once
// instead of twice.
//
// A block owns a tile of 64 cells along z (i2, contiguous) by TB_Y rows along
// y (i1) and sweeps a chunk of planes along x (i0), one wavefront per row of
// the tile plus one halo row on each side.  Per plane x it
//  A: computes the first step S1 of plane x + 1 on all rows;
//  B: computes the second step S2 of plane x on the interior rows from the
//     S1 values of planes x - 1, x, x + 1 and writes it to HBM, while the two
//     halo-row wavefronts compute S1 of plane x + 2 on the z-edge cells
//     (z0 - 1 and z0 + 64) of every row.
// Where S2 pulls population q from S1 at (x - e_x, y - e_y, z - e_z):
//  e_y = e_z = 0   the thread's own registers (kept for up to two planes);
//  e_y = 0, e_z    the neighbouring lane's register, moved by a wavefront
//                  lane shift (DPP), the row ends from the z-edge cells;
//  e_y != 0        LDS, where each such population of S1 lives in a ring of
//                  2 + e_x planes (it is read by the second step 1 + e_x
//                  planes after it was computed; z-edge values, computed one
//                  plane ahead, in rings of 3 + e_x).
// S1 values are rounded to the storage format as a stored step rounds them,
// so the result equals two step_kernel launches up to the rounding of the
// collision arithmetic (bit for bit without collision).
#include "lbm_tb2.h"

namespace lbm {

template <class S>
__host__ __device__ constexpr int tb_kind(int q) {
  return go<S>(q, 1) != 0 ? 2 : (go<S>(q, 2) != 0 ? 1 : 0);
}
// LDS ring depth of population q for row values and for z-edge values
template <class S>
__host__ __device__ constexpr int tb_ring(int q) {
  return 2 + go<S>(q, 0);
}
template <class S>
__host__ __device__ constexpr int tb_ring_edge(int q) {
  return 3 + go<S>(q, 0);
}
// first ring slot of population q among the populations selected by `which`
// (0: e_y != 0, row rings; 1: e_z > 0, edge rings; 2: e_z < 0, edge rings)
// and the total slot count (q == Q)
template <class S>
__host__ __device__ constexpr int tb_base(int q, int which) {
  int b = 0;
  for (int p = 0; p < q && p < S::Q; ++p) {
    if (which == 0 && go<S>(p, 1) != 0) b += tb_ring<S>(p);
    if (which == 1 && go<S>(p, 2) > 0) b += tb_ring_edge<S>(p);
    if (which == 2 && go<S>(p, 2) < 0) b += tb_ring_edge<S>(p);
  }
  return b;
}

// first step (pull + collide, rounded to storage) of the cell at plane offset
// table ax and row/column tables ay, az
template <class S, class ST, int COLL>
__device__ __forceinline__ void tb_first_step(const KParams& p, const CollParams<typename ST::C>& cp,
                                              const int (&ax)[3], const int (&ay)[3],
                                              const int (&az)[3], typename ST::C (&f)[S::Q],
                                              typename ST::T (&t)[S::Q]) {
  using C = typename ST::C;
  using T = typename ST::T;
  const T* fin = (const T*)p.fin;
#pragma unroll
  for (int q = 0; q < S::Q; ++q) {
    const C sh = shift_of<ST>(S::w[q], p);
    f[q] = ST::ld(fin + q * p.sq, ax[1 - go<S>(q, 0)] + ay[1 - go<S>(q, 1)] + az[1 - go<S>(q, 2)], sh);
  }
  collide<S, C, COLL>(f, cp);
#pragma unroll
  for (int q = 0; q < S::Q; ++q) {
    const C sh = shift_of<ST>(S::w[q], p);
    t[q] = ST::enc(f[q], sh);
    f[q] = ST::dec(t[q], sh);
  }
}

__device__ __forceinline__ int tb_wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

template <class S, class ST, int COLL>
__global__ void __launch_bounds__(TB_THREADS) tb2_kernel(const TbParams P) {
  using C = typename ST::C;
  using T = typename ST::T;
  constexpr int NS2 = tb_base<S>(S::Q, 0);
  constexpr int NSL = tb_base<S>(S::Q, 1);
  constexpr int NSR = tb_base<S>(S::Q, 2);
  __shared__ T lds2[NS2][TB_ROWS][64];
  __shared__ T ldsL[NSL][TB_ROWS];
  __shared__ T ldsR[NSR][TB_ROWS];

  const KParams& p = P.k;
  const int row = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int b = blockIdx.x;
  const int zt = b % P.nzt, yt = (b / P.nzt) % P.nyt, xt = b / (P.nzt * P.nyt);
  const int z0 = zt * 64, y0 = yt * TB_Y, x0 = xt * P.cx;
  const int cx = min(P.cx, p.n0 - x0);

  int ay[3], az[3];
  axis_offsets(tb_wrap(y0 - 1 + row, p.n1), p.n1, p.s1, ay);
  axis_offsets(z0 + lane, p.n2, 1, az);
  const bool interior = row >= 1 && row <= TB_Y;
  // the halo-row wavefronts also compute the z-edge cells: row 0 the cells
  // z0 - 1 of every row (lanes 0..TB_ROWS-1), the last row the cells z0 + 64
  const bool edge_wave = !interior;
  const int side = row == 0 ? 0 : 1;
  const bool edge_lane = edge_wave && lane < TB_ROWS;
  int ey_[3], ez_[3];
  axis_offsets(tb_wrap(y0 - 1 + (edge_lane ? lane : 0), p.n1), p.n1, p.s1, ey_);
  axis_offsets(tb_wrap(side ? z0 + 64 : z0 - 1, p.n2), p.n2, 1, ez_);

  const CollParams<C> cp = coll_params<C>(p);
  T* fout = (T*)p.fout;
  C cur[S::Q], h1[S::Q], h2[S::Q];
#pragma unroll
  for (int q = 0; q < S::Q; ++q) h1[q] = h2[q] = C(0);

  // z-edge cells of counter c (plane x0 - 1 + c) into their rings
  auto edge_step = [&](int c) {
    if (!edge_lane) return;
    int ax[3];
    axis_offsets(tb_wrap(x0 - 1 + c, p.n0), p.n0, p.s0, ax);
    C f[S::Q];
    T t[S::Q];
    tb_first_step<S, ST, COLL>(p, cp, ax, ey_, ez_, f, t);
#pragma unroll
    for (int q = 0; q < S::Q; ++q) {
      if (side == 0 && go<S>(q, 2) > 0) ldsL[tb_base<S>(q, 1) + c % tb_ring_edge<S>(q)][lane] = t[q];
      if (side == 1 && go<S>(q, 2) < 0) ldsR[tb_base<S>(q, 2) + c % tb_ring_edge<S>(q)][lane] = t[q];
    }
  };

  edge_step(0);
  for (int it = 0; it < cx + 2; ++it) {
    // ---- A: S1 of plane x0 - 1 + it on every row (ring counter it) ----
    {
      int ax[3];
      axis_offsets(tb_wrap(x0 - 1 + it, p.n0), p.n0, p.s0, ax);
      T t[S::Q];
      tb_first_step<S, ST, COLL>(p, cp, ax, ay, az, cur, t);
#pragma unroll
      for (int q = 0; q < S::Q; ++q)
        if (tb_kind<S>(q) == 2) lds2[tb_base<S>(q, 0) + it % tb_ring<S>(q)][row][lane] = t[q];
    }
    __syncthreads();

    // ---- B: S2 of plane x0 + it - 2 from S1 counters it-2 .. it on the
    // interior rows; the edge cells of counter it + 1 on the halo rows ----
    if (interior) {
      if (it >= 2) {
        const int xb = x0 + it - 2;
        C g[S::Q];
#pragma unroll
        for (int q = 0; q < S::Q; ++q) {
          const int ex = go<S>(q, 0), ey = go<S>(q, 1), ez = go<S>(q, 2);
          const int c = it - 1 - ex;
          const int slot = c % tb_ring<S>(q), eslot = c % tb_ring_edge<S>(q);
          const C sh = shift_of<ST>(S::w[q], p);
          const C reg = ex < 0 ? cur[q] : (ex == 0 ? h1[q] : h2[q]);
          if (tb_kind<S>(q) == 0) {
            g[q] = reg;
          } else if (tb_kind<S>(q) == 1) {
            if (ez > 0) {
              g[q] = wave_shift<+1>(reg);
              if (lane == 0) g[q] = ST::dec(ldsL[tb_base<S>(q, 1) + eslot][row], sh);
            } else {
              g[q] = wave_shift<-1>(reg);
              if (lane == 63) g[q] = ST::dec(ldsR[tb_base<S>(q, 2) + eslot][row], sh);
            }
          } else {
            const int rr = row - ey, ll = lane - ez;
            T t;
            if (ll < 0) t = ldsL[tb_base<S>(q, 1) + eslot][rr];
            else if (ll > 63) t = ldsR[tb_base<S>(q, 2) + eslot][rr];
            else t = lds2[tb_base<S>(q, 0) + slot][rr][ll];
            g[q] = ST::dec(t, sh);
          }
        }
        collide<S, C, COLL>(g, cp);
        const int cell = xb * p.s0 + ay[1] + az[1];
#pragma unroll
        for (int q = 0; q < S::Q; ++q) ST::st(fout + q * p.sq, cell, g[q], shift_of<ST>(S::w[q], p));
      }
    } else if (it + 1 < cx + 2) {
      edge_step(it + 1);
    }
#pragma unroll
    for (int q = 0; q < S::Q; ++q) {
      h2[q] = h1[q];
      h1[q] = cur[q];
    }
    __syncthreads();
  }
}

template <class S, class ST>
TbFn pick_tb2_coll(int coll) {
  switch (coll) {
    case LBM_COLL_NONE: return tb2_kernel<S, ST, LBM_COLL_NONE>;
    case LBM_COLL_BGK: return tb2_kernel<S, ST, LBM_COLL_BGK>;
    case LBM_COLL_TRT: return tb2_kernel<S, ST, LBM_COLL_TRT>;
  }
  return nullptr;
}

template <class S>
TbFn pick_tb2(int storage, int coll) {
  switch (storage) {
    case LBM_F32: return pick_tb2_coll<S, StF32>(coll);
    case LBM_F16_SHIFTED: return pick_tb2_coll<S, StF16>(coll);
    case LBM_BF16_SHIFTED: return pick_tb2_coll<S, StBF16>(coll);
  }
  return nullptr;
}

TbFn pick_tb2_kernel(int stencil, int storage, int coll) {
  if (stencil == LBM_D3Q19) return pick_tb2<D3Q19>(storage, coll);
  return nullptr;
}

}  // namespace lbm

// Two fused pull steps per pass over memory (temporal blocking) for periodic,
// boundary-free 3D lattices.  f_out = K(K(f_in)) with K the fused
// stream + collide step of step_kernel; the intermediate populations never
// leave the chip, so a pair of steps reads and writes every population once
// instead of twice.
//
// A block owns a tile of 64 cells along z (i2, contiguous) by TB_Y rows along
// y (i1) and sweeps a chunk of planes along x (i0).  Per plane x it
//  A: computes the first step S1 of plane x + 1 on the tile plus one halo row
//     on each side (one wavefront per row) and on the two z-edge cells of
//     each row (one extra wavefront);
//  B: computes the second step S2 of plane x on the interior rows from the
//     S1 values of planes x - 1, x, x + 1 and writes it to HBM.
// Where S2 pulls population q from S1 at (x - e_x, y - e_y, z - e_z):
//  e_y = e_z = 0   the thread's own registers (kept for up to two planes);
//  e_y = 0, e_z    the neighbouring lane's register, moved by a wavefront
//                  lane shift (DPP), the row ends from the z-edge cells;
//  e_y != 0        LDS, where each such population of S1 lives in a ring of
//                  2 + e_x planes (it is read by the second step 1 + e_x
//                  planes after it was computed), 3 + e_x with one barrier
//                  per plane.
// S1 values are rounded to the storage format as a stored step rounds them,
// so the result equals two step_kernel launches up to the rounding of the
// collision arithmetic (bit for bit without collision).  A variant that
// computed the z-edge cells on the idle halo-row wavefronts during B, one
// plane ahead, measured slower (46.7k vs 54.4k MLUPS, D3Q19 FP32 256^3): its
// edge loads put a second memory latency on every plane's critical path.
#include "lbm_tb2.h"

namespace lbm {

template <class S>
__host__ __device__ constexpr int tb_kind(int q) {
  return go<S>(q, 1) != 0 ? 2 : (go<S>(q, 2) != 0 ? 1 : 0);
}
// ring depth of population q: S1 of counter c is read by phase B of
// iterations up to c + 1 + e_x; with a second barrier after phase B the
// slot can be rewritten at iteration c + 2 + e_x (depth 2 + e_x), with only
// the barrier after phase A at c + 3 + e_x (depth 3 + e_x: a wavefront that
// runs ahead writes phase A of the next plane while others still read)
template <class S, int NBAR>
__host__ __device__ constexpr int tb_ring(int q) {
  return (NBAR == 2 ? 2 : 3) + go<S>(q, 0);
}
// first ring slot of population q among the populations selected by `which`
// (0: e_y != 0; 1: e_z > 0; 2: e_z < 0) and the total slot count (q == Q)
template <class S, int NBAR>
__host__ __device__ constexpr int tb_base(int q, int which) {
  int b = 0;
  for (int p = 0; p < q && p < S::Q; ++p) {
    const bool sel = which == 0 ? go<S>(p, 1) != 0 : (which == 1 ? go<S>(p, 2) > 0 : go<S>(p, 2) < 0);
    if (sel) b += tb_ring<S, NBAR>(p);
  }
  return b;
}

// WIDE: cell offsets as a 64-bit plane offset plus a 32-bit offset inside
// the plane, for arrays whose offsets (in bytes, for formats of up to 4
// bytes) do not fit 32 bits; otherwise one 32-bit offset per access, which
// keeps the per-population bases loop-invariant in scalar registers.
template <class S, class ST, int COLL, int NBAR, bool WIDE>
__global__ void __launch_bounds__(TB_THREADS) tb2_kernel(const TbParams P) {
  using C = typename ST::C;
  using T = typename ST::T;
  constexpr int NS2 = tb_base<S, NBAR>(S::Q, 0);
  constexpr int NSL = tb_base<S, NBAR>(S::Q, 1);
  constexpr int NSR = tb_base<S, NBAR>(S::Q, 2);
  __shared__ T lds2[NS2][TB_ROWS][64];
  __shared__ T ldsL[NSL][TB_ROWS];
  __shared__ T ldsR[NSR][TB_ROWS];

  const KParams& p = P.k;
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // tile order: with per_xcd, block b runs on XCD b % 8 and takes tile
  // (b % 8) * per_xcd + b / 8, so each XCD sweeps a contiguous range of
  // tiles (z fastest, then y) and the halo rows and z-edge cache lines a
  // tile shares with its neighbours stay in that XCD's L2
  int b = blockIdx.x;
  if (P.per_xcd) {
    b = (b & 7) * P.per_xcd + (b >> 3);
    if (b >= P.ntiles) return;
  }
  const int zt = b % P.nzt, yt = (b / P.nzt) % P.nyt, xt = b / (P.nzt * P.nyt);
  const int z0 = zt * 64, y0 = yt * TB_Y, x0 = xt * P.cx;
  const int cx = min(P.cx, p.n0 - x0);

  // the cell this thread computes S1 for: a row cell, or a z-edge cell
  const bool edge = wid == TB_ROWS;
  const int row = edge ? lane % TB_ROWS : wid;
  const int side = edge ? lane / TB_ROWS : 0;  // 0: z0 - 1, 1: z0 + 64
  const bool active = !edge || lane < 2 * TB_ROWS;
  int y = y0 - 1 + row;
  y = y < 0 ? y + p.n1 : (y >= p.n1 ? y - p.n1 : y);
  int z = edge ? (side ? z0 + 64 : z0 - 1) : z0 + lane;
  z = z < 0 ? z + p.n2 : (z >= p.n2 ? z - p.n2 : z);
  int ay[3], az[3];
  axis_offsets(y, p.n1, p.s1, ay);
  axis_offsets(z, p.n2, 1, az);
  const bool interior = !edge && row >= 1 && row <= TB_Y;
  // D3Q19 (two blocks per CU) uses wave priorities; D3Q27, whose registers
  // leave one block per CU, runs faster without them (KBC FP32 256^3:
  // 29.4k with, 30.8k MLUPS without).
  constexpr bool PRIORITIES = S::Q == 19;
  // The z-edge wave runs at raised priority throughout: phase B of the
  // interior rows waits at the barrier for its 20 edge cells, and at the
  // default priority it issues behind the 10 row waves.  Measured on D3Q19
  // BGK 256^3 FP32 (two runs): 60.6k -> 63.8k MLUPS.
  auto base_prio = [&] {
    if (!PRIORITIES) return;
    if (edge) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
  };
  base_prio();

  const CollParams<C> cp = coll_params<C>(p);
  const T* fin = (const T*)p.fin;
  T* fout = (T*)p.fout;
  // stochastic rounding (BF16 storage): the keys of the launch's two steps
  constexpr bool sr = ST::stochastic;
  const uint32_t key1 = sr ? sr_step_key(p.sr_seed, p.sr_step) : 0u;
  const uint32_t key2 = sr ? sr_step_key(p.sr_seed, p.sr_step + 1) : 0u;
  C cur[S::Q], h1[S::Q], h2[S::Q];
#pragma unroll
  for (int q = 0; q < S::Q; ++q) h1[q] = h2[q] = C(0);

  using X = std::conditional_t<WIDE, long long, int>;
  // D3Q27 (one block per CU at up to 168 VGPRs) issues plane it + 1's loads
  // right after plane it's decode, into the same registers, so they stay in
  // flight over S1, the barrier and phase B.  Measured on 256^3 KBC FP32:
  // 37.1k -> 39.5k MLUPS.  D3Q19 keeps the loads in phase A: 19 more VGPRs
  // leave one block per CU instead of two (FP32 63.9k -> 62.4k, FP16 106k ->
  // 73k), or spill at two (FP32 50.9k).
  constexpr bool PF = S::Q == 27 && !WIDE;
  T raw[S::Q];
  auto issue = [&](int itl) {
    // the prefetch after the chunk's last plane reloads that plane (unused):
    // one plane further could lie past the array where n0 == 1
    if constexpr (PF) itl = min(itl, cx + 1);
    int xa = x0 - 1 + itl;
    xa = xa < 0 ? xa + p.n0 : (xa >= p.n0 ? xa - p.n0 : xa);
    X ax[3];
    axis_offsets(xa, p.n0, (X)p.s0, ax);
#pragma unroll
    for (int q = 0; q < S::Q; ++q) {
      const int in = ay[1 - go<S>(q, 1)] + az[1 - go<S>(q, 2)];
      if constexpr (WIDE)
        raw[q] = elem(scalar_base(fin + q * p.sq + ax[1 - go<S>(q, 0)]), in);
      else
        raw[q] = elem(fin + q * p.sq, ax[1 - go<S>(q, 0)] + in);
    }
  };
  if (PF && active) issue(0);

  for (int it = 0; it < cx + 2; ++it) {
    // ---- A: S1 of plane xa = x0 - 1 + it (ring counter it) ----
    if (active) {
      int xa = x0 - 1 + it;
      xa = xa < 0 ? xa + p.n0 : (xa >= p.n0 ? xa - p.n0 : xa);
      X ax[3];
      axis_offsets(xa, p.n0, (X)p.s0, ax);
      // Without PF the plane's loads go out together, ahead of their
      // decodes, and (for 16-bit formats) at raised wave priority, so that a
      // wave reaching this point issues them before other waves' arithmetic.
      // Measured on D3Q19 BGK 256^3: FP16 89.7k -> 94.4k MLUPS from the
      // batching and -> 98.3k with the priority; FP32 59.6k -> 59.8k from
      // the batching but 58.2k with the priority, so FP32 runs without it.
      constexpr bool PRIO = PRIORITIES && sizeof(T) == 2;
      if constexpr (!PF) {
        if (PRIO) __builtin_amdgcn_s_setprio(3);
        __builtin_amdgcn_sched_barrier(0);
        issue(it);
        __builtin_amdgcn_sched_barrier(0);
        if (PRIO) base_prio();
      }
#pragma unroll
      for (int q = 0; q < S::Q; ++q) cur[q] = ST::dec(raw[q], shift_of<ST>(S::w[q], p));
      if constexpr (PF) {
        __builtin_amdgcn_sched_barrier(0);
        issue(it + 1);
        __builtin_amdgcn_sched_barrier(0);
      }
      collide<S, C, COLL>(cur, cp);
      // S1 rounded to the storage format as a single step would store it,
      // stochastic rounding keyed to the first step and this cell
      const uint32_t hs = sr ? sr_cell(key1, (uint32_t)(ax[1] + ay[1] + az[1])) : 0u;
#pragma unroll
      for (int q = 0; q < S::Q; ++q) {
        const C sh = shift_of<ST>(S::w[q], p);
        const T t = sr ? ST::enc_sr(cur[q], sh, sr_bits(hs, q)) : ST::enc(cur[q], sh);
        cur[q] = ST::dec(t, sh);
        const int slot_of = it % tb_ring<S, NBAR>(q);
        if (!edge) {
          if (tb_kind<S>(q) == 2) lds2[tb_base<S, NBAR>(q, 0) + slot_of][row][lane] = t;
        } else if (side == 0) {
          if (go<S>(q, 2) > 0) ldsL[tb_base<S, NBAR>(q, 1) + slot_of][row] = t;
        } else {
          if (go<S>(q, 2) < 0) ldsR[tb_base<S, NBAR>(q, 2) + slot_of][row] = t;
        }
      }
    }
    __syncthreads();

    // ---- B: S2 of plane xb = x0 + it - 2 from S1 counters it-2 .. it ----
    if (it >= 2 && interior) {
      // phase B at raised wave priority: the interior waves' step 2 and its
      // stores go ahead of other waves' phase A arithmetic.  Measured on
      // D3Q19 BGK 256^3 (two runs each): FP32 59.9k -> 60.7k MLUPS, BF16
      // 97.9k -> 98.9k, FP16 unchanged (99.5k -> 99.9k).
      if (PRIORITIES) __builtin_amdgcn_s_setprio(2);
      const int xb = x0 + it - 2;
      C g[S::Q];
#pragma unroll
      for (int q = 0; q < S::Q; ++q) {
        const int ex = go<S>(q, 0), ey = go<S>(q, 1), ez = go<S>(q, 2);
        const int slot = (it - 1 - ex) % tb_ring<S, NBAR>(q);
        const C sh = shift_of<ST>(S::w[q], p);
        const C reg = ex < 0 ? cur[q] : (ex == 0 ? h1[q] : h2[q]);
        if (tb_kind<S>(q) == 0) {
          g[q] = reg;
        } else if (tb_kind<S>(q) == 1) {
          if (ez > 0) {
            g[q] = wave_shift<+1>(reg);
            if (lane == 0) g[q] = ST::dec(ldsL[tb_base<S, NBAR>(q, 1) + slot][row], sh);
          } else {
            g[q] = wave_shift<-1>(reg);
            if (lane == 63) g[q] = ST::dec(ldsR[tb_base<S, NBAR>(q, 2) + slot][row], sh);
          }
        } else {
          const int rr = row - ey, ll = lane - ez;
          T t;
          if (ll < 0) t = ldsL[tb_base<S, NBAR>(q, 1) + slot][rr];
          else if (ll > 63) t = ldsR[tb_base<S, NBAR>(q, 2) + slot][rr];
          else t = lds2[tb_base<S, NBAR>(q, 0) + slot][rr][ll];
          g[q] = ST::dec(t, sh);
        }
      }
      collide<S, C, COLL>(g, cp);
      const long long plane = xb * p.s0;
      // D3Q27: 27 loop-invariant store bases next to the 27 load bases
      // would not fit the scalar registers, and the compiler would keep
      // them in vector-register lanes, read back (v_readlane) in the load
      // phase.  One base stepped by the population stride instead; the asm
      // keeps each step in place.
      constexpr bool STEP = S::Q == 27 && !WIDE;
      uint64_t sb = (uint64_t)fout;
      const uint64_t sqb = (uint64_t)p.sq * sizeof(T);
      const uint32_t elem2 = (uint32_t)(plane + ay[1] + az[1]);
#pragma unroll
      for (int q = 0; q < S::Q; ++q) {
        const C sh = shift_of<ST>(S::w[q], p);
        if constexpr (WIDE) {
          store_elem<ST>(p, scalar_base(fout + q * p.sq + plane), ay[1] + az[1], g[q], sh, key2,
                         elem2, q);
        } else if constexpr (STEP) {
          store_elem<ST>(p, (T*)sb, (int)plane + ay[1] + az[1], g[q], sh, key2, elem2, q);
          sb += sqb;
          asm volatile("" : "+s"(sb));
        } else {
          store_elem<ST>(p, fout + q * p.sq, (int)plane + ay[1] + az[1], g[q], sh, key2, elem2, q);
        }
      }
      base_prio();
    }
#pragma unroll
    for (int q = 0; q < S::Q; ++q) {
      h2[q] = h1[q];
      h1[q] = cur[q];
    }
    if (NBAR == 2) __syncthreads();
  }
}

// wide grids run the one-barrier variant
template <class S, class ST, int COLL>
TbFn pick_tb2_bar(int barriers, bool wide) {
  if (wide) return tb2_kernel<S, ST, COLL, 1, true>;
  return barriers == 1 ? tb2_kernel<S, ST, COLL, 1, false> : tb2_kernel<S, ST, COLL, 2, false>;
}

template <class S, class ST>
TbFn pick_tb2_coll(int coll, int barriers, bool wide) {
  switch (coll) {
    case LBM_COLL_NONE: return pick_tb2_bar<S, ST, LBM_COLL_NONE>(barriers, wide);
    case LBM_COLL_BGK: return pick_tb2_bar<S, ST, LBM_COLL_BGK>(barriers, wide);
    case LBM_COLL_TRT: return pick_tb2_bar<S, ST, LBM_COLL_TRT>(barriers, wide);
    case LBM_COLL_KBC:
      if constexpr (S::ID == LBM_D3Q27) return pick_tb2_bar<S, ST, LBM_COLL_KBC>(barriers, wide);
      else return nullptr;
  }
  return nullptr;
}

template <class S>
TbFn pick_tb2(int storage, int coll, int barriers, bool wide) {
  switch (storage) {
    case LBM_F32: return pick_tb2_coll<S, StF32>(coll, barriers, wide);
    case LBM_F16_SHIFTED: return pick_tb2_coll<S, StF16>(coll, barriers, wide);
    case LBM_BF16_SHIFTED: return pick_tb2_coll<S, StBF16>(coll, barriers, wide);
    case kCodecBF16SR: return pick_tb2_coll<S, StBF16SR>(coll, barriers, wide);
  }
  return nullptr;
}

TbFn pick_tb2_kernel(int stencil, int storage, int coll, int barriers, bool wide) {
  if (stencil == LBM_D3Q19) return pick_tb2<D3Q19>(storage, coll, barriers, wide);
  if (stencil == LBM_D3Q27) return pick_tb2<D3Q27>(storage, coll, barriers, wide);
  return nullptr;
}

}  // namespace lbm

// Shared declarations of the x-slab kernels: a lattice split along i0 (the
// slowest storage axis) into slabs, each periodic along i1 and i2, that
// exchange the populations crossing their two i0 faces once per step.
#pragma once

#include "lbm_step_impl.h"

namespace lbm {

// Halo buffers of one slab, each [k][n1][n2] elements of the storage type
// (k = populations crossing one face, in stencil order):
//  recv_lo  plane -1:  populations with e_x = +1 of the lower neighbour's
//           last plane (its send_hi);
//  recv_hi  plane n0:  populations with e_x = -1 of the upper neighbour's
//           first plane (its send_lo);
//  send_lo  this slab's plane 0, populations with e_x = -1, after the step;
//  send_hi  this slab's plane n0 - 1, populations with e_x = +1, after the step.
struct SlabHalo {
  const void* recv_lo;
  const void* recv_hi;
  void* send_lo;
  void* send_hi;
};
using FaceFn = void (*)(const KParams, const SlabHalo);

// populations that cross one i0 face: 5 (D3Q19), 9 (D3Q27), else -1
int halo_planes(int stencil);

// the step of the two face planes (reads recv_*, writes f_out and send_*)
FaceFn pick_face_kernel(int stencil, int storage, int coll);

// copies the crossing populations of planes 0 and n0 - 1 of p.fin into
// send_lo / send_hi (raw storage bits)
FaceFn pick_pack_kernel(int stencil, int storage);

}  // namespace lbm

"""One device's share of a lattice split into x-slabs.

lettuce steps a whole lattice on one device (lettuce/_simulation.py:241-256
rolls every axis periodically).  A lattice too large for one device splits
along x, the slowest axis of [Q, X, Y, Z], into slabs that stay periodic
along y and z.  SlabStep runs the pull steps of one slab through
lbm_slab_step: the planes that need no halo first, then the two face planes,
which pull the entering populations from packed receive buffers and leave
the next step's send buffers behind.  Moving the buffers between devices
(torch.distributed send / recv over RCCL, a peer copy) is the caller's.
"""
import ctypes

import torch

from .. import _native as N
from .._ops import layout
from .._simulation import NativeStep, StreamingStrategy

__all__ = ["SlabStep"]

_PARTS = {"all": N.SLAB_ALL, "interior": N.SLAB_INTERIOR, "faces": N.SLAB_FACES}


class SlabStep:
    """Pull steps of one x-slab.

    `simulation` supplies stencil, storage, collision and its parameters
    (pull streaming, no boundaries); `state` is the slab's raw population
    tensor [Q, nx, ny, nz] in the simulation's storage dtype on its device
    (the layout of ``flow._state``).  The halo buffers are [k, ny, nz]
    tensors of the same dtype, k = 5 (D3Q19) or 9 (D3Q27):

      ``send_hi``  plane nx-1, e_x = +1   -> the upper neighbour's ``recv_lo``
      ``send_lo``  plane 0,    e_x = -1   -> the lower neighbour's ``recv_hi``

    Before the first step call :meth:`pack` (it fills the send buffers from
    ``state``); afterwards every face step leaves the next ones behind.
    """

    def __init__(self, simulation, state: torch.Tensor):
        native = simulation._collide_and_stream
        if not isinstance(native, NativeStep):
            raise NotImplementedError("x-slabs need the native step")
        if simulation.streaming_strategy != StreamingStrategy.PRE_STREAMING:
            raise ValueError("x-slabs use pull streaming (StreamingStrategy.PRE_STREAMING)")
        if simulation.cell_classes is not None:
            raise NotImplementedError("x-slabs run without boundaries")
        stencil = simulation.flow.stencil
        if stencil.d != 3:
            raise ValueError("x-slabs split 3D lattices")
        ref = simulation.flow._state
        if state.dtype != ref.dtype or state.device != ref.device:
            raise ValueError(f"state must be {ref.dtype} on {ref.device}, "
                             f"got {state.dtype} on {state.device}")
        self.stencil = stencil
        self.state = state
        self.state_next = torch.empty_like(state)
        n, s = layout(state, stencil.q, stencil.d)
        if tuple(self.state_next.stride()) != tuple(state.stride()):
            raise ValueError("state must be dense (contiguous or the rows layout)")
        self.args = N.LbmStepArgs.from_buffer_copy(native.args)
        self.args.n[:] = n
        self.args.stride[:] = s
        k = int(N.lib().lbm_halo_planes(stencil.native_id))
        shape = (k, n[1], n[2])
        self.send_lo = torch.empty(shape, dtype=state.dtype, device=state.device)
        self.send_hi = torch.empty_like(self.send_lo)
        self.recv_lo = torch.empty_like(self.send_lo)
        self.recv_hi = torch.empty_like(self.send_lo)
        self.device = state.device

    def pack(self):
        """send_lo / send_hi <- the leaving populations of the current state."""
        a = self.args
        N.check(N.lib().lbm_halo_pack(a.stencil, a.storage, a.n, a.stride,
                                      self.state.data_ptr(), self.send_lo.data_ptr(),
                                      self.send_hi.data_ptr(),
                                      N.stream_handle(self.device)), "lbm_halo_pack")

    def step(self, part: str = "all"):
        """One pull step, or its 'interior' / 'faces' part; the state
        advances (state <-> state_next) after the face part."""
        code = _PARTS[part]
        self.args.f_in = self.state.data_ptr()
        self.args.f_out = self.state_next.data_ptr()
        N.check(N.lib().lbm_slab_step(ctypes.byref(self.args), code,
                                      self.recv_lo.data_ptr(), self.recv_hi.data_ptr(),
                                      self.send_lo.data_ptr(), self.send_hi.data_ptr(),
                                      N.stream_handle(self.device)), "lbm_slab_step")
        if code != N.SLAB_INTERIOR:
            self.state, self.state_next = self.state_next, self.state

    def wrap(self):
        """recv <- own send: the slab as a whole periodic lattice."""
        self.recv_lo.copy_(self.send_hi)
        self.recv_hi.copy_(self.send_lo)

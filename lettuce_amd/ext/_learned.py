"""Learned collision: KBC decomposition with MLP-predicted higher-order rates.

There is no lettuce counterpart; it plugs into lettuce's Collision interface
(lettuce/_simulation.py:17-28) like MRTCollision
(lettuce/ext/_collision/mrt_collision.py:8-35) and reuses KBC's split of
f - feq into a shear part and a higher-order part
(lettuce/ext/_collision/kbc_collision.py:76-160).  The kernel and its exact
formula are described in csrc/lbm_learned.hip and include/lettuce_amd.h.
"""
import ctypes
import os
from typing import Optional

import numpy as np
import torch

from .. import _native as N
from .._simulation import Collision
from .._stencil import D2Q9, D3Q27
from ._operators import NativeCollisionDesc, apply_collision

__all__ = ["LearnedParameters", "LearnedCollision", "TRAINED_D3Q27_KBC"]

# D3Q27 weights trained on decaying-turbulence states of KBC runs
# (python -m lettuce_amd.studies train3d --out ...): 64^3, u_rms 0.05,
# nu 1e-4, tau 0.5003, 16-step rollouts against KBC's velocity
TRAINED_D3Q27_KBC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                 "data", "learned_d3q27_kbc.npz")


class LearnedParameters:
    """Weights of the 6 -> 32 -> 32 -> 3 SiLU MLP (FP32 tensors; the kernel
    rounds the weights to bf16) and the rate limits."""

    def __init__(self, w1, b1, w2, b2, w3, b3, in_scale=100.0, rate_lo=0.05,
                 rate_hi=1.95, stable_lo=0.0, stable_hi=2.0):
        H = N.LEARNED_HIDDEN
        shapes = {"w1": (H, 6), "b1": (H,), "w2": (H, H), "b2": (H,),
                  "w3": (3, H), "b3": (3,)}
        vals = dict(w1=w1, b1=b1, w2=w2, b2=b2, w3=w3, b3=b3)
        for k, shape in shapes.items():
            t = torch.as_tensor(vals[k], dtype=torch.float32).detach().cpu()
            if tuple(t.shape) != shape:
                raise ValueError(f"{k} must have shape {shape}, got {tuple(t.shape)}")
            setattr(self, k, t.contiguous())
        if not rate_lo <= rate_hi:
            raise ValueError("rate_lo must not exceed rate_hi")
        self.in_scale = float(in_scale)
        self.rate_lo, self.rate_hi = float(rate_lo), float(rate_hi)
        self.stable_lo, self.stable_hi = float(stable_lo), float(stable_hi)

    @staticmethod
    def seeded(seed: int = 0, out_scale: float = 0.1, **kw) -> "LearnedParameters":
        """Weights drawn like torch.nn.Linear's default initialisation,
        U(-1/sqrt(fan_in), 1/sqrt(fan_in)), from torch.Generator(seed), in
        the order w1, b1, w2, b2, w3, b3; the last layer scaled by
        out_scale."""
        g = torch.Generator().manual_seed(seed)
        H = N.LEARNED_HIDDEN

        def lin(o, i):
            bound = 1.0 / np.sqrt(i)
            w = (torch.rand(o, i, generator=g) * 2 - 1) * bound
            b = (torch.rand(o, generator=g) * 2 - 1) * bound
            return w, b
        w1, b1 = lin(H, 6)
        w2, b2 = lin(H, H)
        w3, b3 = lin(3, H)
        return LearnedParameters(w1, b1, w2, b2, w3 * out_scale, b3 * out_scale, **kw)

    def tensors(self):
        return [self.w1, self.b1, self.w2, self.b2, self.w3, self.b3]

    _LIMITS = ("in_scale", "rate_lo", "rate_hi", "stable_lo", "stable_hi")

    def save(self, path):
        """weights (FP32) and limits to an .npz file"""
        arrays = {k: t.numpy() for k, t in zip(("w1", "b1", "w2", "b2", "w3", "b3"),
                                               self.tensors())}
        arrays.update({k: np.array(getattr(self, k), dtype=np.float64) for k in self._LIMITS})
        with open(path, "wb") as fh:
            np.savez(fh, **arrays)

    @staticmethod
    def load(path, **overrides) -> "LearnedParameters":
        """parameters written by save(); keyword arguments replace limits"""
        with np.load(path, allow_pickle=False) as z:
            w = [torch.from_numpy(z[k].astype(np.float32)) for k in
                 ("w1", "b1", "w2", "b2", "w3", "b3")]
            kw = {k: float(z[k]) for k in LearnedParameters._LIMITS}
        kw.update(overrides)
        return LearnedParameters(*w, **kw)

    def packed(self, device) -> torch.Tensor:
        """The MFMA operand buffer of lbm_learned_pack, on `device`."""
        host = np.zeros(N.LEARNED_PACKED_BYTES, dtype=np.uint8)
        arrays = [np.ascontiguousarray(t.numpy(), dtype=np.float32)
                  for t in self.tensors()]
        ptrs = [a.ctypes.data_as(ctypes.c_void_p) for a in arrays]
        N.check(N.lib().lbm_learned_pack(*ptrs, host.ctypes.data_as(ctypes.c_void_p)),
                "lbm_learned_pack")
        return torch.from_numpy(host).to(device)


class LearnedCollision(Collision):
    """D2Q9 / D3Q27 learned collision; cells whose predicted rates leave the
    stable range relax with KBC.  tau is the flow's relaxation_parameter_lu,
    like KBC.  D2Q9 rates act on the moments xxy, xyy, xxyy; D3Q27 rates on
    the raw moments of order 3, of order 4 and of orders 5-6."""

    def __init__(self, params: Optional[LearnedParameters] = None):
        self.params = params or LearnedParameters.seeded(0)

    def __call__(self, flow):
        return apply_collision(self, flow)

    def native_available(self) -> bool:
        return True

    def native_generator(self, index: int):
        return NativeCollisionDesc(N.COLL_LEARNED, index=index, owner=self)

    def _bind(self, desc, flow):
        if not isinstance(flow.stencil, (D2Q9, D3Q27)):
            raise NotImplementedError("the learned collision is defined for D2Q9 and D3Q27")
        desc.tau = flow.units.relaxation_parameter_lu
        # one packed weight buffer per binding, owned by the descriptor: a
        # second binding of this collision (another Simulation, or
        # collision(flow)) gets its own and leaves this one in place
        packed = self.params.packed(flow.context.device)
        s = N.LbmLearned()
        s.packed = packed.data_ptr()
        p = self.params
        s.in_scale, s.rate_lo, s.rate_hi = p.in_scale, p.rate_lo, p.rate_hi
        s.stable_lo, s.stable_hi = p.stable_lo, p.stable_hi
        desc.learned = s
        desc.keep = [packed]

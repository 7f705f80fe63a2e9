"""Equilibrium, collision and boundary operators.

Each class keeps lettuce's constructor, mask methods and
``native_available()`` / ``native_generator(index)`` hooks
(lettuce/_flow.py:17-52, lettuce/_simulation.py:17-28).  The generator
returns a small descriptor that the native step turns into kernel arguments
or cell classes, instead of CUDA source text.  Calling an operator directly
(``collision(flow)``, ``boundary(flow)``) also runs on the GPU through the
same kernels.
"""
import ctypes
import numbers
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple, Union

import numpy as np
import torch

from .. import _native as N
from .. import _ops
from .._classes import OpSpec
from .._flow import Boundary, Equilibrium
from .._simulation import Collision, StreamingStrategy
from .._stencil import D2Q9, D3Q27

__all__ = ["QuadraticEquilibrium", "NoCollision", "BGKCollision",
           "TRTCollision", "KBCCollision", "BounceBackBoundary",
           "EquilibriumBoundaryPU", "EquilibriumOutletP",
           "HalfwayBounceBackBoundary", "AntiBounceBackOutlet",
           "NativeCollisionDesc", "NativeBoundaryDesc", "apply_collision"]


# ---------------------------------------------------------------------------
# equilibrium
# ---------------------------------------------------------------------------
def _full(t: torch.Tensor, shape) -> torch.Tensor:
    return t.expand(tuple(shape)).contiguous()


class QuadraticEquilibrium(Equilibrium):
    """feq = w rho [(2 e.u - u.u)/(2 cs^2) + (e.u/cs^2)^2/2 + 1]
    (lettuce/ext/_equilibrium/quadratic_equilibrium.py:11-24), evaluated by
    the lbm_equilibrium kernel."""

    def __call__(self, flow, rho=None, u=None):
        ctx = flow.context
        res = list(flow.resolution)
        rho = flow.rho() if rho is None else ctx.convert_to_tensor(rho)
        u = flow.u() if u is None else ctx.convert_to_tensor(u)
        if rho.dim() == len(res):
            rho = rho[None]
        rho = _full(rho, [1] + res)[0]
        u = _full(u, [flow.stencil.d] + res)
        out = ctx.empty_tensor([flow.stencil.q] + res)
        return _ops.equilibrium(flow.stencil, out, rho, u)

    def native_available(self) -> bool:
        return True

    def native_generator(self):
        return "quadratic"


# ---------------------------------------------------------------------------
# collisions
# ---------------------------------------------------------------------------
@dataclass
class NativeCollisionDesc:
    """Kernel-side description of a collision operator."""
    kind: int
    tau: float = 0.0
    tau_minus: float = 0.0
    index: int = 0
    owner: object = None
    force_tau: float = 0.0                  # BGK_GUO
    accel: Tuple[float, ...] = (0.0, 0.0, 0.0)
    # LEARNED: the lbm_learned struct of this binding and the device tensors
    # it points into; they live as long as the descriptor, which the native
    # step holds, so binding the same collision again cannot free them
    learned: object = None
    keep: List[torch.Tensor] = field(default_factory=list)

    def bind(self, flow):
        if self.kind == N.COLL_KBC:
            # KBC takes tau from the flow's units, not from its constructor
            # (lettuce/ext/_collision/kbc_collision.py:97-99)
            self.tau = flow.units.relaxation_parameter_lu
            if not isinstance(flow.stencil, (D2Q9, D3Q27)):
                raise NotImplementedError(
                    "KBC collision is defined for D2Q9 and D3Q27 only")
        if self.owner is not None and hasattr(self.owner, "_bind"):
            self.owner._bind(self, flow)

    def learned_struct(self):
        return self.learned


def apply_collision(collision, flow) -> torch.Tensor:
    """Post-collision populations of the whole grid, without streaming and
    without boundaries: what lettuce's ``collision(flow)`` returns."""
    from .._simulation import NativeStep

    class _Plain:
        pass

    sim = _Plain()
    sim.flow = flow
    sim.collision = collision
    sim.collision_index = 0
    sim.streaming_strategy = StreamingStrategy.NO_STREAMING
    sim.cell_classes = None
    step = NativeStep(sim)
    out = torch.empty_like(flow._state)
    step.args.f_in = flow._state.data_ptr()
    step.args.f_out = out.data_ptr()
    N.check(N.lib().lbm_step(ctypes.byref(step.args),
                             N.stream_handle(flow.context.device)), "lbm_step")
    if flow.context.shifted:
        return _ops.decode(flow.stencil, out, flow.rho0_shift, flow.context.dtype)
    return out


class NoCollision(Collision):
    """f unchanged (lettuce/ext/_collision/no_collision.py:10-11)."""

    def __call__(self, flow):
        return flow.f

    def native_available(self) -> bool:
        return True

    def native_generator(self, index: int):
        return NativeCollisionDesc(N.COLL_NONE, index=index)


class BGKCollision(Collision):
    """f - (f - feq) / tau (lettuce/ext/_collision/bgk_collision.py:12-35)."""

    def __init__(self, tau, force=None):
        self.tau = tau
        self.force = force

    def __call__(self, flow):
        return apply_collision(self, flow)

    def name(self) -> str:
        return type(self).__name__

    def native_available(self) -> bool:
        return self.force is None or self.force.native_available()

    def native_generator(self, index: int):
        if self.force is not None:
            if not self.force.native_available():
                return None
            spec = self.force.native_generator()
            return NativeCollisionDesc(N.COLL_BGK_GUO, tau=float(self.tau), index=index,
                                       force_tau=spec["tau"], accel=spec["acceleration"])
        return NativeCollisionDesc(N.COLL_BGK, tau=float(self.tau), index=index)


class TRTCollision(Collision):
    """Two-relaxation-time collision (lettuce/ext/_collision/trt_collision.py:
    6-33): symmetric parts relax with tau, antisymmetric with tau_minus."""

    def __init__(self, tau, tau_minus=1.0):
        self.tau_plus = tau
        self.tau_minus = tau_minus

    def __call__(self, flow):
        return apply_collision(self, flow)

    def native_available(self) -> bool:
        return True

    def native_generator(self, index: int):
        return NativeCollisionDesc(N.COLL_TRT, tau=float(self.tau_plus),
                                   tau_minus=float(self.tau_minus), index=index)


class KBCCollision(Collision):
    """Entropic KBC collision (lettuce/ext/_collision/kbc_collision.py:11-160)
    for D2Q9 and D3Q27; like lettuce it relaxes with the flow's
    relaxation_parameter_lu whatever tau it was given."""

    def __init__(self, tau: float = None):
        self.tau = tau
        self.beta = None

    def __call__(self, flow):
        return apply_collision(self, flow)

    def native_available(self) -> bool:
        return True

    def native_generator(self, index: int):
        return NativeCollisionDesc(N.COLL_KBC, index=index, owner=self)

    def _bind(self, desc, flow):
        self.tau = desc.tau
        self.beta = 1.0 / (2.0 * desc.tau)


# ---------------------------------------------------------------------------
# boundaries
# ---------------------------------------------------------------------------
@dataclass
class NativeBoundaryDesc:
    """Kernel-side description of a boundary: the cell operator it applies
    on its no-collision cells (or OP_KEEP plus a solid mask for halfway
    bounce-back)."""
    op: int
    index: int
    owner: object = None
    solid_mask: Optional[torch.Tensor] = None

    def op_spec(self, flow) -> OpSpec:
        return self.owner._op_spec(flow)


class BounceBackBoundary(Boundary):
    """Fullway bounce-back: f_q <- f_opp(q) on the masked cells in place of
    collision (lettuce/ext/_boundary/bounce_back_boundary.py:10-32)."""

    def __init__(self, mask: torch.Tensor):
        self._mask = mask

    def __call__(self, flow):
        return flow.f[flow.stencil.opposite]

    def make_no_streaming_mask(self, shape: List[int], context):
        return None

    def make_no_collision_mask(self, shape: List[int], context):
        return self._mask

    def native_available(self) -> bool:
        return True

    def native_generator(self, index: int):
        return NativeBoundaryDesc(N.OP_BOUNCE_BACK, index, owner=self)

    def _op_spec(self, flow):
        return OpSpec(op=N.OP_BOUNCE_BACK)


class EquilibriumBoundaryPU(Boundary):
    """f <- feq(rho(p), u) on the masked cells, with p and u broadcast over
    the grid (lettuce/ext/_boundary/equilibrium_boundary_pu.py:15-98)."""

    @staticmethod
    def checked_tensor(t, context, flow) -> torch.Tensor:
        if not torch.is_tensor(t):
            if isinstance(t, numbers.Number):
                t = torch.tensor(t)
            elif isinstance(t, (np.ndarray, list, tuple)):
                t = torch.as_tensor(t)
            else:
                raise TypeError(f"Cannot convert {type(t)} to tensor")
        d = flow.stencil.d
        spatial = list(flow.resolution)
        full_ndim = 1 + len(spatial)
        if t.ndim == 0:
            t = t.reshape([1] * full_ndim)
        elif t.ndim == 1 and t.shape[0] == d:
            t = t.reshape([d] + [1] * len(spatial))
        elif t.ndim == len(spatial) and list(t.shape) == spatial:
            t = t.unsqueeze(0)
        elif t.ndim != full_ndim:
            raise ValueError(f"Tensor has wrong rank: expected {full_ndim}, "
                             f"got {t.ndim}")
        for i in range(full_ndim):
            size = t.shape[i]
            allowed = (1, d) if i == 0 else (1, spatial[i - 1])
            if size not in allowed:
                raise ValueError(f"dim {i} must be one of {allowed}, got {size}")
        return context.convert_to_tensor(t)

    def __init__(self, context, flow, mask, velocity, pressure=0):
        self.velocity = self.checked_tensor(velocity, context, flow)
        self.pressure = self.checked_tensor(pressure, context, flow)
        self._mask = mask

    def _rho_u(self, flow):
        rho = flow.units.convert_pressure_pu_to_density_lu(self.pressure)
        u = flow.units.convert_velocity_to_lu(self.velocity)
        return rho, u

    def __call__(self, flow):
        rho, u = self._rho_u(flow)
        return flow.equilibrium(flow, rho, u)

    def make_no_collision_mask(self, shape: List[int], context):
        return self._mask

    def make_no_streaming_mask(self, shape: List[int], context):
        return None

    def native_available(self) -> bool:
        return True

    def native_generator(self, index: int):
        return NativeBoundaryDesc(N.OP_EQUILIBRIUM, index, owner=self)

    def _op_spec(self, flow):
        d = flow.stencil.d
        rho, u = self._rho_u(flow)
        rho = rho.to(torch.float64)
        u = u.to(torch.float64)
        if rho.numel() == 1 and all(s == 1 for s in u.shape[1:]):
            uu = u.reshape(-1).tolist()
            if len(uu) == 1:
                uu = uu * d
            return OpSpec(op=N.OP_EQUILIBRIUM, rho=float(rho.reshape(-1)[0]),
                          u=tuple(uu) + (0.0,) * (3 - d))
        res = list(flow.resolution)
        field = torch.zeros([4] + res, dtype=torch.float64, device=rho.device)
        field[0] = rho.expand([1] + res)[0]
        field[1:1 + d] = u.expand([d] + res)
        return OpSpec(op=N.OP_EQUILIBRIUM, bc_field=field)


def _outlet_geometry(direction, flow):
    if len(direction) != flow.stencil.d:
        raise ValueError(f"direction needs {flow.stencil.d} entries")
    if not (direction.count(0) == len(direction) - 1
            and ((1 in direction) ^ (-1 in direction))):
        raise ValueError(f"direction must have one +-1 entry, got {direction}")
    velocities = np.concatenate(np.argwhere(
        np.matmul(np.array(flow.stencil.e), np.array(direction)) > 1 - 1e-6),
        axis=0)
    index = []
    for c in direction:
        index.append(slice(None) if c == 0 else (-1 if c == 1 else 0))
    return velocities, index


class EquilibriumOutletP(Boundary):
    """Constant-pressure equilibrium outlet
    (lettuce/ext/_boundary/equilibrium_outlet_p.py:12-91): the outlet column
    is set to feq(rho_outlet, u of the neighbouring column); populations not
    leaving through the outlet are not streamed there."""

    def __init__(self, direction: List[int], flow, rho_outlet: float = 1.0):
        self.direction = [int(round(float(c))) for c in direction]
        self.context = flow.context
        self.rho_outlet = float(rho_outlet)
        self.velocities, self.index = _outlet_geometry(self.direction, flow)

    def __call__(self, flow):
        raise NotImplementedError(
            "EquilibriumOutletP acts inside the native step only")

    def make_no_streaming_mask(self, shape: List[int], context):
        m = context.zero_tensor(shape, dtype=torch.bool)
        keep = np.setdiff1d(np.arange(shape[0]), self.velocities)
        m[tuple([torch.as_tensor(keep)] + self.index)] = True
        return m

    def make_no_collision_mask(self, shape: List[int], context):
        m = context.zero_tensor(shape, dtype=torch.bool)
        m[tuple(self.index)] = True
        return m

    def native_available(self) -> bool:
        return True

    def native_generator(self, index: int):
        return NativeBoundaryDesc(N.OP_OUTLET, index, owner=self)

    def _op_spec(self, flow):
        d = flow.stencil.d
        axes = ([0] + self.direction) if d == 2 else list(self.direction)
        column = self.make_no_collision_mask(list(flow.resolution), flow.context)
        return OpSpec(op=N.OP_OUTLET, rho=self.rho_outlet, direction=tuple(axes),
                      column_mask=column, lattice_direction=tuple(self.direction))


class HalfwayBounceBackBoundary(Boundary):
    """Halfway bounce-back on a solid mask
    (examples/advanced_projects/efficient_bounce_back_obstacle/boundary/
    halfway_bounce_back_boundary.py:10-255): solid cells neither collide nor
    stream, and a fluid cell x whose neighbour x - e_q is solid receives its
    own post-collision f_opp(q) after streaming.  It is a post-streaming
    boundary (return it from ``flow.post_streaming_boundaries``).  With
    ``calc_force`` the momentum-exchange force is available through
    ``force(simulation)``."""

    def __init__(self, context, flow, solid_mask, calc_force: bool = False):
        self.context = context
        self.mask = context.convert_to_tensor(solid_mask, dtype=torch.bool)
        self.calc_force = calc_force
        self.force_sum = None

    def __call__(self, flow):
        raise NotImplementedError(
            "HalfwayBounceBackBoundary acts inside the native step only")

    def make_no_collision_mask(self, shape: List[int], context):
        return self.mask

    def make_no_streaming_mask(self, shape: List[int], context):
        return self.mask

    def native_available(self) -> bool:
        return True

    def native_generator(self, index: int):
        return NativeBoundaryDesc(N.OP_KEEP, index, owner=self,
                                  solid_mask=self.mask)

    def force(self, simulation) -> torch.Tensor:
        """FP64 momentum-exchange force on the solid, F = 2 sum f*_d(x) e_d
        over links from fluid x into the solid (halfway_bounce_back_boundary.
        py:211-218), from the post-collision populations of the last step."""
        flow = simulation.flow
        post = simulation.streaming_strategy.post_streaming()
        # the link list depends on the flow (stencil, grid, device) and the
        # streaming order, so it is cached per (flow, order): one boundary
        # object may serve several simulations
        key = (id(flow), type(flow.stencil).__name__, tuple(flow.resolution),
               str(flow.context.device), post)
        cache = self.__dict__.setdefault("_link_cache", {})
        if key not in cache:
            from .._classes import halfway_links
            links = halfway_links(self.mask, flow.stencil.e, flow.stencil.opposite)
            if post:
                # after a push step f*_d(x) sits in slot opp(d) of x
                opp = torch.tensor(flow.stencil.opposite, dtype=torch.int32)
                links = torch.stack([links[:, 0], opp[links[:, 1].long()]], 1)
            cache[key] = (links.to(flow.context.device), -1.0 if post else 1.0)
        links, sign = cache[key]
        F = _ops.link_force_f64(flow.stencil, flow._state, links, flow.rho0_shift)
        self.force_sum = F[:flow.stencil.d] * sign
        return self.force_sum


class AntiBounceBackOutlet(Boundary):
    """lettuce's anti-bounce-back outlet
    (lettuce/ext/_boundary/anti_bounce_back_outlet.py:13-109): on the outlet
    column, for the populations v leaving the domain,
    f_opp(v) <- -f_v + w_v rho (2 + (e_v.u_w)^2 / cs^4 - |u_w|^2 / cs^2) with
    u_w = u + (u - u_neighbour) / 2; the incoming populations opp(v) are not
    streamed there.  As in lettuce, `collision` is accepted and unused."""

    def __init__(self, direction: List[int], flow, collision=None):
        self.direction = [int(round(float(c))) for c in direction]
        self.velocities, self.index = _outlet_geometry(self.direction, flow)
        self._opposite = list(flow.stencil.opposite)

    def __call__(self, flow):
        raise NotImplementedError(
            "AntiBounceBackOutlet acts inside the native step only")

    def make_no_streaming_mask(self, shape: List[int], context):
        m = context.zero_tensor(shape, dtype=torch.bool)
        bounced = np.array(self._opposite)[self.velocities]
        m[tuple([torch.as_tensor(bounced)] + self.index)] = True
        return m

    def make_no_collision_mask(self, shape: List[int], context):
        m = context.zero_tensor(shape, dtype=torch.bool)
        m[tuple(self.index)] = True
        return m

    def native_available(self) -> bool:
        return True

    def native_generator(self, index: int):
        return NativeBoundaryDesc(N.OP_ANTI_BOUNCE_BACK, index, owner=self)

    def _op_spec(self, flow):
        d = flow.stencil.d
        axes = ([0] + self.direction) if d == 2 else list(self.direction)
        column = self.make_no_collision_mask(list(flow.resolution), flow.context)
        return OpSpec(op=N.OP_ANTI_BOUNCE_BACK, direction=tuple(axes),
                      column_mask=column, lattice_direction=tuple(self.direction))

"""Time-step driver.

Mirrors lettuce's ``Simulation`` / ``BreakableSimulation`` and the
``Collision`` / ``Reporter`` interfaces (lettuce/_simulation.py:17-344) and
lettuce's ``StreamingStrategy`` (lettuce/cuda_native/_default_code_gen.py:
13-25).  The masks are built exactly as lettuce builds them; the step itself
is always the fused HIP kernel: ``_collide_and_stream`` is bound to a
``NativeStep`` that calls ``lbm_step`` and swaps f / f_next, which is the
job lettuce's generated ``invoke(simulation)`` does
(lettuce/cuda_native/_template.py:35-40).  A configuration without a native
kernel raises ``NotImplementedError`` when the simulation is built.
"""
import ctypes
import enum
from abc import ABC, abstractmethod
from timeit import default_timer as timer
from typing import List

import torch

from . import _native as N
from ._classes import (OpSpec, build_cell_classes, build_masks,
                       halfway_links)
from ._ops import STORAGE_OF, layout

__all__ = ["Collision", "Reporter", "Simulation", "BreakableSimulation",
           "StreamingStrategy", "NativeStep"]


class StreamingStrategy(enum.Enum):
    NO_STREAMING = 0b00
    PRE_STREAMING = 0b10
    POST_STREAMING = 0b01
    DOUBLE_STREAMING = 0b11

    def pre_streaming(self) -> bool:
        return bool(self.value & StreamingStrategy.PRE_STREAMING.value)

    def post_streaming(self) -> bool:
        return bool(self.value & StreamingStrategy.POST_STREAMING.value)


class Collision(ABC):
    @abstractmethod
    def __call__(self, flow):
        ...

    @abstractmethod
    def native_available(self) -> bool:
        ...

    @abstractmethod
    def native_generator(self, index: int):
        ...


class Reporter(ABC):
    interval: int

    def __init__(self, interval: int):
        self.interval = interval

    @abstractmethod
    def __call__(self, simulation: "Simulation"):
        ...


class NativeStep:
    """The per-simulation native step: argument block + class table, bound
    once; each call launches one fused kernel on torch's current stream and
    swaps the A/B buffers."""

    def __init__(self, simulation: "Simulation"):
        flow = simulation.flow
        ctx = flow.context
        if ctx.device.type != "cuda":
            raise NotImplementedError(
                "lettuce_amd steps only on a ROCm device; the flow lives on "
                f"{ctx.device}")
        stencil = flow.stencil
        if stencil.native_id < 0:
            raise NotImplementedError(
                f"no native kernel for {type(stencil).__name__}")
        desc = simulation.collision.native_generator(simulation.collision_index)
        if desc is None:
            raise NotImplementedError(
                f"{type(simulation.collision).__name__} has no native kernel")
        desc.bind(flow)
        self.flow = flow
        self.collision_desc = desc
        self.device = ctx.device
        a = N.LbmStepArgs()
        a.abi_version = N.ABI_VERSION
        a.stencil = stencil.native_id
        a.storage = STORAGE_OF[ctx.storage_dtype]
        a.collision = desc.kind
        a.streaming = simulation.streaming_strategy.value
        a.tau = desc.tau
        a.tau_minus = desc.tau_minus
        a.force_tau = desc.force_tau
        a.accel[:] = list(desc.accel) + [0.0] * (3 - len(desc.accel))
        a.rho0_shift = flow.rho0_shift
        a.sr_seed = getattr(ctx, "rounding_seed", 0)
        # the descriptor owns the learned struct and its packed weights; it
        # stays alive with this step (self.collision_desc)
        self._learned = desc.learned_struct()
        if self._learned is not None:
            a.learned = ctypes.pointer(self._learned)
        n, s = layout(flow._state, stencil.q, stencil.d)
        a.n[:] = n
        a.stride[:] = s
        self._keep = []   # device tensors the argument block points into
        if simulation.cell_classes is not None:
            cc = simulation.cell_classes
            cls = cc.cell_class.to(self.device).contiguous()
            table = cc.table_tensor(self.device)
            self._keep += [cls, table]
            a.cell_class = cls.data_ptr()
            a.classes = table.data_ptr()
            a.n_classes = cc.n_classes
            if cc.bc_field is not None:
                bcf = cc.bc_field.to(self.device).contiguous()
                self._keep.append(bcf)
                a.bc_field = bcf.data_ptr()
        self.args = a

    def _bind_buffers(self):
        flow = self.flow
        f_in, f_out = flow._state, flow.f_next
        n, s = layout(f_in, flow.stencil.q, flow.stencil.d)
        if tuple(s) != tuple(self.args.stride) or tuple(n) != tuple(self.args.n):
            raise RuntimeError("population tensor layout changed after the "
                               "simulation was built")
        if f_out.stride() != f_in.stride():
            raise RuntimeError("f and f_next must share one layout")
        self.args.f_in = f_in.data_ptr()
        self.args.f_out = f_out.data_ptr()
        # the index of the next step: stochastic rounding draws its bits from it
        self.args.sr_step = int(flow.i) & 0xFFFFFFFF

    def __call__(self, simulation=None):
        self._bind_buffers()
        N.check(N.lib().lbm_step(ctypes.byref(self.args),
                                 N.stream_handle(self.device)), "lbm_step")
        self.flow.swap()

    def run(self, nsteps: int):
        """nsteps fused steps in one call (no reporter in between)."""
        if nsteps <= 0:
            return
        self._bind_buffers()
        N.check(N.lib().lbm_run(ctypes.byref(self.args), int(nsteps),
                                N.stream_handle(self.device)), "lbm_run")
        if nsteps % 2 == 1:
            self.flow.swap()

    def plan(self, nsteps: int):
        """(two-step launches, single-step launches) that run(nsteps) makes"""
        self._bind_buffers()
        out = (ctypes.c_int64 * 2)()
        N.check(N.lib().lbm_run_plan(ctypes.byref(self.args), int(nsteps), out),
                "lbm_run_plan")
        return int(out[0]), int(out[1])


class Simulation:
    def __init__(self, flow, collision, reporter: List["Reporter"],
                 streaming_strategy=StreamingStrategy.POST_STREAMING):
        self.flow = flow
        self.flow.collision = collision
        self.context = flow.context
        self.collision = collision
        self.reporter = reporter
        self.streaming_strategy = streaming_strategy
        # lettuce's Obstacle builds new boundary objects on every property
        # access (lettuce/ext/_flows/obstacle.py:107-122): read them once
        self.pre_boundaries = list(flow.pre_boundaries or [])
        self.post_boundaries = list(flow.post_boundaries or [])
        self.post_streaming_boundaries = list(
            getattr(flow, "post_streaming_boundaries", []) or [])
        self.collision_index = len(self.pre_boundaries)
        self.transformer = (self.pre_boundaries + [collision]
                            + self.post_boundaries)

        self.no_collision_mask, self.no_streaming_mask = build_masks(
            flow.stencil.q, flow.resolution, self.pre_boundaries,
            self.post_boundaries, self.post_streaming_boundaries, self.context)
        self.cell_classes = None
        self.solid_mask = None
        if self.no_collision_mask is not None:
            self._build_classes()
        self._collide_and_stream = NativeStep(self)

    # -- cell classes -------------------------------------------------------
    def _build_classes(self):
        flow = self.flow
        ops = {}
        for i, b in enumerate(self.pre_boundaries):
            ops[i] = self._boundary_spec(b, i)
        ops[self.collision_index] = OpSpec(op=N.OP_COLLIDE)
        for i, b in enumerate(self.post_boundaries,
                              start=self.collision_index + 1):
            ops[i] = self._boundary_spec(b, i)
        solid = None
        first = self.collision_index + 1 + len(self.post_boundaries)
        for i, b in enumerate(self.post_streaming_boundaries, start=first):
            desc = b.native_generator(i)
            if desc is None or desc.op != N.OP_KEEP:
                raise NotImplementedError(
                    f"{type(b).__name__} is not a native post-streaming "
                    f"boundary")
            m = desc.solid_mask.to(device=self.context.device, dtype=torch.bool)
            solid = m if solid is None else (solid | m)
        self.solid_mask = solid
        self.cell_classes = build_cell_classes(
            self.no_collision_mask, self.no_streaming_mask, ops,
            flow.stencil.e, solid)
        self._check_outlets(ops)

    def _boundary_spec(self, boundary, index) -> OpSpec:
        desc = boundary.native_generator(index)
        if desc is None:
            raise NotImplementedError(
                f"{type(boundary).__name__} has no native implementation")
        return desc.op_spec(self.flow)

    def _check_outlets(self, ops):
        """The outlets write their whole column in place
        (lettuce/ext/_boundary/equilibrium_outlet_p.py:69-72,
        anti_bounce_back_outlet.py:85-92) and read the velocity of the
        neighbouring column; both are reproduced only when no later boundary
        claims part of the column and the neighbours collide with a
        momentum-conserving operator or keep their values."""
        ncm = self.no_collision_mask
        for i, spec in ops.items():
            if spec.op not in (N.OP_OUTLET, N.OP_ANTI_BOUNCE_BACK):
                continue
            column = spec.column_mask
            if not bool((ncm[column] == i).all()):
                raise NotImplementedError(
                    "an outlet column overlapped by a later "
                    "boundary is not supported natively")
            # ncm of the neighbour x - direction, seen from x
            nb_idx = torch.roll(ncm, shifts=tuple(spec.lattice_direction),
                                dims=tuple(range(ncm.dim())))[column]
            for idx in torch.unique(nb_idx).tolist():
                nb_op = ops.get(idx, OpSpec()).op
                if nb_op not in (N.OP_COLLIDE, N.OP_KEEP, N.OP_OUTLET,
                                 N.OP_ANTI_BOUNCE_BACK):
                    raise NotImplementedError(
                        "an outlet's neighbour column must be fluid or solid cells")

    # -- driver -------------------------------------------------------------
    @property
    def units(self):
        return self.flow.units

    def step(self, num_steps: int):
        return self(num_steps)

    def _report(self):
        for reporter in self.reporter:
            reporter(self)

    def _steps_until_report(self, remaining: int) -> int:
        """How many steps may run before some reporter has to look."""
        n = remaining
        for r in self.reporter:
            interval = getattr(r, "interval", 1) or 1
            if not getattr(r, "periodic_interval", False):
                return 1
            nxt = interval - (self.flow.i % interval)
            n = min(n, nxt)
        return max(1, n)

    def __call__(self, num_steps: int) -> float:
        beg = timer()
        if self.flow.i == 0:
            self._report()
        done = 0
        step = self._collide_and_stream
        while done < num_steps:
            k = self._steps_until_report(num_steps - done)
            if k == 1:
                step(self)
            else:
                step.run(k)
            self.flow.i += k
            done += k
            self._report()
        self.context.synchronize()
        end = timer()
        return num_steps * self._cells() / 1e6 / (end - beg)

    def _cells(self) -> int:
        n = 1
        for k in self.flow.resolution:
            n *= k
        return n

    def halfway_links(self) -> torch.Tensor:
        """(cell, d) links from fluid cells into halfway-bounce-back solids."""
        if self.solid_mask is None:
            return torch.zeros((0, 2), dtype=torch.int32)
        return halfway_links(self.solid_mask, self.flow.stencil.e,
                             self.flow.stencil.opposite)


class BreakableSimulation(Simulation):
    """Stops when a reporter pushes ``flow.i`` past the target
    (lettuce/_simulation.py:325-344)."""

    def __init__(self, flow, collision, reporter: List["Reporter"],
                 streaming_strategy=StreamingStrategy.POST_STREAMING):
        super().__init__(flow, collision, reporter, streaming_strategy)

    def __call__(self, num_steps: int) -> float:
        beg = timer()
        if self.flow.i == 0:
            self._report()
        step = self._collide_and_stream
        while self.flow.i < num_steps:
            k = self._steps_until_report(num_steps - self.flow.i)
            if k == 1:
                step(self)
            else:
                step.run(k)
            self.flow.i += k
            self._report()
        self.context.synchronize()
        end = timer()
        return num_steps * self._cells() / 1e6 / (end - beg)

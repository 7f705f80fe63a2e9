"""Flow: the simulation state and its plugin interfaces.

Mirrors lettuce's ``Equilibrium``/``Boundary``/``Flow`` (lettuce/_flow.py:17-
268).  The populations live in HBM in the context's storage format
(``context.storage``); ``flow.f`` always presents the true populations in
the arithmetic dtype.  For FP32/FP64 storage ``flow.f`` *is* the state
tensor (no copy), exactly like lettuce; for shifted 16-bit storage it is a
decoded copy and assigning ``flow.f`` re-encodes.

Moments (rho, j, u) are computed by the native FP64 moments kernel.
"""
import warnings
from abc import ABC, abstractmethod
from typing import List, Optional, Union

import numpy as np
import torch

from ._stencil import TorchStencil
from . import _ops

__all__ = ["Equilibrium", "Boundary", "Flow"]


class Equilibrium(ABC):
    @abstractmethod
    def __call__(self, flow: "Flow", rho=None, u=None) -> torch.Tensor:
        ...

    @abstractmethod
    def native_available(self) -> bool:
        ...

    @abstractmethod
    def native_generator(self):
        ...


class Boundary(ABC):
    @abstractmethod
    def __call__(self, flow: "Flow"):
        ...

    @abstractmethod
    def make_no_collision_mask(self, shape: List[int], context
                               ) -> Optional[torch.Tensor]:
        ...

    @abstractmethod
    def make_no_streaming_mask(self, shape: List[int], context
                               ) -> Optional[torch.Tensor]:
        ...

    @abstractmethod
    def native_available(self) -> bool:
        ...

    @abstractmethod
    def native_generator(self, index: int):
        ...


class Flow(ABC):
    initialize_pressure: bool = False
    initialize_fneq: bool = False

    def __init__(self, context, resolution: List[int], units, stencil,
                 equilibrium):
        self.context = context
        self.resolution = list(resolution)
        self.units = units
        self.stencil = stencil
        self.torch_stencil = TorchStencil(stencil, context)
        self.equilibrium = equilibrium
        self.i = 0
        self._state = context.population_tensor(stencil.q, self.resolution)
        self._state_next = None
        self.initialize()

    # -- boundaries ---------------------------------------------------------
    @property
    def pre_boundaries(self) -> List["Boundary"]:
        return []

    @property
    def post_boundaries(self) -> List["Boundary"]:
        warnings.warn(
            "post_boundaries is not defined by the flow class; falling back "
            "to the deprecated `boundaries` property", stacklevel=2)
        return self.boundaries

    @property
    def boundaries(self) -> List["Boundary"]:
        return []

    @property
    def post_streaming_boundaries(self) -> List["Boundary"]:
        """Boundaries applied after streaming (halfway bounce-back), as in
        EbbSimulation (examples/.../simulation/ebb_simulation.py:21-25)."""
        return []

    # -- initial state ------------------------------------------------------
    @abstractmethod
    def initial_pu(self):
        """initial (pressure, velocity) in physical units"""
        ...

    def initialize(self):
        """Equilibrium initialisation (lettuce/_flow.py:127-143)."""
        p, u = self.initial_pu()
        rho = self.context.convert_to_tensor(
            self.units.convert_pressure_pu_to_density_lu(p))
        u = self.context.convert_to_tensor(self.units.convert_velocity_to_lu(u))
        if self.initialize_pressure:
            from .util import pressure_poisson
            rho = pressure_poisson(self.units, u, rho)
        self.f = self.equilibrium(self, rho=rho, u=u)
        if self.initialize_fneq:
            from .util import initialize_f_neq
            self.f = initialize_f_neq(self)

    # -- populations --------------------------------------------------------
    @property
    def rho0_shift(self) -> float:
        return 1.0

    @property
    def f(self) -> torch.Tensor:
        if self.context.shifted:
            return _ops.decode(self.stencil, self._state, self.rho0_shift,
                               self.context.dtype)
        return self._state

    @f.setter
    def f(self, value: torch.Tensor):
        value = self.context.convert_to_tensor(value)
        if tuple(value.shape) != tuple(self._state.shape):
            raise ValueError(f"f must have shape {tuple(self._state.shape)}, "
                             f"got {tuple(value.shape)}")
        if self.context.shifted:
            _ops.encode(self.stencil, value, self._state, self.rho0_shift)
        elif self.context.layout == "soa":
            self._state = value.contiguous()
        else:
            state = self.context.population_tensor(self.stencil.q, self.resolution)
            state.copy_(value)
            self._state = state

    @property
    def f_next(self) -> torch.Tensor:
        """The second (B) buffer, created on first use
        (lettuce/_flow.py:145-151); in the storage format."""
        if (self._state_next is None or self._state_next.shape != self._state.shape
                or self._state_next.stride() != self._state.stride()):
            self._state_next = self.context.population_tensor(self.stencil.q,
                                                              self.resolution)
        return self._state_next

    @f_next.setter
    def f_next(self, value: torch.Tensor):
        self._state_next = value

    def swap(self):
        """f, f_next = f_next, f (lettuce/cuda_native/_template.py:39)."""
        self._state, self._state_next = self.f_next, self._state

    # -- moments ------------------------------------------------------------
    def _moments(self, f=None):
        if f is None:
            return _ops.moments_f64(self.stencil, self._state, self.rho0_shift)
        f = self.context.convert_to_tensor(f).contiguous()
        return _ops.moments_f64(self.stencil, f)

    def rho(self, f: Optional[torch.Tensor] = None) -> torch.Tensor:
        rho, _ = self._moments(f)
        return rho[None, ...].to(self.context.dtype)

    @property
    def rho_pu(self) -> torch.Tensor:
        return self.units.convert_density_to_pu(self.rho())

    @property
    def p_pu(self) -> torch.Tensor:
        return self.units.convert_density_lu_to_pressure_pu(self.rho())

    @property
    def u_pu(self):
        return self.units.convert_velocity_to_pu(self.u())

    def j(self, f: Optional[torch.Tensor] = None) -> torch.Tensor:
        rho, u = self._moments(f)
        return (u * rho[None, ...]).to(self.context.dtype)

    def u(self, f: Optional[torch.Tensor] = None, rho=None,
          acceleration=None) -> torch.Tensor:
        """velocity; with `acceleration` the forcing correction a / (2 rho)
        is added (lettuce/_flow.py:178-193)"""
        if rho is None:
            rho_f64, u = self._moments(f)
            u = u.to(self.context.dtype)
            rho = rho_f64[None].to(self.context.dtype)
        else:
            u = self.j(f) / rho
        if acceleration is None:
            return u
        a = self.context.convert_to_tensor(acceleration)
        if a.dim() == 1:
            a = a[(Ellipsis,) + (None,) * self.stencil.d]
        return u + a / (2 * rho)

    @property
    def velocity(self):
        return self.u()

    def incompressible_energy(self, f: Optional[torch.Tensor] = None
                              ) -> torch.Tensor:
        u = self.u(f)
        return 0.5 * (u * u).sum(dim=0)

    # -- checkpoint ---------------------------------------------------------
    def dump(self, filename):
        """Save the true populations (lettuce/_flow.py:258-261) as a .npy
        file (no pickle)."""
        with open(filename, "wb") as fh:
            np.save(fh, self.context.convert_to_ndarray(self.f),
                    allow_pickle=False)

    def load(self, filename):
        with open(filename, "rb") as fh:
            self.f = self.context.convert_to_tensor(
                np.load(fh, allow_pickle=False), dtype=self.context.dtype)
        self._state_next = None

    _UNIT_FIELDS = ("reynolds_number", "mach_number", "characteristic_length_pu",
                    "characteristic_velocity_pu", "characteristic_length_lu",
                    "characteristic_density_lu", "characteristic_density_pu", "cs")

    def save_checkpoint(self, filename):
        """Restart file (.npz, no pickle): the stored populations bit for bit
        in their storage format, the step counter flow.i, the stencil, the
        storage format and the unit parameters.  A superset of dump(), which
        keeps only the decoded populations (lettuce/_flow.py:258-268)."""
        state = self._state.detach()
        if state.dtype in (torch.float16, torch.bfloat16):
            raw = state.view(torch.int16)
        else:
            raw = state
        perm = [0] + list(range(1, state.dim()))
        if self.context.layout == "rows":
            # memory order [*res[:-1], Q, res[-1]]: save as stored
            perm = list(range(1, state.dim() - 1)) + [0, state.dim() - 1]
        raw = raw.permute(perm).contiguous().cpu().numpy()
        meta = {"format": np.array("lettuce_amd-checkpoint-1"),
                "stencil": np.array(type(self.stencil).__name__),
                "storage": np.array(self.context.storage),
                "layout": np.array(self.context.layout),
                "state_dtype": np.array(str(state.dtype).replace("torch.", "")),
                "i": np.array(self.i, dtype=np.int64),
                "resolution": np.array(self.resolution, dtype=np.int64),
                "rho0_shift": np.array(self.rho0_shift, dtype=np.float64)}
        for k in self._UNIT_FIELDS:
            meta["unit_" + k] = np.array(float(getattr(self.units, k)), dtype=np.float64)
        with open(filename, "wb") as fh:
            np.savez(fh, state=raw, **meta)

    def load_checkpoint(self, filename, check_units: bool = True):
        """Restore a save_checkpoint() file into this flow: the stored
        populations bit for bit and flow.i.  The stencil, resolution,
        storage format and layout must match this flow's context; with
        check_units the unit parameters must match too."""
        with np.load(filename, allow_pickle=False) as z:
            if str(z["format"]) != "lettuce_amd-checkpoint-1":
                raise ValueError(f"{filename} is not a lettuce_amd checkpoint")
            for key, mine in (("stencil", type(self.stencil).__name__),
                              ("storage", self.context.storage),
                              ("layout", self.context.layout)):
                if str(z[key]) != mine:
                    raise ValueError(f"checkpoint {key} {str(z[key])!r} differs from "
                                     f"this flow's {mine!r}")
            if list(z["resolution"]) != list(self.resolution):
                raise ValueError(f"checkpoint resolution {list(z['resolution'])} differs "
                                 f"from {self.resolution}")
            # shifted 16-bit states decode against rho0_shift: restoring the
            # bits under another shift would silently change every population
            if self.context.shifted and float(z["rho0_shift"]) != float(self.rho0_shift):
                raise ValueError(f"checkpoint rho0_shift {float(z['rho0_shift'])} differs "
                                 f"from this flow's {self.rho0_shift}")
            if check_units:
                for k in self._UNIT_FIELDS:
                    if float(z["unit_" + k]) != float(getattr(self.units, k)):
                        raise ValueError(f"checkpoint unit {k} = {float(z['unit_' + k])} "
                                         f"differs from {getattr(self.units, k)}")
            raw = torch.from_numpy(np.ascontiguousarray(z["state"]))
            i = int(z["i"])
        state = self.context.population_tensor(self.stencil.q, self.resolution)
        if state.dtype in (torch.float16, torch.bfloat16):
            target = state.view(torch.int16)
        else:
            target = state
        if raw.dtype != target.dtype:
            raise ValueError(f"checkpoint holds {raw.dtype}, this flow stores {target.dtype}")
        perm = [0] + list(range(1, state.dim()))
        if self.context.layout == "rows":
            perm = list(range(1, state.dim() - 1)) + [0, state.dim() - 1]
        target.permute(perm).copy_(raw.to(target.device))
        self._state = state
        self._state_next = None
        self.i = i

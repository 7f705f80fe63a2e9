"""Observables and reporters.

lettuce's observables (lettuce/ext/_reporter/observable_reporter.py:17-200)
with their global sums computed on the GPU in FP64 by the native reduction
kernels; unit conversions are applied on the host afterwards, in the same
order as lettuce.
"""
import sys
from abc import ABC, abstractmethod
from typing import Optional

import numpy as np
import torch

from .. import _ops
from .._simulation import Reporter

__all__ = ["Observable", "ObservableReporter", "MaximumVelocity",
           "IncompressibleKineticEnergy", "Enstrophy", "Mass", "ErrorReporter"]


class Observable(ABC):
    def __init__(self, flow):
        self.context = flow.context
        self.flow = flow

    @abstractmethod
    def __call__(self, f: Optional[torch.Tensor] = None):
        ...


def _state(flow, f):
    """(population tensor, rho0 shift) for the native reductions."""
    if f is None or f is flow._state:
        return flow._state, flow.rho0_shift
    f = flow.context.convert_to_tensor(f).contiguous()
    return f, 1.0


class MaximumVelocity(Observable):
    """max |u| in physical units (observable_reporter.py:27-31)."""

    def __call__(self, f=None):
        _, u = _ops.moments_f64(self.flow.stencil, *_state(self.flow, f))
        u = self.flow.units.convert_velocity_to_pu(u)
        return torch.linalg.vector_norm(u, dim=0).max().to(self.context.dtype)


class IncompressibleKineticEnergy(Observable):
    """sum 1/2 |u|^2 dx^D in physical units (observable_reporter.py:34-42);
    the sum is taken in FP64."""

    def __call__(self, f=None):
        flow = self.flow
        s = _ops.reduce_f64(flow.stencil, *_state(flow, f))[4]
        dx = flow.units.convert_length_to_pu(1.0)
        return flow.units.convert_incompressible_energy_to_pu(s) * dx ** flow.stencil.d


class Enstrophy(Observable):
    """sum |curl u|^2 dx^D in physical units with 6th-order periodic finite
    differences (observable_reporter.py:45-68), in FP64."""

    def __call__(self, f=None):
        flow = self.flow
        _, u = _ops.moments_f64(flow.stencil, *_state(flow, f))
        s = _ops.enstrophy_f64(flow.stencil, u.contiguous())
        dx = flow.units.convert_length_to_pu(1.0)
        scale = flow.units.convert_velocity_to_pu(1.0) / dx
        return s * scale ** 2 * dx ** flow.stencil.d


class Mass(Observable):
    """Total mass in lattice units over f[..., 1:-1, 1:-1] minus the masked
    cells (observable_reporter.py:140-158), in FP64."""

    def __init__(self, flow, no_mass_mask=None):
        super().__init__(flow)
        self.mask = no_mass_mask

    def __call__(self, f=None):
        flow = self.flow
        state, shift = _state(flow, f)
        res = list(flow.resolution)
        edge = torch.ones(res, dtype=torch.bool, device=state.device)
        inner = [slice(None)] * (len(res) - 2) + [slice(1, -1), slice(1, -1)]
        edge[tuple(inner)] = False
        mass = _ops.reduce_f64(flow.stencil, state, shift, exclude=edge)[0]
        if self.mask is not None:
            m = self.context.convert_to_tensor(self.mask, dtype=torch.bool)
            mass = mass - _ops.reduce_f64(flow.stencil, state, shift,
                                          exclude=~m)[0]
        return mass


class ObservableReporter(Reporter):
    """Records or prints [step, time_pu, value] every `interval` steps
    (observable_reporter.py:161-200)."""

    periodic_interval = True

    def __init__(self, observable, interval=1, out=sys.stdout):
        super().__init__(interval)
        self.observable = observable
        self.out = [] if out is None else out
        self._parameter_name = observable.__class__.__name__
        if out is not None:
            print("steps    ", "time    ", self._parameter_name)

    def __call__(self, simulation):
        if simulation.flow.i % self.interval == 0:
            observed = self.observable.context.convert_to_ndarray(
                self.observable(simulation.flow.f
                                if not simulation.flow.context.shifted
                                else None))
            if observed.ndim >= 2:
                raise ValueError("observables must be scalars or vectors")
            observed = [observed.item()] if observed.ndim == 0 else observed.tolist()
            entry = ([simulation.flow.i,
                      simulation.units.convert_time_to_pu(simulation.flow.i)]
                     + observed)
            if isinstance(self.out, list):
                self.out.append(entry)
            else:
                print(*entry, file=self.out)


class ErrorReporter(Reporter):
    """L2 errors of u and p against an analytic solution every `interval`
    steps, each divided by resolution^(d/2)
    (lettuce/ext/_reporter/error_reporter.py:9-42)."""

    periodic_interval = True

    def __init__(self, analytical_solution, interval=1, out=sys.stdout):
        super().__init__(interval)
        self.analytical_solution = analytical_solution
        self.out = [] if out is None else out
        if not isinstance(self.out, list):
            print("#error_u         error_p", file=self.out)

    def __call__(self, simulation):
        flow = simulation.flow
        i = flow.i
        if i % self.interval != 0:
            return
        t = simulation.units.convert_time_to_pu(i)
        pref, uref = self.analytical_solution(t=t)
        pref = flow.context.convert_to_tensor(pref)
        uref = flow.context.convert_to_tensor(uref)
        p, u = flow.p_pu, flow.u_pu
        d = flow.stencil.d
        resolution = torch.pow(torch.prod(flow.context.convert_to_tensor(p.size())), 1 / d)
        err_u = torch.norm(u - uref) / resolution ** (d / 2)
        err_p = torch.norm(p - pref) / resolution ** (d / 2)
        if isinstance(self.out, list):
            self.out.append([err_u.item(), err_p.item()])
        else:
            print(err_u.item(), err_p.item(), file=self.out)

"""Forcing (lettuce/ext/_force/_force.py:6-31, guo.py:9-43).

Guo forcing enters the native step through BGKCollision(tau, force=Guo(...)):
the kernel shifts the equilibrium velocity by a / (2 rho) and adds Guo's
source term in the same pass (collision LBM_COLL_BGK_GUO).  The acceleration
is uniform, a vector in lattice units, as in lettuce's flows
(poiseuille.py:84-86).  source_term() and u_eq() are kept for callers that
evaluate the terms themselves; they are torch expressions of lettuce's
formulas, not part of the time step.
"""
from abc import ABC, abstractmethod

import torch

__all__ = ["Force", "Guo"]


class Force(ABC):
    @abstractmethod
    def __init__(self, flow, tau, acceleration):
        ...

    @abstractmethod
    def source_term(self, u):
        ...

    @abstractmethod
    def u_eq(self, flow=None):
        ...

    @property
    @abstractmethod
    def ueq_scaling_factor(self):
        ...

    @abstractmethod
    def native_available(self) -> bool:
        ...

    @abstractmethod
    def native_generator(self):
        ...


class Guo(Force):
    """Guo forcing with a uniform acceleration (lattice units); `tau` sets the
    source-term prefactor 1 - 1/(2 tau) (guo.py:15-26)."""

    def __init__(self, flow, tau, acceleration):
        self.flow = flow
        self.tau = float(tau)
        a = torch.as_tensor(acceleration, dtype=torch.float64).flatten()
        if a.numel() != flow.stencil.d:
            raise ValueError(f"acceleration needs {flow.stencil.d} components")
        self.acceleration = flow.context.convert_to_tensor(a)

    def _e(self):
        return torch.tensor(self.flow.stencil.e, dtype=self.acceleration.dtype,
                            device=self.acceleration.device)

    def source_term(self, u):
        """(1 - 1/(2 tau)) w_q ((e_q - u) / cs^2 + (e_q.u) e_q / cs^4) . a"""
        e = self._e()
        d = self.flow.stencil.d
        w = torch.tensor(self.flow.stencil.w, dtype=u.dtype, device=u.device)
        cs2 = self.flow.stencil.cs ** 2
        ea = e @ self.acceleration.to(u.dtype)                   # [q]
        ua = torch.einsum("d...,d->...", u, self.acceleration.to(u.dtype))
        eu = torch.einsum("qd,d...->q...", e.to(u.dtype), u)
        ea = ea.reshape([-1] + [1] * d)
        return (1 - 1 / (2 * self.tau)) * w.reshape([-1] + [1] * d) * (
            (ea - ua) / cs2 + eu * ea / cs2 ** 2)

    def u_eq(self, flow=None):
        flow = self.flow if flow is None else flow
        a = self.acceleration.reshape([-1] + [1] * flow.stencil.d)
        return self.ueq_scaling_factor * a / flow.rho()

    @property
    def ueq_scaling_factor(self):
        return 0.5

    def native_available(self) -> bool:
        return True

    def native_generator(self):
        a = [float(x) for x in self.acceleration.tolist()]
        return {"tau": self.tau, "acceleration": tuple(a + [0.0] * (3 - len(a)))}

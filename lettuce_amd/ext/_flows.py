"""Flows (initial conditions and boundary sets).

Same constructors and initial fields as lettuce's ExtFlow,
TaylorGreenVortex, DecayingTurbulence and Obstacle
(lettuce/ext/_flows/_ext_flow.py:8-42, taylorgreen.py:16-98,
decayingturbulence.py:23-189, obstacle.py:16-125).  Initial fields are set
up with torch/numpy on the host or device (setup, not the step); the
equilibrium initialisation runs through the native kernel.

``Obstacle`` additionally accepts ``outlet`` ("antibounceback" as in
lettuce, or "equilibrium" for EquilibriumOutletP), ``solid``
("fullway" as in lettuce, or "halfway" for halfway bounce-back) and
``walls`` (halfway solid rows at the lateral domain edges).
"""
import warnings
from abc import ABC, abstractmethod
from typing import List, Optional, Union

import numpy as np
import torch

from .._flow import Flow
from .._stencil import D2Q9, D3Q19, Stencil
from .._unit import UnitConversion
from ._operators import (AntiBounceBackOutlet, BounceBackBoundary,
                         EquilibriumBoundaryPU, EquilibriumOutletP,
                         HalfwayBounceBackBoundary, QuadraticEquilibrium)

__all__ = ["ExtFlow", "TaylorGreenVortex", "DecayingTurbulence", "Obstacle",
           "PoiseuilleFlow2D"]


class ExtFlow(Flow, ABC):
    def __init__(self, context, resolution: Union[int, List[int]],
                 reynolds_number, mach_number,
                 stencil: Optional[Stencil] = None, equilibrium=None):
        resolution = self.make_resolution(resolution, stencil)
        if len(resolution) not in (2, 3):
            raise ValueError("lettuce_amd supports 2D and 3D flows")
        stencil = stencil or (D2Q9() if len(resolution) == 2 else D3Q19())
        stencil = stencil() if callable(stencil) else stencil
        equilibrium = equilibrium or QuadraticEquilibrium()
        Flow.__init__(self, context, resolution,
                      self.make_units(reynolds_number, mach_number, resolution),
                      stencil, equilibrium)

    @abstractmethod
    def make_resolution(self, resolution, stencil=None) -> List[int]:
        ...

    @abstractmethod
    def make_units(self, reynolds_number, mach_number, resolution
                   ) -> UnitConversion:
        ...


def _periodic_grid(resolution, context):
    """linspace(0, 2 pi) without the endpoint, in the context dtype."""
    axes = tuple(torch.linspace(0, 2 * torch.pi * (1 - 1 / n), steps=n,
                                device=context.device, dtype=context.dtype)
                 for n in resolution)
    return torch.meshgrid(*axes, indexing="ij")


class TaylorGreenVortex(ExtFlow):
    def __init__(self, context, resolution, reynolds_number, mach_number,
                 stencil: Optional[Stencil] = None, equilibrium=None,
                 initialize_fneq: bool = True):
        self.initialize_fneq = initialize_fneq
        if stencil is None and not isinstance(resolution, list):
            warnings.warn("Requiring information about dimensionality! "
                          "Setting dimension to 2.", UserWarning)
            self.stencil = D2Q9()
        else:
            self.stencil = stencil() if callable(stencil) else stencil
        ExtFlow.__init__(self, context, resolution, reynolds_number,
                         mach_number, stencil, equilibrium)

    def make_resolution(self, resolution, stencil=None) -> List[int]:
        if isinstance(resolution, int):
            return [resolution] * self.stencil.d
        return list(resolution)

    def make_units(self, reynolds_number, mach_number, resolution):
        return UnitConversion(
            reynolds_number=reynolds_number, mach_number=mach_number,
            characteristic_length_lu=resolution[0] / (2 * torch.pi),
            characteristic_length_pu=1, characteristic_velocity_pu=1)

    @property
    def grid(self):
        return _periodic_grid(self.resolution, self.context)

    def initial_pu(self):
        return self.analytic_solution(t=0)

    def analytic_solution(self, t: float):
        """2D: u = (cos x sin y, -sin x cos y) e^{-2 nu t},
        p = -(cos 2x + cos 2y)/4 e^{-4 nu t}; 3D initial field
        (taylorgreen.py:66-94)."""
        x = self.grid
        nu = self.context.convert_to_tensor(self.units.viscosity_pu)
        if len(self.resolution) == 2:
            decay = torch.exp(-2 * nu * t)
            u = torch.stack([torch.cos(x[0]) * torch.sin(x[1]) * decay,
                             -torch.sin(x[0]) * torch.cos(x[1]) * decay])
            p = -torch.stack([0.25 * (torch.cos(2 * x[0]) + torch.cos(2 * x[1]))
                              * torch.exp(-4 * nu * t)])
        else:
            u = torch.stack([torch.sin(x[0]) * torch.cos(x[1]) * torch.cos(x[2]),
                             -torch.cos(x[0]) * torch.sin(x[1]) * torch.cos(x[2]),
                             torch.zeros_like(x[0])])
            p = torch.stack([1 / 16. * (torch.cos(2 * x[0]) + torch.cos(2 * x[1]))
                             * (torch.cos(2 * x[2]) + 2)])
        return p, u

    @property
    def post_boundaries(self):
        return []


class PoiseuilleFlow2D(ExtFlow):
    """Channel driven by a uniform acceleration (0.001, 0) between fullway
    bounce-back walls at y = 0 and y = -1 (lettuce/ext/_flows/poiseuille.py:
    18-86); run it with BGKCollision(tau, force=Guo(flow, tau,
    flow.units.convert_acceleration_to_lu(flow.acceleration)))."""

    def __init__(self, context, resolution, reynolds_number, mach_number,
                 stencil: Optional[Stencil] = None, equilibrium=None,
                 initialize_with_zeros: bool = True):
        self.stencil = D2Q9() if stencil is None else (stencil() if callable(stencil)
                                                      else stencil)
        self.initialize_with_zeros = initialize_with_zeros
        ExtFlow.__init__(self, context, resolution, reynolds_number, mach_number,
                         self.stencil, equilibrium)

    def make_resolution(self, resolution, stencil=None) -> List[int]:
        if isinstance(resolution, int):
            return [resolution] * self.stencil.d
        if len(resolution) != self.stencil.d:
            raise ValueError(f"resolution needs {self.stencil.d} entries")
        return list(resolution)

    def make_units(self, reynolds_number, mach_number, resolution):
        return UnitConversion(reynolds_number=reynolds_number, mach_number=mach_number,
                              characteristic_length_lu=resolution[0] - 1,
                              characteristic_length_pu=1, characteristic_velocity_pu=1)

    @property
    def grid(self):
        xyz = tuple(torch.linspace(0, 1, steps=n, device=self.context.device,
                                   dtype=self.context.dtype) for n in self.resolution)
        return torch.meshgrid(*xyz, indexing="ij")

    def analytic_solution(self, t=0):
        """u_x = a / (2 rho nu) (y - h)(1 - h - y), h half a lattice spacing
        (poiseuille.py:31-41)"""
        h = 0.5 / self.resolution[0]
        x, y = self.grid
        nu = self.units.viscosity_pu
        ux = self.acceleration[0] / (2 * 1 * nu) * ((y - h) * (1 - h - y))
        uy = torch.zeros_like(ux)
        p = y * 0 + self.units.convert_density_lu_to_pressure_pu(1)
        return p[None], torch.stack([ux, uy])

    def initial_pu(self):
        if self.initialize_with_zeros:
            z = torch.zeros(self.resolution, device=self.context.device,
                            dtype=self.context.dtype)
            return z[None], torch.stack([z, z])
        return self.analytic_solution()

    @property
    def post_boundaries(self):
        mask = torch.zeros(self.resolution, dtype=torch.bool)
        mask[:, [0, -1]] = True
        return [BounceBackBoundary(mask)]

    @property
    def acceleration(self):
        return self.context.convert_to_tensor([0.001, 0.0])


def turbulence_velocity(resolution, k0, ic_energy, dx, length_pu, randseed, device):
    """Velocity field (PU) of lettuce's DecayingTurbulence
    (lettuce/ext/_flows/decayingturbulence.py:67-155), its shell spectrum and
    shell wavenumbers.  The random phases come from numpy's legacy global
    RNG seeded with `randseed`, as in lettuce (:90-91); the FFTs and the
    projection run in FP64 on `device` with torch, which agrees with
    lettuce's numpy arithmetic to rounding.  Wavenumbers on np.meshgrid's
    default 'xy' grid, shells k - 1/2 < |k| <= k + 1/2 (:67-76);
    E(k) ~ k^4 exp(-2 (k/k0)^2) normalised to ic_energy (:78-86); amplitudes
    sqrt(2/d E / |u_hat|^2) (:102-110); divergence removed with the
    modified wavenumber sin(k dx)/dx (:118-134); energy rescaled to
    ic_energy (:136-143); the inverse transform scaled as in :145-153."""
    dev = torch.device(device)
    f64 = torch.float64
    dims = tuple(int(n) for n in resolution)
    d = len(dims)
    axes = tuple(range(d))
    freqs = [torch.fft.fftfreq(n, d=1.0 / n, dtype=f64, device=dev) for n in dims]
    wavenumber = torch.stack(torch.meshgrid(*freqs, indexing="xy"))
    norms = torch.linalg.vector_norm(wavenumber, dim=0)
    ek = norms ** 4 * torch.exp(-2 * (norms / k0) ** 2)
    ek = ek / ek.sum()
    ek = ek * ic_energy
    # shell k holds the modes with k - 1/2 < |k| <= k + 1/2, k < int(max |k|)
    nshell = int(norms.max())
    shell = torch.ceil(norms - 0.5).long()
    inside = shell < nshell
    spectrum = torch.zeros(nshell, dtype=f64, device=dev).index_add_(
        0, shell[inside], ek[inside])
    del shell, inside
    np.random.seed(randseed)
    phases = np.random.random(tuple(wavenumber.shape)) * 2 * np.pi
    re, im = [], []
    for k in range(d):
        uh = torch.fft.fftn(torch.from_numpy(phases[k]).to(dev).to(torch.complex128), dim=axes)
        re.append(uh.real.contiguous())
        im.append(uh.imag.contiguous())
    del phases
    for k in range(d):
        amp = torch.sqrt(2 / d * ek / (im[k] ** 2 + re[k] ** 2 + 1.e-15))
        re[k] = amp * re[k]
        im[k] = amp * im[k]
        re[k].view(-1)[0] = 0
        im[k].view(-1)[0] = 0
    kmod = [torch.sin(wavenumber[k] * dx) / dx for k in range(d)]
    del wavenumber, norms, ek
    knorm = torch.linalg.vector_norm(torch.stack(kmod), dim=0) + 1e-16
    div_re = torch.zeros(re[0].shape, dtype=f64, device=dev)
    div_im = torch.zeros(re[0].shape, dtype=f64, device=dev)
    for k in range(d):
        div_re += kmod[k] * re[k]
        div_im += kmod[k] * im[k]
    for k in range(d):
        re[k] = re[k] - div_re * kmod[k] / knorm ** 2
        im[k] = im[k] - div_im * kmod[k] / knorm ** 2
        re[k].view(-1)[0] = 0
        im[k].view(-1)[0] = 0
    del div_re, div_im, kmod, knorm
    e_kin = 0.5 * sum(float(torch.sum(re[k] ** 2 + im[k] ** 2)) for k in range(d))
    factor = np.sqrt(ic_energy / e_kin)
    if d == 3:
        norm = dims[0] * dx ** (1 - d) * np.sqrt(length_pu)
    else:
        norm = dims[0] / dx
    u = torch.stack([(torch.fft.ifftn(torch.complex(re[k] * factor, im[k] * factor), dim=axes)
                      * norm).real for k in range(d)])
    return u, spectrum.cpu().numpy(), np.arange(nshell)


class DecayingTurbulence(ExtFlow):
    """Decaying turbulence with E(k) ~ k^4 exp(-2 (k/k0)^2), random phases
    from numpy's legacy RNG, a solenoidal projection with the modified
    wavenumber and velocity scaled to the requested energy
    (decayingturbulence.py:54-170)."""

    def __init__(self, context, resolution, reynolds_number, mach_number,
                 k0=20, ic_energy=0.5, stencil=None, equilibrium=None,
                 initialize_pressure: bool = True,
                 initialize_fneq: bool = True,
                 randseed: Optional[int] = None):
        self.initialize_pressure = initialize_pressure
        self.initialize_fneq = initialize_fneq
        self.randseed = randseed
        self.k0 = k0
        self.ic_energy = ic_energy
        self.wavenumbers = []
        self.spectrum = []
        stencil = stencil or (D2Q9() if len(resolution) == 2 else D3Q19())
        stencil = stencil() if callable(stencil) else stencil
        if stencil.d != 2:
            self.initialize_pressure = False
        super().__init__(context, resolution, reynolds_number, mach_number,
                         stencil, equilibrium)

    def make_resolution(self, resolution, stencil=None) -> List[int]:
        if isinstance(resolution, int):
            return [resolution] * stencil.d
        return list(resolution)

    def make_units(self, reynolds_number, mach_number, resolution):
        return UnitConversion(
            reynolds_number=reynolds_number, mach_number=mach_number,
            characteristic_length_lu=resolution[0],
            characteristic_length_pu=2 * np.pi,
            characteristic_velocity_pu=None)

    def initial_pu(self):
        dx = self.units.convert_length_to_pu(1.0)
        u, self.spectrum, self.wavenumbers = turbulence_velocity(
            self.resolution, self.k0, self.ic_energy, dx,
            self.units.characteristic_length_pu, self.randseed, self.context.device)
        self.units.characteristic_velocity_pu = float(
            torch.linalg.vector_norm(u, dim=0).max())
        p = torch.zeros((1,) + tuple(u.shape[1:]), dtype=torch.float64, device=u.device)
        return p, u

    @property
    def energy_spectrum(self):
        return self.spectrum, self.wavenumbers

    @property
    def grid(self):
        return _periodic_grid(self.resolution, self.context)

    @property
    def post_boundaries(self):
        return []


class Obstacle(ExtFlow):
    """Flow past an obstacle mask: equilibrium inlet at x = 0, an outlet at
    x = -1 and a solid mask (lettuce/ext/_flows/obstacle.py:16-125)."""

    def __init__(self, context, resolution, reynolds_number, mach_number,
                 domain_length_x, char_length=1, char_velocity=1,
                 stencil: Optional[Stencil] = None, equilibrium=None,
                 outlet: str = "antibounceback", solid: str = "fullway",
                 walls: bool = False):
        self.char_length_lu = resolution[0] / domain_length_x * char_length
        self.char_length = char_length
        self.char_velocity = char_velocity
        if outlet not in ("antibounceback", "equilibrium"):
            raise ValueError(f"unknown outlet {outlet!r}")
        if solid not in ("fullway", "halfway"):
            raise ValueError(f"unknown solid boundary {solid!r}")
        self.outlet = outlet
        self.solid = solid
        self.walls = walls
        self.resolution = self.make_resolution(resolution, stencil)
        self._mask = torch.zeros(self.resolution, dtype=torch.bool)
        ExtFlow.__init__(self, context, resolution, reynolds_number,
                         mach_number, stencil, equilibrium)

    def make_units(self, reynolds_number, mach_number, resolution):
        return UnitConversion(
            reynolds_number=reynolds_number, mach_number=mach_number,
            characteristic_length_lu=self.char_length_lu,
            characteristic_length_pu=self.char_length,
            characteristic_velocity_pu=self.char_velocity)

    def make_resolution(self, resolution, stencil=None) -> List[int]:
        if isinstance(resolution, int):
            return [resolution] * (stencil.d if stencil else 2)
        return list(resolution)

    @property
    def mask(self):
        return self._mask

    @mask.setter
    def mask(self, m):
        if not (isinstance(m, (np.ndarray, torch.Tensor))
                and all(m.shape[k] == self.resolution[k]
                        for k in range(len(self.resolution)))):
            raise ValueError("mask must match the resolution")
        self._mask = self.context.convert_to_tensor(m, dtype=torch.bool)

    def wall_mask(self) -> torch.Tensor:
        """Solid rows at both ends of every lateral axis, excluding the inlet
        and outlet columns."""
        m = torch.zeros(self.resolution, dtype=torch.bool,
                        device=self._mask.device)
        if self.walls:
            for a in range(1, len(self.resolution)):
                lo = [slice(1, -1)] + [slice(None)] * (len(self.resolution) - 1)
                hi = list(lo)
                lo[a] = 0
                hi[a] = -1
                m[tuple(lo)] = True
                m[tuple(hi)] = True
        return m

    def solid_mask(self) -> torch.Tensor:
        return self._mask.to(torch.bool) | self.wall_mask().to(self._mask.device)

    def initial_pu(self):
        p = np.zeros(tuple(self.resolution), dtype=float)[None, ...]
        u_char = self.units.characteristic_velocity_pu * self._unit_vector()
        u_char = u_char.reshape([-1] + [1] * len(self.resolution))
        solid = self.solid_mask().cpu()
        u = (~solid).to(u_char.dtype) * u_char
        return p, u

    @property
    def grid(self):
        xyz = tuple(self.units.convert_length_to_pu(torch.arange(n))
                    for n in self.resolution)
        return torch.meshgrid(*xyz, indexing="ij")

    @property
    def post_boundaries(self):
        x = self.grid[0]
        inlet = EquilibriumBoundaryPU(
            flow=self, context=self.context, mask=torch.abs(x) < 1e-6,
            velocity=self.units.characteristic_velocity_pu * self._unit_vector())
        if self.outlet == "antibounceback":
            outlet = AntiBounceBackOutlet(self._unit_vector().tolist(), self)
        else:
            outlet = EquilibriumOutletP([int(c) for c in self._unit_vector().tolist()],
                                        self)
        boundaries = [inlet, outlet]
        if self.solid == "fullway":
            boundaries.append(BounceBackBoundary(self.solid_mask()))
        return boundaries

    @property
    def post_streaming_boundaries(self):
        if self.solid == "halfway":
            return [HalfwayBounceBackBoundary(self.context, self,
                                              self.solid_mask())]
        return []

    def _unit_vector(self, i=0):
        return torch.eye(len(self.resolution))[i]

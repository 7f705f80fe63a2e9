"""Finite differences, regridding and pressure/non-equilibrium
initialisation used to set up flows (lettuce/util/utility.py:37-161,
lettuce/_flow.py:271-367).  These run once per flow with torch on the
flow's device."""
import torch

__all__ = ["LettuceException", "torch_gradient", "grid_fine_to_coarse",
           "torch_jacobi", "pressure_poisson", "initialize_f_neq",
           "append_axes"]


class LettuceException(Exception):
    pass


# central-difference weights for offsets +3..+1, -1..-3 (Fornberg 1988)
_FD_WEIGHTS = {
    2: {1: 0.5},
    4: {1: 2 / 3, 2: -1 / 12},
    6: {1: 3 / 4, 2: -3 / 20, 3: 1 / 60},
}


def torch_gradient(f: torch.Tensor, dx=1, order=2) -> torch.Tensor:
    """[D, *res] periodic first derivatives of order 2, 4 or 6
    (lettuce/util/utility.py:37-99)."""
    if f.ndim not in (2, 3):
        raise LettuceException("Invalid dimension!")
    if order not in _FD_WEIGHTS:
        raise LettuceException(f"order must be 2, 4 or 6, got {order}")
    out = torch.empty((f.ndim,) + tuple(f.shape), dtype=f.dtype, device=f.device)
    with torch.no_grad():
        for a in range(f.ndim):
            acc = torch.zeros_like(f)
            for k, w in _FD_WEIGHTS[order].items():
                # f(x + k) - f(x - k): roll by -k brings x + k to x
                acc = acc + w * (torch.roll(f, -k, dims=a) - torch.roll(f, k, dims=a))
            out[a] = acc * torch.tensor(1.0 / dx, dtype=f.dtype, device=f.device)
    return out


def grid_fine_to_coarse(flow, f_fine, tau_fine, tau_coarse):
    """Injection onto every second node with the non-equilibrium part
    rescaled by 2 tau_coarse / tau_fine (lettuce/util/utility.py:102-116)."""
    if f_fine.dim() == 3:
        sub = f_fine[:, ::2, ::2]
    elif f_fine.dim() == 4:
        sub = f_fine[:, ::2, ::2, ::2]
    else:
        raise LettuceException("Invalid dimension!")
    sub = sub.contiguous()
    f_eq = flow.equilibrium(flow, rho=flow.rho(sub), u=flow.u(sub))
    return f_eq + 2 * tau_coarse / tau_fine * (sub - f_eq)


def torch_jacobi(f, p, dx, dim, tol_abs=1e-10, max_num_steps=100000):
    """Jacobi iterations for lap p = f on a periodic grid
    (lettuce/util/utility.py:119-156)."""
    error, it = 1.0, 0
    axes = tuple(range(dim))
    while error > tol_abs and it < max_num_steps:
        it += 1
        nb = sum(p.roll(shifts=s, dims=a) for a in axes for s in (1, -1))
        p = (f * dx ** 2 - nb) * (-1 / (2 * dim))
        nb = sum(p.roll(shifts=s, dims=a) for a in axes for s in (1, -1))
        residuum = f - (nb - 2 * dim * p) / dx ** 2
        error = torch.mean(residuum ** 2)
    return p


def pressure_poisson(units, u, rho0, tol_abs=1e-10, max_num_steps=100000):
    """Density from the pressure Poisson equation of a velocity field
    (lettuce/_flow.py:271-320)."""
    dx = units.convert_length_to_pu(1.0)
    u = units.convert_velocity_to_pu(u)
    p = units.convert_density_lu_to_pressure_pu(rho0)
    with torch.no_grad():
        u_mod = torch.zeros_like(u[0])
        dim = u.shape[0]
        for i in range(dim):
            for j in range(dim):
                u_mod -= torch_gradient(torch_gradient(u[i] * u[j], dx)[i], dx)[j]
    p_mod = torch_jacobi(u_mod, p[0], dx, dim=2, tol_abs=tol_abs,
                         max_num_steps=max_num_steps)[None, ...]
    return units.convert_pressure_pu_to_density_lu(p_mod)


def initialize_f_neq(flow):
    """feq - fneq with fneq from 6th-order velocity gradients
    (lettuce/_flow.py:341-367)."""
    rho = flow.rho()
    u = flow.u()
    d = flow.stencil.d
    S = torch.cat([torch_gradient(u[k], dx=1, order=6)[None, ...]
                   for k in range(d)])
    cs = flow.torch_stencil.cs
    Pi_1 = 1.0 * flow.units.relaxation_parameter_lu * rho * S / cs ** 2
    e = flow.torch_stencil.e
    Q = (torch.einsum("ia,ib->iab", e, e)
         - torch.eye(d, device=e.device, dtype=e.dtype) * flow.stencil.cs ** 2)
    Pi_1_Q = torch.einsum("ab...,iab->i...", Pi_1, Q)
    fneq = torch.einsum("i,i...->i...", flow.torch_stencil.w, Pi_1_Q)
    return flow.equilibrium(flow, rho, u) - fneq


def append_axes(array, n):
    return array[(Ellipsis,) + (None,) * n]

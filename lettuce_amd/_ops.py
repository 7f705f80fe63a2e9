"""Tensor-level wrappers around the C-ABI utility entry points.

Each function validates shapes, dtypes and devices on the host before it
hands raw pointers to liblettuce_amd.so, launches on torch's current stream
and raises on any reported error.
"""
from typing import Sequence, Tuple

import torch

from . import _native as N

__all__ = ["layout", "equilibrium", "decode", "encode", "moments_f64",
           "reduce_f64", "enstrophy_f64", "link_force_f64", "STORAGE_OF"]

STORAGE_OF = {torch.float32: N.F32, torch.float64: N.F64,
              torch.float16: N.F16_SHIFTED, torch.bfloat16: N.BF16_SHIFTED}


def layout(f: torch.Tensor, q: int, d: int) -> Tuple[tuple, tuple]:
    """(n[3], stride[4]) of a population tensor [Q, *res] in the kernels'
    (i0, i1, i2) convention: 2D grids are (1, X, Y)."""
    if f.dim() != d + 1 or f.shape[0] != q:
        raise ValueError(f"population tensor must be [{q}, *res] with {d} "
                         f"spatial dims, got {tuple(f.shape)}")
    if f.stride(-1) != 1:
        raise ValueError("the last spatial axis must be contiguous")
    if d == 2:
        n = (1, f.shape[1], f.shape[2])
        stride = (f.stride(0), f.shape[1] * f.stride(1), f.stride(1), 1)
    else:
        n = tuple(f.shape[1:])
        stride = (f.stride(0), f.stride(1), f.stride(2), 1)
    return n, stride


def _n(n):
    return N._I64_3(*n)


def _s(s):
    return N._I64_4(*s)


def _check_field(t: torch.Tensor, shape: Sequence[int], dtype, name):
    N.require_device_tensor(t, name)
    if tuple(t.shape) != tuple(shape) or t.dtype != dtype or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous {dtype} tensor of shape "
                         f"{tuple(shape)}, got {t.dtype} {tuple(t.shape)}")


def equilibrium(stencil, f: torch.Tensor, rho: torch.Tensor, u: torch.Tensor,
                rho0: float = 1.0) -> torch.Tensor:
    """Fill population tensor `f` (any storage) with feq(rho, u)."""
    N.require_device_tensor(f, "f")
    n, s = layout(f, stencil.q, stencil.d)
    res = tuple(f.shape[1:])
    arith = torch.float64 if f.dtype == torch.float64 else torch.float32
    _check_field(rho, res, arith, "rho")
    _check_field(u, (stencil.d,) + res, arith, "u")
    N.check(N.lib().lbm_equilibrium(
        stencil.native_id, STORAGE_OF[f.dtype], _n(n), _s(s), rho.data_ptr(),
        u.data_ptr(), rho0, f.data_ptr(), N.stream_handle(f.device)),
        "lbm_equilibrium")
    return f


def decode(stencil, f: torch.Tensor, rho0: float = 1.0,
           out_dtype=torch.float32) -> torch.Tensor:
    """True populations (contiguous [Q, *res]) from any storage."""
    N.require_device_tensor(f, "f")
    n, s = layout(f, stencil.q, stencil.d)
    out = torch.empty(tuple(f.shape), dtype=out_dtype, device=f.device)
    N.check(N.lib().lbm_decode(
        stencil.native_id, STORAGE_OF[f.dtype], _n(n), _s(s), f.data_ptr(),
        rho0, out.data_ptr(), int(out_dtype == torch.float64),
        N.stream_handle(f.device)), "lbm_decode")
    return out


def encode(stencil, values: torch.Tensor, f: torch.Tensor,
           rho0: float = 1.0) -> torch.Tensor:
    """Write true populations `values` into `f`'s storage format."""
    N.require_device_tensor(f, "f")
    n, s = layout(f, stencil.q, stencil.d)
    if values.dtype not in (torch.float32, torch.float64):
        values = values.to(torch.float32)
    values = values.to(f.device).contiguous()
    _check_field(values, tuple(f.shape), values.dtype, "values")
    N.check(N.lib().lbm_encode(
        stencil.native_id, STORAGE_OF[f.dtype], _n(n), _s(s),
        values.data_ptr(), int(values.dtype == torch.float64), rho0,
        f.data_ptr(), N.stream_handle(f.device)), "lbm_encode")
    return f


def moments_f64(stencil, f: torch.Tensor, rho0: float = 1.0):
    """(rho [*res], u [D, *res]) in FP64."""
    N.require_device_tensor(f, "f")
    n, s = layout(f, stencil.q, stencil.d)
    res = tuple(f.shape[1:])
    rho = torch.empty(res, dtype=torch.float64, device=f.device)
    u = torch.empty((stencil.d,) + res, dtype=torch.float64, device=f.device)
    N.check(N.lib().lbm_moments_f64(
        stencil.native_id, STORAGE_OF[f.dtype], _n(n), _s(s), f.data_ptr(),
        rho0, rho.data_ptr(), u.data_ptr(), N.stream_handle(f.device)),
        "lbm_moments_f64")
    return rho, u


def _workspace(device, n):
    nbytes = N.lib().lbm_reduce_workspace_bytes(_n(n))
    return torch.empty((nbytes + 7) // 8, dtype=torch.float64, device=device)


def reduce_f64(stencil, f: torch.Tensor, rho0: float = 1.0,
               exclude: torch.Tensor = None) -> torch.Tensor:
    """FP64 [sum rho, sum jx, sum jy, sum jz, sum |u|^2/2] over the cells
    where `exclude` (uint8/bool [*res]) is zero."""
    N.require_device_tensor(f, "f")
    n, s = layout(f, stencil.q, stencil.d)
    out = torch.empty(5, dtype=torch.float64, device=f.device)
    mask_ptr = None
    if exclude is not None:
        exclude = exclude.to(device=f.device, dtype=torch.uint8).contiguous()
        _check_field(exclude, tuple(f.shape[1:]), torch.uint8, "exclude")
        mask_ptr = exclude.data_ptr()
    ws = _workspace(f.device, n)
    N.check(N.lib().lbm_reduce_f64(
        stencil.native_id, STORAGE_OF[f.dtype], _n(n), _s(s), f.data_ptr(),
        rho0, mask_ptr, out.data_ptr(), ws.data_ptr(),
        N.stream_handle(f.device)), "lbm_reduce_f64")
    return out


def scan_f64(stencil, f: torch.Tensor, rho0: float = 1.0) -> torch.Tensor:
    """FP64 [NaN populations, cells with non-finite |u|^2, max finite |u|^2]
    over all cells, in one pass over the stored populations."""
    N.require_device_tensor(f, "f")
    n, s = layout(f, stencil.q, stencil.d)
    out = torch.empty(3, dtype=torch.float64, device=f.device)
    ws = _workspace(f.device, n)
    N.check(N.lib().lbm_scan_f64(
        stencil.native_id, STORAGE_OF[f.dtype], _n(n), _s(s), f.data_ptr(), rho0,
        out.data_ptr(), ws.data_ptr(), N.stream_handle(f.device)), "lbm_scan_f64")
    return out


def enstrophy_f64(stencil, u: torch.Tensor) -> torch.Tensor:
    """FP64 sum of |curl u|^2 (lattice units, 6th-order periodic FD)."""
    res = tuple(u.shape[1:])
    _check_field(u, (stencil.d,) + res, torch.float64, "u")
    n = (1,) + res if stencil.d == 2 else res
    out = torch.empty(1, dtype=torch.float64, device=u.device)
    ws = _workspace(u.device, n)
    N.check(N.lib().lbm_enstrophy_f64(
        stencil.native_id, _n(n), u.data_ptr(), out.data_ptr(), ws.data_ptr(),
        N.stream_handle(u.device)), "lbm_enstrophy_f64")
    return out[0]


def link_force_f64(stencil, f: torch.Tensor, links: torch.Tensor,
                   rho0: float = 1.0) -> torch.Tensor:
    """FP64 momentum-exchange force 2 sum f_d(x) e_d over the links
    (int32 [L, 2] = (linear cell, d))."""
    N.require_device_tensor(f, "f")
    n, s = layout(f, stencil.q, stencil.d)
    ncell = 1
    for k in f.shape[1:]:
        ncell *= k
    links = links.to(device=f.device, dtype=torch.int32).contiguous()
    if links.dim() != 2 or links.shape[1] != 2:
        raise ValueError("links must be [L, 2]")
    nl = links.shape[0]
    if nl:
        lo = links.min(dim=0).values.tolist()
        hi = links.max(dim=0).values.tolist()
        if lo[0] < 0 or hi[0] >= ncell or lo[1] < 0 or hi[1] >= stencil.q:
            raise ValueError("link cell or direction index out of range")
    out = torch.empty(3, dtype=torch.float64, device=f.device)
    ws = _workspace(f.device, n)
    N.check(N.lib().lbm_link_force_f64(
        stencil.native_id, STORAGE_OF[f.dtype], _n(n), _s(s), f.data_ptr(),
        rho0, links.data_ptr() if nl else None, nl, out.data_ptr(),
        ws.data_ptr(), N.stream_handle(f.device)), "lbm_link_force_f64")
    return out

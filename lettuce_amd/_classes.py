"""Boundary masks and the cell-class map.

lettuce describes boundaries with two integer masks (lettuce/_simulation.py:
100-146):

* ``no_collision_mask`` (ncm, uint8 [*res]) holds, per cell, the index of the
  transformer applied in place of a plain collision: pre-boundaries
  0..P-1, the collision P, post-boundaries P+1...; later boundaries
  overwrite earlier ones;
* ``no_streaming_mask`` (nsm, uint8 [Q, *res]) starts at P and is OR-ed with
  each boundary's boolean mask; the torch step keeps population q at x
  where ``nsm[q](x) == 1`` (lettuce/_simulation.py:254).

The kernels instead read one uint8 class id per cell.  A class is one
distinct (ncm value, streaming bits, halfway-bounce-back flags) triple; the
class table holds the operator and the per-population streaming bits.  The
map is built so that ncm and the ``nsm == 1`` predicate can be recovered
from it bit for bit (``expand_classes``).

Everything here is host-side integer work on torch tensors of any device.
"""
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _native as N

__all__ = ["build_masks", "stream_bits", "near_solid", "CellClasses",
           "build_cell_classes", "expand_classes", "OpSpec"]


def build_masks(q: int, resolution: Sequence[int], pre_boundaries: list,
                post_boundaries: list, streaming_boundaries: list, context):
    """ncm / nsm exactly as lettuce's Simulation builds them
    (lettuce/_simulation.py:100-146), extended by the post-streaming
    boundaries of EbbSimulation
    (examples/.../simulation/ebb_simulation.py:29-54), which take the indices
    after the post-boundaries.  Returns (None, None) without boundaries."""
    pre_boundaries = list(pre_boundaries or [])
    post_boundaries = list(post_boundaries or [])
    streaming_boundaries = list(streaming_boundaries or [])
    if not (pre_boundaries or post_boundaries or streaming_boundaries):
        return None, None
    ci = len(pre_boundaries)
    res = list(resolution)
    ncm = context.full_tensor(res, ci, dtype=torch.uint8)
    nsm = context.full_tensor([q] + res, ci, dtype=torch.uint8)
    ordered = ([(i, b) for i, b in enumerate(pre_boundaries)]
               + [(i, b) for i, b in enumerate(post_boundaries, start=ci + 1)]
               + [(i, b) for i, b in enumerate(
                   streaming_boundaries, start=ci + 1 + len(post_boundaries))])
    for i, boundary in ordered:
        m = boundary.make_no_collision_mask(res, context=context)
        if m is not None:
            ncm[m.to(device=ncm.device, dtype=torch.bool)] = i
        s = boundary.make_no_streaming_mask([q] + res, context=context)
        if s is not None:
            nsm |= s.to(device=nsm.device, dtype=torch.bool)
    return ncm, nsm


def stream_bits(nsm: torch.Tensor) -> torch.Tensor:
    """int64 [*res]: bit q set where population q is kept (nsm[q] == 1)."""
    q = nsm.shape[0]
    bits = torch.zeros(nsm.shape[1:], dtype=torch.int64, device=nsm.device)
    for k in range(q):
        bits |= (nsm[k] == 1).to(torch.int64) << k
    return bits


def near_solid(solid: torch.Tensor, e: Sequence[Sequence[int]]) -> torch.Tensor:
    """bool [*res]: some neighbour x - e_q (periodic) is solid."""
    d = solid.dim()
    near = torch.zeros_like(solid)
    for ei in e:
        if not any(ei):
            continue
        near |= torch.roll(solid, shifts=tuple(ei), dims=tuple(range(d)))
    return near


@dataclass
class OpSpec:
    """What one transformer index does on its cells."""
    op: int = N.OP_KEEP
    rho: float = 1.0
    u: Sequence[float] = (0.0, 0.0, 0.0)
    direction: Sequence[int] = (0, 0, 0)   # OUTLET: storage-axis offset
    bc_field: Optional[torch.Tensor] = None  # EQUILIBRIUM: [4, *res] values
    column_mask: Optional[torch.Tensor] = None  # OUTLET: its column (bool)
    lattice_direction: Sequence[int] = ()       # OUTLET: lettuce direction


@dataclass
class CellClasses:
    cell_class: torch.Tensor              # uint8 [*res]
    table: List[N.LbmCellClass]
    keys: torch.Tensor                    # int64 [n_classes] (ncm, bits, flags)
    ncm_of_class: List[int] = field(default_factory=list)
    bc_field: Optional[torch.Tensor] = None

    @property
    def n_classes(self) -> int:
        return len(self.table)

    def table_tensor(self, device) -> torch.Tensor:
        raw = b"".join(bytes(c) for c in self.table)
        buf = torch.frombuffer(bytearray(raw), dtype=torch.uint8)
        return buf.to(device)


def build_cell_classes(ncm: torch.Tensor, nsm: torch.Tensor,
                       ops: dict, e: Sequence[Sequence[int]],
                       solid: Optional[torch.Tensor] = None) -> CellClasses:
    """Class map and table from lettuce's masks.

    ``ops`` maps a transformer index to its OpSpec (indices without an entry
    keep their populations, like cells whose ncm matches no transformer in
    lettuce's _collide).  ``solid`` marks halfway-bounce-back solid cells.
    """
    q = nsm.shape[0]
    if q > 27:
        raise ValueError("at most 27 populations")
    bits = stream_bits(nsm)
    if solid is None:
        solid = torch.zeros_like(ncm, dtype=torch.bool)
    solid = solid.to(device=ncm.device, dtype=torch.bool)
    near = near_solid(solid, e) & ~solid
    flags = solid.to(torch.int64) * N.CLASS_SOLID + near.to(torch.int64) * N.CLASS_NEAR_SOLID
    key = (ncm.to(torch.int64) << 40) | (bits << 8) | flags
    uniq, inverse = torch.unique(key.flatten(), return_inverse=True)
    if uniq.numel() > 256:
        raise NotImplementedError(
            f"{uniq.numel()} distinct cell classes; at most 256 fit a uint8 map")
    cell_class = inverse.reshape(ncm.shape).to(torch.uint8)

    bc_field = None
    field_ops = {i: s for i, s in ops.items() if s.bc_field is not None}
    if field_ops:
        res = tuple(ncm.shape)
        bc_field = torch.zeros((4,) + res, dtype=torch.float64, device=ncm.device)
        for i, spec in field_ops.items():
            where = ncm == i
            vals = spec.bc_field.to(device=ncm.device, dtype=torch.float64)
            vals = vals.expand((4,) + res)
            bc_field[:, where] = vals[:, where]

    table, ncm_of_class = [], []
    for k in uniq.tolist():
        idx = k >> 40
        cbits = (k >> 8) & ((1 << 32) - 1)
        cflags = k & 0xFF
        spec = ops.get(idx, OpSpec())
        entry = N.LbmCellClass()
        entry.op = spec.op
        entry.flags = cflags | (N.CLASS_BC_FIELD if spec.bc_field is not None else 0)
        for a in range(3):
            entry.dir[a] = int(spec.direction[a])
        entry.no_stream = cbits
        entry.rho = float(spec.rho)
        for a in range(3):
            entry.u[a] = float(spec.u[a]) if a < len(spec.u) else 0.0
        table.append(entry)
        ncm_of_class.append(idx)
    return CellClasses(cell_class=cell_class, table=table, keys=uniq,
                       ncm_of_class=ncm_of_class, bc_field=bc_field)


def expand_classes(cc: CellClasses, q: int):
    """Recover (ncm, nsm == 1 as uint8, solid) from the class map; the
    inverse of build_cell_classes, used to prove the map exact."""
    lut_ncm = torch.tensor(cc.ncm_of_class, dtype=torch.uint8)
    lut_bits = torch.tensor([c.no_stream for c in cc.table], dtype=torch.int64)
    lut_solid = torch.tensor([bool(c.flags & N.CLASS_SOLID) for c in cc.table])
    idx = cc.cell_class.cpu().long()
    ncm = lut_ncm[idx]
    bits = lut_bits[idx]
    nsm = torch.stack([((bits >> k) & 1).to(torch.uint8) for k in range(q)])
    return ncm, nsm, lut_solid[idx]


def halfway_links(solid: torch.Tensor, e: Sequence[Sequence[int]],
                  opposite: Sequence[int]) -> torch.Tensor:
    """int32 [L, 2] (linear fluid cell, d) for every link from a fluid cell x
    to a solid neighbour x + e_d (periodic), ordered by (cell, d).

    The same links as lettuce's halfway bounce-back f_index
    (examples/.../boundary/halfway_bounce_back_boundary.py:71-160), whose
    rows are [opposite(i), fluid coords] for solid p and fluid p + e_i,
    i.e. d = opposite(i) points from the fluid cell into the solid."""
    solid = solid.to(torch.bool).cpu()
    dims = tuple(range(solid.dim()))
    fluid = ~solid
    rows = []
    for d, ed in enumerate(e):
        if not any(ed):
            continue
        # solid at x + e_d  <=>  roll(solid, -e_d)[x]
        hit = fluid & torch.roll(solid, shifts=tuple(-c for c in ed), dims=dims)
        cells = torch.nonzero(hit.flatten(), as_tuple=False).flatten()
        if cells.numel():
            rows.append(torch.stack([cells, torch.full_like(cells, d)], dim=1))
    if not rows:
        return torch.zeros((0, 2), dtype=torch.int32)
    links = torch.cat(rows)
    order = torch.argsort(links[:, 0] * 64 + links[:, 1])
    return links[order].to(torch.int32)

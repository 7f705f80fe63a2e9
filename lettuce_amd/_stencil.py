"""Lattice stencils with lettuce's direction order.

The velocity sets, weights and opposite maps are lettuce's
(lettuce/ext/_stencil/d2q9.py:8-10, d3q19.py:8-13, d3q27.py:8-12;
``Stencil``/``TorchStencil`` at lettuce/_stencil.py:14-50), so population q
of a lettuce_amd flow is population q of the equivalent lettuce flow.  The
same tables are compiled into the kernels (csrc/lbm_device.h).
"""
from functools import cached_property
from typing import List

import numpy as np
import torch

__all__ = ["Stencil", "TorchStencil", "D2Q9", "D3Q19", "D3Q27"]


class Stencil:
    e: List[List[int]]
    w: List[float]
    opposite: List[int]
    cs: float = 1 / np.sqrt(3.0)
    native_id: int = -1

    @cached_property
    def d(self) -> int:
        return len(self.e[0])

    @cached_property
    def q(self) -> int:
        return len(self.e)

    def __eq__(self, other):
        return type(self) is type(other)

    def __hash__(self):
        return hash(type(self).__name__)


def _opposites(e):
    return [e.index([-c for c in ei]) for ei in e]


class D2Q9(Stencil):
    native_id = 0

    def __init__(self):
        self.e = [[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1],
                  [1, 1], [-1, 1], [-1, -1], [1, -1]]
        self.w = [4.0 / 9.0] + [1.0 / 9.0] * 4 + [1.0 / 36.0] * 4
        self.opposite = _opposites(self.e)


class D3Q19(Stencil):
    native_id = 1

    def __init__(self):
        self.e = [[0, 0, 0],
                  [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1],
                  [0, 0, -1],
                  [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1], [1, 0, 1],
                  [-1, 0, -1], [1, 0, -1], [-1, 0, 1], [1, 1, 0], [-1, -1, 0],
                  [1, -1, 0], [-1, 1, 0]]
        self.w = [1.0 / 3.0] + [1.0 / 18.0] * 6 + [1.0 / 36.0] * 12
        self.opposite = _opposites(self.e)


class D3Q27(Stencil):
    native_id = 2

    def __init__(self):
        axes = [[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0],
                [0, 0, 1], [0, 0, -1]]
        edges = [[0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1], [1, 0, 1],
                 [-1, 0, -1], [1, 0, -1], [-1, 0, 1], [1, 1, 0], [-1, -1, 0],
                 [1, -1, 0], [-1, 1, 0]]
        corners = [[1, 1, 1], [-1, -1, -1], [1, 1, -1], [-1, -1, 1],
                   [1, -1, 1], [-1, 1, -1], [1, -1, -1], [-1, 1, 1]]
        self.e = axes + edges + corners
        self.w = ([8.0 / 27.0] + [2.0 / 27.0] * 6 + [1.0 / 54.0] * 12
                  + [1.0 / 216.0] * 8)
        self.opposite = _opposites(self.e)


class TorchStencil:
    cs: float = 1 / np.sqrt(3.0)

    def __init__(self, stencil: Stencil, context):
        self.e = context.convert_to_tensor(stencil.e)
        self.w = context.convert_to_tensor(stencil.w)
        self.opposite = context.convert_to_tensor(stencil.opposite,
                                                  dtype=torch.int64)

    @cached_property
    def d(self) -> int:
        return self.e.shape[1]

    @cached_property
    def q(self) -> int:
        return self.e.shape[0]

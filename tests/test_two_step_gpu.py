"""The two-step kernel (csrc/lbm_tb2.hip, LETTUCE_AMD_TB2=1) gives the same
populations as single fused steps bit for bit: same per-cell arithmetic,
with the intermediate step rounded to the storage format as a stored step
is.  Cases cover several z and y tiles, a partial last x chunk, a single y
tile that wraps onto itself, both memory layouts, odd step counts and an
odd number of pairs (which runs one pair as two single steps)."""
import numpy as np
import pytest
import torch

import lettuce_amd as lt
from oracle import lettuce_torch as O
from tests.test_oracle import _Grid

pytestmark = pytest.mark.gpu

F32, F64 = torch.float32, torch.float64


def _f0(res, seed=5):
    g = torch.Generator().manual_seed(seed)
    oflow = _Grid(O.D3Q19(), res, F64)
    u = 0.05 * (torch.rand([3] + res, generator=g, dtype=F64) - 0.5)
    f0 = oflow.equilibrium(oflow, rho=torch.ones([1] + res, dtype=F64), u=u)
    return (f0 * (1 + 0.02 * torch.rand(f0.shape, generator=g, dtype=F64))).numpy()


class _Flow(lt.ExtFlow):
    def __init__(self, context, f0):
        self._f0 = f0
        super().__init__(context, list(f0.shape[1:]), 100.0, 0.1, lt.D3Q19())

    def make_resolution(self, resolution, stencil=None):
        return list(resolution)

    def make_units(self, reynolds_number, mach_number, resolution):
        return lt.UnitConversion(reynolds_number, mach_number,
                                 characteristic_length_lu=resolution[0])

    def initial_pu(self):
        raise NotImplementedError

    def initialize(self):
        self.f = torch.from_numpy(self._f0)


def _run(f0, storage, layout, coll, steps):
    dtype = F32
    c = lt.Context("cuda:0", dtype=dtype, storage=storage, layout=layout)
    flow = _Flow(c, f0.astype(np.float32))
    sim = lt.Simulation(flow, coll, [], lt.StreamingStrategy.PRE_STREAMING)
    sim(steps)
    torch.cuda.synchronize()
    return flow._state.detach().cpu().contiguous().view(-1).view(torch.uint8).numpy()


@pytest.mark.parametrize("res,planes", [([10, 16, 128], 4), ([7, 8, 64], 16), ([12, 24, 64], 5)])
@pytest.mark.parametrize("storage,layout", [("f32", "soa"), ("f32", "rows"),
                                            ("f16_shifted", "soa"), ("bf16_shifted", "rows")])
@pytest.mark.parametrize("coll", ["bgk", "trt", "none"])
def test_two_step_kernel_is_bitwise_two_single_steps(res, planes, storage, layout, coll,
                                                     monkeypatch):
    f0 = _f0(res)
    make = {"bgk": lambda: lt.BGKCollision(0.62), "trt": lambda: lt.TRTCollision(0.7, 0.9),
            "none": lambda: lt.NoCollision()}[coll]
    for steps in (2, 5, 6, 8):
        monkeypatch.setenv("LETTUCE_AMD_TB2", "0")
        ref = _run(f0, storage, layout, make(), steps)
        monkeypatch.setenv("LETTUCE_AMD_TB2", "1")
        monkeypatch.setenv("LETTUCE_AMD_TB2_PLANES", str(planes))
        got = _run(f0, storage, layout, make(), steps)
        assert np.array_equal(got, ref), (steps, int((got != ref).sum()))

"""The two-step kernel (csrc/lbm_tb2.hip) against single fused steps: bit for
bit without collision (the streaming, the LDS rings and the z-edge cells),
and to the rounding of the collision arithmetic with BGK and TRT (the
compiler may contract the same formulas into FMAs differently in the two
kernels).  Cases cover several z and y tiles, a partial last x chunk, a
single y tile that wraps onto itself, grids of one to three x planes, both memory layouts, the three storage
formats, odd step counts and an odd number of pairs (which runs one pair as
two single steps); and the kernel against the oracle's BGK."""
import numpy as np
import pytest
import torch

import lettuce_amd as lt
from oracle import lettuce_torch as O
from tests.test_oracle import _Grid

pytestmark = pytest.mark.gpu

F32, F64 = torch.float32, torch.float64


def _f0(res, seed=5, stencil=O.D3Q19):
    g = torch.Generator().manual_seed(seed)
    oflow = _Grid(stencil(), res, F64)
    u = 0.05 * (torch.rand([3] + res, generator=g, dtype=F64) - 0.5)
    f0 = oflow.equilibrium(oflow, rho=torch.ones([1] + res, dtype=F64), u=u)
    return (f0 * (1 + 0.02 * torch.rand(f0.shape, generator=g, dtype=F64))).numpy()


class _Flow(lt.ExtFlow):
    def __init__(self, context, f0):
        self._f0 = f0
        stencil = lt.D3Q19() if f0.shape[0] == 19 else lt.D3Q27()
        super().__init__(context, list(f0.shape[1:]), 100.0, 0.1, stencil)

    def make_resolution(self, resolution, stencil=None):
        return list(resolution)

    def make_units(self, reynolds_number, mach_number, resolution):
        return lt.UnitConversion(reynolds_number, mach_number,
                                 characteristic_length_lu=resolution[0])

    def initial_pu(self):
        raise NotImplementedError

    def initialize(self):
        self.f = torch.from_numpy(self._f0)


def _run(f0, storage, layout, coll, steps, pairs=None):
    c = lt.Context("cuda:0", dtype=F32, storage=storage, layout=layout)
    flow = _Flow(c, f0.astype(np.float32))
    sim = lt.Simulation(flow, coll, [], lt.StreamingStrategy.PRE_STREAMING)
    if pairs is not None:
        assert sim._collide_and_stream.plan(steps) == (pairs, steps - 2 * pairs)
    sim(steps)
    torch.cuda.synchronize()
    return (flow._state.detach().cpu().contiguous().view(-1).view(torch.uint8).numpy(),
            flow.f.double().cpu().numpy())


# largest difference of two decoded populations allowed for a rounding
# difference in the collision: a few units in the last place of the stored
# format at |f| <= 0.4 (FP32) or at the stored deviation |f - w| <= 0.05
ULPS = {"f32": 4 * 0.4 * 2.0 ** -23, "f16_shifted": 4 * 0.05 * 2.0 ** -10,
        "bf16_shifted": 4 * 0.05 * 2.0 ** -7}


@pytest.mark.parametrize("res,planes", [([10, 16, 128], 4), ([7, 8, 64], 16), ([12, 24, 64], 5)])
@pytest.mark.parametrize("storage,layout", [("f32", "soa"), ("f32", "rows"),
                                            ("f16_shifted", "soa"), ("bf16_shifted", "rows")])
@pytest.mark.parametrize("coll", ["bgk", "trt", "none"])
@pytest.mark.parametrize("barriers", ["1", "2"])
def test_two_step_kernel_matches_two_single_steps(res, planes, storage, layout, coll, barriers,
                                                     knobs):
    knobs(two_step_barriers=barriers)
    f0 = _f0(res)
    make = {"bgk": lambda: lt.BGKCollision(0.62), "trt": lambda: lt.TRTCollision(0.7, 0.9),
            "none": lambda: lt.NoCollision()}[coll]
    for steps in (2, 5, 6, 8):
        knobs(two_step="0")
        ref_bits, ref = _run(f0, storage, layout, make(), steps, pairs=0)
        knobs(two_step="2")
        knobs(two_step_planes=str(planes))
        pairs = steps // 2 - (steps // 2) % 2
        got_bits, got = _run(f0, storage, layout, make(), steps, pairs=pairs)
        if coll == "none":
            assert np.array_equal(got_bits, ref_bits), (steps, int((got_bits != ref_bits).sum()))
        else:
            err = np.abs(got - ref).max()
            assert err <= steps * ULPS[storage], (steps, err)


@pytest.mark.parametrize("layout", ["soa", "rows"])
def test_two_step_kernel_matches_the_oracle(layout, knobs):
    """FP32 BGK with the default two-step path, 40 steps, against the oracle's
    BGK (lettuce's torch step) to 1e-5"""
    res = [8, 16, 64]
    f0 = _f0(res, seed=2)
    knobs(two_step="1")
    knobs(two_step_planes="3")
    _, got = _run(f0, "f32", layout, lt.BGKCollision(0.58), 40, pairs=20)
    oflow = _Grid(O.D3Q19(), res, F32, f0.astype(np.float32))
    O.Simulation(oflow, O.BGKCollision(0.58), [], O.StreamingStrategy.PRE_STREAMING)(40)
    ref = oflow.f.double().numpy()
    assert np.abs(got - ref).max() / np.abs(ref).max() <= 1e-5


def test_two_step_kernel_applies_only_where_it_can(knobs):
    """D3Q19 in every storage format by default, D3Q27 in FP32 only; no
    two-step launch with the push kernel or extents that do not tile; a
    post-streaming run of 8 steps splits into one collide-only launch, 6
    pull steps (2 two-step launches + 2 single steps, an even number of
    pairs) and one double-streaming launch"""
    knobs(two_step=-1)
    f0 = _f0([4, 16, 64])
    for storage, pairs in (("f32", 4), ("f16_shifted", 4), ("bf16_shifted", 4)):
        c = lt.Context("cuda:0", dtype=F32, storage=storage)
        sim = lt.Simulation(_Flow(c, f0.astype(np.float32)), lt.BGKCollision(0.6), [],
                            lt.StreamingStrategy.PRE_STREAMING)
        assert sim._collide_and_stream.plan(9) == (pairs, 9 - 2 * pairs)
    c = lt.Context("cuda:0", dtype=F32)
    sim = lt.Simulation(_Flow(c, f0.astype(np.float32)), lt.BGKCollision(0.6), [],
                        lt.StreamingStrategy.POST_STREAMING)
    assert sim._collide_and_stream.plan(8) == (2, 4)
    assert sim._collide_and_stream.plan(1) == (0, 1)
    knobs(post_split=0)
    assert sim._collide_and_stream.plan(8) == (0, 8)
    knobs(post_split=-1)
    f1 = _f0([4, 12, 64])
    sim = lt.Simulation(_Flow(c, f1.astype(np.float32)), lt.BGKCollision(0.6), [],
                        lt.StreamingStrategy.PRE_STREAMING)
    assert sim._collide_and_stream.plan(8) == (0, 8)
    f2 = _f0([4, 16, 64], stencil=O.D3Q27)
    sim = lt.Simulation(_Flow(c, f2.astype(np.float32)), lt.KBCCollision(), [],
                        lt.StreamingStrategy.PRE_STREAMING)
    assert sim._collide_and_stream.plan(8) == (4, 0)
    c16 = lt.Context("cuda:0", dtype=F32, storage="bf16_shifted")
    sim = lt.Simulation(_Flow(c16, f2.astype(np.float32)), lt.KBCCollision(), [],
                        lt.StreamingStrategy.PRE_STREAMING)
    assert sim._collide_and_stream.plan(8) == (0, 8)


@pytest.mark.parametrize("coll", ["kbc", "bgk"])
@pytest.mark.parametrize("storage,layout", [("f32", "soa"), ("f32", "rows"), ("f16_shifted", "soa")])
def test_two_step_kernel_d3q27(coll, storage, layout, knobs):
    """D3Q27 (KBC, BGK): two-step launches against single steps"""
    f0 = _f0([6, 16, 64], stencil=O.D3Q27)
    make = {"kbc": lambda: lt.KBCCollision(), "bgk": lambda: lt.BGKCollision(0.62)}[coll]
    for steps in (4, 9):
        knobs(two_step="0")
        _, ref = _run(f0, storage, layout, make(), steps, pairs=0)
        knobs(two_step="2")
        knobs(two_step_planes="4")
        pairs = steps // 2 - (steps // 2) % 2
        _, got = _run(f0, storage, layout, make(), steps, pairs=pairs)
        err = np.abs(got - ref).max()
        # KBC's gamma is a ratio of small sums: a rounding difference in the
        # collision moves it further than in BGK
        tol = steps * ULPS[storage] * (50 if coll == "kbc" else 1)
        assert err <= tol, (steps, err)


@pytest.mark.parametrize("n0", [1, 2, 3])
@pytest.mark.parametrize("stencil,coll", [(O.D3Q19, "none"), (O.D3Q27, "none"), (O.D3Q27, "kbc")])
def test_two_step_kernel_on_thin_x_extents(stencil, coll, n0, knobs):
    """Grids of one to three x planes, one chunk each: the x neighbours wrap
    onto the chunk itself, and D3Q27's load of the plane after the chunk
    stays inside the array"""
    f0 = _f0([n0, 8, 64], stencil=stencil)
    make = {"kbc": lambda: lt.KBCCollision(), "none": lambda: lt.NoCollision()}[coll]
    for steps in (4, 5):
        knobs(two_step="0")
        ref_bits, ref = _run(f0, "f32", "soa", make(), steps, pairs=0)
        knobs(two_step="2")
        pairs = steps // 2 - (steps // 2) % 2
        got_bits, got = _run(f0, "f32", "soa", make(), steps, pairs=pairs)
        if coll == "none":
            assert np.array_equal(got_bits, ref_bits), (steps, int((got_bits != ref_bits).sum()))
        else:
            err = np.abs(got - ref).max()
            assert err <= steps * ULPS["f32"] * 50, (steps, err)


@pytest.mark.parametrize("stencil,coll,storage,layout", [
    (O.D3Q19, "none", "f32", "soa"), (O.D3Q19, "bgk", "f32", "rows"),
    (O.D3Q19, "bgk", "bf16_shifted", "rows"), (O.D3Q27, "kbc", "f32", "soa")])
def test_wide_two_step_kernel_matches_the_narrow_one(stencil, coll, storage, layout, knobs):
    """the wide variant (each cell at a 64-bit plane offset plus a 32-bit
    offset inside the plane, which grids of more than 2^31 elements use)
    forced on a small grid, against the narrow variant (one 32-bit offset):
    bit for bit without collision, else to the rounding of the collision"""
    f0 = _f0([10, 16, 128], stencil=stencil)
    make = {"bgk": lambda: lt.BGKCollision(0.62), "kbc": lambda: lt.KBCCollision(),
            "none": lambda: lt.NoCollision()}[coll]
    knobs(two_step="2")
    knobs(two_step_planes="4")
    out = {}
    for wide in ("0", "1"):
        knobs(wide_offsets=wide)
        out[wide] = _run(f0, storage, layout, make(), 8, pairs=4)
    if coll == "none":
        assert np.array_equal(out["1"][0], out["0"][0])
    else:
        err = np.abs(out["1"][1] - out["0"][1]).max()
        assert err <= 8 * ULPS[storage] * (50 if coll == "kbc" else 1), err

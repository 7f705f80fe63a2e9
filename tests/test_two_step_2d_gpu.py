"""The two-step kernel for 2D grids with boundaries (csrc/lbm_tb2d.hip):
D2Q9 channels past a cylinder with every boundary rule the masked step has
(equilibrium inlet, equilibrium and anti-bounce-back outlets, halfway and
fullway bounce-back, walls) against single masked steps, to the rounding of
the collision and boundary arithmetic (the compiler may contract the same
formulas into FMAs differently in the two kernels); and two-step runs
against the oracle (lettuce's torch step).  Cases cover a partial last row
chunk, chunks of one to sixteen rows, rows of 64 to 1024 cells (two rows
per iteration up to 512, one above), the storage formats, and lettuce's
default post-streaming order (whose pull steps run on the same kernel)."""
import numpy as np
import pytest
import torch

import lettuce_amd as lt
from oracle import lettuce_torch as O

pytestmark = pytest.mark.gpu

F32, F64 = torch.float32, torch.float64
PRE, POST = lt.StreamingStrategy.PRE_STREAMING, lt.StreamingStrategy.POST_STREAMING

# largest difference of two decoded populations allowed per step for a
# rounding difference: a few units in the last place of the stored format
# at |f| <= 0.5 (FP32) or at the stored deviation |f - w| <= 0.1
ULPS = {"f32": 4 * 0.5 * 2.0 ** -23, "f16_shifted": 4 * 0.1 * 2.0 ** -10,
        "bf16_shifted": 4 * 0.1 * 2.0 ** -7}
COLLISIONS = {"bgk": lambda tau: lt.BGKCollision(tau), "kbc": lambda tau: lt.KBCCollision(),
              "trt": lambda tau: lt.TRTCollision(tau, 0.8), "none": lambda tau: lt.NoCollision()}


def _mask(nx, ny):
    yy, xx = np.meshgrid(np.arange(ny), np.arange(nx))
    return (xx - nx / 4) ** 2 + (yy - ny / 2 + 0.5) ** 2 < (min(nx, ny) / 6) ** 2


def _channel(storage, nx, ny, outlet, solid, walls=True, dtype=F32):
    ctx = lt.Context("cuda:0", dtype=dtype, storage=storage)
    flow = lt.Obstacle(ctx, [nx, ny], 100, 0.1, domain_length_x=nx / 10, stencil=lt.D2Q9(),
                       outlet=outlet, solid=solid, walls=walls)
    flow.mask = _mask(nx, ny)
    flow.initialize()
    # a perturbed state, so that every population and rule contributes
    g = torch.Generator().manual_seed(nx * 1000 + ny)
    f = flow.f.double().cpu()
    f = f * (1 + 0.02 * (torch.rand(f.shape, generator=g, dtype=F64) - 0.5))
    flow.f = f.to(dtype)
    return flow


def _run(storage, nx, ny, outlet, solid, coll, steps, strategy, two_step, knobs, planes=0):
    knobs(two_step=1 if two_step else 0, two_step_planes=planes)
    # halfway walls with the equilibrium outlet (an outlet's neighbour
    # column holds fluid or solid cells only, not fullway bounce-back walls
    # or an anti-bounce-back outlet's walls)
    flow = _channel(storage, nx, ny, outlet, solid,
                    walls=outlet == "equilibrium" and solid == "halfway")
    sim = lt.Simulation(flow, COLLISIONS[coll](flow.units.relaxation_parameter_lu), [], strategy)
    plan = sim._collide_and_stream.plan(steps)
    sim(steps)
    torch.cuda.synchronize()
    return flow.f.double().cpu().numpy(), plan


CASES = [
    # storage, nx, ny, outlet, solid, collision, rows per block (0: default)
    ("f32", 100, 64, "equilibrium", "halfway", "bgk", 0),
    ("f32", 96, 128, "antibounceback", "fullway", "bgk", 0),
    ("f32", 70, 64, "equilibrium", "halfway", "kbc", 3),
    ("f32", 64, 192, "antibounceback", "halfway", "trt", 16),
    ("f32", 48, 64, "equilibrium", "fullway", "none", 1),
    ("f32", 130, 512, "equilibrium", "halfway", "bgk", 0),
    ("f16_shifted", 100, 64, "equilibrium", "halfway", "kbc", 0),
    ("bf16_shifted", 100, 128, "antibounceback", "halfway", "bgk", 5),
    # rows of more than 512 cells: one row per iteration
    ("f32", 40, 640, "equilibrium", "halfway", "bgk", 0),
    ("f16_shifted", 36, 1024, "antibounceback", "fullway", "kbc", 7),
]


@pytest.mark.parametrize("storage,nx,ny,outlet,solid,coll,planes", CASES)
@pytest.mark.parametrize("strategy", [PRE, POST])
def test_two_step_2d_matches_single_steps(storage, nx, ny, outlet, solid, coll, planes, strategy,
                                          knobs):
    for steps in (2, 5, 8):
        ref, plan0 = _run(storage, nx, ny, outlet, solid, coll, steps, strategy, False, knobs)
        assert plan0 == (0, steps)
        got, plan = _run(storage, nx, ny, outlet, solid, coll, steps, strategy, True, knobs,
                         planes)
        pull = steps - 2 if strategy == POST else steps
        pairs = pull // 2 - (pull // 2) % 2 if pull >= 2 else 0
        assert plan == (pairs, steps - 2 * pairs), (steps, plan)
        err = np.abs(got - ref).max()
        tol = steps * ULPS[storage] * (50 if coll == "kbc" else 1)
        assert err <= tol, (steps, err, tol)


def test_two_step_2d_is_used_for_the_channel_workload():
    """2048 x 512 FP32 channel with BGK: every pair of steps on the kernel"""
    flow = _channel("f32", 2048, 512, "equilibrium", "halfway")
    sim = lt.Simulation(flow, lt.BGKCollision(flow.units.relaxation_parameter_lu), [], PRE)
    assert sim._collide_and_stream.plan(100) == (50, 0)
    assert sim._collide_and_stream.plan(6) == (2, 2)


def test_grids_the_kernel_does_not_take_run_single_steps():
    """rows that are not a multiple of 64 cells, FP64 storage: single steps"""
    for storage, dtype, ny in (("f32", F32, 40), (None, F64, 64)):
        ctx = lt.Context("cuda:0", dtype=dtype, storage=storage)
        flow = lt.Obstacle(ctx, [64, ny], 100, 0.1, domain_length_x=6.4, stencil=lt.D2Q9(),
                           outlet="equilibrium", solid="halfway", walls=True)
        flow.mask = _mask(64, ny)
        flow.initialize()
        sim = lt.Simulation(flow, lt.BGKCollision(flow.units.relaxation_parameter_lu), [], PRE)
        assert sim._collide_and_stream.plan(8) == (0, 8)


@pytest.mark.parametrize("strategy", ["PRE_STREAMING", "POST_STREAMING"])
def test_two_step_2d_channel_vs_oracle(strategy):
    """FP32 channel with a cylinder (equilibrium inlet and outlet, halfway
    bounce-back on the cylinder and the walls), 128 x 64, 200 steps on the
    two-step kernel, against the oracle in FP64: 1e-5 relative"""
    nx, ny = 128, 64
    mask = _mask(nx, ny)
    oflow = O.ObstacleChannel(O.Context("cpu", F64), [nx, ny], 100, 0.1,
                              domain_length_x=nx / 10.0, mask=mask, stencil=O.D2Q9())
    f0 = oflow.f.numpy().copy()
    osim = O.Simulation(oflow, O.BGKCollision(oflow.units.relaxation_parameter_lu), [],
                        getattr(O.StreamingStrategy, strategy))
    flow = lt.Obstacle(lt.Context("cuda:0", dtype=F32), [nx, ny], 100, 0.1,
                       domain_length_x=nx / 10.0, stencil=lt.D2Q9(), outlet="equilibrium",
                       solid="halfway", walls=True)
    flow.mask = mask
    flow.f = torch.from_numpy(f0).to(F32)
    sim = lt.Simulation(flow, lt.BGKCollision(flow.units.relaxation_parameter_lu), [],
                        getattr(lt.StreamingStrategy, strategy))
    assert sim._collide_and_stream.plan(200)[0] > 0
    osim(200)
    sim(200)
    out = flow.f.double().cpu().numpy()
    ref = oflow.f.numpy()
    assert np.abs(out - ref).max() / np.abs(ref).max() <= 1e-5


class _OracleObstacle(O.Flow):
    """the oracle flow with lt.Obstacle's units and lettuce's default
    Obstacle boundaries (equilibrium inlet, anti-bounce-back outlet,
    fullway bounce-back)"""

    def __init__(self, res, mask, f0, length_x):
        units = O.UnitConversion(reynolds_number=100, mach_number=0.1,
                                 characteristic_length_lu=res[0] / length_x,
                                 characteristic_length_pu=1, characteristic_velocity_pu=1)
        super().__init__(O.Context("cpu", F64), res, units, O.D2Q9())
        self.f = torch.as_tensor(f0).to(F64).clone()
        x = torch.meshgrid(*[units.convert_length_to_pu(torch.arange(n)) for n in res],
                           indexing="ij")[0]
        self.post_boundaries = [
            O.EquilibriumBoundaryPU(self.context, self, torch.abs(x) < 1e-6,
                                    units.characteristic_velocity_pu * torch.eye(2)[0]),
            O.AntiBounceBackOutlet([1, 0], self),
            O.BounceBackBoundary(torch.as_tensor(mask)),
        ]


def test_two_step_2d_default_obstacle_vs_oracle():
    """lettuce's default Obstacle (anti-bounce-back outlet, fullway
    bounce-back) at 128 x 64, FP32, 100 steps on the two-step kernel,
    against the oracle in FP64: 1e-5 relative"""
    nx, ny = 128, 64
    mask = _mask(nx, ny)
    flow = lt.Obstacle(lt.Context("cuda:0", dtype=F32), [nx, ny], 100, 0.1,
                       domain_length_x=nx / 10.0, stencil=lt.D2Q9())
    flow.mask = mask
    flow.initialize()
    f0 = flow.f.double().cpu().numpy()
    tau = flow.units.relaxation_parameter_lu
    sim = lt.Simulation(flow, lt.BGKCollision(tau), [], PRE)
    assert sim._collide_and_stream.plan(100) == (50, 0)
    oflow = _OracleObstacle([nx, ny], mask, f0, nx / 10.0)
    osim = O.Simulation(oflow, O.BGKCollision(tau), [], O.StreamingStrategy.PRE_STREAMING)
    assert np.array_equal(sim.no_collision_mask.cpu().numpy(), osim.no_collision_mask.numpy())
    sim(100)
    osim(100)
    out = flow.f.double().cpu().numpy()
    ref = oflow.f.numpy()
    assert np.abs(out - ref).max() / np.abs(ref).max() <= 1e-5


def test_two_step_2d_with_stochastic_rounding(knobs):
    """BF16 storage with stochastic rounding on a 100 x 128 channel (halfway
    bounce-back, equilibrium inlet and outlet), BGK, 8 steps: the two-step
    kernel rounds its first step with that step's bits, as a single step
    stores it, so it differs from single steps only where the two kernels
    round the collision arithmetic differently (a small fraction of the
    values, under a fifth of what another seed changes) and by at most the
    two-step bound with rounding to nearest"""
    def run(two_step, seed):
        knobs(two_step=1 if two_step else 0)
        ctx = lt.Context("cuda:0", dtype=F32, storage="bf16_shifted", rounding="stochastic",
                         rounding_seed=seed)
        flow = lt.Obstacle(ctx, [100, 128], 100, 0.1, domain_length_x=10.0, stencil=lt.D2Q9(),
                           outlet="equilibrium", solid="halfway", walls=True)
        flow.mask = _mask(100, 128)
        flow.initialize()
        sim = lt.Simulation(flow, lt.BGKCollision(flow.units.relaxation_parameter_lu), [], PRE)
        assert sim._collide_and_stream.plan(8) == ((4, 0) if two_step else (0, 8))
        sim(8)
        torch.cuda.synchronize()
        bits = flow._state.detach().cpu().contiguous().view(-1).view(torch.uint8).numpy()
        return (bits.view(np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    ref, got, other = run(False, 7), run(True, 7), run(False, 8)
    diff = np.abs(got - ref)
    assert diff.max() <= 8 * ULPS["bf16_shifted"], diff.max()
    assert (diff > 0).mean() < 0.2 * (other != ref).mean(), ((diff > 0).mean(), (other != ref).mean())


@pytest.mark.parametrize("dtype", [F32, torch.float16])
def test_two_step_2d_equilibrium_boundary_field_vs_oracle(dtype, knobs):
    """a spatially varying equilibrium boundary (per-cell FP64 field) on the
    x = 0 row and an interior block of a 24 x 64 D2Q9 grid, BGK, 20 pull
    steps on the two-step kernel, against the oracle: 1e-5 relative in FP32
    storage, 5e-3 in shifted FP16 storage (whose stored deviations round to
    2^-11 relative each step)"""
    from tests.test_oracle import _Grid, _perturbed
    from tests.test_parity_gpu import ArrayFlow
    res = [24, 64]
    g = torch.Generator().manual_seed(6)
    velocity = torch.zeros([2, 1, 64], dtype=F64)
    velocity[0, 0] = 0.05 * torch.sin(torch.linspace(0, np.pi, 64, dtype=F64))
    pressure = 0.002 * torch.rand([1, 24, 64], generator=g, dtype=F64)
    mask = torch.zeros(res, dtype=torch.bool)
    mask[0] = True
    mask[10:13, 20:30] = True
    f0 = np.ascontiguousarray(_perturbed(O.D2Q9(), 64, F64, seed=9).f.numpy()[:, :24, :])
    storage = None if dtype == F32 else "f16_shifted"
    flow = ArrayFlow(lt.Context("cuda:0", dtype=F32, storage=storage), f0, lt.D2Q9())
    flow._post = [lt.EquilibriumBoundaryPU(flow.context, flow, mask, velocity, pressure)]
    sim = lt.Simulation(flow, lt.BGKCollision(0.7), [], PRE)
    assert sim.cell_classes.bc_field is not None
    assert sim._collide_and_stream.plan(20) == (10, 0)
    oflow = _Grid(O.D2Q9(), res, F64, f0)
    oflow.post_boundaries = [O.EquilibriumBoundaryPU(oflow.context, oflow, mask, velocity,
                                                     pressure)]
    osim = O.Simulation(oflow, O.BGKCollision(0.7), [], O.StreamingStrategy.PRE_STREAMING)
    sim(20)
    osim(20)
    got = flow.f.double().cpu().numpy()
    ref = oflow.f.numpy()
    tol = 1e-5 if dtype == F32 else 5e-3   # FP16: a few storage ulps per step
    assert np.abs(got - ref).max() / np.abs(ref).max() <= tol


def _periodic_f0(nx, ny, seed=4):
    """a perturbed equilibrium of a random velocity field on a periodic
    nx x ny D2Q9 grid"""
    from tests.test_oracle import _Grid
    g = torch.Generator().manual_seed(seed)
    oflow = _Grid(O.D2Q9(), [nx, ny], F64)
    u = 0.05 * (torch.rand([2, nx, ny], generator=g, dtype=F64) - 0.5)
    f0 = oflow.equilibrium(oflow, rho=torch.ones([1, nx, ny], dtype=F64), u=u)
    return (f0 * (1 + 0.02 * torch.rand(f0.shape, generator=g, dtype=F64))).numpy()


@pytest.mark.parametrize("storage,layout,nx,ny,coll", [
    ("f32", "soa", 40, 64, "bgk"), ("f32", "rows", 37, 128, "kbc"),
    ("f32", "soa", 24, 640, "trt"), ("f16_shifted", "soa", 33, 512, "bgk"),
    ("bf16_shifted", "rows", 20, 256, "none"), ("f32", "soa", 16, 128, "none")])
def test_two_step_2d_periodic_matches_single_steps(storage, layout, nx, ny, coll, knobs):
    """periodic D2Q9 grids without a class map run the same kernel with
    every cell a plain collision: bit for bit without collision, otherwise
    to the rounding of the collision arithmetic, in both memory layouts"""
    from tests.test_parity_gpu import ArrayFlow
    f0 = _periodic_f0(nx, ny)

    def run(two_step, steps):
        knobs(two_step=1 if two_step else 0)
        ctx = lt.Context("cuda:0", dtype=F32, storage=storage, layout=layout)
        flow = ArrayFlow(ctx, f0, lt.D2Q9())
        sim = lt.Simulation(flow, COLLISIONS[coll](0.6), [], PRE)
        plan = sim._collide_and_stream.plan(steps)
        sim(steps)
        torch.cuda.synchronize()
        bits = flow._state.detach().cpu().contiguous().view(-1).view(torch.uint8).numpy()
        return bits, flow.f.double().cpu().numpy(), plan
    for steps in (4, 7):
        rb, ref, p0 = run(False, steps)
        gb, got, p1 = run(True, steps)
        assert p0 == (0, steps)
        pairs = steps // 2 - (steps // 2) % 2
        assert p1 == (pairs, steps - 2 * pairs)
        if coll == "none":
            assert np.array_equal(gb, rb), steps
        else:
            tol = steps * ULPS[storage] * (50 if coll == "kbc" else 1)
            assert np.abs(got - ref).max() <= tol, (steps, np.abs(got - ref).max())

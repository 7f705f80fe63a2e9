"""GPU parity of the anti-bounce-back outlet (LBM_OP_ANTI_BOUNCE_BACK)
against the oracle's restatement of lettuce's AntiBounceBackOutlet
(lettuce/ext/_boundary/anti_bounce_back_outlet.py:13-109), inside lettuce's
default Obstacle setup: equilibrium inlet, anti-bounce-back outlet and
fullway bounce-back on the obstacle (lettuce/ext/_flows/obstacle.py:107-121)."""
import numpy as np
import pytest
import torch

import lettuce_amd as lt
from oracle import lettuce_torch as O
from tests.conftest import golden
from tests.test_parity_gpu import ctx, F32, F64

pytestmark = pytest.mark.gpu


class _OracleObstacle(O.Flow):
    """the oracle flow with lt.Obstacle's units and lettuce's default
    Obstacle boundaries"""

    def __init__(self, dtype, res, mask, f0, length_x):
        units = O.UnitConversion(reynolds_number=50, mach_number=0.1,
                                 characteristic_length_lu=res[0] / length_x,
                                 characteristic_length_pu=1, characteristic_velocity_pu=1)
        super().__init__(O.Context("cpu", dtype), res, units, O.D2Q9())
        self.f = torch.as_tensor(f0).to(dtype).clone()
        x = torch.meshgrid(*[units.convert_length_to_pu(torch.arange(n)) for n in res],
                           indexing="ij")[0]
        unit = torch.eye(2)[0]
        self.post_boundaries = [
            O.EquilibriumBoundaryPU(self.context, self, torch.abs(x) < 1e-6,
                                    units.characteristic_velocity_pu * unit),
            O.AntiBounceBackOutlet([1, 0], self),
            O.BounceBackBoundary(torch.as_tensor(mask)),
        ]


@pytest.mark.parametrize("dtype", [F32, F64])
@pytest.mark.parametrize("strategy", ["PRE_STREAMING", "POST_STREAMING"])
def test_obstacle_with_anti_bounce_back_outlet_vs_oracle(dtype, strategy):
    g = golden("g6_obstacle_eqoutlet.npz")
    res = [48, 24]
    flow = lt.Obstacle(ctx(dtype), res, reynolds_number=50, mach_number=0.1,
                       domain_length_x=4.0, stencil=lt.D2Q9(), outlet="antibounceback")
    flow.mask = g["mask"]
    flow.f = torch.from_numpy(g["f0"]).to(dtype)
    tau = flow.units.relaxation_parameter_lu
    sim = lt.Simulation(flow, lt.BGKCollision(tau), [], getattr(lt.StreamingStrategy, strategy))
    oflow = _OracleObstacle(dtype, res, g["mask"], g["f0"], 4.0)
    osim = O.Simulation(oflow, O.BGKCollision(tau), [], getattr(O.StreamingStrategy, strategy))
    assert np.array_equal(sim.no_collision_mask.cpu().numpy(), osim.no_collision_mask.numpy())
    assert np.array_equal(sim.no_streaming_mask.cpu().numpy(), osim.no_streaming_mask.numpy())
    for _ in range(4):
        sim(25)
        osim(25)
        ref = oflow.f.numpy().astype(np.float64)
        got = flow.f.cpu().numpy().astype(np.float64)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        assert err <= (1e-12 if dtype == F64 else 1e-5), err


def test_uniform_flow_passes_the_outlet_unchanged():
    """a uniform equilibrium flow is a fixed point of the outlet
    (feq_v + feq_opp(v) = w rho (2 + 9 (e.u)^2 - 3 |u|^2)): on a periodic
    grid whose only boundary is the outlet, every population stays feq(rho, u)
    to rounding"""
    from tests.test_parity_gpu import ArrayFlow
    res = [16, 8]
    of = O.Flow(O.Context("cpu", F64), res, O.UnitConversion(50, 0.1), O.D2Q9())
    u = torch.zeros([2] + res, dtype=F64)
    u[0], u[1] = 0.05, -0.01
    f0 = of.equilibrium(of, rho=torch.full([1] + res, 1.01, dtype=F64), u=u).numpy()
    flow = ArrayFlow(ctx(F64), f0, lt.D2Q9())
    flow._post = [lt.AntiBounceBackOutlet([1, 0], flow)]
    sim = lt.Simulation(flow, lt.BGKCollision(0.6), [], lt.StreamingStrategy.PRE_STREAMING)
    sim(20)
    err = float((flow.f.cpu() - torch.from_numpy(f0)).abs().max())
    assert err <= 1e-14, err


G8_CASES = {"D2Q9": (lt.D2Q9, [48, 24], 4.0, "bgk"), "D3Q27": (lt.D3Q27, [16, 8, 8], 2.0, "kbc")}


@pytest.mark.parametrize("name", list(G8_CASES))
@pytest.mark.parametrize("tag", ["post", "pre"])
@pytest.mark.parametrize("dtype", [F32, F64])
def test_default_obstacle_matches_lettuce(name, tag, dtype):
    """lettuce's default Obstacle (equilibrium inlet, anti-bounce-back outlet,
    fullway bounce-back; obstacle.py:107-122) against lettuce's own run
    (golden g8): D2Q9 BGK and D3Q27 KBC, masks bit for bit, 10 steps in
    either order; FP64 to 1e-12, FP32 to 1e-5 of the FP64 golden"""
    g = golden("g8_obstacle_abb.npz")
    stencil, res, length_x, coll = G8_CASES[name]
    flow = lt.Obstacle(ctx(dtype), res, reynolds_number=50, mach_number=0.1,
                       domain_length_x=length_x, stencil=stencil())
    flow.mask = g[f"{name}_mask"]
    flow.f = torch.from_numpy(g[f"{name}_f0"]).to(dtype)
    tau = flow.units.relaxation_parameter_lu
    assert tau == float(g[f"{name}_tau"])
    collision = lt.BGKCollision(tau) if coll == "bgk" else lt.KBCCollision()
    strategy = (lt.StreamingStrategy.POST_STREAMING if tag == "post"
                else lt.StreamingStrategy.PRE_STREAMING)
    sim = lt.Simulation(flow, collision, [], strategy)
    assert np.array_equal(sim.no_collision_mask.cpu().numpy(), g[f"{name}_ncm"])
    assert np.array_equal(sim.no_streaming_mask.cpu().numpy(), g[f"{name}_nsm"])
    sim(10)
    ref = g[f"{name}_f10_{tag}"]
    got = flow.f.cpu().numpy().astype(np.float64)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    elem = (np.abs(got - ref) / np.abs(ref)).max()
    assert err <= (1e-12 if dtype == F64 else 1e-5), err
    assert elem <= (1e-11 if dtype == F64 else 1e-4), elem

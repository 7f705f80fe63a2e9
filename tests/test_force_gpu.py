"""GPU checks of BGK with Guo forcing (LBM_COLL_BGK_GUO): parity with the
oracle's restatement of lettuce's BGKCollision(force=Guo)
(lettuce/ext/_collision/bgk_collision.py:12-22, ext/_force/guo.py:9-43),
the momentum invariant, and lettuce's own Poiseuille force test
(tests/collision/test_force.py:5-52)."""
import numpy as np
import pytest
import torch

import lettuce_amd as lt
from oracle import lettuce_torch as O
from tests.test_oracle import _perturbed
from tests.test_parity_gpu import ArrayFlow, ctx, F32, F64

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [F32, F64])
@pytest.mark.parametrize("stencil", ["D2Q9", "D3Q19", "D3Q27"])
@pytest.mark.parametrize("strategy", ["PRE_STREAMING", "POST_STREAMING"])
def test_bgk_guo_matches_oracle(dtype, stencil, strategy):
    os_ = getattr(O, stencil)()
    n = 16 if os_.d == 2 else 8
    oflow = _perturbed(os_, n, dtype, seed=4)
    f0 = oflow.f.numpy().copy()
    tau = 0.61
    a = [0.001, -0.0004, 0.0002][:os_.d]
    osim = O.Simulation(oflow, O.BGKCollision(tau, force=O.Guo(oflow, tau, a)), [],
                        getattr(O.StreamingStrategy, strategy))
    flow = ArrayFlow(ctx(dtype), f0, getattr(lt, stencil)())
    sim = lt.Simulation(flow, lt.BGKCollision(tau, force=lt.Guo(flow, tau, a)), [],
                        getattr(lt.StreamingStrategy, strategy))
    sim(50)
    osim(50)
    ref = oflow.f.numpy()
    err = np.abs(flow.f.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err <= (1e-13 if dtype == F64 else 1e-5), err


def test_guo_momentum_grows_by_the_acceleration():
    """periodic grid, FP64: total momentum after n steps = j0 + n a cells"""
    oflow = _perturbed(O.D2Q9(), 32, F64, seed=1)
    flow = ArrayFlow(ctx(F64), oflow.f.numpy(), lt.D2Q9())
    a = [0.0007, 0.0003]
    j0 = flow.j().sum(dim=(1, 2))
    sim = lt.Simulation(flow, lt.BGKCollision(0.7, force=lt.Guo(flow, 0.7, a)), [],
                        lt.StreamingStrategy.PRE_STREAMING)
    sim(40)
    j = flow.j().sum(dim=(1, 2))
    expect = j0 + 40 * 32 * 32 * torch.tensor(a, dtype=F64, device=j.device)
    assert torch.allclose(j, expect, rtol=0, atol=1e-10)


def test_poiseuille_force_like_lettuce():
    """lettuce's test_force: Poiseuille 17 x 17, Re 1, Ma 0.02, 1000
    post-streaming steps, velocity (with the a/(2 rho) correction) within 1 %
    of the analytic profile"""
    context = lt.Context("cuda:0", dtype=torch.float64)
    flow = lt.PoiseuilleFlow2D(context, resolution=17, reynolds_number=1, mach_number=0.02,
                               initialize_with_zeros=True)
    acceleration_lu = flow.units.convert_acceleration_to_lu(flow.acceleration)
    force = lt.Guo(flow, tau=flow.units.relaxation_parameter_lu, acceleration=acceleration_lu)
    collision = lt.BGKCollision(tau=flow.units.relaxation_parameter_lu, force=force)
    errors = lt.ErrorReporter(flow.analytic_solution, interval=100, out=None)
    simulation = lt.Simulation(flow, collision, [errors])
    simulation(1000)
    u_sim = flow.units.convert_velocity_to_pu(flow.u(acceleration=acceleration_lu))
    _, u_ref = flow.analytic_solution()
    fluid = torch.eq(simulation.no_collision_mask, 0).to(u_sim.device)
    for dim in range(2):
        a = u_sim[dim][fluid].cpu().numpy()
        b = u_ref[dim][fluid].cpu().numpy()
        assert a.max() == pytest.approx(b.max(), rel=0.01)
        assert a == pytest.approx(b, rel=None, abs=0.01 * float(u_ref[0].max()))
    assert len(errors.out) == 11 and errors.out[-1][0] < errors.out[1][0]


def test_poiseuille_guo_matches_lettuce_trajectory():
    """lettuce's force test case run by lettuce itself (golden g9):
    PoiseuilleFlow2D 17^2 with the Guo force and bounce-back walls, FP64,
    post-streaming; the state after 10 and 100 steps and the velocity with
    the a / (2 rho) shift"""
    from tests.conftest import golden
    g = golden("g9_poiseuille_guo.npz")
    flow = lt.PoiseuilleFlow2D(lt.Context("cuda:0", dtype=torch.float64), resolution=17,
                               reynolds_number=1, mach_number=0.02,
                               initialize_with_zeros=True)
    assert np.array_equal(flow.f.cpu().numpy(), g["f0"])
    acc = flow.units.convert_acceleration_to_lu(flow.acceleration)
    tau = flow.units.relaxation_parameter_lu
    sim = lt.Simulation(flow, lt.BGKCollision(tau, force=lt.Guo(flow, tau, acc)), [])
    assert np.array_equal(sim.no_collision_mask.cpu().numpy(), g["ncm"])
    assert np.array_equal(sim.no_streaming_mask.cpu().numpy(), g["nsm"])
    sim(10)
    assert np.abs(flow.f.cpu().numpy() - g["f10"]).max() <= 1e-14
    sim(90)
    assert np.abs(flow.f.cpu().numpy() - g["f100"]).max() <= 1e-14
    u = flow.u(acceleration=acc).cpu().numpy()
    assert np.abs(u - g["u100"]).max() <= 1e-15

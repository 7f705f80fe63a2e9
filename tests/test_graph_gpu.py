"""HIP graph capture of the native step: lbm_run launches only kernels on
torch's current stream (no allocation, no synchronisation, no memset), so a
run of steps can be captured once and replayed.  Each replay is checked
against an eager run of the same steps, with eager work and a host
synchronisation between the replays."""
import numpy as np
import pytest
import torch

import lettuce_amd as lt
from tests.test_parity_gpu import ArrayFlow

pytestmark = pytest.mark.gpu

F32 = torch.float32


def _capture(sim, steps):
    """a graph of sim(steps); warm-up and capture on a side stream of
    default priority, as torch requires"""
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        sim._collide_and_stream.plan(steps)  # binds the buffers before capture
    torch.cuda.current_stream().wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        sim._collide_and_stream.run(steps)
    return graph


def _case(kind):
    if kind == "d3q19_two_step":
        ctx = lt.Context("cuda:0", dtype=F32, layout="rows")
        flow = lt.TaylorGreenVortex(ctx, 64, 1600, 0.1, lt.D3Q19(), initialize_fneq=False)
        coll = lt.BGKCollision(flow.units.relaxation_parameter_lu)
        return flow, coll, None
    if kind == "d2q9_learned_channel":
        nx, ny = 128, 64
        yy, xx = np.meshgrid(np.arange(ny), np.arange(nx))
        mask = (xx - 32) ** 2 + (yy - 31.5) ** 2 < 8.0 ** 2
        ctx = lt.Context("cuda:0", dtype=F32, storage="f16_shifted")
        flow = lt.Obstacle(ctx, [nx, ny], 100, 0.1, domain_length_x=nx / 16.0,
                           stencil=lt.D2Q9(), outlet="equilibrium", solid="halfway", walls=True)
        flow.mask = mask
        flow.initialize()
        coll = lt.LearnedCollision(lt.LearnedParameters.seeded(0, out_scale=0.1))
        return flow, coll, mask
    if kind == "d2q9_kbc_channel_two_step":
        # rows of 512 cells: the 2D two-step kernel with 155.6 KB of LDS
        nx, ny = 48, 512
        yy, xx = np.meshgrid(np.arange(ny), np.arange(nx))
        mask = (xx - 12) ** 2 + (yy - 255.5) ** 2 < 6.0 ** 2
        ctx = lt.Context("cuda:0", dtype=F32)
        flow = lt.Obstacle(ctx, [nx, ny], 100, 0.1, domain_length_x=nx / 16.0,
                           stencil=lt.D2Q9(), outlet="equilibrium", solid="halfway", walls=True)
        flow.mask = mask
        flow.initialize()
        return flow, lt.KBCCollision(), mask
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["d3q19_two_step", "d2q9_learned_channel",
                                  "d2q9_kbc_channel_two_step"])
def test_captured_steps_replay_like_eager_steps(kind):
    steps = 4  # even: the result stays in the state buffer the graph reads
    flow, coll, mask = _case(kind)
    sim = lt.Simulation(flow, coll, [], lt.StreamingStrategy.PRE_STREAMING)
    f0 = flow._state.clone()
    graph = _capture(sim, steps)
    # capture launches nothing: the state is untouched
    torch.cuda.synchronize()
    assert torch.equal(flow._state, f0)

    ref_flow, ref_coll, _ = _case(kind)
    ref_flow._state.copy_(f0)
    ref = lt.Simulation(ref_flow, ref_coll, [], lt.StreamingStrategy.PRE_STREAMING)
    for replay in range(3):
        graph.replay()
        ref(steps)                      # eager work between the replays
        torch.cuda.synchronize()
        assert torch.equal(flow._state, ref_flow._state), replay

"""x-slabs (csrc/lbm_slab.hip) against the periodic lattice.

A lattice split along x into slabs of 20, 1, 2 and 25 planes (a slab with
a single plane is both of its faces; one with two planes has no interior),
whose halo buffers are passed around the ring by device copies, must step
exactly as the whole periodic lattice: bit for bit without collision (the
halo packing, the face pulls and the send buffers), and to the rounding of
the collision arithmetic with BGK / TRT / KBC, whose formulas the face
kernel and step_kernel each compile.  The same holds for one slab wrapped
onto itself.  The packing order of the halo buffers is checked against the
stencil's velocities."""
import numpy as np
import pytest
import torch

import lettuce_amd as lt
from tests.test_two_step_gpu import ULPS, _f0, _Flow

pytestmark = pytest.mark.gpu

F32 = torch.float32

COLLS = {"none": lambda: lt.NoCollision(), "bgk": lambda: lt.BGKCollision(0.62),
         "trt": lambda: lt.TRTCollision(0.7, 0.9), "kbc": lambda: lt.KBCCollision(0.55)}


def _sim(f0, storage, layout, coll):
    dtype = torch.float64 if storage == "f64" else F32
    c = lt.Context("cuda:0", dtype=dtype, storage=storage, layout=layout)
    flow = _Flow(c, f0.astype(np.float64 if storage == "f64" else np.float32))
    return lt.Simulation(flow, COLLS[coll](), [], lt.StreamingStrategy.PRE_STREAMING)


def _periodic(f0, storage, layout, coll, steps, knobs):
    # single fused steps (the two-step kernel rounds its collisions apart)
    knobs(two_step="0")
    sim = _sim(f0, storage, layout, coll)
    sim(steps)
    torch.cuda.synchronize()
    return sim.flow


def _decoded(flow, state):
    """state decoded to float64 through the flow's own decoder"""
    keep = flow._state
    flow._state = state
    try:
        return flow.f.double().cpu().numpy()
    finally:
        flow._state = keep


def _run_slabs(f0, storage, layout, coll, steps, sizes, split):
    sim = _sim(f0, storage, layout, coll)
    full = sim.flow._state
    bounds = np.cumsum([0] + list(sizes))
    assert bounds[-1] == full.shape[1]
    slabs = [lt.SlabStep(sim, full[:, a:b].clone()) for a, b in zip(bounds[:-1], bounds[1:])]
    for s in slabs:
        s.pack()
    k = len(slabs)
    for _ in range(steps):
        if split:
            for s in slabs:
                s.step("interior")
        for i, s in enumerate(slabs):   # the exchange around the ring
            s.recv_lo.copy_(slabs[(i - 1) % k].send_hi)
            s.recv_hi.copy_(slabs[(i + 1) % k].send_lo)
        for s in slabs:
            s.step("faces" if split else "all")
    out = torch.cat([s.state for s in slabs], dim=1)
    torch.cuda.synchronize()
    return sim.flow, out


def _compare(flow_ref, flow, out, storage, coll):
    ref = flow_ref._state
    if coll == "none":
        assert torch.equal(out.contiguous().cpu(), ref.contiguous().cpu())
        return
    a = _decoded(flow, out.contiguous())
    b = _decoded(flow_ref, ref.contiguous())
    err = float(np.abs(a - b).max())
    assert err <= 8 * ULPS[storage], err


@pytest.mark.parametrize("stencil", ["D3Q19", "D3Q27"])
@pytest.mark.parametrize("storage,layout", [("f32", "soa"), ("f32", "rows"),
                                            ("f16_shifted", "soa"), ("bf16_shifted", "rows")])
@pytest.mark.parametrize("coll", ["none", "bgk", "trt", "kbc"])
def test_slabs_step_as_the_periodic_lattice(stencil, storage, layout, coll, knobs):
    if stencil == "D3Q19" and coll == "kbc":
        pytest.skip("KBC is defined for D2Q9 and D3Q27")
    from oracle import lettuce_torch as O
    res = [48, 16, 24]
    f0 = _f0(res, seed=7, stencil=getattr(O, stencil))
    steps = 6
    flow_ref = _periodic(f0, storage, layout, coll, steps, knobs)
    for split in (False, True):
        flow, out = _run_slabs(f0, storage, layout, coll, steps, [20, 1, 2, 25], split)
        _compare(flow_ref, flow, out, storage, coll)


@pytest.mark.parametrize("stencil,coll", [("D3Q19", "none"), ("D3Q19", "bgk"), ("D3Q27", "kbc")])
def test_one_slab_wrapped_onto_itself(stencil, coll, knobs):
    from oracle import lettuce_torch as O
    res = [12, 8, 64]
    f0 = _f0(res, seed=3, stencil=getattr(O, stencil))
    steps = 9
    flow_ref = _periodic(f0, "f32", "soa", coll, steps, knobs)
    sim = _sim(f0, "f32", "soa", coll)
    slab = lt.SlabStep(sim, sim.flow._state.clone())
    slab.pack()
    for _ in range(steps):
        slab.wrap()
        slab.step()
    torch.cuda.synchronize()
    _compare(flow_ref, sim.flow, slab.state, "f32", coll)


@pytest.mark.parametrize("stencil", ["D3Q19", "D3Q27"])
@pytest.mark.parametrize("storage", ["f32", "f64", "bf16_shifted"])
def test_halo_packing_order(stencil, storage):
    """send_hi[k] is plane nx-1 of the k-th population with e_x = +1 in
    stencil order, send_lo[k] plane 0 of the k-th with e_x = -1; the face
    step leaves the same values of the new state behind"""
    from oracle import lettuce_torch as O
    res = [5, 8, 16]
    f0 = _f0(res, seed=11, stencil=getattr(O, stencil))
    sim = _sim(f0, storage, "soa", "bgk")
    e = np.asarray(sim.flow.stencil.e)
    plus = [q for q in range(len(e)) if e[q][0] == 1]
    minus = [q for q in range(len(e)) if e[q][0] == -1]
    slab = lt.SlabStep(sim, sim.flow._state.clone())
    assert slab.send_hi.shape == (len(plus), 8, 16) == slab.send_lo.shape
    for check_step in (False, True):
        if check_step:
            slab.wrap()
            slab.step()
        else:
            slab.pack()
        torch.cuda.synchronize()
        st = slab.state
        assert torch.equal(slab.send_hi.cpu(), st[plus, -1].cpu())
        assert torch.equal(slab.send_lo.cpu(), st[minus, 0].cpu())

"""Checks at the bench workloads' full sizes, through properties that do not
need a CPU reference of that size: the two-step kernel against single steps
on the 256^3 headline grid, conservation of mass and momentum (FP64
reductions), the shifted storage's decode -> encode round trip, the learned
collision in BF16 against FP32, and an x-slab of one of 8 devices at its real
shape against the periodic step of the same array."""
import ctypes
import math

import numpy as np
import pytest
import torch

import lettuce_amd as lt
from lettuce_amd import _native as N
from lettuce_amd import _ops

pytestmark = pytest.mark.gpu

F32 = torch.float32


def _tgv(stencil, n, storage="f32", layout="rows"):
    ctx = lt.Context("cuda:0", dtype=F32, storage=storage, layout=layout)
    return lt.TaylorGreenVortex(ctx, n, 1600, 0.1, stencil, initialize_fneq=False)


def _sums(flow):
    return _ops.reduce_f64(flow.stencil, flow._state, flow.rho0_shift).cpu().numpy()


def test_headline_two_step_matches_single_steps_at_256(knobs):
    """D3Q19 BGK TGV 256^3 FP32 (the bench workload), 20 steps: the default
    path (10 two-step launches) against 20 single-step launches, and mass
    and momentum conserved"""
    S = lt.D3Q19()
    runs = {}
    for mode in ("1", "0"):
        knobs(two_step=mode)
        flow = _tgv(S, 256)
        sim = lt.Simulation(flow, lt.BGKCollision(flow.units.relaxation_parameter_lu), [],
                            lt.StreamingStrategy.PRE_STREAMING)
        assert sim._collide_and_stream.plan(20) == ((10, 0) if mode == "1" else (0, 20))
        s0 = _sums(flow)
        sim(20)
        s1 = _sums(flow)
        # mass and momentum to FP32 rounding of 16.8 M cells
        assert abs(s1[0] - s0[0]) <= 1e-6 * s0[0], (s0[0], s1[0])
        assert np.all(np.abs(s1[1:4] - s0[1:4]) <= 1e-9 * s0[0]), (s0[1:4], s1[1:4])
        runs[mode] = flow._state.clone()
        del sim, flow
    # the collision arithmetic may round apart by a few ulp per step
    diff = (runs["1"] - runs["0"]).abs().max().item()
    assert diff <= 2e-6, diff


def test_d3q27_kbc_conserves_mass_and_momentum_at_256():
    S = lt.D3Q27()
    flow = _tgv(S, 256)
    sim = lt.Simulation(flow, lt.KBCCollision(flow.units.relaxation_parameter_lu), [],
                        lt.StreamingStrategy.PRE_STREAMING)
    s0 = _sums(flow)
    sim(10)
    s1 = _sums(flow)
    assert abs(s1[0] - s0[0]) <= 1e-6 * s0[0]
    assert np.all(np.abs(s1[1:4] - s0[1:4]) <= 1e-9 * s0[0]), (s0[1:4], s1[1:4])
    scan = _ops.scan_f64(S, flow._state, flow.rho0_shift).cpu().numpy()
    assert scan[0] == 0 and scan[1] == 0


@pytest.mark.parametrize("storage", ["f16_shifted", "bf16_shifted"])
def test_shifted_storage_round_trip_is_idempotent_at_256(storage):
    """decode -> encode of a stepped 256^3 D3Q19 state returns the same
    stored values (bit for bit but for the sign of zero: a stored -0 decodes
    to w rho0 and encodes back as +0)"""
    S = lt.D3Q19()
    flow = _tgv(S, 256, storage)
    sim = lt.Simulation(flow, lt.BGKCollision(flow.units.relaxation_parameter_lu), [],
                        lt.StreamingStrategy.PRE_STREAMING)
    sim(4)
    raw = flow._state.clone()
    decoded = _ops.decode(S, flow._state, flow.rho0_shift)
    again = torch.empty_like(flow._state)
    _ops.encode(S, decoded, again, flow.rho0_shift)
    assert torch.equal(again, raw)
    nonzero = raw != 0
    assert torch.equal(again.view(torch.int16)[nonzero], raw.view(torch.int16)[nonzero])


def test_learned_bf16_tracks_fp32_at_256():
    """the C5 pairing at 256^3: learned collision with shifted BF16 storage
    against FP32 storage, 10 steps from the same TGV; mass conserved to the
    16-bit rounding"""
    S = lt.D3Q27()
    params = lt.LearnedParameters.seeded(0, out_scale=0.1)
    out = {}
    for storage in ("f32", "bf16_shifted"):
        flow = _tgv(S, 256, storage)
        sim = lt.Simulation(flow, lt.LearnedCollision(params), [],
                            lt.StreamingStrategy.PRE_STREAMING)
        s0 = _sums(flow)
        sim(10)
        s1 = _sums(flow)
        assert abs(s1[0] - s0[0]) <= (1e-6 if storage == "f32" else 1e-4) * s0[0]
        out[storage] = flow.f.to(F32)
        del sim, flow
    err = ((out["bf16_shifted"] - out["f32"]).abs().max() / out["f32"].abs().max()).item()
    assert err <= 2e-2, err


def test_slab_of_eight_devices_wrapped_is_the_periodic_step():
    """a 48 x 384 x 384 D3Q27 KBC slab (384^3 over 8 devices) whose halos are
    its own: 10 steps bit for bit equal to the periodic step of the array"""
    S = lt.D3Q27()
    flow = _tgv(S, 384, layout="soa")
    sim = lt.Simulation(flow, lt.KBCCollision(flow.units.relaxation_parameter_lu), [],
                        lt.StreamingStrategy.PRE_STREAMING)
    start = flow._state[:, :48].clone()
    slab = lt.SlabStep(sim, start.clone())
    args = N.LbmStepArgs.from_buffer_copy(slab.args)
    a, b = start.clone(), torch.empty_like(start)
    for _ in range(10):
        args.f_in, args.f_out = a.data_ptr(), b.data_ptr()
        N.check(N.lib().lbm_step(ctypes.byref(args), N.stream_handle(a.device)), "lbm_step")
        a, b = b, a
    slab.pack()
    for _ in range(10):
        slab.wrap()
        slab.step()
    torch.cuda.synchronize()
    assert torch.equal(slab.state, a)


class _TiledFlow(lt.ExtFlow):
    """D3Q19 at n^3 whose initial populations repeat with period h along x
    and y: a TGV-like equilibrium built on one h x h x n block and copied
    into the other three, so the copies are bit for bit equal"""

    def __init__(self, context, n, h):
        self._n, self._h = n, h
        super().__init__(context, [n, n, n], 100.0, 0.05, lt.D3Q19())

    def make_resolution(self, resolution, stencil=None):
        return list(resolution)

    def make_units(self, reynolds_number, mach_number, resolution):
        return lt.UnitConversion(reynolds_number, mach_number,
                                 characteristic_length_lu=resolution[0])

    def initial_pu(self):
        raise NotImplementedError

    def initialize(self):
        n, h, f = self._n, self._h, self._state
        dev = f.device
        a = torch.arange(h, device=dev, dtype=F32) * (2 * math.pi / h)
        c = torch.cos(torch.arange(n, device=dev, dtype=F32) * (2 * math.pi / n))
        u = torch.zeros(3, h, h, n, device=dev, dtype=F32)
        u[0] = 0.05 * torch.sin(a)[:, None, None] * torch.cos(a)[None, :, None] * c
        u[1] = -0.05 * torch.cos(a)[:, None, None] * torch.sin(a)[None, :, None] * c
        rho = torch.ones(h, h, n, device=dev, dtype=F32)
        _ops.equilibrium(self.stencil, f[:, :h, :h], rho, u)
        del u, rho
        for i, j in ((1, 0), (0, 1), (1, 1)):
            f[:, i * h:(i + 1) * h, j * h:(j + 1) * h].copy_(f[:, :h, :h])


@pytest.mark.parametrize("layout", ["soa", "rows"])
def test_largest_d3q19_grid_keeps_its_translation_symmetry(layout):
    """D3Q19 BGK FP32 at 1024^3 (1.07 G cells, two 76 GB population
    arrays): a state that repeats with period 512 along x and y keeps doing
    so bit for bit through 8 steps of the two-step kernel, since every cell
    runs the same arithmetic on the same neighbourhood.  This checks the
    64-bit plane offsets (2.0e10 elements in `rows`, whose planes interleave
    all 19 populations), the 32-bit offsets inside a plane and the tiling at
    a grid near the largest one device holds; mass is conserved as well."""
    n, h = 1024, 512
    flow = _TiledFlow(lt.Context("cuda:0", dtype=F32, layout=layout), n, h)
    sim = lt.Simulation(flow, lt.BGKCollision(0.51), [], lt.StreamingStrategy.PRE_STREAMING)
    assert sim._collide_and_stream.plan(8) == (4, 0)
    s0 = _sums(flow)
    sim(8)
    s1 = _sums(flow)
    assert abs(s1[0] - s0[0]) <= 1e-6 * s0[0], (s0[0], s1[0])
    f = flow._state
    assert torch.isfinite(f[:, :h, :h]).all()
    for i, j in ((1, 0), (0, 1), (1, 1)):
        assert torch.equal(f[:, i * h:(i + 1) * h, j * h:(j + 1) * h], f[:, :h, :h]), (i, j)


def _turbulence(ctx, res=384):
    """the 384^3 D3Q27 decaying-turbulence workload: u_rms 0.05, nu 1e-4"""
    from lettuce_amd.studies import turbulence_units
    re, ma = turbulence_units(res)
    return lt.DecayingTurbulence(ctx, [res] * 3, re, ma, k0=4, stencil=lt.D3Q27(),
                                 initialize_pressure=False, initialize_fneq=False, randseed=0)


def test_turbulence_384_kbc_and_trained_learned_at_full_size():
    """The 384^3 D3Q27 decaying-turbulence workload at its real size: the
    field has the requested u_rms and tau; KBC (FP32, two-step kernel)
    conserves mass and momentum over 20 steps; the trained learned
    collision with shifted FP16 storage conserves mass, leaves at most 1e-3
    of the cells on the KBC fallback, and its kinetic energy tracks KBC's
    to 1e-3."""
    flow = _turbulence(lt.Context("cuda:0", dtype=F32))
    u = flow.u()
    assert abs(float(torch.sqrt((u * u).sum(0).mean())) / 0.05 - 1) <= 1e-6
    assert abs(flow.units.relaxation_parameter_lu - 0.5003) <= 1e-12
    del u
    sim = lt.Simulation(flow, lt.KBCCollision(), [], lt.StreamingStrategy.PRE_STREAMING)
    assert sim._collide_and_stream.plan(20) == (10, 0)
    s0 = _sums(flow)
    sim(20)
    s1 = _sums(flow)
    cells = 384 ** 3
    assert abs(s1[0] - s0[0]) <= 1e-6 * cells
    assert np.all(np.abs(s1[1:4] - s0[1:4]) <= 1e-7 * cells)
    ke_kbc = float(lt.IncompressibleKineticEnergy(flow)())
    del sim, flow
    torch.cuda.empty_cache()
    params = lt.LearnedParameters.load(lt.TRAINED_D3Q27_KBC)
    flow = _turbulence(lt.Context("cuda:0", dtype=F32, storage="f16_shifted"))
    free = lt.LearnedParameters(*params.tensors(), in_scale=params.in_scale,
                                rate_lo=params.rate_lo, rate_hi=params.rate_hi,
                                stable_lo=-1e30, stable_hi=1e30)
    differ = lt.LearnedCollision(params)(flow) != lt.LearnedCollision(free)(flow)
    assert float(differ.any(0).float().mean()) <= 1e-3
    sim = lt.Simulation(flow, lt.LearnedCollision(params), [], lt.StreamingStrategy.PRE_STREAMING)
    m0 = _sums(flow)[0]
    sim(20)
    assert abs(_sums(flow)[0] - m0) <= 1e-4 * cells
    ke = float(lt.IncompressibleKineticEnergy(flow)())
    assert abs(ke / ke_kbc - 1) <= 1e-3, (ke, ke_kbc)


def test_channel_2048x512_post_streaming_split_and_forces_at_full_size(knobs):
    """The 2048 x 512 channel with a D = 64 cylinder (halfway bounce-back on
    the cylinder and walls, equilibrium inlet and outlet) at its real size,
    KBC FP32, 200 steps in lettuce's default order: the split post-streaming
    run (its pull steps on the 2D two-step kernel) equals push steps to the
    rounding of the collision, no NaN, and the
    drag on the cylinder is positive and the same in both runs"""
    nx, ny = 2048, 512
    yy, xx = np.meshgrid(np.arange(ny), np.arange(nx))
    mask = (xx - nx / 4) ** 2 + (yy - ny / 2 - 1) ** 2 < 32.0 ** 2
    out, forces = [], []
    for split in (0, 1):
        knobs(post_split=split)
        flow = lt.Obstacle(lt.Context("cuda:0", dtype=F32), [nx, ny], 200, 0.05 * 3 ** 0.5,
                           domain_length_x=nx / 64, stencil=lt.D2Q9(), outlet="equilibrium",
                           solid="halfway", walls=True)
        flow.mask = mask
        flow.initialize()
        sim = lt.Simulation(flow, lt.KBCCollision(), [])
        sim(200)
        out.append(flow.f.cpu().numpy())
        forces.append(flow.post_streaming_boundaries[0].force(sim).cpu().numpy())
    assert np.isfinite(out[1]).all()
    assert np.abs(out[0] - out[1]).max() <= 1e-4 * np.abs(out[0]).max()
    assert forces[1][0] > 0 and abs(forces[1][0] / forces[0][0] - 1) <= 1e-3

"""GPU parity: the native step (through liblettuce_amd.so) against lettuce's
own trajectories (tests/golden) and the CPU oracle, in the pattern of
lettuce's tests/native/*: build the same flow twice, step both, compare f.

Tolerances: FP64 runs to 1e-12 absolute; FP32 runs to 1e-5 relative max
error on f (max |f_gpu - f_ref| / max |f_ref|) after up to 1000 steps, and
to 5e-5 entrywise (max |f_gpu - f_ref| / |f_ref|, which weighs D3Q27's
small corner populations fully); masks, class maps and link lists
bit-exact."""
import os

import numpy as np
import pytest
import torch

import lettuce_amd as lt
from lettuce_amd._classes import expand_classes
from oracle import lettuce_torch as O
from tests.conftest import golden

pytestmark = pytest.mark.gpu

F32, F64 = torch.float32, torch.float64
REL32 = 1e-5
# the entrywise error (each population relative to itself) that the FP32
# comparisons meet; measured on MI355X at most 2.5e-5 (1000 D3Q27 steps:
# 2.1e-5, KBC golden: 2.5e-5, halfway channel: 8.6e-6)
ELEM32 = 5e-5


# (test, max-normalised error, entrywise error) of every rel_err comparison;
# conftest writes them to the file LBM_ERR_LOG names, when it is set
ERR_LOG = []


def elem_err(a, b):
    """max |a - b| / |b| over the entries: the error of each population
    relative to itself (D3Q27's corner populations are 1/64 of its rest
    population)"""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float((np.abs(a - b) / np.abs(b)).max())


def rel_err(a, b):
    """max |a - b| / max |b|"""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    rel = float(np.abs(a - b).max() / np.abs(b).max())
    ERR_LOG.append((os.environ.get("PYTEST_CURRENT_TEST", ""), rel, elem_err(a, b)))
    return rel


class ArrayFlow(lt.ExtFlow):
    """A flow whose initial populations are given; lettuce's TestFlow units
    (Re=100, Ma=0.1, L=resolution[0])."""

    def __init__(self, context, f0, stencil, post=(), streaming=()):
        self._f0 = np.asarray(f0)
        self._post = list(post)
        self._psb = list(streaming)
        lt.ExtFlow.__init__(self, context, list(self._f0.shape[1:]), 100.0, 0.1,
                            stencil)

    def make_resolution(self, resolution, stencil=None):
        return list(resolution)

    def make_units(self, reynolds_number, mach_number, resolution):
        return lt.UnitConversion(reynolds_number, mach_number,
                                 characteristic_length_lu=resolution[0])

    def initial_pu(self):
        raise NotImplementedError

    def initialize(self):
        self.f = torch.from_numpy(self._f0)

    @property
    def post_boundaries(self):
        return self._post

    @property
    def post_streaming_boundaries(self):
        return self._psb


def ctx(dtype=F32, storage=None):
    return lt.Context("cuda:0", dtype=dtype, storage=storage)


def native_run(flow, collision, steps, strategy):
    sim = lt.Simulation(flow, collision, [], streaming_strategy=strategy)
    sim(steps)
    return flow.f.cpu().numpy()


def _stencils():
    return {"D2Q9": (lt.D2Q9(), O.D2Q9()), "D3Q19": (lt.D3Q19(), O.D3Q19()),
            "D3Q27": (lt.D3Q27(), O.D3Q27())}


def _perturbed(name, n, dtype=F64, seed=0):
    """random positive populations near equilibrium (oracle-built)"""
    from tests.test_oracle import _perturbed as make
    return make(_stencils()[name][1], n, dtype, seed).f.numpy()


# ------------------------------------------------------------ BGK / stream
@pytest.mark.parametrize("tag,dtype", [("f32", F32), ("f64", F64)])
@pytest.mark.parametrize("strategy", ["PRE_STREAMING", "POST_STREAMING",
                                      "NO_STREAMING", "DOUBLE_STREAMING"])
def test_native_bgk_spike_matches_lettuce(tag, dtype, strategy):
    """lettuce's tests/native/test_native_bgk_collision.py and
    test_native_streaming_strategy.py case, against lettuce's own values."""
    g = golden("g1_spike_bgk_d2q9.npz")
    for steps in (1, 10):
        f0 = np.ones([9, 16, 16])
        f0[:, 2, 2] = 2.0
        flow = ArrayFlow(ctx(dtype), f0, lt.D2Q9())
        out = native_run(flow, lt.BGKCollision(2.0), steps,
                         getattr(lt.StreamingStrategy, strategy))
        ref = g[f"{tag}_{strategy}_{steps}"]
        tol = 1e-6 if dtype == F32 else 1e-13
        assert np.allclose(out, ref, rtol=tol, atol=0), (strategy, steps)


@pytest.mark.parametrize("name", ["D2Q9", "D3Q19", "D3Q27"])
def test_native_streaming_one_hot(name):
    s, _ = _stencils()[name]
    n = 6
    f0 = np.zeros([s.q] + [n] * s.d)
    x0 = [n - 1] + [2] * (s.d - 1)
    for q in range(s.q):
        f0[(q,) + tuple(x0)] = q + 1.0
    for strategy in (lt.StreamingStrategy.PRE_STREAMING,
                     lt.StreamingStrategy.POST_STREAMING):
        flow = ArrayFlow(ctx(F64), f0, s)
        out = native_run(flow, lt.NoCollision(), 1, strategy)
        for q in range(s.q):
            dst = tuple((x0[a] + s.e[q][a]) % n for a in range(s.d))
            assert out[(q,) + dst] == q + 1.0
            assert out[q].sum() == q + 1.0


def test_tgv2d_1000_steps_matches_lettuce_fp32():
    """D2Q9 TGV 128^2, Re 100, u0 0.05, BGK, 1000 steps in FP32: native vs
    lettuce's torch path (golden) to <= 1e-5 relative max error on f."""
    g = golden("g3_tgv2d_128_bgk_1000.npz")
    flow = ArrayFlow(ctx(F32), g["f0"], lt.D2Q9())
    out = native_run(flow, lt.BGKCollision(float(g["tau"])), 1000,
                     lt.StreamingStrategy.PRE_STREAMING)
    assert rel_err(out, g["f1000"]) <= REL32
    assert elem_err(out, g["f1000"]) <= ELEM32


def test_tgv2d_native_flow_setup_matches_lettuce():
    g = golden("g3_tgv2d_128_bgk_1000.npz")
    flow = lt.TaylorGreenVortex(ctx(F32), 128, 100, 0.05 * np.sqrt(3), lt.D2Q9(),
                                initialize_fneq=False)
    assert flow.units.relaxation_parameter_lu == float(g["tau"])
    assert rel_err(flow.f.cpu().numpy(), g["f0"]) < 1e-6


@pytest.mark.parametrize("name,n", [("D3Q19", 32), ("D3Q27", 16)])
def test_tgv3d_bgk_1000_steps_vs_oracle_fp32(name, n):
    """Re 100 keeps BGK stable at these small grids (at Re 1600, tau - 1/2
    is 5e-4 and both sides diverge)."""
    s, os_ = _stencils()[name]
    oflow = O.TaylorGreenVortex(O.Context("cpu", F32), n, 100, 0.1, os_)
    f0 = oflow.f.numpy().copy()
    tau = oflow.units.relaxation_parameter_lu
    O.Simulation(oflow, O.BGKCollision(tau), [], O.StreamingStrategy.PRE_STREAMING)(1000)
    flow = ArrayFlow(ctx(F32), f0, s)
    out = native_run(flow, lt.BGKCollision(tau), 1000, lt.StreamingStrategy.PRE_STREAMING)
    assert rel_err(out, oflow.f.numpy()) <= REL32
    assert elem_err(out, oflow.f.numpy()) <= ELEM32


@pytest.mark.parametrize("layout", ["rows", "soa"])
def test_tgv3d_two_step_kernel_1000_steps_vs_oracle_fp32(layout):
    """The default D3Q19 path at 64^3 runs the two-step kernel (500
    launches of two fused steps); after 1000 steps it meets the same 1e-5
    bound against the oracle as the single-step kernel."""
    s, os_ = _stencils()["D3Q19"]
    oflow = O.TaylorGreenVortex(O.Context("cpu", F32), 64, 100, 0.1, os_)
    f0 = oflow.f.numpy().copy()
    tau = oflow.units.relaxation_parameter_lu
    O.Simulation(oflow, O.BGKCollision(tau), [], O.StreamingStrategy.PRE_STREAMING)(1000)
    c = lt.Context("cuda:0", dtype=F32, layout=layout)
    flow = ArrayFlow(c, f0, s)
    sim = lt.Simulation(flow, lt.BGKCollision(tau), [], lt.StreamingStrategy.PRE_STREAMING)
    assert sim._collide_and_stream.plan(1000) == (500, 0)
    sim(1000)
    assert rel_err(flow.f.cpu().numpy(), oflow.f.numpy()) <= REL32
    assert elem_err(flow.f.cpu().numpy(), oflow.f.numpy()) <= ELEM32


def test_tgv3d_golden_10_steps():
    g = golden("g2_tgv3d_bgk.npz")
    for name in ("D3Q19", "D3Q27"):
        s, _ = _stencils()[name]
        flow = ArrayFlow(ctx(F32), g[f"{name}_f0"], s)
        out = native_run(flow, lt.BGKCollision(float(g[f"{name}_tau"])), 10,
                         lt.StreamingStrategy.PRE_STREAMING)
        assert rel_err(out, g[f"{name}_f10"]) <= REL32
        assert elem_err(out, g[f"{name}_f10"]) <= ELEM32


# --------------------------------------------------------------- TRT / KBC
def test_trt_matches_lettuce():
    g = golden("g5_trt.npz")
    for name in ("D2Q9", "D3Q19", "D3Q27"):
        s, _ = _stencils()[name]
        flow = ArrayFlow(ctx(F64), g[f"{name}_f0"], s)
        out = native_run(flow, lt.TRTCollision(0.8, 1.3), 10,
                         lt.StreamingStrategy.PRE_STREAMING)
        assert np.allclose(out, g[f"{name}_f10"], rtol=0, atol=1e-13), name


@pytest.mark.parametrize("tag,dtype", [("f32", F32), ("f64", F64)])
def test_kbc_matches_lettuce(tag, dtype):
    g = golden("g4_kbc.npz")
    for name in ("D2Q9", "D3Q27"):
        s, _ = _stencils()[name]
        flow = ArrayFlow(ctx(dtype), g[f"{name}_{tag}_f0"], s)
        assert flow.units.relaxation_parameter_lu == float(g[f"{name}_tau"])
        post = lt.KBCCollision()(flow).cpu().numpy()
        ref = g[f"{name}_{tag}_coll"]
        tol = 1e-5 if dtype == F32 else 1e-12
        etol = ELEM32 if dtype == F32 else 1e-12
        assert rel_err(post, ref) <= tol, name
        assert elem_err(post, ref) <= etol, name
        out = native_run(flow, lt.KBCCollision(), 10, lt.StreamingStrategy.PRE_STREAMING)
        assert rel_err(out, g[f"{name}_{tag}_f10"]) <= tol, name
        assert elem_err(out, g[f"{name}_{tag}_f10"]) <= etol, name


@pytest.mark.parametrize("name,n", [("D2Q9", 64), ("D3Q27", 12)])
def test_kbc_tgv_vs_oracle_fp64(name, n):
    """300 KBC steps of a TGV in FP64.  KBC's stabiliser gamma is a ratio of
    sums of the higher-order part dh, which starts at rounding level, and it
    jumps to 2 below 1e-15 (kbc_collision.py:150-157); long FP32 runs from an
    equilibrium state therefore follow rounding noise and are compared in
    FP64 here (FP32 parity is pinned on well-conditioned states above)."""
    s, os_ = _stencils()[name]
    oflow = O.TaylorGreenVortex(O.Context("cpu", F64), n, 3000, 0.1, os_)
    f0 = oflow.f.numpy().copy()
    steps = 300
    O.Simulation(oflow, O.KBCCollision(), [], O.StreamingStrategy.PRE_STREAMING)(steps)
    flow = lt.TaylorGreenVortex(ctx(F64), n, 3000, 0.1, s, initialize_fneq=False)
    flow.f = torch.from_numpy(f0)
    out = native_run(flow, lt.KBCCollision(), steps, lt.StreamingStrategy.PRE_STREAMING)
    assert rel_err(out, oflow.f.numpy()) <= 1e-8


@pytest.mark.parametrize("name", ["D2Q9", "D3Q19", "D3Q27"])
@pytest.mark.parametrize("coll", ["bgk", "trt", "kbc", "none"])
def test_collision_operator_call_vs_oracle(name, coll):
    """collision(flow) without streaming, FP64, every stencil."""
    if coll == "kbc" and name == "D3Q19":
        pytest.skip("KBC is defined for D2Q9 and D3Q27")
    s, os_ = _stencils()[name]
    f0 = _perturbed(name, 5, F64, seed=11)
    flow = ArrayFlow(ctx(F64), f0, s)
    from tests.test_oracle import _Grid
    oflow = _Grid(os_, [5] * s.d, F64, f0)
    tau = oflow.units.relaxation_parameter_lu
    ours = {"bgk": lt.BGKCollision(tau), "trt": lt.TRTCollision(tau, 0.7),
            "kbc": lt.KBCCollision(), "none": lt.NoCollision()}[coll]
    ref = {"bgk": O.BGKCollision(tau), "trt": O.TRTCollision(tau, 0.7),
           "kbc": O.KBCCollision(), "none": O.NoCollision()}[coll]
    out = ours(flow)
    out = out.cpu().numpy() if isinstance(out, torch.Tensor) else out
    assert np.allclose(out, ref(oflow).numpy(), rtol=0, atol=1e-13)


# --------------------------------------------------------------- boundaries
def test_bounce_back_matches_oracle():
    """lettuce's tests/native/test_native_bounce_back.py case."""
    f0 = np.zeros([9, 16, 16])
    f0[:, 1, 1] = 1.0
    mask = torch.zeros(16, 16, dtype=torch.bool)
    mask[0, :] = mask[:, 0] = True
    mask[2:, :] = mask[:, 2:] = True
    for strategy in ("POST_STREAMING", "PRE_STREAMING"):
        flow = ArrayFlow(ctx(F64), f0, lt.D2Q9(), post=[lt.BounceBackBoundary(mask)])
        sim = lt.Simulation(flow, lt.NoCollision(), [],
                            getattr(lt.StreamingStrategy, strategy))
        from tests.test_oracle import _Grid
        oflow = _Grid(O.D2Q9(), [16, 16], F64, f0)
        oflow.post_boundaries = [O.BounceBackBoundary(mask)]
        osim = O.Simulation(oflow, O.NoCollision(), [],
                            getattr(O.StreamingStrategy, strategy))
        for _ in range(2):
            sim(1)
            osim(1)
            assert np.array_equal(flow.f.cpu().numpy(), oflow.f.numpy())


def _obstacle_native(g, strategy=lt.StreamingStrategy.POST_STREAMING):
    flow = lt.Obstacle(ctx(F64), [48, 24], reynolds_number=50, mach_number=0.1,
                       domain_length_x=4.0, stencil=lt.D2Q9(), outlet="equilibrium")
    flow.mask = g["mask"]
    flow.f = torch.from_numpy(g["f0"])
    sim = lt.Simulation(flow, lt.BGKCollision(float(g["tau"])), [], strategy)
    return flow, sim


def test_obstacle_masks_classes_and_steps_match_lettuce():
    """Obstacle with EquilibriumBoundaryPU inlet, EquilibriumOutletP and
    fullway bounce-back: masks bit-exact, 10 POST steps vs lettuce."""
    g = golden("g6_obstacle_eqoutlet.npz")
    flow, sim = _obstacle_native(g)
    assert np.array_equal(sim.no_collision_mask.cpu().numpy(), g["ncm"])
    assert np.array_equal(sim.no_streaming_mask.cpu().numpy(), g["nsm"])
    ncm, nsm1, _ = expand_classes(sim.cell_classes, 9)
    assert np.array_equal(ncm.numpy(), g["ncm"])
    assert np.array_equal(nsm1.numpy(), (g["nsm"] == 1).astype(np.uint8))
    sim(10)
    assert np.allclose(flow.f.cpu().numpy(), g["f10_post"], rtol=0, atol=1e-12)


def _channel_case(dtype, nx=64, ny=32):
    yy, xx = np.meshgrid(np.arange(ny), np.arange(nx))
    mask = (xx - 16) ** 2 + (yy - 15.5) ** 2 < 5.0 ** 2
    oflow = O.ObstacleChannel(O.Context("cpu", dtype), [nx, ny], 100, 0.1,
                              domain_length_x=nx / 10.0, mask=mask, stencil=O.D2Q9())
    return mask, oflow


@pytest.mark.parametrize("dtype", [F32, F64])
@pytest.mark.parametrize("strategy", ["PRE_STREAMING", "POST_STREAMING"])
def test_halfway_channel_cylinder_vs_oracle(dtype, strategy):
    """Channel with a cylinder: equilibrium inlet, equilibrium outlet,
    halfway bounce-back on the cylinder and the walls; 200 steps; the
    momentum-exchange force too."""
    mask, oflow = _channel_case(dtype)
    f0 = oflow.f.numpy().copy()
    osim = O.Simulation(oflow, O.BGKCollision(oflow.units.relaxation_parameter_lu), [],
                        getattr(O.StreamingStrategy, strategy))
    flow = lt.Obstacle(ctx(dtype), [64, 32], 100, 0.1, domain_length_x=6.4,
                       stencil=lt.D2Q9(), outlet="equilibrium", solid="halfway",
                       walls=True)
    flow.mask = mask
    flow.f = torch.from_numpy(f0)
    hw = flow.post_streaming_boundaries[0]
    sim = lt.Simulation(flow, lt.BGKCollision(flow.units.relaxation_parameter_lu), [],
                        getattr(lt.StreamingStrategy, strategy))
    # the native solid cells are the oracle's
    _, _, solid = expand_classes(sim.cell_classes, 9)
    assert torch.equal(solid, oflow.solid)
    osim(200)
    sim(200)
    out = flow.f.cpu().numpy()
    tol = REL32 if dtype == F32 else 1e-12
    assert rel_err(out, oflow.f.numpy()) <= tol
    assert elem_err(out, oflow.f.numpy()) <= (ELEM32 if dtype == F32 else 1e-12)
    if strategy == "PRE_STREAMING":
        # the FP64 link force of the stored post-collision state against the
        # momentum-exchange sum over the same links and populations (the
        # populations themselves were checked against the oracle above)
        hw_links = sim.halfway_links()
        F = hw.force(sim).cpu().numpy()
        fpost = torch.from_numpy(out).to(F64)
        e = torch.tensor(O.D2Q9.e, dtype=F64)
        cells = hw_links[:, 0].long()
        d = hw_links[:, 1].long()
        vals = fpost.reshape(9, -1)[d, cells]
        ref = (2 * vals[:, None] * e[d]).sum(0).numpy()
        scale = np.abs(ref).max()
        assert np.all(np.abs(F - ref) <= 1e-12 * scale)
        # and against the oracle's populations, to the FP32 trajectory error
        # summed over the links
        vals_o = oflow.f.to(F64).reshape(9, -1)[d, cells]
        ref_o = (2 * vals_o[:, None] * e[d]).sum(0).numpy()
        tol = 1e-9 if dtype == F64 else 3e-5
        assert np.all(np.abs(F - ref_o) <= tol * scale)


# ---------------------------------------------------------- shifted storage
@pytest.mark.parametrize("storage,tdt", [("f16_shifted", torch.float16),
                                         ("bf16_shifted", torch.bfloat16)])
def test_16_bit_encoding_rounds_to_nearest_even(storage, tdt):
    """lbm_encode with rho0 = 0 is the plain FP32 -> 16-bit conversion: bit
    for bit torch's round to nearest even, on random values, exact ties
    (a 16-bit value plus half its unit in the last place, both parities),
    signed zeros and FP32 subnormals."""
    g = torch.Generator().manual_seed(4)
    v = (torch.randn(9, 32, 32, generator=g) * 10.0 ** torch.randint(-6, 2, (9, 32, 32),
                                                                      generator=g)).to(F32)
    mant = 8 if tdt == torch.bfloat16 else 11
    base = v[:4].to(tdt).to(F32)
    half = torch.where(base == 0, torch.zeros_like(base),
                       2.0 ** (torch.floor(torch.log2(base.abs())) - mant))
    v[:4] = base + half                                   # ties
    v[4, 0, :6] = torch.tensor([0.0, -0.0, 1e-39, -3e-40, 1.4e-45, -1e-38])
    f = torch.empty(9, 32, 32, dtype=tdt, device="cuda:0")
    lt._ops.encode(lt.D2Q9(), v, f, 0.0)
    torch.cuda.synchronize()
    ref = v.to(tdt)
    assert torch.equal(f.cpu().view(torch.int16), ref.view(torch.int16))


@pytest.mark.parametrize("storage,tdt", [("f16_shifted", torch.float16),
                                         ("bf16_shifted", torch.bfloat16)])
def test_shifted_storage_one_step_rounding(storage, tdt):
    """One step from a representable state: native = FP32 oracle step on the
    decoded populations, re-encoded with round-to-nearest-even, up to one
    unit in the last place of the 16-bit format."""
    s = lt.D3Q19()
    oflow = O.TaylorGreenVortex(O.Context("cpu", F32), 16, 1600, 0.1, O.D3Q19())
    w = torch.tensor(O.D3Q19.w, dtype=F32).reshape(-1, 1, 1, 1)
    g0 = (oflow.f - w).to(tdt)
    oflow.f = g0.to(F32) + w
    f0 = oflow.f.numpy().copy()
    tau = oflow.units.relaxation_parameter_lu
    O.Simulation(oflow, O.BGKCollision(tau), [], O.StreamingStrategy.PRE_STREAMING)(1)
    ref = ((oflow.f - w).to(tdt).to(F32) + w).numpy()
    flow = ArrayFlow(ctx(F32, storage), f0, s)
    assert np.array_equal(flow.f.cpu().numpy(), f0)   # exact round trip
    out = native_run(flow, lt.BGKCollision(tau), 1, lt.StreamingStrategy.PRE_STREAMING)
    dev = np.abs(ref - w.numpy())
    ulp = np.maximum(dev, 1e-30) * (2.0 ** -10 if tdt == torch.float16 else 2.0 ** -7)
    # plus two FP32 ulps of f ~ 0.4: the arithmetic before the rounding
    floor = 6e-8 + (2.0 ** -24 if tdt == torch.float16 else 0.0)
    assert np.all(np.abs(out - ref) <= ulp + floor)


@pytest.mark.parametrize("storage,bound", [("f16_shifted", 2e-3), ("bf16_shifted", 2e-2)])
def test_shifted_storage_kinetic_energy_vs_fp64(storage, bound):
    """D3Q19 TGV 32^3, 300 steps: kinetic energy of shifted 16-bit storage
    against an FP64 run of the same case."""
    runs = {}
    for key, c in (("ref", ctx(F64)), ("low", ctx(F32, storage))):
        flow = lt.TaylorGreenVortex(c, 32, 1600, 0.1, lt.D3Q19(), initialize_fneq=False)
        ke = lt.IncompressibleKineticEnergy(flow)
        rep = lt.ObservableReporter(ke, interval=100, out=None)
        lt.Simulation(flow, lt.BGKCollision(flow.units.relaxation_parameter_lu), [rep],
                      lt.StreamingStrategy.PRE_STREAMING)(300)
        runs[key] = np.array([r[2] for r in rep.out])
    rel = np.abs(runs["low"] - runs["ref"]) / runs["ref"]
    assert rel.max() < bound


# -------------------------------------------------------------- observables
@pytest.mark.parametrize("name,n", [("D2Q9", 32), ("D3Q19", 16), ("D3Q27", 12)])
def test_observables_fp64_vs_oracle(name, n):
    s, os_ = _stencils()[name]
    oflow = O.TaylorGreenVortex(O.Context("cpu", F64), n, 100, 0.1, os_)
    flow = lt.TaylorGreenVortex(ctx(F64), n, 100, 0.1, s, initialize_fneq=False)
    flow.f = oflow.f
    assert np.isclose(float(lt.IncompressibleKineticEnergy(flow)()),
                      float(O.incompressible_kinetic_energy(oflow)), rtol=1e-12)
    assert np.isclose(float(lt.Enstrophy(flow)()), float(O.enstrophy(oflow)), rtol=1e-11)
    assert np.isclose(float(lt.Mass(flow)()), float(O.mass(oflow)), rtol=1e-13)
    assert np.allclose(flow.u().cpu().numpy(), oflow.u().numpy(), rtol=0, atol=1e-15)
    assert np.allclose(flow.rho().cpu().numpy(), oflow.rho().numpy(), rtol=0, atol=1e-14)


def test_equilibrium_matches_oracle():
    s = lt.D3Q27()
    f0 = _perturbed("D3Q27", 5, F64, seed=2)
    flow = ArrayFlow(ctx(F64), f0, s)
    from tests.test_oracle import _Grid
    oflow = _Grid(O.D3Q27(), [5, 5, 5], F64, f0)
    assert np.allclose(flow.equilibrium(flow).cpu().numpy(),
                       oflow.equilibrium(oflow).numpy(), rtol=0, atol=1e-15)


def test_run_batches_equal_single_steps():
    """lbm_run (many steps per call) equals step-by-step launching."""
    s = lt.D3Q19()
    f0 = _perturbed("D3Q19", 8, F64, seed=4)
    a = ArrayFlow(ctx(F64), f0, s)
    b = ArrayFlow(ctx(F64), f0, s)
    sa = lt.Simulation(a, lt.BGKCollision(0.7), [], lt.StreamingStrategy.PRE_STREAMING)
    sb = lt.Simulation(b, lt.BGKCollision(0.7), [], lt.StreamingStrategy.PRE_STREAMING)
    sa(7)
    for _ in range(7):
        sb._collide_and_stream(sb)
    assert np.array_equal(a.f.cpu().numpy(), b.f.cpu().numpy())


@pytest.mark.parametrize("name,res", [("D3Q19", [8, 6, 10]), ("D2Q9", [12, 2]),
                                      ("D3Q27", [4, 4, 6]), ("D2Q9", [5, 7]),
                                      ("D3Q19", [4, 16, 64]), ("D3Q27", [6, 16, 64])])
@pytest.mark.parametrize("storage,dtype", [("f32", F32), ("f64", F64),
                                           ("bf16_shifted", F32)])
def test_split_post_streaming_run_matches_push_steps(name, res, storage, dtype, knobs):
    """lettuce's default order (collide, then stream) run as one collide-only
    launch, pull steps (two-step kernel where it applies: the 16 x 64 grids)
    and one double-streaming launch gives the populations of the same number
    of push steps: bit for bit without collision, else to the rounding of
    the collision arithmetic, which the compiler may contract differently in
    different kernels"""
    s, os_ = _stencils()[name]
    g = torch.Generator().manual_seed(9)
    from tests.test_oracle import _Grid
    oflow = _Grid(os_, res, F64)
    u = 0.05 * (torch.rand([s.d] + res, generator=g, dtype=F64) - 0.5)
    f0 = oflow.equilibrium(oflow, rho=torch.ones([1] + res, dtype=F64), u=u)
    f0 = (f0 * (1 + 0.02 * torch.rand(f0.shape, generator=g, dtype=F64))).numpy()
    ulp = {"f32": 2e-7, "f64": 1e-15, "bf16_shifted": 1e-4}[storage]
    for make in (lambda: lt.NoCollision(),
                 lambda: lt.TRTCollision(0.7, 0.9) if name == "D3Q19" else lt.KBCCollision()):
        outs = []
        for split in (0, 1):
            knobs(post_split=split)
            flow = ArrayFlow(ctx(dtype, storage), f0, s)
            sim = lt.Simulation(flow, make(), [], lt.StreamingStrategy.POST_STREAMING)
            pairs, singles = sim._collide_and_stream.plan(9)
            assert 2 * pairs + singles == 9 and (pairs == 0 or split == 1)
            sim(9)
            outs.append(flow._state.float().cpu().numpy() if storage == "bf16_shifted"
                        else flow._state.cpu().numpy())
        if isinstance(make(), lt.NoCollision):
            assert np.array_equal(outs[0], outs[1])
        else:
            # KBC's gamma is a ratio of small sums: a rounding difference moves
            # it further than in BGK / TRT
            tol = 9 * ulp * (50 if name != "D3Q19" else 1)
            assert np.abs(outs[0] - outs[1]).max() <= tol


def test_split_post_streaming_channel_matches_push_steps(knobs):
    """A channel with a cylinder (halfway bounce-back, equilibrium inlet and
    outlet), 12 post-streaming steps: split run against push steps"""
    outs = []
    for split in (0, 1):
        knobs(post_split=split)
        flow = lt.Obstacle(lt.Context("cuda:0", dtype=F64), [96, 40], 100, 0.1,
                           domain_length_x=96 / 16, stencil=lt.D2Q9(), outlet="equilibrium",
                           solid="halfway", walls=True)
        yy, xx = np.meshgrid(np.arange(40), np.arange(96))
        flow.mask = (xx - 24) ** 2 + (yy - 21) ** 2 < 6.0 ** 2
        flow.initialize()
        sim = lt.Simulation(flow, lt.BGKCollision(flow.units.relaxation_parameter_lu), [],
                            lt.StreamingStrategy.POST_STREAMING)
        sim(12)
        outs.append(flow.f.cpu().numpy())
    assert np.abs(outs[0] - outs[1]).max() <= 1e-14


@pytest.mark.parametrize("name,storage,dtype", [("D3Q19", "f32", F32), ("D2Q9", "f64", F64),
                                                ("D3Q27", "f16_shifted", F32)])
def test_rows_layout_matches_soa(name, storage, dtype):
    """Population-interleaved rows give the same populations bit for bit,
    observables included; boundaries are refused in that layout."""
    s, os_ = _stencils()[name]
    res = [10, 12] if s.d == 2 else [6, 8, 10]
    outs, kes = [], []
    for layout in ("soa", "rows"):
        c = lt.Context("cuda:0", dtype=dtype, storage=storage, layout=layout)
        flow = lt.TaylorGreenVortex(c, res, 100, 0.1, s, initialize_fneq=False)
        if layout == "rows":
            assert flow._state.stride(0) == res[-1]
        lt.Simulation(flow, lt.TRTCollision(0.7, 0.9), [], lt.StreamingStrategy.PRE_STREAMING)(7)
        outs.append(flow.f.cpu().numpy())
        kes.append(float(lt.IncompressibleKineticEnergy(flow)()))
    assert np.array_equal(outs[0], outs[1])
    assert kes[0] == kes[1]
    c = lt.Context("cuda:0", layout="rows")
    flow = lt.Obstacle(c, [64, 16], 100, 0.1, domain_length_x=4.0, stencil=lt.D2Q9(),
                       outlet="equilibrium")
    with pytest.raises(RuntimeError, match="contiguous"):
        lt.Simulation(flow, lt.BGKCollision(0.6), [])(1)

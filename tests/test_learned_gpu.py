"""GPU parity of the learned collision (csrc/lbm_learned.hip) against the
oracle's LearnedCollision, plus its two reductions: an MLP that outputs 0
gives BGK, and rates outside the stable range give KBC."""
import numpy as np
import pytest
import torch

import lettuce_amd as lt
from oracle import lettuce_torch as O
from tests.test_parity_gpu import ArrayFlow, ctx, elem_err, rel_err, F32, F64
from tests.test_oracle import _Grid, _perturbed

pytestmark = pytest.mark.gpu


def pair(seed=0, out_scale=1.0, **kw):
    p = lt.LearnedParameters.seeded(seed, out_scale=out_scale, **kw)
    op = O.LearnedParameters(*[t.clone() for t in p.tensors()], in_scale=p.in_scale,
                             rate_lo=p.rate_lo, rate_hi=p.rate_hi,
                             stable_lo=p.stable_lo, stable_hi=p.stable_hi)
    return p, op


def zero_params(**kw):
    H = 32
    z = torch.zeros
    return lt.LearnedParameters(z(H, 6), z(H), z(H, H), z(H), z(3, H), z(3), **kw)


def test_packing_roundtrip_shapes():
    p, _ = pair()
    buf = p.packed(torch.device("cuda:0"))
    assert buf.numel() == 17408 and buf.data_ptr() % 16 == 0


def test_zero_mlp_is_bgk():
    """o = 0 gives every higher-order rate 1/tau: the update is BGK."""
    f0 = _perturbed(O.D2Q9(), 64, F32, seed=3).f.numpy()
    flow = ArrayFlow(ctx(F32), f0, lt.D2Q9())
    tau = flow.units.relaxation_parameter_lu
    out = lt.LearnedCollision(zero_params())(flow).cpu().numpy()
    ref = lt.BGKCollision(tau)(flow).cpu().numpy()
    assert rel_err(out, ref) < 2e-6


def test_unstable_rates_fall_back_to_kbc():
    p = zero_params()
    p.b3 = torch.full((3,), 5.0)  # raw rates 1/tau + 5 > stable_hi
    f0 = _perturbed(O.D2Q9(), 64, F32, seed=4).f.numpy()
    flow = ArrayFlow(ctx(F32), f0, lt.D2Q9())
    out = lt.LearnedCollision(p)(flow).cpu().numpy()
    ref = lt.KBCCollision()(flow).cpu().numpy()
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("seed", [0, 1])
def test_learned_collision_vs_oracle(seed):
    p, op = pair(seed)
    oflow = _perturbed(O.D2Q9(), 64, F32, seed=5 + seed)
    f0 = oflow.f.numpy().copy()
    ref_coll = O.LearnedCollision(op)
    ref = ref_coll(oflow).numpy()
    _, fb = ref_coll.rates(oflow)
    flow = ArrayFlow(ctx(F32), f0, lt.D2Q9())
    out = lt.LearnedCollision(p)(flow).cpu().numpy()
    assert rel_err(out, ref) <= 1e-5
    assert elem_err(out, ref) <= 5e-5
    print("fallback fraction", float(fb.float().mean()))


def test_learned_tgv_100_steps_vs_oracle():
    p, op = pair(0, out_scale=0.1)
    oflow = O.TaylorGreenVortex(O.Context("cpu", F32), 64, 300, 0.1, O.D2Q9())
    f0 = oflow.f.numpy().copy()
    O.Simulation(oflow, O.LearnedCollision(op), [], O.StreamingStrategy.PRE_STREAMING)(100)
    flow = lt.TaylorGreenVortex(ctx(F32), 64, 300, 0.1, lt.D2Q9(), initialize_fneq=False)
    flow.f = torch.from_numpy(f0)
    lt.Simulation(flow, lt.LearnedCollision(p), [], lt.StreamingStrategy.PRE_STREAMING)(100)
    assert rel_err(flow.f.cpu().numpy(), oflow.f.numpy()) <= 1e-5
    assert elem_err(flow.f.cpu().numpy(), oflow.f.numpy()) <= 5e-5


@pytest.mark.parametrize("storage", ["f32", "f16_shifted"])
def test_learned_channel_cylinder_vs_oracle(storage):
    """halfway bounce-back, equilibrium inlet and outlet around the learned
    collision; FP32 storage against the oracle, shifted FP16 storage
    against the FP32 native run."""
    p, op = pair(0, out_scale=0.1)
    nx, ny = 128, 64
    yy, xx = np.meshgrid(np.arange(ny), np.arange(nx))
    mask = (xx - 32) ** 2 + (yy - 31.5) ** 2 < 8.0 ** 2
    oflow = O.ObstacleChannel(O.Context("cpu", F32), [nx, ny], 100, 0.1,
                              domain_length_x=nx / 16.0, mask=mask, stencil=O.D2Q9())
    f0 = oflow.f.numpy().copy()
    O.Simulation(oflow, O.LearnedCollision(op), [], O.StreamingStrategy.PRE_STREAMING)(100)
    c = ctx(F32, storage)
    flow = lt.Obstacle(c, [nx, ny], 100, 0.1, domain_length_x=nx / 16.0, stencil=lt.D2Q9(),
                       outlet="equilibrium", solid="halfway", walls=True)
    flow.mask = mask
    flow.f = torch.from_numpy(f0)
    lt.Simulation(flow, lt.LearnedCollision(p), [], lt.StreamingStrategy.PRE_STREAMING)(100)
    err = rel_err(flow.f.cpu().numpy(), oflow.f.numpy())
    assert err <= (1e-5 if storage == "f32" else 2e-3)
    # entrywise: measured 3.9e-6 (FP32), 7.7e-4 (FP16-shifted)
    assert elem_err(flow.f.cpu().numpy(), oflow.f.numpy()) <= (5e-5 if storage == "f32" else 5e-3)


# ------------------------------------------------------------------ D3Q27
def _f27(res=(6, 8, 64), seed=3, amp=0.05):
    g = torch.Generator().manual_seed(seed)
    res = list(res)
    oflow = _Grid(O.D3Q27(), res, F64)
    u = 0.05 * (torch.rand([3] + res, generator=g, dtype=F64) - 0.5)
    rho = 1 + 0.02 * (torch.rand([1] + res, generator=g, dtype=F64) - 0.5)
    f0 = oflow.equilibrium(oflow, rho=rho, u=u)
    return (f0 * (1 + amp * (torch.rand(f0.shape, generator=g, dtype=F64) * 2 - 1))).float().numpy()


def test_d3q27_zero_mlp_is_bgk():
    flow = ArrayFlow(ctx(F32), _f27(), lt.D3Q27())
    tau = flow.units.relaxation_parameter_lu    # 0.51: 1/tau above the default rate_hi
    out = lt.LearnedCollision(zero_params(rate_hi=1.999))(flow).cpu().numpy()
    ref = lt.BGKCollision(tau)(flow).cpu().numpy()
    # FP32 rounding of the 3 x 3-point moment transforms and back
    assert rel_err(out, ref) < 5e-6


@pytest.mark.parametrize("storage", ["f32", "bf16_shifted"])
def test_d3q27_unstable_rates_fall_back_to_kbc(storage):
    """every cell takes the KBC update: native KBC to the rounding of the
    collision arithmetic, which the two kernels each compile (the learned
    one may contract the same formulas into FMAs differently).  With
    shifted BF16 storage the learned kernel works on the stored values and
    decodes them for the fallback: the two then differ by at most one BF16
    unit in the last place of a stored value (2^-7 of the largest, ~0.018
    here: ~4e-4 of the largest population)."""
    p = zero_params()
    p.b3 = torch.full((3,), 5.0)
    flow = ArrayFlow(ctx(F32, storage), _f27(seed=4), lt.D3Q27())
    out = lt.LearnedCollision(p)(flow).cpu().numpy()
    ref = lt.KBCCollision()(flow).cpu().numpy()
    err = rel_err(out, ref)
    print(f"learned fallback vs KBC ({storage}): {err:.3e}")
    assert err <= (2e-6 if storage == "f32" else 1e-3)


def test_d3q27_zero_mlp_is_bgk_in_bf16():
    """the stored-value path of shifted BF16 storage (moments of g = f - w
    rho0 plus the moments of the shift) against the BGK kernel on the same
    stored values, to one BF16 unit in the last place of the largest stored
    value (~4e-4 of the largest population here)"""
    flow = ArrayFlow(ctx(F32, "bf16_shifted"), _f27(), lt.D3Q27())
    tau = flow.units.relaxation_parameter_lu
    out = lt.LearnedCollision(zero_params(rate_hi=1.999))(flow).cpu().numpy()
    ref = lt.BGKCollision(tau)(flow).cpu().numpy()
    err = rel_err(out, ref)
    print(f"learned zero MLP vs BGK (bf16): {err:.3e}")
    assert err <= 1e-3


@pytest.mark.parametrize("seed", [0, 1])
def test_d3q27_learned_collision_vs_oracle(seed):
    p, op = pair(seed, out_scale=0.3)
    f0 = _f27(seed=5 + seed)
    oflow = _Grid(O.D3Q27(), list(f0.shape[1:]), F32, f0)
    ref = O.LearnedCollision(op)(oflow).numpy()
    flow = ArrayFlow(ctx(F32), f0, lt.D3Q27())
    out = lt.LearnedCollision(p)(flow).cpu().numpy()
    assert rel_err(out, ref) <= 1e-5
    assert elem_err(out, ref) <= 5e-5


@pytest.mark.parametrize("storage", ["f32", "bf16_shifted"])
def test_d3q27_learned_30_steps_vs_oracle(storage):
    """pull streaming + the learned collision for 30 steps; FP32 storage
    against the oracle, shifted BF16 storage against the native FP32 run"""
    p, op = pair(0, out_scale=0.1)
    f0 = _f27(seed=8, amp=0.01)
    oflow = _Grid(O.D3Q27(), list(f0.shape[1:]), F32, f0)
    O.Simulation(oflow, O.LearnedCollision(op), [], O.StreamingStrategy.PRE_STREAMING)(30)
    flow = ArrayFlow(ctx(F32, storage), f0, lt.D3Q27())
    lt.Simulation(flow, lt.LearnedCollision(p), [], lt.StreamingStrategy.PRE_STREAMING)(30)
    err = rel_err(flow.f.cpu().numpy(), oflow.f.numpy())
    assert err <= (1e-5 if storage == "f32" else 2e-2), err
    # entrywise: measured at most 1.1e-5 (FP32) and 2.2e-3 (BF16-shifted)
    assert elem_err(flow.f.cpu().numpy(), oflow.f.numpy()) <= (5e-5 if storage == "f32" else 1e-2)


@pytest.mark.parametrize("storage", ["f32", "bf16_shifted"])
@pytest.mark.parametrize("strategy", ["PRE_STREAMING", "POST_STREAMING"])
def test_d3q27_learned_with_bounce_back_vs_oracle(strategy, storage):
    """the masked learned kernel (D3Q27 works in raw-moment space): a
    fullway bounce-back sphere in a periodic box, 20 steps against the
    oracle's learned collision with lettuce's BounceBackBoundary; shifted
    BF16 storage (whose boundary cells decode and re-encode around the
    bounce-back) to the BF16 tolerance of the unmasked test"""
    p, op = pair(1, out_scale=0.1)
    res = [8, 16, 64]
    f0 = _f27(res=res, seed=12, amp=0.01)
    x, y, z = torch.meshgrid(*[torch.arange(n, dtype=F64) for n in res], indexing="ij")
    mask = (x - 3.5) ** 2 + (y - 7.5) ** 2 + (z - 20.0) ** 2 < 3.2 ** 2
    oflow = _Grid(O.D3Q27(), res, F32, f0)
    oflow.post_boundaries = [O.BounceBackBoundary(mask)]
    O.Simulation(oflow, O.LearnedCollision(op), [],
                 getattr(O.StreamingStrategy, strategy))(20)
    flow = ArrayFlow(ctx(F32, storage), f0, lt.D3Q27(), post=[lt.BounceBackBoundary(mask)])
    lt.Simulation(flow, lt.LearnedCollision(p), [], getattr(lt.StreamingStrategy, strategy))(20)
    err = rel_err(flow.f.cpu().numpy(), oflow.f.numpy())
    assert err <= (1e-5 if storage == "f32" else 2e-2), err
    # entrywise: measured at most 1.1e-5 (FP32) and 2.2e-3 (BF16-shifted)
    assert elem_err(flow.f.cpu().numpy(), oflow.f.numpy()) <= (5e-5 if storage == "f32" else 1e-2)


@pytest.mark.parametrize("case", ["d3q27_bf16", "d3q27_f32", "d2q9_channel_f16"])
def test_wide_learned_kernel_matches_the_narrow_one(case, knobs):
    """the wide learned kernel (64-bit plane offsets, which grids of more
    than 2^31 elements use) forced on a small grid, against the narrow one:
    equal to the rounding of the collision arithmetic"""
    p, _ = pair(0, out_scale=0.1)
    out = {}
    for wide in ("0", "1"):
        knobs(wide_offsets=wide)
        if case.startswith("d3q27"):
            storage = "f32" if case.endswith("f32") else "bf16_shifted"
            flow = ArrayFlow(ctx(F32, storage), _f27(seed=8, amp=0.01), lt.D3Q27())
        else:
            nx, ny = 128, 64
            yy, xx = np.meshgrid(np.arange(ny), np.arange(nx))
            c = ctx(F32, "f16_shifted")
            flow = lt.Obstacle(c, [nx, ny], 100, 0.1, domain_length_x=nx / 16.0,
                               stencil=lt.D2Q9(), outlet="equilibrium", solid="halfway",
                               walls=True)
            flow.mask = (xx - 32) ** 2 + (yy - 31.5) ** 2 < 8.0 ** 2
            flow.initialize()
        lt.Simulation(flow, lt.LearnedCollision(p), [], lt.StreamingStrategy.PRE_STREAMING)(10)
        out[wide] = flow.f.cpu().numpy()
    err = rel_err(out["1"], out["0"])
    print(f"wide vs narrow learned ({case}): {err:.3e}")
    assert err <= (1e-6 if case == "d3q27_f32" else 1e-3), err


def test_rebinding_a_learned_collision_keeps_the_first_binding_alive():
    """A second binding of one LearnedCollision (collision(flow), another
    Simulation) must not free the weight buffer the first Simulation's
    kernels read: after the rebind and fresh allocations that could reuse
    freed memory, the first simulation steps bit for bit like one built from
    a fresh collision."""
    p, _ = pair(0, out_scale=0.5)

    def tgv():
        return lt.TaylorGreenVortex(ctx(F32), 64, 300, 0.1, lt.D2Q9(), initialize_fneq=False)

    coll = lt.LearnedCollision(p)
    a = tgv()
    sim_a = lt.Simulation(a, coll, [], lt.StreamingStrategy.PRE_STREAMING)
    coll(tgv())                                  # second binding, then dropped
    lt.Simulation(tgv(), coll, [], lt.StreamingStrategy.PRE_STREAMING)
    torch.cuda.synchronize()
    junk = [torch.full((17408 // 4,), float(k + 7), device="cuda:0") for k in range(64)]
    sim_a(20)
    b = tgv()
    lt.Simulation(b, lt.LearnedCollision(lt.LearnedParameters.seeded(0, out_scale=0.5)), [],
                  lt.StreamingStrategy.PRE_STREAMING)(20)
    assert torch.equal(a._state.cpu(), b._state.cpu())
    del junk

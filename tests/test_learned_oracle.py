"""The oracle's learned collision and its differentiable training step,
pinned on the CPU: the dual vectors against the moment rows, the reductions
to BGK (rates 1/tau, zero MLP output), forced rates against the inference
formula, conservation, and autograd against central differences."""
import numpy as np
import pytest
import torch

from oracle import lettuce_torch as O

F64 = torch.float64


def _grid(n=6, batch=None, seed=0, amp=0.05):
    """D2Q9 populations near equilibrium [9, (batch,) n, n] (FP64)"""
    g = torch.Generator().manual_seed(seed)
    shape = ([batch] if batch else []) + [n, n]
    u = 0.05 * (torch.rand([2] + shape, generator=g, dtype=F64) - 0.5)
    rho = 1 + 0.02 * (torch.rand(shape, generator=g, dtype=F64) - 0.5)
    e = torch.tensor(O.D2Q9.e, dtype=F64)
    w = torch.tensor(O.D2Q9.w, dtype=F64).reshape([9] + [1] * len(shape))
    eu = torch.einsum("qd,d...->q...", e, u)
    feq = w * rho * (1 + 3 * eu - 1.5 * (u * u).sum(0) + 4.5 * eu * eu)
    r = torch.rand(feq.shape, generator=g, dtype=F64) * 2 - 1
    return feq * (1 + amp * r)


def test_mlp_in_the_log2_domain_is_the_plain_mlp():
    """The oracle evaluates the MLP as the inference kernel does, with every
    pre-activation z carried as -log2(e) z: without the bf16 roundings it is
    W3 silu(W2 silu(W1 x + b1) + b2) + b3 evaluated directly (FP64, to
    rounding), and with them it stays within their error (measured 2.8e-3
    of the largest output, 4.8e-3 for the same roundings on unscaled
    values)."""
    p = O.LearnedParameters.seeded(3, out_scale=1.0)
    lc = O.LearnedCollision(p)
    x = torch.randn(6, 4000, generator=torch.Generator().manual_seed(0), dtype=F64) * 3
    w1, b1, w2, b2, w3, b3 = [t.double() for t in (p.w1, p.b1, p.w2, p.b2, p.w3, p.b3)]
    silu = torch.nn.functional.silu
    ref = w3 @ silu(w2 @ silu(w1 @ x + b1[:, None]) + b2[:, None]) + b3[:, None]
    scale = ref.abs().max()
    assert (lc.mlp(x, rnd=lambda t: t) - ref).abs().max() <= 1e-14 * scale
    assert (lc.mlp(x.float()).double() - ref).abs().max() <= 1e-2 * scale


class _Units:
    def __init__(self, tau):
        self.relaxation_parameter_lu = tau


class _Flow(O.Flow):
    """bare D2Q9 flow whose units only carry tau"""

    def __init__(self, f, tau):
        super().__init__(O.Context("cpu", F64), list(f.shape[1:]), _Units(tau), O.D2Q9())
        self.f = f.clone()


def _theta(seed=0, out_scale=0.1):
    p = O.LearnedParameters.seeded(seed, out_scale=out_scale)
    return torch.cat([t.reshape(-1) for t in (p.w1, p.b1, p.w2, p.b2, p.w3, p.b3)]).to(F64)


def test_phi_is_dual_to_the_moment_rows():
    rows = O.d2q9_moment_rows(F64)
    phi = torch.tensor([O.PHI_XXY, O.PHI_XYY, O.PHI_XXYY], dtype=F64)
    e = torch.tensor(O.D2Q9.e, dtype=F64)
    conserved = torch.stack([torch.ones(9, dtype=F64), e[:, 0], e[:, 1]])
    assert torch.equal(rows[3:] @ phi.T, torch.eye(3, dtype=F64))
    assert torch.equal(rows[:3] @ phi.T, torch.zeros(3, 3, dtype=F64))
    assert torch.equal(conserved @ phi.T, torch.zeros(3, 3, dtype=F64))


def test_forced_rates_one_over_tau_is_bgk():
    tau = 0.61
    flow = _Flow(_grid(), tau)
    lc = O.LearnedCollision(O.LearnedParameters.seeded(0))
    lc.force_rates = [1 / tau] * 3
    torch.testing.assert_close(lc(flow), O.BGKCollision(tau)(flow), rtol=0, atol=1e-15)


def test_train_collide_zero_mlp_is_bgk():
    tau = 0.57
    f = _grid(batch=2)
    out = O.learned_train_collide(f, torch.zeros(O.LEARNED_PARAMS, dtype=F64), tau)
    for b in range(2):
        bgk = O.BGKCollision(tau)(_Flow(f[:, b], tau))
        torch.testing.assert_close(out[:, b], bgk, rtol=0, atol=1e-15)


def test_train_collide_constant_rates_match_inference_formula():
    """w3 = 0, b3 = c: the MLP output is c exactly, so the training step must
    equal the inference formula with forced rates 1/tau + c"""
    tau = 0.6
    c = torch.tensor([0.1, -0.2, 0.25], dtype=F64)   # rates inside [0.05, 1.95]
    theta = _theta(3)
    theta[-3 - 3 * 32:-3] = 0
    theta[-3:] = c
    f = _grid(seed=4)
    lc = O.LearnedCollision(O.LearnedParameters.seeded(0))
    lc.force_rates = (1 / tau + c).tolist()
    ref = lc(_Flow(f, tau))
    out = O.learned_train_collide(f, theta, tau)
    torch.testing.assert_close(out, ref, rtol=0, atol=1e-15)


def test_train_collide_conserves_mass_and_momentum():
    f = _grid(batch=3, seed=2)
    out = O.learned_train_collide(f, _theta(1, out_scale=1.0), 0.53)
    e = torch.tensor(O.D2Q9.e, dtype=F64)
    torch.testing.assert_close(out.sum(0), f.sum(0), rtol=0, atol=1e-14)
    torch.testing.assert_close(torch.einsum("qd,q...->d...", e, out),
                               torch.einsum("qd,q...->d...", e, f), rtol=0, atol=1e-14)


def test_train_step_streams_like_the_simulation():
    """learned_train_step = pull streaming + collision; with theta = 0 it is
    the oracle Simulation's PRE_STREAMING BGK step"""
    tau = 0.7
    f = _grid(n=5, seed=7)
    flow = _Flow(f, tau)
    O.Simulation(flow, O.BGKCollision(tau), [], O.StreamingStrategy.PRE_STREAMING)(1)
    out = O.learned_train_step(f[:, None], torch.zeros(O.LEARNED_PARAMS, dtype=F64), tau)
    torch.testing.assert_close(out[:, 0], flow.f, rtol=0, atol=1e-15)


@pytest.mark.parametrize("clamped", [False, True])
def test_train_step_autograd_matches_central_differences(clamped):
    tau = 0.58
    theta = _theta(5, out_scale=0.3)
    if clamped:   # push the first rate above rate_hi: its gradient is zero
        theta[-3] = 5.0
    f0 = _grid(n=5, batch=2, seed=9)
    g = torch.Generator().manual_seed(1)
    wgt = torch.rand(f0.shape, generator=g, dtype=F64)

    def loss(f, th):
        for _ in range(3):
            f = O.learned_train_step(f, th, tau)
        return (wgt * f).sum() + 10 * (f.sum(0) ** 2).sum()

    th = theta.clone().requires_grad_(True)
    ff = f0.clone().requires_grad_(True)
    loss(ff, th).backward()
    for which, x, grad in (("theta", theta, th.grad), ("f", f0, ff.grad)):
        d = torch.randn(x.shape, generator=g, dtype=F64)
        h = 1e-6
        num = (loss(*((f0 + h * d, theta) if which == "f" else (f0, theta + h * d)))
               - loss(*((f0 - h * d, theta) if which == "f" else (f0, theta - h * d)))) / (2 * h)
        ana = (grad * d).sum()
        assert abs(float(num - ana)) <= 1e-6 * max(1.0, abs(float(ana))), (which, num, ana)
    if clamped:
        assert float(th.grad[-3]) == 0.0
        assert torch.all(th.grad[-3 - 96:-3 - 64] == 0)


# ------------------------------------------------------------------ D3Q27
def _grid27(n=4, seed=3, amp=0.05):
    g = torch.Generator().manual_seed(seed)
    shape = [n, n, n]
    u = 0.05 * (torch.rand([3] + shape, generator=g, dtype=F64) - 0.5)
    rho = 1 + 0.02 * (torch.rand(shape, generator=g, dtype=F64) - 0.5)
    e = torch.tensor(O.D3Q27.e, dtype=F64)
    w = torch.tensor(O.D3Q27.w, dtype=F64).reshape([27, 1, 1, 1])
    eu = torch.einsum("qd,d...->q...", e, u)
    feq = w * rho * (1 + 3 * eu - 1.5 * (u * u).sum(0) + 4.5 * eu * eu)
    r = torch.rand(feq.shape, generator=g, dtype=F64) * 2 - 1
    return feq * (1 + amp * r)


class _Flow27(O.Flow):
    def __init__(self, f, tau):
        super().__init__(O.Context("cpu", F64), list(f.shape[1:]), _Units(tau), O.D3Q27())
        self.f = f.clone()


def test_d3q27_dual_basis_is_the_tensor_product_of_1d_duals():
    """inverse of the raw-moment matrix = products of the 1D duals of
    (1, x, x^2) at x = -1, 0, 1: (x^2 - x)/2, 1 - x^2, (x^2 + x)/2"""
    M, orders = O.d3q27_raw_moments()
    phi = torch.linalg.inv(M)
    W = {-1: [0.0, -0.5, 0.5], 0: [1.0, 0.0, -1.0], 1: [0.0, 0.5, 0.5]}
    e = O.D3Q27.e
    closed = torch.tensor([[W[e[q][0]][k // 9] * W[e[q][1]][(k // 3) % 3] * W[e[q][2]][k % 3]
                            for k in range(27)] for q in range(27)], dtype=F64)
    torch.testing.assert_close(phi, closed, rtol=0, atol=1e-14)
    assert orders.count(3) == 7 and orders.count(4) == 6 and orders.count(5) == 3


def test_d3q27_shear_part_carries_exactly_the_second_order_moments():
    """ds (the KBC shear part from raw moments) has the second-order moments
    of df and no moment of order 0, 1 or >= 3, so dh = df - ds lives in the
    orders 3..6 that the three rates act on"""
    f = _grid27()
    flow = _Flow27(f, 0.6)
    params = O.LearnedParameters.seeded(0)
    params.rate_lo = 0.0
    lc = O.LearnedCollision(params)
    lc.force_rates = [0.0, 0.0, 0.0]
    tau = 0.6
    # with all rates 0 the update is f - ds / tau: recover ds
    ds = (f - lc(flow)) * tau
    feq = flow.equilibrium(flow)
    M, orders = O.d3q27_raw_moments()
    m_ds = torch.einsum("kq,q...->k...", M, ds)
    m_df = torch.einsum("kq,q...->k...", M, f - feq)
    for k, o in enumerate(orders):
        if o == 2:
            torch.testing.assert_close(m_ds[k], m_df[k], rtol=0, atol=1e-14)
        else:
            assert float(m_ds[k].abs().max()) <= 1e-14, (k, o)


def test_d3q27_rates_one_over_tau_is_bgk_and_beta_gamma_is_kbc():
    tau = 0.57
    f = _grid27(seed=5)
    flow = _Flow27(f, tau)
    lc = O.LearnedCollision(O.LearnedParameters.seeded(0))
    lc.force_rates = [1 / tau] * 3
    torch.testing.assert_close(lc(flow), O.BGKCollision(tau)(flow), rtol=0, atol=1e-14)
    kbc = O.KBCCollision()
    kbc.force_gamma = 1.7
    beta = 1 / (2 * tau)
    lc.force_rates = [beta * 1.7] * 3
    torch.testing.assert_close(lc(flow), kbc(flow), rtol=0, atol=1e-13)


def test_d3q27_learned_conserves_mass_and_momentum():
    f = _grid27(seed=7)
    flow = _Flow27(f, 0.6)
    lc = O.LearnedCollision(O.LearnedParameters.seeded(2, out_scale=0.5))
    post = lc(flow)
    e = torch.tensor(O.D3Q27.e, dtype=F64)
    torch.testing.assert_close(post.sum(0), f.sum(0), rtol=0, atol=1e-14)
    torch.testing.assert_close(torch.einsum("qd,q...->d...", e, post),
                               torch.einsum("qd,q...->d...", e, f), rtol=0, atol=1e-14)


def test_d3q27_train_collide_zero_mlp_is_bgk():
    tau = 0.57
    f = _grid27(seed=6)
    out = O.learned_train_collide_d3q27(f, torch.zeros(O.LEARNED_PARAMS, dtype=F64), tau)
    torch.testing.assert_close(out, O.BGKCollision(tau)(_Flow27(f, tau)), rtol=0, atol=1e-15)


def test_d3q27_train_collide_constant_rates_match_inference_formula():
    """w3 = 0, b3 = c: the training step equals the D3Q27 inference formula
    with forced rates 1/tau + c (that formula reduces to KBC at beta gamma)"""
    tau = 0.6
    c = torch.tensor([0.1, -0.2, 0.25], dtype=F64)
    theta = _theta(3)
    theta[-3 - 3 * 32:-3] = 0
    theta[-3:] = c
    f = _grid27(seed=8)
    lc = O.LearnedCollision(O.LearnedParameters.seeded(0))
    lc.force_rates = (1 / tau + c).tolist()
    ref = lc(_Flow27(f, tau))
    torch.testing.assert_close(O.learned_train_collide_d3q27(f, theta, tau), ref, rtol=0,
                               atol=1e-15)


def test_d3q27_train_collide_conserves_mass_and_momentum():
    f = _grid27(seed=2)
    out = O.learned_train_collide_d3q27(f, _theta(1, out_scale=1.0), 0.53)
    e = torch.tensor(O.D3Q27.e, dtype=F64)
    torch.testing.assert_close(out.sum(0), f.sum(0), rtol=0, atol=1e-14)
    torch.testing.assert_close(torch.einsum("qd,q...->d...", e, out),
                               torch.einsum("qd,q...->d...", e, f), rtol=0, atol=1e-14)


def test_d3q27_train_step_streams_like_the_simulation():
    tau = 0.7
    f = _grid27(n=3, seed=7)
    flow = _Flow27(f, tau)
    O.Simulation(flow, O.BGKCollision(tau), [], O.StreamingStrategy.PRE_STREAMING)(1)
    out = O.learned_train_step_d3q27(f, torch.zeros(O.LEARNED_PARAMS, dtype=F64), tau)
    torch.testing.assert_close(out, flow.f, rtol=0, atol=1e-15)


def test_d3q27_train_step_autograd_matches_central_differences():
    tau = 0.58
    theta = _theta(5, out_scale=0.3)
    f0 = _grid27(n=3, seed=9)
    g = torch.Generator().manual_seed(1)
    wgt = torch.rand(f0.shape, generator=g, dtype=F64)

    def loss(f, th):
        for _ in range(2):
            f = O.learned_train_step_d3q27(f, th, tau)
        return (wgt * f).sum() + 10 * (f.sum(0) ** 2).sum()

    th = theta.clone().requires_grad_(True)
    ff = f0.clone().requires_grad_(True)
    loss(ff, th).backward()
    for which, x, grad in (("theta", theta, th.grad), ("f", f0, ff.grad)):
        d = torch.randn(x.shape, generator=g, dtype=F64)
        h = 1e-6
        num = (loss(*((f0 + h * d, theta) if which == "f" else (f0, theta + h * d)))
               - loss(*((f0 - h * d, theta) if which == "f" else (f0, theta - h * d)))) / (2 * h)
        ana = (grad * d).sum()
        assert abs(float(num - ana)) <= 1e-6 * max(1.0, abs(float(ana))), (which, num, ana)

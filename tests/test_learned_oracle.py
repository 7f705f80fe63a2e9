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

"""GPU checks of the learned-collision training path (csrc/lbm_train.hip and
lettuce_amd.ext.LearnedRollout): the forward step against the oracle's
learned_train_step, the hand-written adjoint against torch autograd through
the oracle, FP64 central differences on a 16-step unroll with checkpoints
every 4, the clamp's zero gradient, and checkpoint-independence."""
import pytest
import torch

import lettuce_amd as lt
from oracle import lettuce_torch as O
from tests.test_learned_oracle import _grid, _theta

pytestmark = pytest.mark.gpu

F32, F64 = torch.float32, torch.float64
DEV = torch.device("cuda:0")


def _cfg(tau):
    return lt.TrainConfig(tau=tau, in_scale=100.0, rate_lo=0.05, rate_hi=1.95)


def _rollout(f0, theta, tau, steps, every=4):
    return lt.learned_rollout(f0.to(DEV).contiguous(), theta.to(DEV), _cfg(tau), steps, every)


def _oracle_rollout(f0, theta, tau, steps):
    f = f0
    for _ in range(steps):
        f = O.learned_train_step(f, theta, tau)
    return f


@pytest.mark.parametrize("dtype,tol", [(F64, 1e-13), (F32, 1e-5)])
@pytest.mark.parametrize("cells", [2048, 5000, 70000])
def test_gradient_reduction_matches_the_products(dtype, tol, cells):
    """lbm_learned_grad_reduce against the FP64 products on the CPU, in the
    parameter layout [W1, b1, W2, b2, W3, b3]: dW1 = dz1 x^T, dW2 = dz2 a1^T,
    dW3 = do a2^T and the biases the row sums of dz1, dz2, do, over the
    cells of a random record, onto a non-zero gradient, for cell counts
    that fill whole 2048-cell blocks, leave a partial block, and give many
    blocks; two calls in a row with the same workspace add the sums twice"""
    from lettuce_amd import _native as N
    g = torch.Generator().manual_seed(cells)
    rec = torch.randn(N.LEARNED_RECORD, cells, generator=g, dtype=F64).to(dtype)
    grad0 = torch.randn(N.LEARNED_PARAMS, generator=g, dtype=F64).to(dtype)
    r = rec.double()
    x, a1, a2 = r[0:6], r[6:38], r[38:70]
    dz1, dz2, do = r[70:102], r[102:134], r[134:137]
    sums = torch.cat([(dz1 @ x.t()).reshape(-1), dz1.sum(1), (dz2 @ a1.t()).reshape(-1),
                      dz2.sum(1), (do @ a2.t()).reshape(-1), do.sum(1)])
    assert sums.numel() == N.LEARNED_PARAMS
    storage = N.F64 if dtype == F64 else N.F32
    lib = N.lib()
    nbytes = lib.lbm_learned_grad_workspace(storage, cells)
    ws = torch.empty(nbytes // rec.element_size(), dtype=dtype, device=DEV)
    rec_d = rec.to(DEV).contiguous()
    grad = grad0.to(DEV).contiguous()
    stream = N.stream_handle(DEV)
    for _ in range(2):
        N.check(lib.lbm_learned_grad_reduce(storage, cells, rec_d.data_ptr(), ws.data_ptr(),
                                            grad.data_ptr(), stream), "lbm_learned_grad_reduce")
    ref = grad0.double() + 2 * sums
    err = (grad.cpu().double() - ref).abs().max() / ref.abs().max()
    assert err <= tol, err


@pytest.mark.parametrize("steps", [1, 7])
def test_forward_matches_oracle_f64(steps):
    tau = 0.58
    f0 = _grid(n=16, batch=2, seed=1)[..., :12].contiguous()    # 16 x 12, not square
    theta = _theta(2, out_scale=0.5)
    got = _rollout(f0, theta, tau, steps).cpu()
    ref = _oracle_rollout(f0, theta, tau, steps)
    err = (got - ref).abs().max().item()
    assert err <= 1e-14, err


def test_forward_f32_close_to_f64_oracle():
    tau = 0.58
    f0 = _grid(n=32, batch=3, seed=2)
    theta = _theta(4, out_scale=0.5)
    got = _rollout(f0.to(F32), theta.to(F32), tau, 16).cpu().double()
    ref = _oracle_rollout(f0, theta, tau, 16)
    rel = ((got - ref).abs().max() / ref.abs().max()).item()
    assert rel <= 1e-5, rel


def test_unpull_is_the_transpose_of_pull():
    f = torch.rand(9, 2, 8, 12, dtype=F64, device=DEV)
    from lettuce_amd.ext._training import _Problem
    prob = _Problem(f, torch.zeros(1379, dtype=F64, device=DEV), _cfg(0.6))
    lam = torch.empty_like(f)
    prob.unpull(f, lam)
    e = O.D2Q9.e
    ref = torch.stack([torch.roll(f[q], shifts=(-e[q][0], -e[q][1]), dims=(-2, -1))
                       for q in range(9)])
    assert torch.equal(lam, ref)


def _loss(f, wgt):
    rho, u = lt.velocity_d2q9(f)
    return (wgt * f).sum() + 10.0 * (rho ** 2).sum() + 100.0 * (u ** 2).sum()


@pytest.mark.parametrize("steps,every", [(6, 4), (5, 1), (3, 8)])
def test_adjoint_matches_torch_autograd(steps, every):
    tau = 0.56
    f0 = _grid(n=12, batch=2, seed=5)[..., :10].contiguous()
    theta = _theta(6, out_scale=0.4)
    wgt = torch.rand(f0.shape, dtype=F64, generator=torch.Generator().manual_seed(3))
    # oracle: torch autograd through the CPU restatement
    f_r = f0.clone().requires_grad_(True)
    th_r = theta.clone().requires_grad_(True)
    _loss(_oracle_rollout(f_r, th_r, tau, steps), wgt).backward()
    # GPU: the hand-written adjoint
    f_g = f0.to(DEV).requires_grad_(True)
    th_g = theta.to(DEV).requires_grad_(True)
    _loss(lt.learned_rollout(f_g, th_g, _cfg(tau), steps, every), wgt.to(DEV)).backward()
    for name, a, b in (("theta", th_g.grad.cpu(), th_r.grad), ("f0", f_g.grad.cpu(), f_r.grad)):
        rel = ((a - b).abs().max() / b.abs().max()).item()
        assert rel <= 1e-10, (name, rel)


def test_adjoint_central_differences_16_steps():
    """FP64, 16-step unroll with checkpoints every 4; central differences
    with a relative step of 1e-4 agree with the adjoint to 1e-3"""
    tau = 0.55
    f0 = _grid(n=16, batch=2, seed=8).to(DEV)
    theta = _theta(7, out_scale=0.5).to(DEV)
    target = _grid(n=16, batch=2, seed=9).to(DEV)
    _, u_t = lt.velocity_d2q9(target)

    def loss(th, f=f0):
        fN = lt.learned_rollout(f, th, _cfg(tau), 16, 4)
        _, u = lt.velocity_d2q9(fN)
        return ((u - u_t) ** 2).mean() * 1e4

    th = theta.clone().requires_grad_(True)
    f = f0.clone().requires_grad_(True)
    loss(th, f).backward()
    g = torch.Generator(device=DEV).manual_seed(11)
    h = 1e-4
    # a direction that perturbs every parameter by 1e-4 of its own size
    d = theta * torch.randn(theta.shape, generator=g, dtype=F64, device=DEV)
    with torch.no_grad():
        num = (loss(theta + h * d) - loss(theta - h * d)) / (2 * h)
    ana = (th.grad * d).sum()
    assert abs((num - ana) / ana).item() <= 1e-3, (num.item(), ana.item())
    # single parameters: first-layer weight, second-layer weight, output bias
    for i in (0, 6 * 32 + 32 + 5 * 32 + 7, 1379 - 2):
        e = torch.zeros_like(theta)
        e[i] = theta[i].abs().clamp_min(1e-3)
        with torch.no_grad():
            num = (loss(theta + h * e) - loss(theta - h * e)) / (2 * h)
        ana = th.grad[i] * e[i]
        assert abs(num - ana).item() <= 1e-3 * max(abs(ana).item(), 1e-8), (i, num, ana)
    # the initial populations
    d = f0 * 1e-2 * torch.randn(f0.shape, generator=g, dtype=F64, device=DEV)
    with torch.no_grad():
        num = (loss(theta, f0 + h * d) - loss(theta, f0 - h * d)) / (2 * h)
    ana = (f.grad * d).sum()
    assert abs((num - ana) / ana).item() <= 1e-3, (num.item(), ana.item())


def test_clamped_rate_has_zero_gradient():
    tau = 0.6
    f0 = _grid(n=8, batch=1, seed=3).to(DEV)
    theta = _theta(1, out_scale=0.2).to(DEV)
    theta[-3] = 5.0          # raw rate of xxy far above rate_hi everywhere
    th = theta.clone().requires_grad_(True)
    _loss(lt.learned_rollout(f0, th, _cfg(tau), 4, 2), torch.ones_like(f0)).backward()
    assert th.grad[-3].item() == 0.0
    assert torch.all(th.grad[-3 - 96:-3 - 64] == 0)    # row 0 of w3
    assert th.grad[-2].item() != 0.0


def test_checkpoint_interval_does_not_change_gradients():
    tau = 0.57
    f0 = _grid(n=16, batch=2, seed=4).to(DEV)
    theta = _theta(2, out_scale=0.3).to(DEV)
    wgt = torch.rand(f0.shape, dtype=F64, device=DEV)
    grads = []
    for every in (1, 4, 16):
        th = theta.clone().requires_grad_(True)
        _loss(lt.learned_rollout(f0, th, _cfg(tau), 16, every), wgt).backward()
        grads.append(th.grad)
    assert torch.equal(grads[0], grads[1]) and torch.equal(grads[1], grads[2])


def test_f32_gradient_close_to_f64():
    tau = 0.57
    f0 = _grid(n=32, batch=4, seed=6)
    theta = _theta(3, out_scale=0.3)
    wgt = torch.rand(f0.shape, dtype=F64)
    out = {}
    for dt in (F64, F32):
        th = theta.to(DEV, dt).requires_grad_(True)
        fN = lt.learned_rollout(f0.to(DEV, dt), th, _cfg(tau), 16, 4)
        _loss(fN, wgt.to(DEV, dt)).backward()
        out[dt] = th.grad.double().cpu()
    rel = ((out[F32] - out[F64]).norm() / out[F64].norm()).item()
    assert rel <= 1e-3, rel


def test_training_reduces_the_loss():
    """a few Adam steps on a fixed target lower the loss (the gradient
    points downhill)"""
    tau = 0.56
    f0 = _grid(n=32, batch=2, seed=10).to(DEV, F32)
    target = lt.learned_rollout(f0, _theta(9, out_scale=1.0).to(DEV, F32), _cfg(tau), 8, 4)
    _, u_t = lt.velocity_d2q9(target)
    th = _theta(3, out_scale=0.1).to(DEV, F32).requires_grad_(True)
    opt = torch.optim.Adam([th], lr=3e-3)
    losses = []
    for _ in range(30):
        opt.zero_grad()
        _, u = lt.velocity_d2q9(lt.learned_rollout(f0, th, _cfg(tau), 8, 4))
        loss = ((u - u_t) ** 2).mean() * 1e6
        loss.backward()
        opt.step()
        losses.append(loss.item())
    assert losses[-1] < 0.5 * losses[0], losses


# ------------------------------------------------------------------ D3Q27
from tests.test_learned_oracle import _grid27  # noqa: E402


def _oracle_rollout27(f0, theta, tau, steps):
    f = f0
    for _ in range(steps):
        f = O.learned_train_step_d3q27(f, theta, tau)
    return f


@pytest.mark.parametrize("steps", [1, 5])
def test_d3q27_forward_matches_oracle_f64(steps):
    """the D3Q27 training forward (raw-moment formulation) against the
    oracle's (KBC shear split and dual basis), FP64, a 6 x 5 x 4 field"""
    tau = 0.58
    f0 = _grid27(n=6, seed=1)[:, :, :5, :4].contiguous()
    theta = _theta(2, out_scale=0.5)
    got = _rollout(f0, theta, tau, steps).cpu()
    ref = _oracle_rollout27(f0, theta, tau, steps)
    err = (got - ref).abs().max().item()
    assert err <= 1e-14, err


def test_d3q27_unpull_is_the_transpose_of_pull():
    f = torch.rand(27, 5, 4, 6, dtype=F64, device=DEV)
    from lettuce_amd.ext._training import _Problem
    prob = _Problem(f, torch.zeros(1379, dtype=F64, device=DEV), _cfg(0.6))
    lam = torch.empty_like(f)
    prob.unpull(f, lam)
    e = O.D3Q27.e
    ref = torch.stack([torch.roll(f[q], shifts=tuple(-c for c in e[q]), dims=(0, 1, 2))
                       for q in range(27)])
    assert torch.equal(lam, ref)


@pytest.mark.parametrize("steps,every", [(4, 4), (5, 2)])
def test_d3q27_adjoint_matches_torch_autograd(steps, every):
    """the D3Q27 hand-written adjoint against torch autograd through the
    oracle's training step, FP64"""
    tau = 0.56
    f0 = _grid27(n=5, seed=5)[:, :, :, :4].contiguous()
    theta = _theta(6, out_scale=0.4)
    wgt = torch.rand(f0.shape, dtype=F64, generator=torch.Generator().manual_seed(3))

    def loss(f, w):
        e = torch.tensor(O.D3Q27.e, dtype=F64, device=f.device)
        rho = f.sum(0)
        u = torch.einsum("qd,q...->d...", e, f) / rho
        return (w * f).sum() + 10.0 * (rho ** 2).sum() + 100.0 * (u ** 2).sum()

    f_r = f0.clone().requires_grad_(True)
    th_r = theta.clone().requires_grad_(True)
    loss(_oracle_rollout27(f_r, th_r, tau, steps), wgt).backward()
    f_g = f0.to(DEV).requires_grad_(True)
    th_g = theta.to(DEV).requires_grad_(True)
    loss(lt.learned_rollout(f_g, th_g, _cfg(tau), steps, every), wgt.to(DEV)).backward()
    for name, a, b in (("theta", th_g.grad.cpu(), th_r.grad), ("f0", f_g.grad.cpu(), f_r.grad)):
        rel = ((a - b).abs().max() / b.abs().max()).item()
        assert rel <= 1e-10, (name, rel)


def test_d3q27_adjoint_central_differences_16_steps():
    """FP64, 16-step unroll with checkpoints every 4, an 8^3 field: central
    differences with a relative step of 1e-4 agree with the adjoint to 1e-3"""
    tau = 0.55
    f0 = _grid27(n=8, seed=8).to(DEV)
    theta = _theta(7, out_scale=0.5).to(DEV)
    target = _grid27(n=8, seed=9).to(DEV)
    e = torch.tensor(O.D3Q27.e, dtype=F64, device=DEV)

    def vel(f):
        return torch.einsum("qd,q...->d...", e, f) / f.sum(0)

    u_t = vel(target)

    def loss(th):
        return ((vel(lt.learned_rollout(f0, th, _cfg(tau), 16, 4)) - u_t) ** 2).mean() * 1e4

    th = theta.clone().requires_grad_(True)
    loss(th).backward()
    g = torch.Generator(device=DEV).manual_seed(11)
    h = 1e-4
    d = theta * torch.randn(theta.shape, generator=g, dtype=F64, device=DEV)
    with torch.no_grad():
        num = (loss(theta + h * d) - loss(theta - h * d)) / (2 * h)
    ana = (th.grad * d).sum()
    assert abs((num - ana) / ana).item() <= 1e-3, (num.item(), ana.item())

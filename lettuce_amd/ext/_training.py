"""Training of the learned collision through a hand-written adjoint.

lettuce has no trainable collision; this is the gradient path of the
learned collision of ._learned (the MLP 6 -> 32 -> 32 -> 3 that sets the
relaxation rates of the moments xxy, xyy, xxyy on top of KBC's shear split,
lettuce/ext/_collision/kbc_collision.py:76-160).

LearnedRollout is a torch.autograd.Function over `steps` periodic pull
steps f_{n+1} = C(S f_n) of a batch of D2Q9 fields [9, B, X, Y] or of one
D3Q27 field [27, X, Y, Z].  Its backward
does not record a graph; it runs the adjoint kernels of csrc/lbm_train.hip:

  mu_n          = (dC/dg)^T lambda_{n+1}     per cell (lbm_learned_train_backward)
  lambda_n,q(x) = mu_n,q(x + e_q)            transpose of the pull, fused into
                                             the next backward launch
  dL/dtheta    += sum over cells of outer products of the per-cell record,
                  one GEMM per step

States are kept every `checkpoint_every` steps and the steps between two
kept states are recomputed during the backward pass, so the memory is
(steps / checkpoint_every + checkpoint_every) population arrays instead of
`steps`.

The forward formula is the inference kernel's with the MLP evaluated in
the populations' dtype (FP32, or FP64 for gradient checks) instead of bf16
and without the per-cell KBC fallback; the clamp of the rates to
[rate_lo, rate_hi] passes the gradient inside the interval and blocks it
outside, as torch.clamp does.
"""
import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from .. import _native as N
from ._learned import LearnedParameters

__all__ = ["LearnedRollout", "learned_rollout", "flatten_parameters",
           "unflatten_parameters", "TrainConfig", "velocity_d2q9",
           "equilibrium_d2q9", "spectral_velocity_2d", "spectral_downsample"]

H = N.LEARNED_HIDDEN
# offsets in the flat parameter vector and in the per-cell record
_P_SIZES = [H * 6, H, H * H, H, 3 * H, 3]
_P_OFF = np.concatenate([[0], np.cumsum(_P_SIZES)]).tolist()
_R_X, _R_A1, _R_A2, _R_DZ1, _R_DZ2, _R_DO = 0, 6, 6 + H, 6 + 2 * H, 6 + 3 * H, 6 + 4 * H
assert _P_OFF[-1] == N.LEARNED_PARAMS and _R_DO + 3 == N.LEARNED_RECORD


def flatten_parameters(params: LearnedParameters, dtype=torch.float32, device="cpu"):
    """[w1, b1, w2, b2, w3, b3] of `params` as one vector"""
    return torch.cat([t.reshape(-1) for t in params.tensors()]).to(dtype=dtype, device=device)


def unflatten_parameters(theta, like: LearnedParameters = None) -> LearnedParameters:
    """LearnedParameters (FP32, CPU) from a flat vector; rate limits and
    in_scale are taken from `like` when given"""
    t = theta.detach().to("cpu", torch.float32)
    parts = [t[_P_OFF[i]:_P_OFF[i + 1]] for i in range(6)]
    w1, b1, w2, b2, w3, b3 = parts
    kw = {}
    if like is not None:
        kw = dict(in_scale=like.in_scale, rate_lo=like.rate_lo, rate_hi=like.rate_hi,
                  stable_lo=like.stable_lo, stable_hi=like.stable_hi)
    return LearnedParameters(w1.reshape(H, 6), b1, w2.reshape(H, H), b2, w3.reshape(3, H),
                             b3, **kw)


@dataclass(frozen=True)
class TrainConfig:
    tau: float
    in_scale: float = 100.0
    rate_lo: float = 0.05
    rate_hi: float = 1.95

    @staticmethod
    def of(params: LearnedParameters, tau: float) -> "TrainConfig":
        return TrainConfig(float(tau), params.in_scale, params.rate_lo, params.rate_hi)


_STORAGE = {torch.float32: N.F32, torch.float64: N.F64}


class _Problem:
    """stencil, extents, storage code and stream of one rollout"""

    def __init__(self, f: torch.Tensor, theta: torch.Tensor, cfg: TrainConfig):
        N.require_device_tensor(f, "f")
        if f.dtype not in _STORAGE:
            raise RuntimeError("training runs on float32 or float64 populations")
        if f.shape[0] == 9 and f.dim() in (3, 4):
            self.stencil = N.D2Q9
            shape = f.shape[1:] if f.dim() == 4 else (1,) + tuple(f.shape[1:])
        elif f.shape[0] == 27 and f.dim() == 4:
            self.stencil = N.D3Q27
            shape = tuple(f.shape[1:])
        else:
            raise RuntimeError("populations must be [9, X, Y] or [9, B, X, Y] (D2Q9) "
                               "or [27, X, Y, Z] (D3Q27)")
        if not f.is_contiguous():
            raise RuntimeError("populations must be contiguous")
        if theta.dtype != f.dtype or theta.device != f.device or theta.numel() != N.LEARNED_PARAMS:
            raise RuntimeError(f"theta must hold {N.LEARNED_PARAMS} values of the populations' "
                               "dtype on their device")
        self.n = (ctypes.c_int64 * 3)(*[int(v) for v in shape])
        self.cells = int(shape[0] * shape[1] * shape[2])
        self.storage = _STORAGE[f.dtype]
        self.cfg = cfg
        self.stream = N.stream_handle(f.device)
        self.theta = theta.detach().contiguous()

    def step(self, src, dst):
        c = self.cfg
        N.check(N.lib().lbm_learned_train_forward(
            self.stencil, self.storage, self.n, c.tau, self.theta.data_ptr(), c.in_scale, c.rate_lo,
            c.rate_hi, src.data_ptr(), dst.data_ptr(), self.stream),
            "lbm_learned_train_forward")

    def adjoint(self, f, lam, gathered, mu, record):
        c = self.cfg
        N.check(N.lib().lbm_learned_train_backward(
            self.stencil, self.storage, self.n, c.tau, self.theta.data_ptr(), c.in_scale, c.rate_lo,
            c.rate_hi, f.data_ptr(), lam.data_ptr(), int(gathered), mu.data_ptr(),
            record.data_ptr() if record is not None else None, self.stream),
            "lbm_learned_train_backward")

    def grad_workspace(self):
        """the per-block partial sums of lbm_learned_grad_reduce"""
        nbytes = N.lib().lbm_learned_grad_workspace(self.storage, self.cells)
        if nbytes < 0:
            raise RuntimeError("lbm_learned_grad_workspace rejected the arguments")
        return torch.empty(nbytes // self.theta.element_size(), dtype=self.theta.dtype,
                           device=self.theta.device)

    def grad_reduce(self, record, workspace, grad):
        """grad [LEARNED_PARAMS] += the parameter gradient of one record"""
        N.check(N.lib().lbm_learned_grad_reduce(self.storage, self.cells, record.data_ptr(),
                                                workspace.data_ptr(), grad.data_ptr(),
                                                self.stream), "lbm_learned_grad_reduce")

    def unpull(self, mu, lam):
        N.check(N.lib().lbm_unpull(self.stencil, self.storage, self.n, mu.data_ptr(), lam.data_ptr(),
                                   self.stream), "lbm_unpull")


class LearnedRollout(torch.autograd.Function):
    """f_steps = (C S)^steps f0 with gradients for f0 and theta."""

    @staticmethod
    def forward(ctx, f0, theta, cfg: TrainConfig, steps: int, checkpoint_every: int = 4):
        if steps < 1 or checkpoint_every < 1:
            raise ValueError("steps and checkpoint_every must be positive")
        f0 = f0.detach().contiguous()
        prob = _Problem(f0, theta, cfg)
        kept = []
        a, b = f0, torch.empty_like(f0)
        for n in range(steps):
            if n % checkpoint_every == 0:
                kept.append(a.clone())
            prob.step(a, b)
            a, b = b, (torch.empty_like(f0) if n == 0 else a)
        ctx.prob, ctx.kept, ctx.steps, ctx.every = prob, kept, steps, checkpoint_every
        return a

    @staticmethod
    def backward(ctx, grad_out):
        prob, steps, every = ctx.prob, ctx.steps, ctx.every
        lam = grad_out.to(prob.theta.dtype).contiguous()
        gathered = False
        mu = [torch.empty_like(lam), torch.empty_like(lam)]
        record = torch.empty((N.LEARNED_RECORD, prob.cells), dtype=lam.dtype, device=lam.device)
        need_theta = ctx.needs_input_grad[1]
        # the parameter gradient, summed over the cells of every step's
        # record by lbm_learned_grad_reduce
        grad_theta = torch.zeros(N.LEARNED_PARAMS, dtype=lam.dtype, device=lam.device) \
            if need_theta else None
        workspace = prob.grad_workspace() if need_theta else None
        k = 0
        for seg in reversed(range(len(ctx.kept))):
            start = seg * every
            length = min(every, steps - start)
            states = [ctx.kept[seg]]
            for _ in range(length - 1):
                nxt = torch.empty_like(states[-1])
                prob.step(states[-1], nxt)
                states.append(nxt)
            for i in reversed(range(length)):
                out = mu[k & 1]
                prob.adjoint(states[i], lam, gathered, out, record if need_theta else None)
                if need_theta:
                    prob.grad_reduce(record, workspace, grad_theta)
                lam, gathered, k = out, True, k + 1
            states.clear()
        grad_f0 = None
        if ctx.needs_input_grad[0]:
            grad_f0 = torch.empty_like(lam)
            prob.unpull(lam, grad_f0)
        return grad_f0, grad_theta, None, None, None


def learned_rollout(f0, theta, cfg: TrainConfig, steps=16, checkpoint_every=4):
    """`steps` learned-collision steps of f0 [9, (B,) X, Y] or [27, X, Y, Z]
    with the hand-written adjoint as backward"""
    return LearnedRollout.apply(f0, theta, cfg, steps, checkpoint_every)


_E = [[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1], [1, 1], [-1, 1], [-1, -1], [1, -1]]
_W = [4 / 9] + [1 / 9] * 4 + [1 / 36] * 4


def velocity_d2q9(f):
    """(rho [...], u [2, ...]) of D2Q9 populations f [9, ...] (differentiable)"""
    e = torch.tensor(_E, dtype=f.dtype, device=f.device)
    rho = f.sum(0)
    return rho, torch.einsum("qd,q...->d...", e, f) / rho


def equilibrium_d2q9(rho, u):
    """D2Q9 quadratic equilibrium (lettuce/ext/_equilibrium/quadratic_equilibrium.py:11-24)
    for rho [...] and u [2, ...]"""
    e = torch.tensor(_E, dtype=u.dtype, device=u.device)
    w = torch.tensor(_W, dtype=u.dtype, device=u.device).reshape([9] + [1] * rho.dim())
    eu = torch.einsum("qd,d...->q...", e, u)
    uu = (u * u).sum(0)
    return w * rho * (1 + 3 * eu - 1.5 * uu + 4.5 * eu * eu)


def spectral_velocity_2d(n, seed, k0=4.0, u_rms=0.04):
    """Periodic solenoidal 2D velocity [2, n, n] (FP64 numpy) with energy
    spectrum E(k) ~ k^4 exp(-2 (k/k0)^2) (lettuce's DecayingTurbulence
    spectrum, lettuce/ext/_flows/decayingturbulence.py:80), random phases from
    numpy.random.default_rng(seed), scaled to sqrt(<|u|^2>) = u_rms.

    The field is the curl of a stream function, u = (d psi/dy, -d psi/dx),
    with |psi_hat(k)| = sqrt(E(k) / (pi k)) / k, so that the shell energy
    pi k |u_hat|^2 follows E(k).  Its spectrum falls below 1e-100 of its peak
    beyond |k| = 12 k0, so for n >= 32 k0 the field on the grid n / 2 is
    exactly its every-other-point sample (see spectral_downsample)."""
    rng = np.random.default_rng(seed)
    k = np.fft.fftfreq(n, 1.0 / n)
    kx, ky = np.meshgrid(k, k, indexing="ij")
    kk = np.hypot(kx, ky)
    amp = np.zeros_like(kk)
    nz = kk > 0
    e = kk[nz] ** 4 * np.exp(-2.0 * (kk[nz] / k0) ** 2)
    amp[nz] = np.sqrt(e / (np.pi * kk[nz])) / kk[nz]
    psi_hat = amp * np.exp(2j * np.pi * rng.random((n, n)))
    ux = np.real(np.fft.ifft2(1j * ky * psi_hat))
    uy = np.real(np.fft.ifft2(-1j * kx * psi_hat))
    u = np.stack([ux, uy])
    return u * (u_rms / np.sqrt((u * u).sum(0).mean()))


def spectral_downsample(u, factor=2):
    """Filter a periodic field [..., N, N] to the wavenumbers the grid
    N / factor resolves (|k_x|, |k_y| < N / (2 factor); the coarse Nyquist
    modes are dropped) and sample it there: the sharp spectral cut-off
    filter followed by exact downsampling."""
    n_f = u.shape[-1]
    if u.shape[-2] != n_f or n_f % (2 * factor):
        raise ValueError("square fields with N divisible by 2 * factor")
    n_c = n_f // factor
    U = torch.fft.fft2(u)
    kc = torch.fft.fftfreq(n_c, 1.0 / n_c, device=u.device).round().long()
    idx = kc % n_f
    Uc = U.index_select(-2, idx).index_select(-1, idx) / factor ** 2
    keep = (kc.abs() < n_c // 2).to(Uc.real.dtype)
    Uc = Uc * keep[:, None] * keep[None, :]
    return torch.fft.ifft2(Uc).real

"""Accuracy studies on the synthetic flows, each against an FP64 run of the
same case on the GPU.  Prints one JSON object per study.

    python -m lettuce_amd.studies tgv2d        # D2Q9 TGV 128^2 vs analytic decay
    python -m lettuce_amd.studies tgv3d        # D3Q19 TGV 256^3, FP32 / FP16 / BF16 storage
    python -m lettuce_amd.studies cylinder     # D2Q9 channel + cylinder, KBC vs learned
    python -m lettuce_amd.studies turbulence3d # D3Q27 KBC decaying turbulence 384^3
    python -m lettuce_amd.studies train2d      # training of the learned collision
    python -m lettuce_amd.studies train3d      # D3Q27 training on KBC states
    python -m lettuce_amd.studies slab         # one device's x-slab vs the periodic step
    python -m lettuce_amd.studies all

Options: --steps N overrides a study's step count; --res N its resolution.
"""
import argparse
import ctypes
import json
import math
import sys
import time

import numpy as np
import torch

import lettuce_amd as lt
from lettuce_amd import _native as N


def _ke_series(sim, flow, steps, every):
    """(steps, KE, enstrophy) every `every` steps, FP64 reductions."""
    ke = lt.IncompressibleKineticEnergy(flow)
    en = lt.Enstrophy(flow)
    out = [(0, float(ke()), float(en()))]
    t = 0
    while t < steps:
        k = min(every, steps - t)
        sim(k)
        t += k
        out.append((t, float(ke()), float(en())))
    return np.array(out)


def tgv2d(steps=1000, res=128):
    """D2Q9 BGK Taylor-Green vortex, Re 100, u0 = 0.05 (Ma = 0.05 sqrt 3),
    FP32; kinetic energy against the analytic exp(-4 nu t) (lettuce's 2D
    TGV velocity decays as exp(-2 nu t), taylorgreen.py:72-81) and against an
    FP64 run."""
    res_ = {}
    for name, dtype in (("f32", torch.float32), ("f64", torch.float64)):
        flow = lt.TaylorGreenVortex(lt.Context("cuda:0", dtype=dtype), res, 100,
                                    0.05 * math.sqrt(3), lt.D2Q9(), initialize_fneq=False)
        sim = lt.Simulation(flow, lt.BGKCollision(flow.units.relaxation_parameter_lu), [],
                            lt.StreamingStrategy.PRE_STREAMING)
        res_[name] = _ke_series(sim, flow, steps, 100)
        units = flow.units
    t_pu = units.convert_time_to_pu(res_["f32"][:, 0])
    analytic = res_["f32"][0, 1] * np.exp(-4 * units.viscosity_pu * t_pu)
    return {"study": "tgv2d_d2q9_bgk", "grid": [res, res], "steps": steps,
            "ke_ratio_f32": float(res_["f32"][-1, 1] / res_["f32"][0, 1]),
            "ke_ratio_analytic": float(analytic[-1] / analytic[0]),
            "ke_rel_err_vs_analytic_max": float(np.max(np.abs(res_["f32"][:, 1] / analytic - 1))),
            "ke_rel_err_f32_vs_f64_max": float(np.max(np.abs(res_["f32"][:, 1] / res_["f64"][:, 1] - 1)))}


def tgv3d(steps=2000, res=256):
    """D3Q19 BGK TGV, Re 1600, Ma 0.1: FP32 storage and shifted FP16/BF16
    storage (FP32 arithmetic; BF16 rounded to nearest and stochastically)
    against FP64; KE and enstrophy every 50 steps."""
    series, timing = {}, {}
    for name, dtype, storage, rounding in (
            ("f64", torch.float64, "f64", "nearest"), ("f32", torch.float32, "f32", "nearest"),
            ("f16_shifted", torch.float32, "f16_shifted", "nearest"),
            ("bf16_shifted", torch.float32, "bf16_shifted", "nearest"),
            ("bf16_shifted_sr", torch.float32, "bf16_shifted", "stochastic")):
        ctx = lt.Context("cuda:0", dtype=dtype, storage=storage, rounding=rounding)
        flow = lt.TaylorGreenVortex(ctx, res, 1600, 0.1, lt.D3Q19(), initialize_fneq=False)
        sim = lt.Simulation(flow, lt.BGKCollision(flow.units.relaxation_parameter_lu), [],
                            lt.StreamingStrategy.PRE_STREAMING)
        t0 = time.perf_counter()
        series[name] = _ke_series(sim, flow, steps, 50)
        torch.cuda.synchronize()
        timing[name] = time.perf_counter() - t0
        del sim, flow
        torch.cuda.empty_cache()
    ref = series["f64"]
    out = {"study": "tgv3d_d3q19_bgk", "grid": [res] * 3, "steps": steps, "every": 50}
    for name in ("f32", "f16_shifted", "bf16_shifted", "bf16_shifted_sr"):
        s = series[name]
        out[name] = {"ke_rel_err_max": float(np.max(np.abs(s[:, 1] / ref[:, 1] - 1))),
                     "enstrophy_rel_err_max": float(np.max(np.abs(s[:, 2] / ref[:, 2] - 1))),
                     "ke_final": float(s[-1, 1]), "enstrophy_final": float(s[-1, 2]),
                     "wall_s_incl_reporting": round(timing[name], 3)}
    out["f64"] = {"ke_final": float(ref[-1, 1]), "enstrophy_final": float(ref[-1, 2])}
    return out


def _strouhal(lift, every, diameter, u_lu):
    """St = f D / U from the lift signal (samples `every` steps apart) over
    its second half: the peak of the zero-padded Hann-windowed spectrum,
    refined by a least-squares sine fit a sin(2 pi f t) + b cos(2 pi f t) + c
    over frequencies within one spectral bin of the peak.  NaN for a signal
    shorter than 16 samples or without variation."""
    x = np.asarray(lift[len(lift) // 2:], dtype=np.float64)
    n = len(x)
    if n < 16 or not np.all(np.isfinite(x)) or np.ptp(x) == 0:
        return float("nan")
    x = x - x.mean()
    pad = 16 * n
    spec = np.abs(np.fft.rfft(x * np.hanning(n), pad))
    k = 1 + int(np.argmax(spec[1:]))
    f0 = k / pad                                      # cycles per sample
    t = np.arange(n, dtype=np.float64)

    def residual(f):
        A = np.stack([np.sin(2 * np.pi * f * t), np.cos(2 * np.pi * f * t), np.ones(n)], 1)
        coef = np.linalg.lstsq(A, x, rcond=None)[0]
        return float(np.sum((A @ coef - x) ** 2))
    grid = np.linspace(max(f0 - 1.0 / n, 0.5 / n), f0 + 1.0 / n, 401)
    f = grid[int(np.argmin([residual(g) for g in grid]))]
    return float(f / every * diameter / u_lu)


def cylinder(steps=200000, nx=2048, ny=512, diameter=64, every=50, storage="f32",
             collision="kbc", seed=0, dtype=torch.float32):
    """D2Q9 channel with a cylinder of `diameter` cells centred at
    (nx/4, ny/2 + 1): halfway bounce-back on cylinder and walls, equilibrium
    inlet u = (0.05, 0), equilibrium outlet, Re 200 on the diameter.
    Drag and lift coefficients by momentum exchange every `every` steps,
    Strouhal number from the lift signal."""
    ctx = lt.Context("cuda:0", dtype=dtype, storage=storage)
    u_lu = 0.05
    ma = u_lu * math.sqrt(3)
    flow = lt.Obstacle(ctx, [nx, ny], 200, ma, domain_length_x=nx / diameter,
                       stencil=lt.D2Q9(), outlet="equilibrium", solid="halfway", walls=True)
    yy, xx = np.meshgrid(np.arange(ny), np.arange(nx))
    cx, cy = nx / 4.0, ny / 2.0 + 1.0
    flow.mask = (xx - cx) ** 2 + (yy - cy) ** 2 < (diameter / 2.0) ** 2
    flow.initialize()
    coll = lt.KBCCollision() if collision == "kbc" else lt.LearnedCollision(
        lt.LearnedParameters.seeded(seed, out_scale=0.1))
    sim = lt.Simulation(flow, coll, [], lt.StreamingStrategy.PRE_STREAMING)
    # the force on the cylinder only: links into the cylinder's cells
    cyl = torch.as_tensor(flow.mask, dtype=torch.bool)
    from lettuce_amd._classes import halfway_links
    links = halfway_links(cyl, flow.stencil.e, flow.stencil.opposite)
    norm = 0.5 * 1.0 * u_lu ** 2 * diameter
    cd, cl = [], []
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(every, steps - done)
        sim(k)
        done += k
        F = lt._ops.link_force_f64(flow.stencil, flow._state, links, flow.rho0_shift)
        cd.append(float(F[0]) / norm)
        cl.append(float(F[1]) / norm)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    print(f"cylinder {collision} {storage}: {wall:.1f} s", file=sys.stderr, flush=True)
    half = len(cd) // 2
    return {"study": "cylinder_d2q9", "collision": collision, "storage": storage,
            "grid": [nx, ny], "diameter": diameter, "steps": steps,
            "tau": flow.units.relaxation_parameter_lu,
            "cd_mean": float(np.mean(cd[half:])) if cd else None,
            "cl_amplitude": float(0.5 * (np.max(cl[half:]) - np.min(cl[half:]))) if cl else None,
            "strouhal": _strouhal(cl, every, diameter, u_lu),
            "every": every, "cd_series": cd, "cl_series": cl,
            "finite": bool(np.all(np.isfinite(cd + cl))),
            "wall_s_incl_forces": round(wall, 3),
            "mlups_incl_forces": round(steps * nx * ny / 1e6 / wall, 1)}


def turbulence3d(steps=5000, res=384, every=50, u_rms=0.05, nu=1e-4, k0=4, seed=0,
                 learned=True):
    """D3Q27 decaying isotropic turbulence, E(k) ~ k^4 exp(-2 (k/k0)^2),
    u_rms = 0.05 and nu = 1e-4 in lattice units: KBC with FP32 storage and
    shifted BF16 / FP16 storage, and the learned collision (the weights
    trained on KBC states, lettuce_amd/data/learned_d3q27_kbc.npz, KBC
    fallback) with FP32 and shifted FP16 / BF16 storage, BF16 also with
    stochastic rounding (`_sr`), each against KBC in FP64: kinetic energy,
    the dissipation rate -dE/dt and the enstrophy, every `every` steps."""
    out = {"study": "turbulence3d_d3q27_kbc", "grid": [res] * 3, "steps": steps,
           "every": every, "nu_lu": nu, "u_rms_lu": u_rms}
    re, ma = turbulence_units(res, u_rms, nu, k0, seed)
    out["ma"], out["re"] = ma, re
    series, timing = {}, {}
    runs = [("f64", torch.float64, "f64", "kbc"), ("f32", torch.float32, "f32", "kbc"),
            ("bf16_shifted", torch.float32, "bf16_shifted", "kbc"),
            ("bf16_shifted_sr", torch.float32, "bf16_shifted", "kbc"),
            ("f16_shifted", torch.float32, "f16_shifted", "kbc")]
    if learned:
        runs += [("learned_f32", torch.float32, "f32", "learned"),
                 ("learned_f16_shifted", torch.float32, "f16_shifted", "learned"),
                 ("learned_bf16_shifted", torch.float32, "bf16_shifted", "learned"),
                 ("learned_bf16_shifted_sr", torch.float32, "bf16_shifted", "learned")]
        params = lt.LearnedParameters.load(lt.TRAINED_D3Q27_KBC)
    for name, dtype, storage, coll in runs:
        ctx = lt.Context("cuda:0", dtype=dtype, storage=storage,
                         rounding="stochastic" if name.endswith("_sr") else "nearest")
        flow = lt.DecayingTurbulence(ctx, [res] * 3, re, ma, k0=k0, stencil=lt.D3Q27(),
                                     initialize_pressure=False, initialize_fneq=False,
                                     randseed=seed)
        collision = lt.KBCCollision() if coll == "kbc" else lt.LearnedCollision(params)
        sim = lt.Simulation(flow, collision, [], lt.StreamingStrategy.PRE_STREAMING)
        t0 = time.perf_counter()
        series[name] = _ke_series(sim, flow, steps, every)
        torch.cuda.synchronize()
        timing[name] = time.perf_counter() - t0
        out["tau"] = flow.units.relaxation_parameter_lu
        out["nu_lu_check"] = flow.units.viscosity_lu
        print(f"turbulence3d {name}: {timing[name]:.1f} s", file=sys.stderr, flush=True)
        del sim, flow
        torch.cuda.empty_cache()
    ref = series["f64"]
    for name in [r[0] for r in runs if r[0] != "f64"]:
        s = series[name]
        ke_err = np.abs(s[:, 1] / ref[:, 1] - 1)
        diss = -np.diff(s[:, 1])
        diss_ref = -np.diff(ref[:, 1])
        # the dissipation error relative to the peak dissipation rate: the
        # pointwise ratio is meaningless where -dE/dt of the reference
        # passes through zero in the initial acoustic transient
        out[name] = {"ke_rel_err_max": float(ke_err.max()),
                     "dissipation_err_max_over_peak": float(np.max(np.abs(diss - diss_ref))
                                                            / np.max(np.abs(diss_ref))),
                     "enstrophy_rel_err_max": float(np.max(np.abs(s[:, 2] / ref[:, 2] - 1))),
                     "ke_final_over_initial": float(s[-1, 1] / s[0, 1]),
                     "wall_s_incl_reporting": round(timing[name], 2)}
    out["f64"] = {"ke_final_over_initial": float(ref[-1, 1] / ref[0, 1]),
                  "wall_s_incl_reporting": round(timing["f64"], 2)}
    out["series"] = {name: series[name].tolist() for name in series}
    return out


class _FieldFlow(lt.ExtFlow):
    """periodic D2Q9 flow that starts from given populations f0 [9, X, Y],
    with units whose relaxation parameter is tau (characteristic velocity
    u_lu, length X)"""

    def __init__(self, context, f0, tau, u_lu):
        self._f0 = f0
        self._tau, self._u = tau, u_lu
        super().__init__(context, list(f0.shape[1:]), 1.0, 1.0, lt.D2Q9())

    def make_resolution(self, resolution, stencil=None):
        return list(resolution)

    def make_units(self, reynolds_number, mach_number, resolution):
        nu = (self._tau - 0.5) / 3.0
        return lt.UnitConversion(self._u * resolution[0] / nu, self._u * math.sqrt(3.0),
                                 characteristic_length_lu=resolution[0])

    def initial_pu(self):
        raise NotImplementedError

    def initialize(self):
        self.f = self._f0


def train2d(iters=300, n=256, fields=32, holdout=8, steps=16, every=4, u_rms=0.04,
            tau=0.506, k0=4.0, lr=1e-3, seed=0, eval_every=25, rate_hi=1.99):
    """Training of the learned collision on decaying 2D turbulence.

    Fields: `fields` training and `holdout` test velocity fields (seeds
    0, 1, ...), E(k) ~ k^4 exp(-2 (k/k0)^2), u_rms = 0.04, generated at 2n x 2n.
    Target: BGK on the 2n grid (tau chosen for the same physical viscosity,
    nu_fine = 2 nu) for 2 * steps steps, spectrally filtered and downsampled
    to n x n.  Model: the learned collision on the n grid for `steps` steps,
    started from the every-other-point sample of the same field, both grids
    at rest density with equilibrium populations.  Loss: mean |u - u_target|^2
    over the cells, divided by the same mean for BGK on the n grid (the
    training kernel with a zero MLP output).  Adam on
    all fields at once, the hand-written adjoint with checkpoints every
    `every` steps.  The trained weights are then run in the bf16 MFMA
    inference kernel (lettuce_amd.LearnedCollision) on the held-out fields."""
    dev = torch.device("cuda:0")
    t_all = time.perf_counter()
    total = fields + holdout
    u_f = torch.from_numpy(np.stack([lt.spectral_velocity_2d(2 * n, s, k0, u_rms)
                                     for s in range(seed, seed + total)], 1)).to(dev)
    tau_f = 0.5 + 2.0 * (tau - 0.5)
    zero = torch.zeros(N.LEARNED_PARAMS, dtype=torch.float32, device=dev)
    with torch.no_grad():
        f_fine = lt.equilibrium_d2q9(torch.ones_like(u_f[0]), u_f).float().contiguous()
        # the training step with a zero MLP is BGK exactly
        fine_end = lt.learned_rollout(f_fine, zero, lt.TrainConfig(tau=tau_f), 2 * steps,
                                      2 * steps)
        u_target = lt.spectral_downsample(lt.velocity_d2q9(fine_end.double())[1])
        del f_fine, fine_end
        u_c = u_f[:, :, ::2, ::2]
        f_c = lt.equilibrium_d2q9(torch.ones_like(u_c[0]), u_c).float().contiguous()
    del u_f
    torch.cuda.empty_cache()
    tr, te = slice(0, fields), slice(fields, total)
    f_tr, f_te = f_c[:, tr].contiguous(), f_c[:, te].contiguous()
    ut_tr, ut_te = u_target[:, tr].float(), u_target[:, te].float()

    def err(u, ut):
        return ((u - ut) ** 2).sum(0).mean()

    # rate_hi above 1/tau: a zero MLP output is then BGK exactly, and the
    # clamp does not block the gradient of rates near 1/tau
    # the training step has no fallback: the inference kernel runs with an
    # unbounded stable range so that it applies the same formula
    base = lt.LearnedParameters.seeded(seed, out_scale=0.1, rate_hi=rate_hi,
                                       stable_lo=-1e30, stable_hi=1e30)
    cfg = lt.TrainConfig.of(base, tau)
    with torch.no_grad():
        bgk_tr = err(lt.velocity_d2q9(lt.learned_rollout(f_tr, zero, cfg, steps, every))[1], ut_tr)
        bgk_te = err(lt.velocity_d2q9(lt.learned_rollout(f_te, zero, cfg, steps, every))[1], ut_te)
    theta = lt.flatten_parameters(base, torch.float32, dev).requires_grad_(True)
    opt = torch.optim.Adam([theta], lr=lr)
    history = []

    def evaluate(it):
        with torch.no_grad():
            e = err(lt.velocity_d2q9(lt.learned_rollout(f_te, theta, cfg, steps, every))[1], ut_te)
        history.append({"iter": it, "train": float(loss), "test": float(e / bgk_te)})
        print(f"train2d iter {it}: train {float(loss):.4f} test {float(e / bgk_te):.4f}",
              file=sys.stderr, flush=True)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss = torch.tensor(float("nan"))
    for it in range(iters):
        opt.zero_grad()
        u = lt.velocity_d2q9(lt.learned_rollout(f_tr, theta, cfg, steps, every))[1]
        loss = err(u, ut_tr) / bgk_tr
        loss.backward()
        if it == 0:
            first = float(loss)
        if it % eval_every == 0:
            evaluate(it)
        opt.step()
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    evaluate(iters)
    trained = lt.unflatten_parameters(theta, like=base)

    # the same fields through the production kernels (one simulation per field)
    def production(make_collision, params_storage="f32"):
        total_err = 0.0
        for b in range(holdout):
            ctx = lt.Context("cuda:0", dtype=torch.float32, storage=params_storage)
            flow = _FieldFlow(ctx, f_te[:, b].clone(), tau, u_rms)
            sim = lt.Simulation(flow, make_collision(), [], lt.StreamingStrategy.PRE_STREAMING)
            sim(steps)
            u = flow.u().float()
            total_err += float(((u - ut_te[:, b]) ** 2).sum(0).mean())
        return total_err / holdout / float(bgk_te)

    prod = {"bgk": production(lambda: lt.BGKCollision(tau)),
            "kbc": production(lambda: lt.KBCCollision()),
            "learned_initial": production(lambda: lt.LearnedCollision(base)),
            "learned_trained": production(lambda: lt.LearnedCollision(trained)),
            "learned_trained_f16s": production(lambda: lt.LearnedCollision(trained),
                                               "f16_shifted")}
    cells = fields * n * n
    # forward: steps; backward: steps adjoint launches + (steps - steps/every) recomputed steps
    launches = steps + steps + (steps - (steps + every - 1) // every)
    return {"study": "train2d_d2q9_learned", "grid": [n, n], "fields": fields,
            "holdout": holdout, "steps": steps, "checkpoint_every": every, "tau": tau,
            "tau_fine": tau_f, "u_rms": u_rms, "k0": k0, "iters": iters, "lr": lr,
            "rate_lo": base.rate_lo, "rate_hi": base.rate_hi,
            "loss_initial": first, "loss_final_train": history[-1]["train"],
            "loss_final_test": history[-1]["test"],
            "note": "losses are mean |u - u_target|^2 divided by BGK's on the same fields",
            "production_kernels_test_loss": prod, "history": history,
            "train_s": round(t_train, 2), "ms_per_iter": round(1e3 * t_train / iters, 3),
            "cell_launches_per_s": round(cells * launches * iters / t_train / 1e6, 1),
            "wall_s": round(time.perf_counter() - t_all, 2)}


def turbulence_units(res, u_rms=0.05, nu=1e-4, k0=4, seed=0):
    """(Re, Ma) that give lettuce's DecayingTurbulence (D3Q27, res^3, seed)
    the lattice-unit rms velocity u_rms and viscosity nu: lettuce scales the
    field to max |u| = Ma cs (decayingturbulence.py:167-169), so Ma follows
    from the field's rms / max ratio"""
    probe = lt.DecayingTurbulence(lt.Context("cuda:0", dtype=torch.float64), [res] * 3, 1.0,
                                  0.1, k0=k0, stencil=lt.D3Q27(), initialize_pressure=False,
                                  initialize_fneq=False, randseed=seed)
    u = probe.u()
    ratio = float(torch.sqrt((u * u).sum(0).mean()) / torch.linalg.vector_norm(u, dim=0).max())
    del probe, u
    torch.cuda.empty_cache()
    u_max = u_rms / ratio
    return res * u_max / nu, u_max / lt.D3Q27().cs


class _StateFlow(lt.ExtFlow):
    """periodic flow that starts from given populations f0 [Q, *res] and
    uses a given UnitConversion"""

    def __init__(self, context, f0, units, stencil):
        self._f0, self._units = f0, units
        super().__init__(context, list(f0.shape[1:]), 1.0, 1.0, stencil)

    def make_resolution(self, resolution, stencil=None):
        return list(resolution)

    def make_units(self, reynolds_number, mach_number, resolution):
        return self._units

    def initial_pu(self):
        raise NotImplementedError

    def initialize(self):
        self.f = self._f0


def _enstrophy(u):
    """mean |curl u|^2 of a periodic velocity [3, X, Y, Z], vorticity by
    second-order central differences (differentiable)"""
    def d(c, axis):
        return 0.5 * (torch.roll(u[c], -1, axis) - torch.roll(u[c], 1, axis))
    w = torch.stack([d(2, 1) - d(1, 2), d(0, 2) - d(2, 0), d(1, 0) - d(0, 1)])
    return (w * w).sum(0).mean()


def train3d(iters=500, n=64, steps=64, every=8, u_rms=0.05, nu=1e-4, k0=4, seed=0,
            starts=(0, 200, 400, 600, 800, 1000), lr=2e-3, eval_every=25, rate_hi=1.999,
            stable_hi=3.0, enstrophy_weight=0.0, out=None):
    """Training of the D3Q27 learned collision on states that KBC generates.

    lettuce's DecayingTurbulence (D3Q27, E(k) ~ k^4 exp(-2 (k/k0)^2), u_rms
    and nu in lattice units as in the 384^3 benchmark, hence the same tau)
    at n^3 runs with KBC (FP32, the production kernel).  Its states after
    `starts` steps are the rollouts' starts and its states `steps` steps
    later the targets.  Model: `steps` steps of the learned collision's
    training formula from the same starts.  Loss: mean |u - u_KBC|^2 over the
    cells, divided by the same for BGK (a zero MLP), plus enstrophy_weight
    times the squared error of the mean enstrophy (vorticity by central
    differences), divided by BGK's, which weighs the dissipation at the
    grid scale; averaged over the starts.  Adam, the hand-written D3Q27 adjoint with checkpoints every
    `every` steps.  Held out: the same from the field of seed + 1.  The
    trained weights then run in the bf16 MFMA inference kernel on the
    held-out starts: loss and the fraction of cells that fall back to KBC.
    The production limits clamp every rate into [0.05, rate_hi], inside the
    stable interval (0, 2), and send a cell to KBC when its raw rate is NaN
    or outside [0, stable_hi]: at tau = 0.5003 (1/tau = 1.9988) a fallback
    bound of 2 would catch every raw rate a hair above 1/tau, and every
    64-cell wavefront with one such cell runs both collisions.
    With `out`, the trained parameters are saved there (LearnedParameters.save)."""
    dev = torch.device("cuda:0")
    t_all = time.perf_counter()
    re, ma = turbulence_units(n, u_rms, nu, k0, seed)

    ew = float(enstrophy_weight)

    def kbc_states(field_seed):
        flow = lt.DecayingTurbulence(lt.Context("cuda:0", dtype=torch.float32), [n] * 3, re,
                                     ma, k0=k0, stencil=lt.D3Q27(), initialize_pressure=False,
                                     initialize_fneq=False, randseed=field_seed)
        sim = lt.Simulation(flow, lt.KBCCollision(), [], lt.StreamingStrategy.PRE_STREAMING)
        pairs, t = [], 0
        for s0 in starts:
            sim(s0 - t)
            f0 = flow.f.contiguous().clone()
            sim(steps)
            t = s0 + steps
            ut = flow.u().float().clone()
            pairs.append((f0, (ut, _enstrophy(ut))))
        return flow.units, pairs

    units, train = kbc_states(seed)
    _, test = kbc_states(seed + 1)
    tau = units.relaxation_parameter_lu
    zero = torch.zeros(N.LEARNED_PARAMS, dtype=torch.float32, device=dev)
    e = torch.tensor(lt.D3Q27().e, dtype=torch.float32, device=dev)

    def vel(f):
        return torch.einsum("qd,q...->d...", e, f) / f.sum(0)

    def err(u, target):
        """(velocity error, enstrophy error) against a target (u, enstrophy)"""
        ut, wt = target
        e_u = ((u - ut) ** 2).sum(0).mean()
        if ew == 0.0:
            return e_u, None
        return e_u, (_enstrophy(u) - wt) ** 2

    def ratio(e, b):
        """the loss of an error pair against BGK's pair b"""
        return e[0] / b[0] + (ew * e[1] / b[1] if ew else 0.0)

    base = lt.LearnedParameters.seeded(seed, out_scale=0.1, rate_hi=rate_hi,
                                       stable_lo=-1e30, stable_hi=1e30)
    cfg = lt.TrainConfig.of(base, tau)
    with torch.no_grad():
        bgk_tr = [err(vel(lt.learned_rollout(f0, zero, cfg, steps, every)), ut)
                  for f0, ut in train]
        bgk_te = [err(vel(lt.learned_rollout(f0, zero, cfg, steps, every)), ut)
                  for f0, ut in test]
    theta = lt.flatten_parameters(base, torch.float32, dev).requires_grad_(True)
    opt = torch.optim.Adam([theta], lr=lr)
    history = []

    def test_loss(th):
        with torch.no_grad():
            return float(sum(ratio(err(vel(lt.learned_rollout(f0, th, cfg, steps, every)), ut), b)
                             for (f0, ut), b in zip(test, bgk_te)) / len(test))

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    first = None
    for it in range(iters + 1):
        opt.zero_grad()
        loss = sum(ratio(err(vel(lt.learned_rollout(f0, theta, cfg, steps, every)), ut), b)
                   for (f0, ut), b in zip(train, bgk_tr)) / len(train)
        if first is None:
            first = float(loss)
        if it % eval_every == 0 or it == iters:
            history.append({"iter": it, "train": float(loss), "test": test_loss(theta)})
            print(f"train3d iter {it}: train {history[-1]['train']:.4f} "
                  f"test {history[-1]['test']:.4f}", file=sys.stderr, flush=True)
        if it == iters:
            break
        loss.backward()
        opt.step()
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    # the production limits: rates clamped into [0.05, rate_hi], the KBC
    # fallback for raw rates outside [0, stable_hi]
    trained = lt.unflatten_parameters(theta, like=base)
    trained.stable_lo, trained.stable_hi = 0.0, stable_hi
    if out:
        trained.save(out)

    def production(make_collision):
        total, fallback = 0.0, 0.0
        for (f0, ut), b in zip(test, bgk_te):
            flow = _StateFlow(lt.Context("cuda:0", dtype=torch.float32), f0.clone(), units,
                              lt.D3Q27())
            coll = make_collision()
            if isinstance(coll, lt.LearnedCollision):
                p = coll.params
                free = lt.LearnedParameters(*p.tensors(), in_scale=p.in_scale,
                                            rate_lo=p.rate_lo, rate_hi=p.rate_hi,
                                            stable_lo=-1e30, stable_hi=1e30)
                a = lt.LearnedCollision(p)(flow)
                b_ = lt.LearnedCollision(free)(flow)
                fallback += float((a != b_).any(0).float().mean())
            sim = lt.Simulation(flow, coll, [], lt.StreamingStrategy.PRE_STREAMING)
            sim(steps)
            total += float(ratio(err(flow.u().float(), ut), b))
        return total / len(test), fallback / len(test)

    prod = {"kbc": production(lambda: lt.KBCCollision())[0]}
    for name, params in (("learned_initial", lt.LearnedParameters.seeded(seed, out_scale=0.1)),
                         ("learned_trained", trained),
                         ("learned_trained_fallback_above_2",
                          lt.LearnedParameters(*trained.tensors(), in_scale=trained.in_scale,
                                               rate_lo=trained.rate_lo, rate_hi=trained.rate_hi,
                                               stable_lo=0.0, stable_hi=2.0))):
        loss_p, fb = production(lambda: lt.LearnedCollision(params))
        prod[name] = loss_p
        prod[name + "_fallback_cell_frac"] = fb
    return {"study": "train3d_d3q27_learned_kbc_targets", "grid": [n] * 3, "starts": list(starts),
            "steps": steps, "checkpoint_every": every, "tau": tau, "u_rms": u_rms, "nu": nu,
            "enstrophy_weight": ew,
            "k0": k0, "iters": iters, "lr": lr, "rate_hi": rate_hi,
            "loss_initial": first, "loss_final_train": history[-1]["train"],
            "loss_final_test": history[-1]["test"],
            "note": "losses are mean |u - u_KBC|^2 (plus enstrophy_weight times the squared "
                    "error of the mean enstrophy) after `steps` steps, each divided by BGK's",
            "production_kernels_test_loss": prod, "history": history,
            "train_s": round(t_train, 2), "ms_per_iter": round(1e3 * t_train / max(iters, 1), 2),
            "weights": out, "wall_s": round(time.perf_counter() - t_all, 2)}


def slab(steps=100, stencil="D3Q27", collision="kbc", res=384, planes=48, storage="f32"):
    """One device's share of a res^3 lattice split into x-slabs of `planes`
    planes (48 x 384 x 384: 384^3 over 8 devices), initialised from the
    Taylor-Green vortex.  The slab steps through lbm_slab_step with its
    halos wrapped onto itself (a device copy per face stands in for the
    transfer); after `steps` steps it must equal the periodic step of the
    same array.  Times, with HIP events on the launch stream: the periodic
    step, the slab step (interior + faces + the two copies) and each part."""
    S = getattr(lt, stencil)()
    ctx = lt.Context("cuda:0", dtype=torch.float32, storage=storage)
    flow = lt.TaylorGreenVortex(ctx, res, 1600, 0.1, S, initialize_fneq=False)
    coll = lt.KBCCollision(flow.units.relaxation_parameter_lu) if collision == "kbc" else \
        lt.BGKCollision(flow.units.relaxation_parameter_lu)
    sim = lt.Simulation(flow, coll, [], lt.StreamingStrategy.PRE_STREAMING)
    start = flow._state[:, :planes].clone()
    sl = lt.SlabStep(sim, start.clone())
    ref_a, ref_b = start.clone(), torch.empty_like(start)
    args = N.LbmStepArgs.from_buffer_copy(sl.args)
    stream = N.stream_handle(ctx.device)

    def periodic(n):
        # single fused steps of the slab as a periodic lattice
        nonlocal ref_a, ref_b
        args.f_in, args.f_out = ref_a.data_ptr(), ref_b.data_ptr()
        for _ in range(n):
            N.check(N.lib().lbm_step(ctypes.byref(args), stream), "lbm_step")
            ref_a, ref_b = ref_b, ref_a
            args.f_in, args.f_out = ref_a.data_ptr(), ref_b.data_ptr()

    def slab_steps(n, part="all"):
        for _ in range(n):
            if part in ("all", "faces"):
                sl.wrap()
            sl.step(part)

    periodic(steps)
    sl.pack()
    slab_steps(steps)
    torch.cuda.synchronize()
    diff = (sl.state.double() - ref_a.double()).abs().max().item()

    def timed(fn, n=50):
        fn(3)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn(n)
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) * 1e3 / n

    def copies(n):
        for _ in range(n):
            sl.wrap()

    t_per = timed(periodic)
    t_all = timed(slab_steps)
    t_int = timed(lambda n: slab_steps(n, "interior"))
    t_face = timed(lambda n: [sl.step("faces") for _ in range(n)])
    t_copy = timed(copies)
    cells = planes * res * res
    halo = sl.send_lo.numel() * sl.send_lo.element_size()
    return {"study": "slab", "stencil": stencil, "collision": collision.upper(),
            "storage": storage, "slab": [planes, res, res], "of_lattice": [res] * 3,
            "devices": res // planes, "steps": steps,
            "max_abs_diff_vs_periodic": diff,
            "periodic_step_us": round(t_per, 2), "slab_step_us": round(t_all, 2),
            "interior_us": round(t_int, 2), "faces_us": round(t_face, 2),
            "halo_copies_us": round(t_copy, 2),
            "mlups_periodic": round(cells / t_per, 1), "mlups_slab": round(cells / t_all, 1),
            "halo_bytes_per_face": halo,
            "xgmi_us_per_face_at_153GBs": round(halo / 153e9 * 1e6, 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("study", choices=["tgv2d", "tgv3d", "cylinder", "turbulence3d", "train2d",
                                      "train3d", "slab", "all"])
    ap.add_argument("--out", default=None, help="train3d: file for the trained parameters")
    ap.add_argument("--iters", type=int, default=None)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--res", type=int, default=None)
    args = ap.parse_args()
    kw = {}
    if args.steps:
        kw["steps"] = args.steps
    if args.study in ("tgv2d", "all"):
        print(json.dumps(tgv2d(**kw, **({"res": args.res} if args.res else {}))), flush=True)
    if args.study in ("tgv3d", "all"):
        print(json.dumps(tgv3d(**kw, **({"res": args.res} if args.res else {}))), flush=True)
    if args.study in ("cylinder", "all"):
        for coll, storage in (("kbc", "f32"), ("kbc", "f16_shifted"), ("learned", "f32"),
                              ("learned", "f16_shifted")):
            print(json.dumps(cylinder(collision=coll, storage=storage, **kw)), flush=True)
        ref = cylinder(collision="kbc", storage="f64", dtype=torch.float64, **kw)
        print(json.dumps(ref), flush=True)
    if args.study in ("train2d", "all"):
        kw2 = dict(kw, **({"iters": args.iters} if args.iters else {}))
        print(json.dumps(train2d(**kw2, **({"n": args.res} if args.res else {}))), flush=True)
    if args.study in ("train3d", "all"):
        kw3 = dict(kw, **({"iters": args.iters} if args.iters else {}))
        print(json.dumps(train3d(**kw3, out=args.out, **({"n": args.res} if args.res else {}))),
              flush=True)
    if args.study in ("turbulence3d", "all"):
        print(json.dumps(turbulence3d(**kw, **({"res": args.res} if args.res else {}))),
              flush=True)
    if args.study in ("slab", "all"):
        for stencil, coll, res, planes in (("D3Q27", "kbc", 384, 48), ("D3Q19", "bgk", 256, 32)):
            print(json.dumps(slab(stencil=stencil, collision=coll,
                                  res=args.res or res, planes=planes, **kw)), flush=True)


if __name__ == "__main__":
    main()

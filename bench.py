"""Benchmark of the fused lattice-Boltzmann step on one MI355X.

Default workload: D3Q19 BGK Taylor-Green vortex, 256^3 periodic, FP32
populations, Re 1600, Ma 0.1, pre-streaming (pull) fused kernel.  One
"step" is one fused stream + collide pass over the whole grid.

Prints one JSON line with
 - MLUPS (cells * steps / 1e6 / s, HIP events on the launch stream);
 - the roofline of the dominant kernel: algorithmic HBM bytes per launch
   (one read and one write of every population, 2 Q sizeof(storage) per
   cell and memory pass, + 1 B per cell for a class map) over the measured
   launch time against 8 TB/s, next to the HBM bytes rocprofv3 measured for
   that kernel (profiles/traffic_<workload>.json, used only when it was
   measured on the current kernel sources);
 - learned_vs_plain: the D3Q27 decaying-turbulence workload at 384^3 with
   plain FP32 KBC, the learned collision with shifted FP16 / BF16 storage,
   their MLUPS ratio and each run's kinetic-energy error against FP64 KBC;
 - a CPU baseline: the torch CPU restatement of lettuce's step
   (oracle/lettuce_torch.py) on the workload's own grid, a few steps after
   two warm-ups, on this host's cores.

    python bench.py [--steps K] [--warmup W] [--workload NAME] [--no-cpu]
                    [--no-learned] [--c5-steps N] [--c5-res N]
    python bench.py --traffic-from DIR --workload NAME   (no GPU: writes
        profiles/traffic_NAME.json from rocprofv3 FETCH_SIZE / WRITE_SIZE
        csv output of the workload and of the copy calibration)
"""
import argparse
import csv
import glob
import hashlib
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0

# workloads: (stencil, collision, storage, resolution, Re, Ma)
WORKLOADS = {
    "tgv3d_d3q19_bgk_256_f32": ("D3Q19", "bgk", "f32", 256, 1600, 0.1),
    "tgv3d_d3q19_bgk_256_f16s": ("D3Q19", "bgk", "f16_shifted", 256, 1600, 0.1),
    "tgv3d_d3q19_bgk_256_bf16s": ("D3Q19", "bgk", "bf16_shifted", 256, 1600, 0.1),
    # names ending in _sr: BF16 storage with stochastic rounding
    "tgv3d_d3q19_bgk_256_bf16s_sr": ("D3Q19", "bgk", "bf16_shifted", 256, 1600, 0.1),
    "tgv3d_d3q27_kbc_256_f32": ("D3Q27", "kbc", "f32", 256, 1600, 0.1),
    "tgv3d_d3q27_kbc_256_bf16s": ("D3Q27", "kbc", "bf16_shifted", 256, 1600, 0.1),
    "tgv3d_d3q27_kbc_256_bf16s_sr": ("D3Q27", "kbc", "bf16_shifted", 256, 1600, 0.1),
    "tgv3d_d3q27_learned_256_f32": ("D3Q27", "learned", "f32", 256, 1600, 0.1),
    "tgv3d_d3q27_learned_256_bf16s": ("D3Q27", "learned", "bf16_shifted", 256, 1600, 0.1),
    "tgv3d_d3q27_learned_256_bf16s_sr": ("D3Q27", "learned", "bf16_shifted", 256, 1600, 0.1),
    "tgv3d_d3q19_none_256_f32": ("D3Q19", "none", "f32", 256, 1600, 0.1),
    # lettuce's default streaming order (collide, then stream)
    "tgv3d_d3q19_bgk_256_f32_post": ("D3Q19", "bgk", "f32", 256, 1600, 0.1),
    # a D3Q19 FP32 grid near the largest one device holds (two 76 GB arrays)
    "tgv3d_d3q19_bgk_1024_f32": ("D3Q19", "bgk", "f32", 1024, 1600, 0.1),
    "tgv2d_d2q9_bgk_2048x512_f32": ("D2Q9", "bgk", "f32", [2048, 512], 200, 0.1),
    # channel + cylinder (D = 64, halfway bounce-back, equilibrium in/outlet)
    "cyl2d_d2q9_kbc_2048x512_f32": ("D2Q9", "kbc", "f32", [2048, 512], 200, 0.05 * 3 ** 0.5),
    "cyl2d_d2q9_kbc_2048x512_f16s": ("D2Q9", "kbc", "f16_shifted", [2048, 512], 200,
                                     0.05 * 3 ** 0.5),
    "cyl2d_d2q9_learned_2048x512_f32": ("D2Q9", "learned", "f32", [2048, 512], 200,
                                        0.05 * 3 ** 0.5),
    "cyl2d_d2q9_learned_2048x512_f16s": ("D2Q9", "learned", "f16_shifted", [2048, 512], 200,
                                         0.05 * 3 ** 0.5),
    "cyl2d_d2q9_bgk_2048x512_f32": ("D2Q9", "bgk", "f32", [2048, 512], 200, 0.05 * 3 ** 0.5),
}
BYTES = {"f32": 4, "f64": 8, "f16_shifted": 2, "bf16_shifted": 2}


def build_sim(lt, name, layout="soa"):
    stencil, coll, storage, res, re, ma = WORKLOADS[name]
    S = {"D2Q9": lt.D2Q9, "D3Q19": lt.D3Q19, "D3Q27": lt.D3Q27}[stencil]()
    ctx = lt.Context("cuda:0", dtype=torch.float32, storage=storage, layout=layout,
                     rounding="stochastic" if name.endswith("_sr") else "nearest")
    if name.startswith("cyl2d"):
        nx, ny = res
        flow = lt.Obstacle(ctx, [nx, ny], re, ma, domain_length_x=nx / 64, stencil=S,
                           outlet="equilibrium", solid="halfway", walls=True)
        yy, xx = np.meshgrid(np.arange(ny), np.arange(nx))
        flow.mask = (xx - nx / 4) ** 2 + (yy - ny / 2 - 1) ** 2 < 32.0 ** 2
        flow.initialize()
    else:
        flow = lt.TaylorGreenVortex(ctx, res, re, ma, S, initialize_fneq=False)
    tau = flow.units.relaxation_parameter_lu
    collision = {"bgk": lambda: lt.BGKCollision(tau), "kbc": lambda: lt.KBCCollision(),
                 "none": lambda: lt.NoCollision(),
                 "trt": lambda: lt.TRTCollision(tau, 1.0),
                 "learned": lambda: lt.LearnedCollision(
                     lt.LearnedParameters.seeded(0, out_scale=0.1))}[coll]()
    strategy = (lt.StreamingStrategy.POST_STREAMING if name.endswith("_post")
                else lt.StreamingStrategy.PRE_STREAMING)
    sim = lt.Simulation(flow, collision, [], strategy)
    return sim, S


def time_gpu(sim, steps, warmup):
    step = sim._collide_and_stream
    step.run(warmup)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(stream)
    step.run(steps)
    stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return start.elapsed_time(stop) / 1e3, wall


def cpu_baseline(name, budget_s=20.0, min_steps=5, max_steps=20):
    """The oracle (torch CPU restatement of lettuce's step: roll streaming,
    einsum equilibrium, torch.where masks) on the workload's own grid and
    collision in FP32, with all host threads: 2 warm-up steps, then steps
    until `budget_s` has passed (at least min_steps, at most max_steps)."""
    from oracle import lettuce_torch as O
    stencil, coll, storage, res, re, ma = WORKLOADS[name]
    S = {"D2Q9": O.D2Q9, "D3Q19": O.D3Q19, "D3Q27": O.D3Q27}[stencil]()
    # the host share of one GPU (OMP_NUM_THREADS where the host sets it)
    threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "0")) or 1 << 30)
    torch.set_num_threads(threads)
    if name.startswith("cyl2d"):
        nx, ny = res
        yy, xx = np.meshgrid(np.arange(ny), np.arange(nx))
        mask = (xx - nx / 4) ** 2 + (yy - ny / 2 - 1) ** 2 < 32.0 ** 2
        flow = O.ObstacleChannel(O.Context("cpu", torch.float32), [nx, ny], re, ma,
                                 domain_length_x=nx / 64, mask=mask, stencil=S)
    else:
        flow = O.TaylorGreenVortex(O.Context("cpu", torch.float32), res, re, ma, S)
    tau = flow.units.relaxation_parameter_lu
    c = {"bgk": lambda: O.BGKCollision(tau), "kbc": lambda: O.KBCCollision(),
         "none": lambda: O.NoCollision(),
         "trt": lambda: O.TRTCollision(tau, 1.0),
         "learned": lambda: O.LearnedCollision(O.LearnedParameters.seeded(0, out_scale=0.1))
         }[coll]()
    strategy = (O.StreamingStrategy.POST_STREAMING if name.endswith("_post")
                else O.StreamingStrategy.PRE_STREAMING)
    sim = O.Simulation(flow, c, [], strategy)
    sim(2)  # warm-up
    cells = int(np.prod(flow.resolution))
    steps, t0 = 0, time.perf_counter()
    while True:
        sim(1)
        steps += 1
        el = time.perf_counter() - t0
        if steps >= max_steps or (steps >= min_steps and el > budget_s):
            break
    return {"value": round(steps * cells / 1e6 / el, 3), "unit": "MLUPS",
            "cores": torch.get_num_threads(), "kind": "port",
            "sample": f"{stencil} {coll.upper()} {type(flow).__name__} "
                      f"{'x'.join(map(str, flow.resolution))} (the workload's grid) "
                      f"FP32, {steps} steps after 2 warm-ups, {el:.1f} s, torch CPU "
                      f"({torch.get_num_threads()} threads)"}


def copy_bytes(nbytes):
    """bytes one copy moves each way for a population array of `nbytes`: at
    most 4 GiB (the ceiling does not change with size past a few hundred MB,
    and a 1024^3 array next to the simulation's two would not fit), a
    multiple of 16"""
    nbytes = min(nbytes, 4 << 30)
    return nbytes - nbytes % 16


def copy_ceiling(nbytes, steps, warmup):
    """Device-to-device copies of one population array: the HIP runtime's
    copy (torch's copy_, the reference kernel for scaling the FETCH_SIZE /
    WRITE_SIZE counters) and lbm_stream_copy (16 B per lane, plain and
    non-temporal stores).  GB/s counts read + write bytes."""
    import ctypes
    from lettuce_amd import _native as N
    nbytes = copy_bytes(nbytes)
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    src.uniform_()
    dst = torch.empty_like(src)
    stream = N.stream_handle(src.device)

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(steps):
            fn()
        stop.record()
        torch.cuda.synchronize()
        s = start.elapsed_time(stop) / 1e3 / steps
        return round(2 * nbytes / s / 1e9, 1)

    def ours(nt):
        return lambda: N.check(N.lib().lbm_stream_copy(
            ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()), nbytes, nt,
            stream), "lbm_stream_copy")
    out = {"bytes_per_copy": 2 * nbytes, "runtime_copy_GBps": timed(lambda: dst.copy_(src)),
           "stream_copy_GBps": timed(ours(0)), "stream_copy_nt_GBps": timed(ours(1))}
    out["best_GBps"] = max(out["runtime_copy_GBps"], out["stream_copy_GBps"],
                           out["stream_copy_nt_GBps"])
    return out


# the sources a launch kind's kernel and its launch configuration come from
KERNEL_SOURCES = {
    "two-step": ["lbm_tb2.hip", "lbm_tb2.h", "lbm_step_impl.h", "lbm_device.h", "lbm_capi.hip"],
    "single-step": ["step_d2q9.hip", "step_d3q19.hip", "step_d3q27.hip", "lbm_step_impl.h",
                    "lbm_device.h", "lbm_capi.hip"],
    "learned": ["lbm_learned.hip", "lbm_step_impl.h", "lbm_device.h", "lbm_capi.hip"],
}


def kernel_sources_sha(kind):
    """sha256 over the kernel sources of a launch kind (lettuce_amd/csrc),
    which a traffic measurement must have been taken on to be used"""
    h = hashlib.sha256()
    for name in KERNEL_SOURCES[kind]:
        with open(os.path.join(ROOT, "lettuce_amd", "csrc", name), "rb") as fh:
            h.update(name.encode() + b"\0" + fh.read())
    return h.hexdigest()[:16]


def load_traffic(name, kernel):
    """(HBM bytes per launch of `kernel` on workload `name` from
    profiles/traffic_<name>.json, or None; whether a file exists but was
    measured on other kernel sources or another kernel)"""
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if not os.path.exists(path):
        return None, False
    with open(path) as fh:
        d = json.load(fh)
    if d.get("launch") != kernel or d.get("kernel_sources_sha") != kernel_sources_sha(kernel):
        return None, True
    return d.get("hbm_bytes_per_launch"), False


def _pmc_rows(directory, counter):
    """(kernel name, counter value in KiB) per dispatch from the rocprofv3
    counter_collection csv files under `directory`"""
    rows = []
    for path in glob.glob(os.path.join(directory, "**", "*counter_collection.csv"),
                          recursive=True):
        with open(path) as fh:
            for r in csv.DictReader(fh):
                if r.get("Counter_Name") == counter:
                    rows.append((r["Kernel_Name"], float(r["Counter_Value"])))
    return rows


def write_traffic(name, directory):
    """profiles/traffic_<name>.json from rocprofv3 --pmc output under
    `directory`: subdirectories fetch_kernel, write_kernel (the workload's
    bench run) and fetch_copy, write_copy (bench.py --copy).  The copy
    moves a known number of bytes each way, which scales each counter."""
    def per_kernel(sub, counter, match):
        vals = [v for k, v in _pmc_rows(os.path.join(directory, sub), counter) if match(k)]
        if not vals:
            raise SystemExit(f"no {counter} rows for the kernel under {directory}/{sub}")
        vals.sort()
        return vals[len(vals) // 2]  # median dispatch
    stencil, _, storage, res, _, _ = WORKLOADS[name]
    q = {"D2Q9": 9, "D3Q19": 19, "D3Q27": 27}[stencil]
    cells = int(np.prod(res if isinstance(res, list) else [res] * (2 if q == 9 else 3)))
    copy_rows = _pmc_rows(os.path.join(directory, "fetch_copy"), "FETCH_SIZE")
    copy_name = next(k for k, _ in copy_rows if "k_stream_copy<false>" in k)
    copy_kib = copy_bytes(q * cells * BYTES[storage]) / 1024   # each way
    fetch_scale = copy_kib / per_kernel("fetch_copy", "FETCH_SIZE", lambda k: k == copy_name)
    write_scale = copy_kib / per_kernel("write_copy", "WRITE_SIZE", lambda k: k == copy_name)
    rows = _pmc_rows(os.path.join(directory, "fetch_kernel"), "FETCH_SIZE")
    kind = ("two-step" if any("tb2_kernel" in k for k, _ in rows) else
            "learned" if any("learned" in k for k, _ in rows) else "single-step")
    marker = {"two-step": "tb2_kernel", "learned": "learned", "single-step": "step_kernel"}[kind]
    kname = next(k for k, _ in rows if marker in k)
    fetch = per_kernel("fetch_kernel", "FETCH_SIZE", lambda k: k == kname)
    write = per_kernel("write_kernel", "WRITE_SIZE", lambda k: k == kname)
    out = {"workload": name, "launch": kind, "kernel": kname,
           "kernel_sources_sha": kernel_sources_sha(kind),
           "fetch_kib": round(fetch, 1), "fetch_scale": round(fetch_scale, 4),
           "write_kib": round(write, 1), "write_scale": round(write_scale, 4),
           "hbm_bytes_per_launch": int(round((fetch * fetch_scale + write * write_scale) * 1024)),
           "algorithmic_bytes_per_launch": 2 * q * BYTES[storage] * cells}
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
        fh.write("\n")
    print(json.dumps(out))


class _SharedTurbulence:
    """lettuce's DecayingTurbulence initial field, computed once and handed
    to every flow of the learned_vs_plain comparison"""

    def __init__(self, lt, res, u_rms, nu, k0, seed):
        from lettuce_amd.studies import turbulence_units
        # lettuce scales the field to max |u| = Ma cs (decayingturbulence.py:
        # 167-169): Ma and Re follow from the requested u_rms and nu
        self.re, self.ma = turbulence_units(res, u_rms, nu, k0, seed)
        flow = lt.DecayingTurbulence(lt.Context("cuda:0", dtype=torch.float64), [res] * 3,
                                     self.re, self.ma, k0=k0, stencil=lt.D3Q27(),
                                     initialize_pressure=False, initialize_fneq=False,
                                     randseed=seed)
        self.pu = flow.initial_pu()
        self.char_velocity_pu = flow.units.characteristic_velocity_pu
        self.args = (res, k0, seed)
        del flow
        torch.cuda.empty_cache()

    def flow(self, lt, ctx):
        shared = self
        res, k0, seed = self.args

        class Flow(lt.DecayingTurbulence):
            def initial_pu(self):
                self.units.characteristic_velocity_pu = shared.char_velocity_pu
                return shared.pu

        return Flow(ctx, [res] * 3, self.re, self.ma, k0=k0, stencil=lt.D3Q27(),
                    initialize_pressure=False, initialize_fneq=False, randseed=seed)


def learned_vs_plain(lt, res=384, steps=1000, every=50, timed=100, warmup=10, u_rms=0.05,
                     nu=1e-4, k0=4, seed=0, params=None, weights=None):
    """D3Q27 decaying isotropic turbulence (E(k) ~ k^4 exp(-2 (k/k0)^2),
    u_rms 0.05, nu 1e-4 in lattice units, seeded field): MLUPS of plain KBC
    with FP32 storage and of the learned collision (seed-0 weights, KBC
    fallback) with shifted FP16 and BF16 storage (BF16 rounded to nearest
    and stochastically), timed over `timed` steps (with each run's
    algorithmic fraction of the HBM roofline)
    after `warmup`; and the largest relative kinetic-energy error of each
    against KBC with FP64 storage, every `every` steps over `steps` steps."""
    t_all = time.perf_counter()
    from lettuce_amd.ext._learned import TRAINED_D3Q27_KBC
    if params is not None:
        weights = weights or "given parameters"
    elif os.path.exists(TRAINED_D3Q27_KBC):
        params = lt.LearnedParameters.load(TRAINED_D3Q27_KBC)
        weights = ("trained on KBC decaying-turbulence states at 64^3 "
                   "(lettuce_amd/data/learned_d3q27_kbc.npz)")
    else:
        params = lt.LearnedParameters.seeded(seed, out_scale=0.1)
        weights = f"LearnedParameters.seeded({seed}, out_scale=0.1)"
    free = lt.LearnedParameters(*params.tensors(), in_scale=params.in_scale,
                                rate_lo=params.rate_lo, rate_hi=params.rate_hi,
                                stable_lo=-1e30, stable_hi=1e30)
    field = _SharedTurbulence(lt, res, u_rms, nu, k0, seed)
    runs = [("kbc_f64", torch.float64, "f64", "kbc"), ("kbc_f32", torch.float32, "f32", "kbc"),
            ("learned_f16s", torch.float32, "f16_shifted", "learned"),
            ("learned_bf16s", torch.float32, "bf16_shifted", "learned"),
            ("learned_bf16s_sr", torch.float32, "bf16_shifted", "learned")]
    out = {"workload": f"turb3d_d3q27_{res}", "grid": [res] * 3, "cells": res ** 3,
           "u_rms_lu": u_rms, "nu_lu": nu, "ke_steps": steps, "ke_every": every,
           "timed_steps": timed, "reference": "kbc_f64", "learned_weights": weights,
           "rate_clamp": [params.rate_lo, params.rate_hi],
           "kbc_fallback_outside": [params.stable_lo, params.stable_hi]}
    ke = {}
    for name, dtype, storage, coll in runs:
        rounding = "stochastic" if name.endswith("_sr") else "nearest"
        flow = field.flow(lt, lt.Context("cuda:0", dtype=dtype, storage=storage,
                                         rounding=rounding))
        out["tau"] = flow.units.relaxation_parameter_lu
        collision = lt.KBCCollision() if coll == "kbc" else lt.LearnedCollision(params)
        sim = lt.Simulation(flow, collision, [], lt.StreamingStrategy.PRE_STREAMING)
        energy = lt.IncompressibleKineticEnergy(flow)
        series = [float(energy())]
        for _ in range(steps // every):
            sim._collide_and_stream.run(every)
            series.append(float(energy()))
        ke[name] = np.array(series)
        entry = {"storage": storage, "collision": coll.upper(), "rounding": rounding}
        if coll == "learned":
            # cells that took the KBC fallback in the last state's collision:
            # where the learned collision differs from the same one with an
            # unbounded stable range
            a = lt.LearnedCollision(params)(flow)
            b = lt.LearnedCollision(free)(flow)
            fb = (a != b).any(0)
            entry["kbc_fallback_cell_frac"] = float(fb.float().mean())
            # a wavefront of the learned kernel owns 64 consecutive cells
            # along the contiguous axis and runs the KBC branch if any of
            # them falls back
            entry["kbc_fallback_wave_frac"] = float(
                fb.reshape(-1, 64).any(1).float().mean())
            del a, b, fb
        if name != "kbc_f64":
            gpu_s, _ = time_gpu(sim, timed, warmup)
            entry["mlups"] = round(timed * res ** 3 / 1e6 / gpu_s, 1)
            pairs, singles = sim._collide_and_stream.plan(timed)
            two_step = pairs and not singles
            entry["kernel"] = "two-step" if two_step else "single step"
            # algorithmic bytes of one memory pass (one read and one write of
            # every population) per second, as a fraction of HBM peak; a
            # two-step launch makes two updates per pass
            passes = entry["mlups"] * 1e6 / (2 if two_step else 1)
            entry["roofline_frac"] = round(2 * 27 * BYTES[storage] * passes / (PEAK_HBM_GBS * 1e9), 4)
        out[name] = entry
        del sim, flow, energy
        torch.cuda.empty_cache()
    ref = ke["kbc_f64"]
    out["kbc_f64"]["ke_final_over_initial"] = round(float(ref[-1] / ref[0]), 6)
    for name in ("kbc_f32", "learned_f16s", "learned_bf16s", "learned_bf16s_sr"):
        out[name]["ke_rel_err_max_vs_f64"] = float(np.abs(ke[name] / ref - 1).max())
    base = out["kbc_f32"]["mlups"]
    for name in ("learned_f16s", "learned_bf16s", "learned_bf16s_sr"):
        out[name]["speedup_vs_kbc_f32"] = round(out[name]["mlups"] / base, 3)
    out["wall_s"] = round(time.perf_counter() - t_all, 1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="tgv3d_d3q19_bgk_256_f32",
                    choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true",
                    help="leave out the learned_vs_plain block")
    ap.add_argument("--c5-steps", type=int, default=1000,
                    help="steps of the learned_vs_plain kinetic-energy comparison")
    ap.add_argument("--c5-res", type=int, default=384)
    ap.add_argument("--copy", action="store_true",
                    help="only time a device copy of one population array")
    ap.add_argument("--traffic-from", default=None,
                    help="write profiles/traffic_<workload>.json from rocprofv3 output")
    ap.add_argument("--layout", default=None, choices=["soa", "rows"],
                    help="population memory order (see lettuce_amd.Context); default "
                         "'rows' for periodic flows, 'soa' for flows with boundaries")
    args = ap.parse_args()
    if args.traffic_from:
        write_traffic(args.workload, args.traffic_from)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the library steps one device's x-slab "
              "(lettuce_amd.SlabStep) and leaves the halo transfer between "
              "devices to the caller, so there is no multi-GPU bench; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        sys.exit(2)

    if args.copy:
        stencil, _, storage, res, _, _ = WORKLOADS[args.workload]
        q = {"D2Q9": 9, "D3Q19": 19, "D3Q27": 27}[stencil]
        cells = int(np.prod(res if isinstance(res, list) else [res] * (2 if q == 9 else 3)))
        print(json.dumps({"copy": copy_ceiling(q * cells * BYTES[storage], args.steps,
                                               args.warmup)}), flush=True)
        return

    import lettuce_amd as lt
    if args.layout is None:
        args.layout = "soa" if args.workload.startswith("cyl2d") else "rows"
    sim, S = build_sim(lt, args.workload, args.layout)
    stencil, coll, storage, res, re, ma = WORKLOADS[args.workload]
    cells = int(np.prod(sim.flow.resolution))
    gpu_s, wall_s = time_gpu(sim, args.steps, args.warmup)
    mlups = args.steps * cells / 1e6 / gpu_s
    # populations read + written once per memory pass; the uint8 cell class
    # read once
    bytes_per_cell = 2 * S.q * BYTES[storage] + (1 if sim.cell_classes is not None else 0)
    pairs, singles = sim._collide_and_stream.plan(args.steps)
    launches = pairs + singles
    launch_s = gpu_s / launches
    # algorithmic bytes of one launch: every launch reads and writes each
    # population once, whether it makes one update (single step) or two
    # (two-step kernel)
    algorithmic = bytes_per_cell * cells
    achieved = algorithmic / launch_s / 1e9
    # the same launches counted per cell update (two per two-step launch):
    # what single steps would have to move for the same progress
    single_equiv = (2 * pairs + singles) * bytes_per_cell * cells / launches / launch_s / 1e9
    kind = "two-step" if pairs else ("learned" if coll == "learned" else "single-step")
    traffic, stale = load_traffic(args.workload, kind)
    streaming = "post (collide, then stream)" if args.workload.endswith("_post") \
        else "pre (pull), fused"
    line = {
        "metric": "MLUPS",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(gpu_s / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if storage in ("f32", "f16_shifted", "bf16_shifted") else "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "lattice": stencil,
                   "collision": coll.upper(), "storage": storage,
                   "rounding": "stochastic" if args.workload.endswith("_sr") else "nearest",
                   "grid": sim.flow.resolution, "cells": cells,
                   "streaming": streaming, "Re": re, "Ma": ma,
                   "layout": args.layout,
                   "kernel": ("two-step (two fused steps per launch)" if pairs and not singles
                              else "single step" if not pairs
                              else f"{pairs} two-step + {singles} single-step launches")},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1),
                     "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": traffic,
                     "algorithmic_bytes_per_launch": algorithmic,
                     "bytes_per_cell_and_pass": bytes_per_cell,
                     "updates_per_launch": round(args.steps / launches, 3),
                     "launch_us": round(launch_s * 1e6, 2),
                     "single_step_equiv_GBps": round(single_equiv, 1),
                     "single_step_equiv_frac": round(single_equiv / PEAK_HBM_GBS, 4)},
        "wall_s": round(wall_s, 4),
    }
    if traffic:
        # what the HBM actually moved per launch (rocprofv3 FETCH_SIZE and
        # WRITE_SIZE on these kernel sources, calibrated on a copy), at the
        # launch time measured here
        line["roofline"]["hbm_achieved"] = round(traffic / launch_s / 1e9, 1)
        line["roofline"]["hbm_frac"] = round(traffic / launch_s / 1e9 / PEAK_HBM_GBS, 4)
    elif stale:
        line["roofline"]["traffic_note"] = ("profiles/traffic file measured on other kernel "
                                            "sources or another kernel: not used")
    ceiling = copy_ceiling(S.q * cells * BYTES[storage], 20, 3)
    line["roofline"]["measured_copy_ceiling"] = ceiling["best_GBps"]
    line["roofline"]["frac_of_copy_ceiling"] = round(achieved / ceiling["best_GBps"], 4)
    del sim
    torch.cuda.empty_cache()
    if not args.no_learned:
        line["learned_vs_plain"] = learned_vs_plain(lt, res=args.c5_res, steps=args.c5_steps)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.workload)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

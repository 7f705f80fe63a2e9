"""Benchmark of the fused lattice-Boltzmann step on one MI355X.

Default workload: D3Q19 BGK Taylor-Green vortex, 256^3 periodic, FP32
populations, Re 1600, Ma 0.1, pre-streaming (pull) fused kernel.  One
"step" is one fused stream + collide pass over the whole grid.

Prints one JSON line with MLUPS (cells * steps / 1e6 / s), the roofline of
the dominant kernel (algorithmic HBM bytes 2 Q sizeof(storage) per cell
against 8 TB/s), and a CPU baseline: the torch CPU restatement of lettuce's
step (oracle/lettuce_torch.py) timed on a bounded sample of the same
workload on this host.

    python bench.py [--steps K] [--warmup W] [--workload NAME] [--no-cpu]
"""
import argparse
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
    "tgv3d_d3q27_kbc_256_f32": ("D3Q27", "kbc", "f32", 256, 1600, 0.1),
    "tgv3d_d3q27_kbc_256_bf16s": ("D3Q27", "kbc", "bf16_shifted", 256, 1600, 0.1),
    "tgv3d_d3q27_learned_256_f32": ("D3Q27", "learned", "f32", 256, 1600, 0.1),
    "tgv3d_d3q27_learned_256_bf16s": ("D3Q27", "learned", "bf16_shifted", 256, 1600, 0.1),
    "tgv3d_d3q19_none_256_f32": ("D3Q19", "none", "f32", 256, 1600, 0.1),
    # a D3Q19 FP32 grid near the largest one device holds (two 76 GB arrays)
    "tgv3d_d3q19_bgk_1024_f32": ("D3Q19", "bgk", "f32", 1024, 1600, 0.1),
    "tgv2d_d2q9_bgk_2048x512_f32": ("D2Q9", "bgk", "f32", [2048, 512], 200, 0.1),
    # channel + cylinder (D = 64, halfway bounce-back, equilibrium in/outlet)
    "cyl2d_d2q9_kbc_2048x512_f32": ("D2Q9", "kbc", "f32", [2048, 512], 200, 0.05 * 3 ** 0.5),
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
    ctx = lt.Context("cuda:0", dtype=torch.float32, storage=storage, layout=layout)
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
    sim = lt.Simulation(flow, collision, [], lt.StreamingStrategy.PRE_STREAMING)
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


def cpu_baseline(name, budget_s=15.0):
    """The oracle (torch CPU restatement of lettuce's step) on a bounded
    sample: the same lattice, collision and FP32 arithmetic on a 96^3 (3D)
    or 512x128 (2D) grid, as many steps as fit the budget."""
    from oracle import lettuce_torch as O
    stencil, coll, storage, res, re, ma = WORKLOADS[name]
    S = {"D2Q9": O.D2Q9, "D3Q19": O.D3Q19, "D3Q27": O.D3Q27}[stencil]()
    sample = [512, 128] if S.d == 2 else 96
    threads = os.cpu_count() or 1
    torch.set_num_threads(threads)
    if name.startswith("cyl2d"):
        yy, xx = np.meshgrid(np.arange(128), np.arange(512))
        mask = (xx - 128) ** 2 + (yy - 65) ** 2 < 8.0 ** 2
        flow = O.ObstacleChannel(O.Context("cpu", torch.float32), sample, re, ma,
                                 domain_length_x=512 / 16, mask=mask, stencil=S)
    else:
        flow = O.TaylorGreenVortex(O.Context("cpu", torch.float32), sample, re, ma, S)
    tau = flow.units.relaxation_parameter_lu
    c = {"bgk": lambda: O.BGKCollision(tau), "kbc": lambda: O.KBCCollision(),
         "none": lambda: O.NoCollision(),
         "trt": lambda: O.TRTCollision(tau, 1.0),
         "learned": lambda: O.LearnedCollision(O.LearnedParameters.seeded(0, out_scale=0.1))
         }[coll]()
    sim = O.Simulation(flow, c, [], O.StreamingStrategy.PRE_STREAMING)
    sim(1)  # warm-up
    cells = int(np.prod(flow.resolution))
    steps, t0 = 0, time.perf_counter()
    while True:
        sim(1)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or steps >= 200:
            break
    return {"value": round(steps * cells / 1e6 / el, 3), "unit": "MLUPS",
            "cores": threads, "kind": "port",
            "sample": f"{stencil} {coll.upper()} {type(flow).__name__} "
                      f"{'x'.join(map(str, flow.resolution))} "
                      f"FP32, {steps} steps after 1 warm-up, torch CPU "
                      f"({torch.get_num_threads()} threads)"}


def copy_ceiling(nbytes, steps, warmup):
    """Device-to-device copies of one population array: the HIP runtime's
    copy (torch's copy_, the reference kernel for scaling the FETCH_SIZE /
    WRITE_SIZE counters) and lbm_stream_copy (16 B per lane, plain and
    non-temporal stores).  GB/s counts read + write bytes."""
    import ctypes
    from lettuce_amd import _native as N
    # one population array, at most 4 GiB: the ceiling does not change with
    # size past a few hundred MB, and a 1024^3 array next to the simulation's
    # two would not fit
    nbytes = min(nbytes, 4 << 30)
    nbytes -= nbytes % 16
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


def load_traffic(name, kernel):
    """HBM bytes per launch of `kernel` on workload `name` from
    profiles/traffic_<name>.json (None when no profile of that kernel)"""
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    return d.get("hbm_bytes_per_launch") if d.get("launch") == kernel else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="tgv3d_d3q19_bgk_256_f32",
                    choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy", action="store_true",
                    help="only time a device copy of one population array")
    ap.add_argument("--layout", default=None, choices=["soa", "rows"],
                    help="population memory order (see lettuce_amd.Context); default "
                         "'rows' for periodic flows, 'soa' for flows with boundaries")
    args = ap.parse_args()
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
    # populations read + written once per cell update; the uint8 cell class
    # read once
    bytes_per_cell = 2 * S.q * BYTES[storage] + (1 if sim.cell_classes is not None else 0)
    # the launches of the timed region: two-step launches update every cell
    # twice, so their algorithmic bytes are 2 * bytes_per_cell * cells
    pairs, singles = sim._collide_and_stream.plan(args.steps)
    launches = pairs + singles
    launch_s = gpu_s / launches
    achieved = (2 * pairs + singles) * bytes_per_cell * cells / launches / launch_s / 1e9
    traffic = load_traffic(args.workload, "two-step" if pairs else "single-step")
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
                   "grid": sim.flow.resolution, "cells": cells,
                   "streaming": "pre (pull), fused", "Re": re, "Ma": ma,
                   "layout": args.layout,
                   "kernel": ("two-step (two fused steps per launch)" if pairs and not singles
                              else "single step" if not pairs
                              else f"{pairs} two-step + {singles} single-step launches")},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1),
                     "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": traffic,
                     "bytes_per_cell": bytes_per_cell,
                     "steps_per_launch": round(args.steps / launches, 3),
                     "launch_us": round(launch_s * 1e6, 2)},
        "wall_s": round(wall_s, 4),
    }
    if traffic:
        # what the HBM actually moved per launch (rocprofv3 FETCH_SIZE and
        # WRITE_SIZE, calibrated on a copy), at the launch time measured here
        line["roofline"]["hbm_achieved"] = round(traffic / launch_s / 1e9, 1)
        line["roofline"]["hbm_frac"] = round(traffic / launch_s / 1e9 / PEAK_HBM_GBS, 4)
    ceiling = copy_ceiling(S.q * cells * BYTES[storage], 20, 3)
    line["roofline"]["measured_copy_ceiling"] = ceiling["best_GBps"]
    line["roofline"]["frac_of_copy_ceiling"] = round(achieved / ceiling["best_GBps"], 4)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.workload)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

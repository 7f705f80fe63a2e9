"""Write the golden vectors in tests/golden/*.npz from lettuce itself.

Run in a container that has the lettuce sources at /root/reference:

    python tests/golden/make_golden.py

It imports lettuce from /root/reference (never copying its sources) with
stand-in modules for three imports that lettuce pulls in at import time but
that no torch time-step code uses (mmh3: the cuda_native module-name hash,
pyevtk: VTK output, h5py: HDF5 output).  Python writes no bytecode there.

The vectors pin what only lettuce's own arithmetic fixes: FP32/FP64
trajectories of its torch step (stream, collide, boundaries, masks), KBC as
composed, the decaying-turbulence initial field and the observables.
"""
import os
import sys
import types

sys.dont_write_bytecode = True
REF = "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def import_lettuce():
    for name in ("mmh3", "pyevtk", "pyevtk.hl", "h5py"):
        sys.modules.setdefault(name, types.ModuleType(name))
    sys.modules["mmh3"].hash = lambda *a, **k: 0
    hl = sys.modules["pyevtk.hl"]
    hl.imageToVTK = hl.gridToVTK = lambda *a, **k: None
    sys.modules["pyevtk"].hl = hl
    sys.path.insert(0, REF)
    import lettuce
    return lettuce


lt = import_lettuce()
import numpy as np  # noqa: E402
import torch  # noqa: E402
from lettuce.ext import ExtFlow  # noqa: E402

CPU = torch.device("cpu")


def ctx(dtype):
    return lt.Context(CPU, dtype=dtype, use_native=False)


class SpikeFlow(ExtFlow):
    """lettuce's tests/native/test_native_bgk_collision.py:9-33 flow:
    f = 1 everywhere, f = 2 at (2, 2), D2Q9 16^2."""

    def __init__(self, context, n=16):
        ExtFlow.__init__(self, context, n, 1.0, 1.0)

    def make_resolution(self, resolution, stencil=None):
        return [resolution, resolution]

    def make_units(self, reynolds_number, mach_number, _):
        return lt.UnitConversion(reynolds_number=reynolds_number,
                                 mach_number=mach_number)

    def initial_pu(self):
        ...

    def initialize(self):
        self.f[:, :, :] = 1.0
        self.f[:, 2, 2] = 2.0

    @property
    def post_boundaries(self):
        return []


class PerturbedFlow(ExtFlow):
    """feq(rho=1, u=0.05) times (1 + 0.05 r), r uniform in [-1, 1), seeded;
    units of lettuce's tests TestFlow (Re=100, Ma=0.1, L=resolution[0])."""

    def __init__(self, context, n, stencil, seed):
        self.seed = seed
        ExtFlow.__init__(self, context, [n] * stencil.d, 100.0, 0.1, stencil)

    def make_resolution(self, resolution, stencil=None):
        return resolution

    def make_units(self, reynolds_number, mach_number, resolution):
        return lt.UnitConversion(reynolds_number, mach_number,
                                 characteristic_length_lu=resolution[0])

    def initial_pu(self):
        ...

    def initialize(self):
        d = self.stencil.d
        u = torch.full([d] + self.resolution, 0.05, dtype=self.f.dtype)
        u[1:] = -0.02
        rho = torch.ones([1] + self.resolution, dtype=self.f.dtype)
        feq = self.equilibrium(self, rho=rho, u=u)
        g = torch.Generator().manual_seed(self.seed)
        r = torch.rand(self.f.shape, generator=g, dtype=torch.float64) * 2 - 1
        self.f = feq * (1 + 0.05 * r.to(self.f.dtype))

    @property
    def post_boundaries(self):
        return []


def run(flow, collision, steps, strategy):
    sim = lt.Simulation(flow, collision, [], streaming_strategy=strategy)
    sim(steps)
    return sim


def g1_spike():
    out = {}
    for dtype, tag in ((torch.float32, "f32"), (torch.float64, "f64")):
        for strat in (lt.StreamingStrategy.PRE_STREAMING,
                      lt.StreamingStrategy.POST_STREAMING,
                      lt.StreamingStrategy.NO_STREAMING,
                      lt.StreamingStrategy.DOUBLE_STREAMING):
            for steps in (1, 10):
                flow = SpikeFlow(ctx(dtype))
                run(flow, lt.BGKCollision(2.0), steps, strat)
                out[f"{tag}_{strat.name}_{steps}"] = flow.f.numpy()
    np.savez_compressed(os.path.join(OUT, "g1_spike_bgk_d2q9.npz"), **out)


def g2_tgv3d():
    out = {}
    for stencil, n in ((lt.D3Q19(), 16), (lt.D3Q27(), 8)):
        name = type(stencil).__name__
        flow = lt.TaylorGreenVortex(ctx(torch.float32), n, 1600, 0.1, stencil,
                                    initialize_fneq=False)
        out[f"{name}_f0"] = flow.f.numpy().copy()
        tau = flow.units.relaxation_parameter_lu
        out[f"{name}_tau"] = np.array(tau)
        run(flow, lt.BGKCollision(tau), 10, lt.StreamingStrategy.PRE_STREAMING)
        out[f"{name}_f10"] = flow.f.numpy()
    np.savez_compressed(os.path.join(OUT, "g2_tgv3d_bgk.npz"), **out)


def g3_tgv2d():
    """config: D2Q9 TGV 128^2, Re=100, u0 = 0.05 (Ma = 0.05 sqrt 3)."""
    flow = lt.TaylorGreenVortex(ctx(torch.float32), 128, 100, 0.05 * np.sqrt(3),
                                lt.D2Q9(), initialize_fneq=False)
    f0 = flow.f.numpy().copy()
    tau = flow.units.relaxation_parameter_lu
    ke = lt.IncompressibleKineticEnergy(flow)
    ens = lt.Enstrophy(flow)
    rep_ke = lt.ObservableReporter(ke, interval=100, out=None)
    rep_en = lt.ObservableReporter(ens, interval=100, out=None)
    sim = lt.Simulation(flow, lt.BGKCollision(tau), [rep_ke, rep_en],
                        streaming_strategy=lt.StreamingStrategy.PRE_STREAMING)
    sim(1000)
    np.savez_compressed(
        os.path.join(OUT, "g3_tgv2d_128_bgk_1000.npz"), f0=f0, f1000=flow.f.numpy(),
        tau=np.array(tau), ke=np.array(rep_ke.out), enstrophy=np.array(rep_en.out))


def g4_kbc():
    out = {}
    for stencil, n in ((lt.D2Q9(), 16), (lt.D3Q27(), 6)):
        name = type(stencil).__name__
        for dtype, tag in ((torch.float32, "f32"), (torch.float64, "f64")):
            flow = PerturbedFlow(ctx(dtype), n, stencil, seed=7)
            out[f"{name}_{tag}_f0"] = flow.f.numpy().copy()
            out[f"{name}_{tag}_coll"] = lt.KBCCollision()(flow).numpy()
            run(flow, lt.KBCCollision(), 10, lt.StreamingStrategy.PRE_STREAMING)
            out[f"{name}_{tag}_f10"] = flow.f.numpy()
            out[f"{name}_tau"] = np.array(flow.units.relaxation_parameter_lu)
    np.savez_compressed(os.path.join(OUT, "g4_kbc.npz"), **out)


def g5_trt():
    out = {}
    for stencil, n in ((lt.D2Q9(), 16), (lt.D3Q19(), 6), (lt.D3Q27(), 6)):
        name = type(stencil).__name__
        flow = PerturbedFlow(ctx(torch.float64), n, stencil, seed=3)
        out[f"{name}_f0"] = flow.f.numpy().copy()
        run(flow, lt.TRTCollision(0.8, 1.3), 10, lt.StreamingStrategy.PRE_STREAMING)
        out[f"{name}_f10"] = flow.f.numpy()
    np.savez_compressed(os.path.join(OUT, "g5_trt.npz"), **out)


class EqOutletObstacle(lt.Obstacle):
    """lettuce's Obstacle with EquilibriumOutletP in place of the
    anti-bounce-back outlet (the commented-out alternative at
    lettuce/ext/_flows/obstacle.py:117-118)."""

    @property
    def post_boundaries(self):
        x = self.grid[0]
        return [
            lt.EquilibriumBoundaryPU(
                flow=self, context=self.context, mask=torch.abs(x) < 1e-6,
                velocity=self.units.characteristic_velocity_pu * self._unit_vector()),
            lt.EquilibriumOutletP(direction=[1, 0], flow=self),
            lt.BounceBackBoundary(self.mask),
        ]


def g6_obstacle():
    out = {}
    nx, ny = 48, 24
    flow = EqOutletObstacle(ctx(torch.float64), [nx, ny], reynolds_number=50,
                            mach_number=0.1, domain_length_x=4.0)
    xx, yy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    mask = (xx - 14) ** 2 + (yy - 12) ** 2 < 4.5 ** 2
    flow.mask = mask
    flow.initialize()
    out["mask"] = mask
    out["f0"] = flow.f.numpy().copy()
    out["tau"] = np.array(flow.units.relaxation_parameter_lu)
    sim = lt.Simulation(flow, lt.BGKCollision(flow.units.relaxation_parameter_lu), [])
    out["ncm"] = sim.no_collision_mask.numpy()
    out["nsm"] = sim.no_streaming_mask.numpy()
    sim(10)
    out["f10_post"] = flow.f.numpy()
    np.savez_compressed(os.path.join(OUT, "g6_obstacle_eqoutlet.npz"), **out)


def g7_turbulence():
    out = {}
    flow = lt.DecayingTurbulence(ctx(torch.float64), [32, 32], 1000, 0.1,
                                 k0=4, stencil=lt.D2Q9(), initialize_pressure=False,
                                 initialize_fneq=False, randseed=0)
    out["u0_lu"] = flow.u().numpy()
    out["f0"] = flow.f.numpy()
    out["char_velocity_pu"] = np.array(flow.units.characteristic_velocity_pu)
    flow3 = lt.DecayingTurbulence(ctx(torch.float64), [8, 8, 8], 1000, 0.1, k0=4,
                                  stencil=lt.D3Q27(), initialize_pressure=False,
                                  initialize_fneq=False, randseed=0)
    out["u0_lu_3d"] = flow3.u().numpy()
    np.savez_compressed(os.path.join(OUT, "g7_decaying_turbulence.npz"), **out)


def g8_obstacle_abb():
    """lettuce's default Obstacle (lettuce/ext/_flows/obstacle.py:107-122):
    equilibrium inlet, anti-bounce-back outlet, fullway bounce-back on the
    mask; D2Q9 BGK 48 x 24 and D3Q27 KBC 16 x 8 x 8, FP64, 10 steps in
    lettuce's default order (collide, then stream) and 10 pull steps."""
    out = {}
    cases = (("D2Q9", lt.D2Q9(), [48, 24], 4.0, "bgk"),
             ("D3Q27", lt.D3Q27(), [16, 8, 8], 2.0, "kbc"))
    for name, stencil, res, length_x, coll in cases:
        for strat in (lt.StreamingStrategy.POST_STREAMING, lt.StreamingStrategy.PRE_STREAMING):
            flow = lt.Obstacle(ctx(torch.float64), res, reynolds_number=50, mach_number=0.1,
                               domain_length_x=length_x, stencil=stencil)
            grid = np.meshgrid(*[np.arange(n) for n in res], indexing="ij")
            if len(res) == 2:
                mask = (grid[0] - 14) ** 2 + (grid[1] - 12) ** 2 < 4.5 ** 2
            else:
                mask = ((grid[0] - 5) ** 2 + (grid[1] - 3.5) ** 2 + (grid[2] - 4) ** 2
                        < 2.5 ** 2)
            flow.mask = mask
            flow.initialize()
            tau = flow.units.relaxation_parameter_lu
            collision = lt.BGKCollision(tau) if coll == "bgk" else lt.KBCCollision()
            sim = lt.Simulation(flow, collision, [], streaming_strategy=strat)
            tag = "post" if strat == lt.StreamingStrategy.POST_STREAMING else "pre"
            if tag == "post":
                out[f"{name}_mask"] = mask
                out[f"{name}_f0"] = flow.f.numpy().copy()
                out[f"{name}_tau"] = np.array(tau)
                out[f"{name}_ncm"] = sim.no_collision_mask.numpy()
                out[f"{name}_nsm"] = sim.no_streaming_mask.numpy()
            sim(10)
            out[f"{name}_f10_{tag}"] = flow.f.numpy()
    np.savez_compressed(os.path.join(OUT, "g8_obstacle_abb.npz"), **out)


def g9_poiseuille_guo():
    """lettuce's force test case (tests/collision/test_force.py:6-32):
    PoiseuilleFlow2D 17^2, Re 1, Ma 0.02, zero start, BGK with the Guo force
    (lettuce/ext/_force/guo.py:9-43), bounce-back walls, FP64, lettuce's
    default order; the state after 10 and 100 steps and the velocity
    u(acceleration) after 100 steps."""
    flow = lt.PoiseuilleFlow2D(context=ctx(torch.float64), resolution=17, reynolds_number=1,
                               mach_number=0.02, initialize_with_zeros=True)
    acc = flow.units.convert_acceleration_to_lu(flow.acceleration)
    tau = flow.units.relaxation_parameter_lu
    force = lt.Guo(flow=flow, tau=tau, acceleration=acc)
    sim = lt.Simulation(flow=flow, collision=lt.BGKCollision(tau=tau, force=force), reporter=[])
    out = {"f0": flow.f.numpy().copy(), "tau": np.array(tau),
           "acceleration_lu": acc.numpy(), "ncm": sim.no_collision_mask.numpy(),
           "nsm": sim.no_streaming_mask.numpy()}
    sim(10)
    out["f10"] = flow.f.numpy().copy()
    sim(90)
    out["f100"] = flow.f.numpy()
    out["u100"] = flow.u(acceleration=acc).numpy()
    np.savez_compressed(os.path.join(OUT, "g9_poiseuille_guo.npz"), **out)


if __name__ == "__main__":
    only = sys.argv[1:]
    for fn in (g1_spike, g2_tgv3d, g3_tgv2d, g4_kbc, g5_trt, g6_obstacle, g7_turbulence,
               g8_obstacle_abb, g9_poiseuille_guo):
        if only and fn.__name__ not in only:
            continue
        fn()
        print("wrote", fn.__name__)

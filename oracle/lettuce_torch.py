"""CPU oracle: a restatement of lettuce's torch time step.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg may import this module, and only as the checker
(or the timed CPU baseline); the lettuce_amd package never imports it.

It follows lettuce's torch path operation for operation (torch.roll
streaming, torch.where boundary selection, tensordot/einsum equilibrium,
the KBC moment transforms) so that its FP32 rounding is lettuce's.  Each
function cites the lettuce file:line it restates.  The class names are
lettuce's so that parity tests read like lettuce's own tests.

Pinning (what checks each part; see tests/test_oracle.py):
  stencils            symmetry / weights / opposite invariants
  equilibrium         conservation of rho and j, feq(1, 0) = w, golden vectors
  stream              one-hot shift known-answer test, golden vectors
  BGK                 tau = 1/2 applied twice is the identity, tau = 1 gives
                      feq, shear-moment relaxation, golden vectors
  TRT                 tau_minus = tau reduces to BGK, golden vectors
  KBC                 conservation, shear relaxation, entropy >= BGK,
                      gamma = 2 reduces to BGK, golden vectors
  bounce-back         f[opposite], golden vectors
  equilibrium BC      hand-computed feq
  equilibrium outlet  feq(rho_out, u of the neighbour column), golden vectors
  halfway bounce-back brute-force link rule on tiny grids, Couette/Poiseuille
                      wall velocity, exact mass conservation
  masks               golden ncm/nsm from lettuce's Simulation
  TGV / observables   analytic decay, golden vectors
  Guo forcing         a = 0 is BGK; mass kept and momentum raised by exactly
                      a per cell and step (closed form)
  anti-bounce-back    a uniform equilibrium is a fixed point; the wall
                      velocity extrapolation by hand
  learned collision   zero MLP output = BGK, rates 1/tau = BGK (D2Q9, D3Q27),
                      rates beta*gamma = KBC, dual basis = inverse moment
                      matrix (closed form), forced
                      rates against the inference formula; training step's
                      autograd gradient against central differences
"""
import math
from typing import List, Optional

import numpy as np
import torch

__all__ = [
    "Context", "D2Q9", "D3Q19", "D3Q27", "UnitConversion",
    "QuadraticEquilibrium", "NoCollision", "BGKCollision", "TRTCollision",
    "KBCCollision", "BounceBackBoundary", "EquilibriumBoundaryPU",
    "EquilibriumOutletP", "AntiBounceBackOutlet", "HalfwayBounceBackBoundary", "Flow",
    "Guo",
    "TaylorGreenVortex", "DecayingTurbulence", "ObstacleChannel",
    "Simulation", "StreamingStrategy", "torch_gradient",
    "incompressible_kinetic_energy", "enstrophy", "mass",
    "LearnedParameters", "LearnedCollision", "learned_train_collide",
    "learned_train_step", "split_learned_params",
]


# --------------------------------------------------------------------------
# context, stencils, units  (lettuce/_context.py, ext/_stencil/*, _unit.py)
# --------------------------------------------------------------------------
class Context:
    def __init__(self, device="cpu", dtype=torch.float32):
        self.device = torch.device(device)
        self.dtype = dtype

    def convert_to_tensor(self, a, dtype=None):
        if dtype is None:
            if hasattr(a, "dtype") and a.dtype in (bool, torch.bool, torch.uint8,
                                                   np.uint8, np.dtype(bool)):
                dtype = torch.uint8
            else:
                dtype = self.dtype
        if isinstance(a, torch.Tensor):
            return a.to(device=self.device, dtype=dtype)
        return torch.from_numpy(np.ascontiguousarray(np.array(a))).to(
            device=self.device, dtype=dtype)


class _Stencil:
    cs = 1 / np.sqrt(3.0)

    @property
    def d(self):
        return len(self.e[0])

    @property
    def q(self):
        return len(self.e)


class D2Q9(_Stencil):  # lettuce/ext/_stencil/d2q9.py:8-10
    e = [[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1], [1, 1], [-1, 1], [-1, -1], [1, -1]]
    w = [4.0 / 9.0] + [1.0 / 9.0] * 4 + [1.0 / 36.0] * 4
    opposite = [0, 3, 4, 1, 2, 7, 8, 5, 6]


class D3Q19(_Stencil):  # lettuce/ext/_stencil/d3q19.py:8-13
    e = [[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
         [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1], [1, 0, 1], [-1, 0, -1],
         [1, 0, -1], [-1, 0, 1], [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0]]
    w = [1.0 / 3.0] + [1.0 / 18.0] * 6 + [1.0 / 36.0] * 12
    opposite = [0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17]


class D3Q27(_Stencil):  # lettuce/ext/_stencil/d3q27.py:8-12
    e = [[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
         [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1], [1, 0, 1], [-1, 0, -1],
         [1, 0, -1], [-1, 0, 1], [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0],
         [1, 1, 1], [-1, -1, -1], [1, 1, -1], [-1, -1, 1], [1, -1, 1], [-1, 1, -1],
         [1, -1, -1], [-1, 1, 1]]
    w = [8.0 / 27.0] + [2.0 / 27.0] * 6 + [1.0 / 54.0] * 12 + [1.0 / 216.0] * 8
    opposite = [0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17,
                20, 19, 22, 21, 24, 23, 26, 25]


class UnitConversion:  # lettuce/_unit.py:13-145 (the conversions used here)
    def __init__(self, reynolds_number, mach_number=0.05, characteristic_length_pu=1,
                 characteristic_velocity_pu=1, characteristic_length_lu=1,
                 characteristic_density_lu=1, characteristic_density_pu=1,
                 cs=1 / np.sqrt(3.0)):
        self.cs = cs
        self.reynolds_number = reynolds_number
        self.mach_number = mach_number
        self.characteristic_length_pu = characteristic_length_pu
        self.characteristic_velocity_pu = characteristic_velocity_pu
        self.characteristic_length_lu = characteristic_length_lu
        self.characteristic_density_lu = characteristic_density_lu
        self.characteristic_density_pu = characteristic_density_pu

    @property
    def characteristic_velocity_lu(self):
        return self.cs * self.mach_number

    @property
    def characteristic_pressure_pu(self):
        return self.characteristic_density_pu * self.characteristic_velocity_pu ** 2

    @property
    def characteristic_pressure_lu(self):
        return self.characteristic_density_lu * self.characteristic_velocity_lu ** 2

    @property
    def viscosity_lu(self):
        return self.characteristic_length_lu * self.characteristic_velocity_lu / self.reynolds_number

    @property
    def viscosity_pu(self):
        return self.characteristic_length_pu * self.characteristic_velocity_pu / self.reynolds_number

    @property
    def relaxation_parameter_lu(self):
        return self.viscosity_lu / self.cs ** 2 + 0.5

    def convert_velocity_to_pu(self, v):
        return v / self.characteristic_velocity_lu * self.characteristic_velocity_pu

    def convert_velocity_to_lu(self, v):
        return v / self.characteristic_velocity_pu * self.characteristic_velocity_lu

    def convert_time_to_pu(self, t):
        return t / (self.characteristic_length_lu / self.characteristic_velocity_lu) * (
            self.characteristic_length_pu / self.characteristic_velocity_pu)

    def convert_pressure_to_lu(self, p):
        return p / self.characteristic_pressure_pu * self.characteristic_pressure_lu

    def convert_pressure_pu_to_density_lu(self, p):
        return self.convert_pressure_to_lu(p) / self.cs ** 2 + self.characteristic_density_lu

    def convert_length_to_pu(self, x):
        return x * self.characteristic_length_pu / self.characteristic_length_lu

    def convert_incompressible_energy_to_pu(self, e):
        return e * self.characteristic_velocity_pu ** 2 / self.characteristic_velocity_lu ** 2


class _TorchStencil:  # lettuce/_stencil.py:31-50
    def __init__(self, stencil, context):
        self.e = context.convert_to_tensor(stencil.e)
        self.w = context.convert_to_tensor(stencil.w)
        self.cs = 1 / np.sqrt(3.0)


# --------------------------------------------------------------------------
# flow state (lettuce/_flow.py:55-204)
# --------------------------------------------------------------------------
class Flow:
    def __init__(self, context, resolution, units, stencil, equilibrium=None):
        self.context = context
        self.resolution = list(resolution)
        self.units = units
        self.stencil = stencil
        self.torch_stencil = _TorchStencil(stencil, context)
        self.equilibrium = equilibrium or QuadraticEquilibrium()
        self.i = 0
        self.f = torch.empty([stencil.q] + self.resolution, dtype=context.dtype)

    pre_boundaries: List = []
    post_boundaries: List = []
    post_streaming_boundaries: List = []

    def initialize_equilibrium(self, p, u):  # lettuce/_flow.py:127-143
        rho = self.context.convert_to_tensor(self.units.convert_pressure_pu_to_density_lu(p))
        u = self.context.convert_to_tensor(self.units.convert_velocity_to_lu(u))
        self.f = self.equilibrium(self, rho=rho, u=u)

    def rho(self, f=None):  # :157-159
        return torch.sum(self.f if f is None else f, dim=0)[None, ...]

    def j(self, f=None):  # :173-176
        return torch.einsum("qd,q...->d...", [self.torch_stencil.e, self.f if f is None else f])

    def u(self, f=None, rho=None):  # :178-193
        rho = self.rho(f=f) if rho is None else rho
        return self.j(f=f) / rho

    def incompressible_energy(self, f=None):  # :200-204
        f = self.f if f is None else f
        return 0.5 * torch.einsum("d...,d...->...", [self.u(f), self.u(f)])

    def shear_tensor(self, f=None):
        e = self.torch_stencil.e
        shear = torch.einsum("qa,qb->qab", [e, e])
        return torch.einsum("q...,qab->ab...", [self.f if f is None else f, shear])

    def pseudo_entropy_local(self, f=None):
        f = self.f if f is None else f
        f_feq = f / self.equilibrium(self)
        return self.rho(f) - torch.einsum("q...,q...->...", [f, f_feq])


class QuadraticEquilibrium:  # lettuce/ext/_equilibrium/quadratic_equilibrium.py:11-24
    def __call__(self, flow, rho=None, u=None):
        rho = flow.rho() if rho is None else rho
        u = flow.u() if u is None else u
        exu = torch.tensordot(flow.torch_stencil.e, u, dims=1)
        uxu = torch.einsum("d...,d...->...", [u, u])
        cs2 = flow.torch_stencil.cs ** 2
        return torch.einsum("q...,q...->q...", [
            flow.torch_stencil.w,
            rho * ((2 * exu - uxu) / (2 * cs2) + 0.5 * (exu / cs2) ** 2 + 1)])


# --------------------------------------------------------------------------
# collisions (lettuce/ext/_collision/*)
# --------------------------------------------------------------------------
class NoCollision:  # no_collision.py:10-11
    def __call__(self, flow):
        return flow.f


class BGKCollision:  # bgk_collision.py:12-22
    def __init__(self, tau, force=None):
        self.tau = tau
        self.force = force

    def __call__(self, flow):
        u_eq = 0 if self.force is None else self.force.u_eq(flow)
        u = flow.u() + u_eq
        feq = flow.equilibrium(flow, u=u)
        si = self.force.source_term(u) if self.force is not None else 0
        return flow.f - 1.0 / self.tau * (flow.f - feq) + si


class Guo:  # ext/_force/guo.py:9-43
    def __init__(self, flow, tau, acceleration):
        self.flow = flow
        self.tau = tau
        self.acceleration = torch.as_tensor(acceleration, dtype=flow.context.dtype)

    def source_term(self, u):  # :15-26
        e = torch.tensor(self.flow.stencil.e, dtype=u.dtype)
        w = torch.tensor(self.flow.stencil.w, dtype=u.dtype)
        d = self.flow.stencil.d
        cs = self.flow.stencil.cs
        emu = e.reshape(list(e.shape) + [1] * d) - u
        eu = torch.einsum("ib,b...->i...", e, u)
        eeu = torch.einsum("ia,i...->ia...", e, eu)
        emu_eeu = emu / cs ** 2 + eeu / cs ** 4
        emu_eeu_f = torch.einsum("ia...,a->i...", emu_eeu, self.acceleration)
        return (1 - 1 / (2 * self.tau)) * w.reshape([-1] + [1] * d) * emu_eeu_f

    def u_eq(self, flow=None):  # :28-32
        flow = self.flow if flow is None else flow
        a = self.acceleration.reshape([-1] + [1] * flow.stencil.d)
        return 0.5 * a / flow.rho()


class TRTCollision:  # trt_collision.py:16-27
    def __init__(self, tau, tau_minus=1.0):
        self.tau_plus = tau
        self.tau_minus = tau_minus

    def __call__(self, flow):
        feq = flow.equilibrium(flow)
        opp = flow.stencil.opposite
        d = (((flow.f + flow.f[opp]) - (feq + feq[opp])) / (2.0 * self.tau_plus))
        d += (((flow.f - flow.f[opp]) - (feq - feq[opp])) / (2.0 * self.tau_minus))
        return flow.f - d


class KBCCollision:  # kbc_collision.py:11-160
    def __init__(self, tau=None):
        self.tau = tau
        self.beta = None
        self.M = None
        self.force_gamma = None   # test hook: fixed gamma instead of the entropic one

    def _moments(self, f):  # :22-39
        if f.dim() == 4:
            m = torch.einsum("abcq,qmno", self.M, f)
            rho = m[0, 0, 0]
            m = m / rho
            m[0, 0, 0] = rho
        else:
            m = torch.einsum("abq,qmn", self.M, f)
            rho = m[0, 0]
            m = m / rho
            m[0, 0] = rho
        return m

    @staticmethod
    def _s3d(f, m):  # :41-74
        s = torch.zeros_like(f)
        T = m[2, 0, 0] + m[0, 2, 0] + m[0, 0, 2]
        N_xz = m[2, 0, 0] - m[0, 0, 2]
        N_yz = m[0, 2, 0] - m[0, 0, 2]
        Pi_xy, Pi_xz, Pi_yz = m[1, 1, 0], m[1, 0, 1], m[0, 1, 1]
        r = m[0, 0, 0]
        s[0] = r * -T
        s[1] = 1. / 6. * r * (2 * N_xz - N_yz + T)
        s[2] = s[1]
        s[3] = 1. / 6. * r * (2 * N_yz - N_xz + T)
        s[4] = s[3]
        s[5] = 1. / 6. * r * (-N_xz - N_yz + T)
        s[6] = s[5]
        s[7] = 1. / 4 * r * Pi_yz
        s[8] = s[7]
        s[9] = -1. / 4 * r * Pi_yz
        s[10] = s[9]
        s[11] = 1. / 4 * r * Pi_xz
        s[12] = s[11]
        s[13] = -1. / 4 * r * Pi_xz
        s[14] = s[13]
        s[15] = 1. / 4 * r * Pi_xy
        s[16] = s[15]
        s[17] = -1. / 4 * r * Pi_xy
        s[18] = s[17]
        return s

    @staticmethod
    def _s2d(f, m):  # :76-94
        s = torch.zeros_like(f)
        T = m[2, 0] + m[0, 2]
        N = m[2, 0] - m[0, 2]
        Pi_xy = m[1, 1]
        r = m[0, 0]
        s[0] = r * -T
        s[1] = 1. / 2. * r * (0.5 * (T + N))
        s[2] = 1. / 2. * r * (0.5 * (T - N))
        s[3] = 1. / 2. * r * (0.5 * (T + N))
        s[4] = 1. / 2. * r * (0.5 * (T - N))
        s[5] = 1. / 4. * r * Pi_xy
        s[6] = -s[5]
        s[7] = 1. / 4 * r * Pi_xy
        s[8] = -s[7]
        return s

    def _setup(self, flow):  # :97-125
        self.tau = flow.units.relaxation_parameter_lu
        self.beta = 1. / (2 * self.tau)
        e = flow.torch_stencil.e
        if flow.stencil.d == 3:
            assert isinstance(flow.stencil, D3Q27)
            self.M = torch.zeros([3, 3, 3, 27], dtype=flow.context.dtype)
            for i in range(3):
                for j in range(3):
                    for k in range(3):
                        self.M[i, j, k] = e[:, 0] ** i * e[:, 1] ** j * e[:, 2] ** k
            self._s = self._s3d
        else:
            assert isinstance(flow.stencil, D2Q9)
            self.M = torch.zeros([3, 3, 9], dtype=flow.context.dtype)
            for i in range(3):
                for j in range(3):
                    self.M[i, j] = e[:, 0] ** i * e[:, 1] ** j
            self._s = self._s2d

    def __call__(self, flow):  # :126-160
        if self.M is None:
            self._setup(flow)
        feq = flow.equilibrium(flow)
        delta_s = self._s(flow.f, self._moments(flow.f))
        delta_s -= self._s(flow.f, self._moments(feq))
        delta_h = flow.f - feq - delta_s
        sum_s = flow.rho(delta_s * delta_h / feq)
        sum_h = flow.rho(delta_h * delta_h / feq)
        gamma = 1. / self.beta - (2 - 1. / self.beta) * sum_s / sum_h
        gamma[gamma < 1E-15] = 2.0
        gamma[torch.isnan(gamma)] = 2.0
        if self.force_gamma is not None:
            gamma = torch.full_like(gamma, self.force_gamma)
        return flow.f - self.beta * (2 * delta_s + gamma * delta_h)


# --------------------------------------------------------------------------
# boundaries (lettuce/ext/_boundary/*, examples/.../halfway_bounce_back_boundary.py)
# --------------------------------------------------------------------------
class BounceBackBoundary:  # bounce_back_boundary.py:10-32
    def __init__(self, mask):
        self.mask = mask

    def __call__(self, flow):
        return flow.f[flow.stencil.opposite]

    def make_no_collision_mask(self, shape, context):
        return self.mask

    def make_no_streaming_mask(self, shape, context):
        return None


class EquilibriumBoundaryPU:  # equilibrium_boundary_pu.py:71-98 (broadcast tensors)
    def __init__(self, context, flow, mask, velocity, pressure=0):
        d = flow.stencil.d
        nd = len(flow.resolution)
        v = torch.as_tensor(velocity, dtype=context.dtype)
        if v.dim() == 1:
            v = v.reshape([d] + [1] * nd)
        p = torch.as_tensor(pressure, dtype=context.dtype)
        if p.dim() == 0:
            p = p.reshape([1] * (nd + 1))
        elif p.dim() == nd:
            p = p[None]
        self.velocity, self.pressure, self.mask = v, p, mask

    def __call__(self, flow):
        rho = flow.units.convert_pressure_pu_to_density_lu(self.pressure)
        u = flow.units.convert_velocity_to_lu(self.velocity)
        feq = flow.equilibrium(flow, rho, u)
        return torch.einsum("q...,q...->q...", [feq, torch.ones_like(flow.f)])

    def make_no_collision_mask(self, shape, context):
        return self.mask

    def make_no_streaming_mask(self, shape, context):
        return None


class EquilibriumOutletP:  # equilibrium_outlet_p.py:12-91
    def __init__(self, direction, flow, rho_outlet=1.0):
        self.rho_outlet = torch.tensor(rho_outlet, dtype=flow.context.dtype)
        direction = np.array(direction)
        self.velocities = np.concatenate(np.argwhere(
            np.matmul(flow.stencil.e, direction) > 1 - 1e-6), axis=0)
        self.index, self.neighbor = [], []
        for c in direction:
            if c == 0:
                self.index.append(slice(None))
                self.neighbor.append(slice(None))
            if c == 1:
                self.index.append(-1)
                self.neighbor.append(-2)
            if c == -1:
                self.index.append(0)
                self.neighbor.append(1)

    def __call__(self, flow):
        here = tuple([slice(None)] + self.index)
        other = tuple([slice(None)] + self.neighbor)
        rho = flow.rho()
        u = flow.u()
        rho_w = self.rho_outlet * torch.ones_like(rho[here])
        u_w = u[other]
        feq = flow.f   # lettuce writes the outlet column of flow.f in place
        feq[here] = flow.equilibrium(flow, rho_w[..., None], u_w[..., None])[..., 0]
        return torch.einsum("q...,q...->q...", [feq, torch.ones_like(flow.f)])

    def make_no_streaming_mask(self, shape, context):
        m = torch.zeros(shape, dtype=torch.bool)
        m[tuple([np.setdiff1d(np.arange(shape[0]), self.velocities)] + self.index)] = 1
        return m

    def make_no_collision_mask(self, shape, context):
        m = torch.zeros(shape, dtype=torch.bool)
        m[tuple(self.index)] = 1
        return m


class AntiBounceBackOutlet:  # anti_bounce_back_outlet.py:13-109
    """f_opp(v) <- -f_v + w_v rho (2 + (e_v.u_w)^2 / cs^4 - |u_w|^2 / cs^2) on
    the outlet column for the velocities v leaving the domain, with
    u_w = u + (u - u_neighbour) / 2 (:79-92)"""

    def __init__(self, direction, flow):
        direction = np.array(direction)
        self.velocities = np.concatenate(np.argwhere(
            np.matmul(flow.stencil.e, direction) > 1 - 1e-6), axis=0)
        self.opposite = np.array(flow.stencil.opposite)
        self.index, self.neighbor = [], []
        for c in direction:
            if c == 0:
                self.index.append(slice(None))
                self.neighbor.append(slice(None))
            if c == 1:
                self.index.append(-1)
                self.neighbor.append(-2)
            if c == -1:
                self.index.append(0)
                self.neighbor.append(1)
        w = torch.tensor(flow.stencil.w, dtype=flow.context.dtype)[self.velocities]
        nd = len(direction)
        self.w = w.reshape([-1] + [1] * (nd - 1))
        self.e = torch.tensor(flow.stencil.e, dtype=flow.context.dtype)[self.velocities]

    def __call__(self, flow):
        here = tuple([slice(None)] + self.index)
        other = tuple([slice(None)] + self.neighbor)
        u = flow.u()
        u_w = u[here] + 0.5 * (u[here] - u[other])
        cs = flow.stencil.cs
        f = flow.f
        eu = torch.einsum("dc,c...->d...", self.e, u_w)
        f[tuple([self.opposite[self.velocities]] + self.index)] = (
            -flow.f[tuple([self.velocities] + self.index)]
            + self.w * flow.rho()[here]
            * (2 + eu ** 2 / cs ** 4 - (torch.norm(u_w, dim=0) / cs) ** 2))
        return f

    def make_no_streaming_mask(self, shape, context):
        m = torch.zeros(shape, dtype=torch.bool)
        m[tuple([self.opposite[self.velocities]] + self.index)] = 1
        return m

    def make_no_collision_mask(self, shape, context):
        m = torch.zeros(shape, dtype=torch.bool)
        m[tuple(self.index)] = 1
        return m


class HalfwayBounceBackBoundary:
    """Post-streaming boundary (halfway_bounce_back_boundary.py:170-209,
    ebb_simulation.py:60-106): links from solid p to fluid p + e_i store
    f*_opp(i)(p + e_i) before streaming and write it into f_i(p + e_i)
    after streaming.  Links are found with periodic neighbours."""

    def __init__(self, context, flow, mask):
        self.mask = torch.as_tensor(mask, dtype=torch.bool)
        stencil = flow.stencil
        res = list(self.mask.shape)
        rows = []
        for p in torch.nonzero(self.mask).tolist():
            for i in range(stencil.q):
                if not any(stencil.e[i]):
                    continue
                nb = [(p[a] + stencil.e[i][a]) % res[a] for a in range(len(res))]
                if not bool(self.mask[tuple(nb)]):
                    rows.append([stencil.opposite[i]] + nb)
        self.f_index = torch.tensor(rows, dtype=torch.int64).reshape(-1, 1 + len(res))
        self.opposite = torch.tensor(stencil.opposite, dtype=torch.int64)
        self.f_collided = None

    def store_f_collided(self, f):
        idx = [self.f_index[:, k] for k in range(self.f_index.shape[1])]
        self.f_collided = f[tuple(idx)].clone()

    def __call__(self, flow):
        pos = [self.f_index[:, k] for k in range(1, self.f_index.shape[1])]
        flow.f[tuple([self.opposite[self.f_index[:, 0]]] + pos)] = self.f_collided

    def force(self, flow):
        """2 sum f*_d e_d over links (halfway_bounce_back_boundary.py:211-218)"""
        e = torch.tensor(flow.stencil.e, dtype=torch.float64)
        return 2 * torch.einsum("i,id->d", self.f_collided.to(torch.float64),
                                e[self.f_index[:, 0]])

    def make_no_collision_mask(self, shape, context):
        return self.mask

    def make_no_streaming_mask(self, shape, context):
        return self.mask


# --------------------------------------------------------------------------
# simulation (lettuce/_simulation.py:42-344 + ebb_simulation.py:11-106)
# --------------------------------------------------------------------------
class StreamingStrategy:
    NO_STREAMING = 0b00
    PRE_STREAMING = 0b10
    POST_STREAMING = 0b01
    DOUBLE_STREAMING = 0b11


class Simulation:
    def __init__(self, flow, collision, reporter=(), streaming_strategy=StreamingStrategy.POST_STREAMING):
        self.flow = flow
        self.collision = collision
        self.reporter = list(reporter)
        self.strategy = getattr(streaming_strategy, "value", streaming_strategy)
        self.pre_boundaries = list(flow.pre_boundaries)
        self.post_boundaries = list(flow.post_boundaries)
        self.psb = list(flow.post_streaming_boundaries)
        self.collision_index = len(self.pre_boundaries)
        self.no_collision_mask = None
        self.no_streaming_mask = None
        q, res = flow.stencil.q, list(flow.resolution)
        if self.pre_boundaries or self.post_boundaries or self.psb:   # :100-146
            ci = self.collision_index
            self.no_collision_mask = torch.full(res, ci, dtype=torch.uint8)
            self.no_streaming_mask = torch.full([q] + res, ci, dtype=torch.uint8)
            order = (list(enumerate(self.pre_boundaries))
                     + list(enumerate(self.post_boundaries, start=ci + 1))
                     + list(enumerate(self.psb, start=ci + 1 + len(self.post_boundaries))))
            for i, b in order:
                m = b.make_no_collision_mask(res, flow.context)
                if m is not None:
                    self.no_collision_mask[torch.as_tensor(m).to(torch.bool)] = i
                s = b.make_no_streaming_mask([q] + res, flow.context)
                if s is not None:
                    self.no_streaming_mask |= torch.as_tensor(s).to(torch.bool)

    def _stream(self):  # :241-256, plus the post-streaming boundaries
        f = self.flow.f
        for b in self.psb:
            b.store_f_collided(f)
        d = self.flow.stencil.d
        for i in range(1, self.flow.stencil.q):
            new = torch.roll(f[i], shifts=tuple(self.flow.stencil.e[i]), dims=tuple(range(d)))
            if self.no_streaming_mask is None:
                f[i] = new
            else:
                f[i] = torch.where(torch.eq(self.no_streaming_mask[i], 1), f[i], new)
        for b in self.psb:
            b(self.flow)

    def _collide(self):  # :258-305
        flow = self.flow
        if self.no_collision_mask is None:
            flow.f = self.collision(flow)
            return
        ncm = self.no_collision_mask
        for i, b in enumerate(self.pre_boundaries):
            torch.where(torch.eq(ncm, i), b(flow), flow.f, out=flow.f)
        torch.where(torch.eq(ncm, self.collision_index), self.collision(flow), flow.f, out=flow.f)
        for i, b in enumerate(self.post_boundaries, start=self.collision_index + 1):
            torch.where(torch.eq(ncm, i), b(flow), flow.f, out=flow.f)

    def step(self):
        pre, post = bool(self.strategy & 0b10), bool(self.strategy & 0b01)
        if pre:
            self._stream()
        self._collide()
        if post:
            self._stream()

    def __call__(self, num_steps):  # :311-323
        if self.flow.i == 0:
            for r in self.reporter:
                r(self)
        for _ in range(num_steps):
            self.step()
            self.flow.i += 1
            for r in self.reporter:
                r(self)


# --------------------------------------------------------------------------
# flows (lettuce/ext/_flows/*)
# --------------------------------------------------------------------------
def _tgv_units(resolution, reynolds_number, mach_number):  # taylorgreen.py:52-58
    return UnitConversion(reynolds_number=reynolds_number, mach_number=mach_number,
                          characteristic_length_lu=resolution[0] / (2 * torch.pi),
                          characteristic_length_pu=1, characteristic_velocity_pu=1)


class TaylorGreenVortex(Flow):  # taylorgreen.py:16-98
    def __init__(self, context, resolution, reynolds_number, mach_number, stencil):
        res = [resolution] * stencil.d if isinstance(resolution, int) else list(resolution)
        super().__init__(context, res, _tgv_units(res, reynolds_number, mach_number), stencil)
        p, u = self.analytic_solution(0)
        self.initialize_equilibrium(p, u)

    @property
    def grid(self):
        xyz = tuple(torch.linspace(0, 2 * torch.pi * (1 - 1 / n), steps=n, dtype=self.context.dtype)
                    for n in self.resolution)
        return torch.meshgrid(*xyz, indexing="ij")

    def analytic_solution(self, t):
        x = self.grid
        nu = self.context.convert_to_tensor(self.units.viscosity_pu)
        if len(self.resolution) == 2:
            u = torch.stack([torch.cos(x[0]) * torch.sin(x[1]) * torch.exp(-2 * nu * t),
                             -torch.sin(x[0]) * torch.cos(x[1]) * torch.exp(-2 * nu * t)])
            p = -torch.stack([0.25 * (torch.cos(2 * x[0]) + torch.cos(2 * x[1]))
                              * torch.exp(-4 * nu * t)])
        else:
            u = torch.stack([torch.sin(x[0]) * torch.cos(x[1]) * torch.cos(x[2]),
                             -torch.cos(x[0]) * torch.sin(x[1]) * torch.cos(x[2]),
                             torch.zeros_like(x[0])])
            p = torch.stack([1 / 16. * (torch.cos(2 * x[0]) + torch.cos(2 * x[1]))
                             * (torch.cos(2 * x[2]) + 2)])
        return p, u


class DecayingTurbulence(Flow):
    """decayingturbulence.py:54-170 without pressure / f_neq initialisation:
    returns the seeded velocity field (PU) and sets f to equilibrium."""

    def __init__(self, context, resolution, reynolds_number, mach_number, stencil,
                 k0=20, ic_energy=0.5, randseed=None):
        res = [resolution] * stencil.d if isinstance(resolution, int) else list(resolution)
        units = UnitConversion(reynolds_number=reynolds_number, mach_number=mach_number,
                               characteristic_length_lu=res[0],
                               characteristic_length_pu=2 * np.pi,
                               characteristic_velocity_pu=None)
        super().__init__(context, res, units, stencil)
        self.k0, self.ic_energy, self.randseed = k0, ic_energy, randseed
        u = self.initial_velocity()
        units.characteristic_velocity_pu = np.linalg.norm(u, axis=0).max()
        self.initialize_equilibrium(np.zeros([1] + res), u)

    def initial_velocity(self):
        d, dims = self.stencil.d, tuple(self.resolution)
        axes = tuple(range(d))
        freqs = [np.fft.fftfreq(n, d=1 / n) for n in dims]
        k = np.meshgrid(*freqs)   # lettuce's default 'xy' indexing (:60-63)
        kn = np.linalg.norm(k, axis=0)
        ek = kn ** 4 * np.exp(-2 * (kn / self.k0) ** 2)
        ek /= np.sum(ek)
        ek *= self.ic_energy
        dx = self.units.convert_length_to_pu(1.0)
        np.random.seed(self.randseed)
        u = np.random.random(np.array(k).shape) * 2 * np.pi + 0j
        u = [np.fft.fftn(u[a], axes=axes) for a in range(d)]
        ur = [u[a].real for a in range(d)]
        ui = [u[a].imag for a in range(d)]
        for a in range(d):
            ur[a].ravel()[0] = 0
            ui[a].ravel()[0] = 0
        urh = [np.sqrt(2 / d * ek / (ui[a] ** 2 + ur[a] ** 2 + 1.e-15)) * ur[a] for a in range(d)]
        uih = [np.sqrt(2 / d * ek / (ui[a] ** 2 + ur[a] ** 2 + 1.e-15)) * ui[a] for a in range(d)]
        for a in range(d):
            urh[a].ravel()[0] = 0
            uih[a].ravel()[0] = 0
        km = [np.sin(k[a] * dx) / dx for a in range(d)]
        kmn = np.linalg.norm(km, axis=0) + 1e-16
        dr = np.zeros(dims)
        di = np.zeros(dims)
        for a in range(d):
            dr += km[a] * urh[a]
            di += km[a] * uih[a]
        ur = [urh[a] - dr * km[a] / kmn ** 2 for a in range(d)]
        ui = [uih[a] - di * km[a] / kmn ** 2 for a in range(d)]
        for a in range(d):
            ur[a].ravel()[0] = 0
            ui[a].ravel()[0] = 0
        e_kin = np.sum([np.sum(ur[a] ** 2 + ui[a] ** 2) for a in range(d)]) * .5
        factor = np.sqrt(self.ic_energy / e_kin)
        ur = [r * factor for r in ur]
        ui = [i * factor for i in ui]
        norm = (self.resolution[0] * dx ** (1 - d) * np.sqrt(self.units.characteristic_length_pu)
                if d == 3 else self.resolution[0] / dx)
        return np.asarray([(np.fft.ifftn(ur[a] + ui[a] * 1.0j, axes=axes) * norm).real
                           for a in range(d)])


class ObstacleChannel(Flow):
    """Channel past a solid mask (obstacle.py:16-125 geometry and units):
    equilibrium inlet at x = 0 (EquilibriumBoundaryPU), EquilibriumOutletP
    at x = -1 and halfway bounce-back on mask | walls, the walls being the
    first and last row of every lateral axis without the inlet/outlet
    columns."""

    def __init__(self, context, resolution, reynolds_number, mach_number,
                 domain_length_x, mask, stencil, walls=True, char_length=1):
        char_length_lu = resolution[0] / domain_length_x * char_length
        units = UnitConversion(reynolds_number=reynolds_number, mach_number=mach_number,
                               characteristic_length_lu=char_length_lu,
                               characteristic_length_pu=char_length,
                               characteristic_velocity_pu=1)
        super().__init__(context, resolution, units, stencil)
        solid = torch.as_tensor(mask, dtype=torch.bool).clone()
        if walls:
            nd = len(resolution)
            for a in range(1, nd):
                lo = [slice(1, -1)] + [slice(None)] * (nd - 1)
                hi = list(lo)
                lo[a], hi[a] = 0, -1
                solid[tuple(lo)] = True
                solid[tuple(hi)] = True
        self.solid = solid
        unit = torch.eye(len(resolution))[0]
        u = (~solid).to(context.dtype) * (units.characteristic_velocity_pu * unit).reshape(
            [-1] + [1] * len(resolution)).to(context.dtype)
        self.initialize_equilibrium(np.zeros([1] + list(resolution)), u)
        x = torch.meshgrid(*[units.convert_length_to_pu(torch.arange(n)) for n in resolution],
                           indexing="ij")[0]
        self.post_boundaries = [
            EquilibriumBoundaryPU(context, self, torch.abs(x) < 1e-6,
                                  units.characteristic_velocity_pu * unit),
            EquilibriumOutletP([int(c) for c in unit.tolist()], self),
        ]
        self.hwbb = HalfwayBounceBackBoundary(context, self, solid)
        self.post_streaming_boundaries = [self.hwbb]


# --------------------------------------------------------------------------
# observables (observable_reporter.py:34-158, util/utility.py:37-99)
# --------------------------------------------------------------------------
def torch_gradient(f, dx=1, order=6):
    weights = {2: [-1 / 2, 1 / 2, 0, 0, 0, 0],
               4: [1 / 12, -2 / 3, 2 / 3, -1 / 12, 0, 0],
               6: [-1 / 60, 3 / 20, -3 / 4, 3 / 4, -3 / 20, 1 / 60]}[order]
    shifts = {2: [1, -1, 0, 0, 0, 0], 4: [2, 1, -1, -2, 0, 0], 6: [3, 2, 1, -1, -2, -3]}[order]
    dim = f.ndim
    out = torch.cat(dim * [f[None, ...]])
    for i in range(dim):
        acc = 0
        for w, s in zip(weights, shifts):
            sh = [0] * dim
            sh[i] = s
            acc = acc + w * f.roll(shifts=tuple(sh), dims=tuple(range(dim)))
        out[i] = acc * torch.tensor(1.0 / dx, dtype=f.dtype)
    return out


def incompressible_kinetic_energy(flow):
    dx = flow.units.convert_length_to_pu(1.0)
    e = flow.units.convert_incompressible_energy_to_pu(torch.sum(flow.incompressible_energy()))
    return e * dx ** flow.stencil.d


def enstrophy(flow):
    u = flow.units.convert_velocity_to_pu(flow.u())
    dx = flow.units.convert_length_to_pu(1.0)
    g0 = torch_gradient(u[0], dx=dx, order=6)
    g1 = torch_gradient(u[1], dx=dx, order=6)
    w = torch.sum((g0[1] - g1[0]) * (g0[1] - g1[0]))
    if flow.stencil.d == 3:
        g2 = torch_gradient(u[2], dx=dx, order=6)
        w = w + torch.sum((g2[1] - g1[2]) ** 2 + (g0[2] - g2[0]) ** 2)
    return w * dx ** flow.stencil.d


def mass(flow, no_mass_mask=None):
    f = flow.f
    m = f[..., 1:-1, 1:-1].sum()
    if no_mass_mask is not None:
        m -= (f * no_mass_mask.to(dtype=f.dtype)).sum()
    return m


# --------------------------------------------------------------------------
# learned collision (no lettuce counterpart; nearest hook MRTCollision,
# lettuce/ext/_collision/mrt_collision.py:8-35, KBC decomposition of
# kbc_collision.py:76-160)
# --------------------------------------------------------------------------
def bf16_round(t):
    return t.to(torch.bfloat16).to(torch.float32)


LOG2E = 1.4426950408889634
LN2 = 0.6931471805599453


# D2Q9 dual vectors of the three higher-order raw moments of a population
# vector with zero density, momentum and second moments:
# m_xxy -> PHI_XXY, m_xyy -> PHI_XYY, m_xxyy -> PHI_XXYY
PHI_XXY = [0.0, 0.0, -0.5, 0.0, 0.5, 0.25, 0.25, -0.25, -0.25]
PHI_XYY = [0.0, -0.5, 0.0, 0.5, 0.0, 0.25, -0.25, -0.25, 0.25]
PHI_XXYY = [1.0, -0.5, -0.5, -0.5, -0.5, 0.25, 0.25, 0.25, 0.25]


def d2q9_moment_rows(dtype=torch.float64):
    """rows of the raw moments T = xx + yy, N = xx - yy, Pxy, xxy, xyy, xxyy"""
    e = torch.tensor(D2Q9.e, dtype=dtype)
    x, y = e[:, 0], e[:, 1]
    return torch.stack([x * x + y * y, x * x - y * y, x * y, x * x * y, x * y * y,
                        x * x * y * y])


class LearnedParameters:
    """MLP 6 -> 32 -> 32 -> 3 with SiLU; weights stored in bf16."""

    def __init__(self, w1, b1, w2, b2, w3, b3, in_scale=100.0, rate_lo=0.05,
                 rate_hi=1.95, stable_lo=0.0, stable_hi=2.0):
        self.w1 = bf16_round(torch.as_tensor(w1, dtype=torch.float32))
        self.w2 = bf16_round(torch.as_tensor(w2, dtype=torch.float32))
        self.w3 = bf16_round(torch.as_tensor(w3, dtype=torch.float32))
        self.b1 = torch.as_tensor(b1, dtype=torch.float32)
        self.b2 = torch.as_tensor(b2, dtype=torch.float32)
        self.b3 = torch.as_tensor(b3, dtype=torch.float32)
        self.in_scale = float(in_scale)
        self.rate_lo, self.rate_hi = float(rate_lo), float(rate_hi)
        self.stable_lo, self.stable_hi = float(stable_lo), float(stable_hi)

    @staticmethod
    def seeded(seed=0, hidden=32, out_scale=0.1):
        """torch.manual_seed(seed) + nn.Linear default init (U(+-1/sqrt(fan_in)))
        with the last layer scaled by out_scale."""
        g = torch.Generator().manual_seed(seed)

        def lin(o, i):
            bound = 1.0 / np.sqrt(i)
            w = (torch.rand(o, i, generator=g) * 2 - 1) * bound
            b = (torch.rand(o, generator=g) * 2 - 1) * bound
            return w, b
        w1, b1 = lin(hidden, 6)
        w2, b2 = lin(hidden, hidden)
        w3, b3 = lin(3, hidden)
        return LearnedParameters(w1, b1, w2, b2, w3 * out_scale, b3 * out_scale)


def d3q27_raw_moments(dtype=torch.float64):
    """rows x^a y^b z^c (a, b, c in 0..2, index 9a + 3b + c) of the D3Q27
    velocities and the order a + b + c of each row"""
    e = torch.tensor(D3Q27.e, dtype=dtype)
    rows, orders = [], []
    for a in range(3):
        for b in range(3):
            for c in range(3):
                rows.append(e[:, 0] ** a * e[:, 1] ** b * e[:, 2] ** c)
                orders.append(a + b + c)
    return torch.stack(rows), orders


def d3q27_second_moments(df):
    """(xx, yy, zz, xy, xz, yz) raw moments of df [27, ...]"""
    e = torch.tensor(D3Q27.e, dtype=df.dtype)
    x, y, z = e[:, 0], e[:, 1], e[:, 2]
    return [torch.einsum("q,q...->...", w, df) for w in (x * x, y * y, z * z, x * y, x * z, y * z)]


class LearnedCollision:
    """D2Q9 / D3Q27 learned collision.

    D3Q27: inputs T, Nxz, Nyz, Pxy, Pxz, Pyz of df / rho * in_scale; ds the
    KBC shear part from those raw moments (kbc_collision.py:41-74 with
    r = 1); the three rates act on the raw moments x^a y^b z^c of dh of order
    3, of order 4 and of orders 5-6, reconstructed with the dual basis
    (inverse of the 27 x 27 raw-moment matrix):
        f' = f - ds / tau - Phi (rate_of_order(k) * M dh)
    

    D2Q9: with df = f - feq, ds the KBC shear part of df and dh = df - ds
    (which carries only the moments xxy, xyy, xxyy):
        x   = (T, N, Pxy, xxy, xyy, xxyy)(df) / rho * in_scale   (bf16)
        o   = W3 silu(W2 silu(W1 x + b1) + b2) + b3   (bf16 operands, FP32 sums)
        r_g = 1/tau + o_g                             raw rates of xxy, xyy, xxyy
        f'  = f - ds / tau - sum_g clamp(r_g) m_g(dh) phi_g
    A cell whose raw rates are NaN or leave [stable_lo, stable_hi] relaxes
    with KBC instead.  With every rate equal to beta*gamma the update is KBC,
    with every rate 1/tau it is BGK.
    """

    def __init__(self, params: LearnedParameters):
        self.params = params
        self.kbc = KBCCollision()
        self.force_rates = None   # test hook: fixed rates [3] or [3, *res]

    def mlp(self, x, rnd=bf16_round):
        """o = W3 silu(W2 silu(W1 x + b1) + b2) + b3 with bf16 operands and
        sums in x's dtype, evaluated as the inference kernel evaluates it: in
        the log2 domain, where each pre-activation z is carried as
        s = -log2(e) z, SiLU as s / (1 + 2^s) = -log2(e) silu(z) (so the
        weights of layers 2 and 3 take no factor), and the output as -ln(2)
        times its s.  The bf16 roundings (rnd) apply to the scaled inputs
        and activations.  With rnd the identity it is the plain MLP up to
        rounding."""
        p = self.params
        shape = [-1] + [1] * (x.dim() - 1)

        def layer(w, h, b):
            bias = (-LOG2E * b.double()).to(x.dtype).reshape(shape)
            return torch.einsum("oi,i...->o...", w.to(x.dtype), rnd(h)) + bias
        s = layer(p.w1, x * -LOG2E, p.b1)
        s = s / (1 + torch.exp2(s))
        s = layer(p.w2, s, p.b2)
        s = s / (1 + torch.exp2(s))
        return -LN2 * layer(p.w3, s, p.b3)

    def inputs(self, flow, df, rho):
        if isinstance(flow.stencil, D3Q27):
            xx, yy, zz, xy, xz, yz = d3q27_second_moments(df)
            m = torch.stack([xx + yy + zz, xx - zz, yy - zz, xy, xz, yz])
        else:
            rows = d2q9_moment_rows(df.dtype)
            m = torch.einsum("kq,q...->k...", rows, df)
        return (m / rho * self.params.in_scale).to(torch.float32)

    def _call_d3q27(self, flow):
        tau = flow.units.relaxation_parameter_lu
        feq = flow.equilibrium(flow)
        df = flow.f - feq
        xx, yy, zz, xy, xz, yz = d3q27_second_moments(df)
        T, Nxz, Nyz = xx + yy + zz, xx - zz, yy - zz
        ds = torch.zeros_like(df)
        ds[0] = -T
        ds[1] = ds[2] = (2 * Nxz - Nyz + T) / 6
        ds[3] = ds[4] = (2 * Nyz - Nxz + T) / 6
        ds[5] = ds[6] = (-Nxz - Nyz + T) / 6
        ds[7] = ds[8] = 0.25 * yz
        ds[9] = ds[10] = -0.25 * yz
        ds[11] = ds[12] = 0.25 * xz
        ds[13] = ds[14] = -0.25 * xz
        ds[15] = ds[16] = 0.25 * xy
        ds[17] = ds[18] = -0.25 * xy
        dh = df - ds
        M, orders = d3q27_raw_moments(df.dtype)
        phi = torch.linalg.inv(M)
        m = torch.einsum("kq,q...->k...", M, dh)
        omega, fallback = self.rates(flow)
        group = torch.tensor([-1 if o <= 2 else (0 if o == 3 else (1 if o == 4 else 2))
                              for o in orders])
        w = torch.zeros_like(m)
        for k in range(27):
            if group[k] >= 0:
                w[k] = omega[group[k]] * m[k]
        f_learned = flow.f - ds / tau - torch.einsum("qk,k...->q...", phi, w)
        if not fallback.any():
            return f_learned
        return torch.where(fallback[None], self.kbc(flow), f_learned)

    def rates(self, flow):
        """(clamped rates [3, *res], fallback mask [*res])"""
        tau = flow.units.relaxation_parameter_lu
        feq = flow.equilibrium(flow)
        df = flow.f - feq
        x = self.inputs(flow, df, flow.rho())
        raw = (1.0 / tau + self.mlp(x)).to(flow.f.dtype)
        if self.force_rates is not None:
            raw = torch.as_tensor(self.force_rates, dtype=flow.f.dtype)
            raw = raw.reshape([3] + [1] * flow.stencil.d).expand([3] + list(df.shape[1:]))
        p = self.params
        fallback = (torch.isnan(raw) | (raw < p.stable_lo) | (raw > p.stable_hi)).any(0)
        return raw.clamp(p.rate_lo, p.rate_hi), fallback

    def __call__(self, flow):
        if isinstance(flow.stencil, D3Q27):
            return self._call_d3q27(flow)
        assert isinstance(flow.stencil, D2Q9)
        tau = flow.units.relaxation_parameter_lu
        feq = flow.equilibrium(flow)
        df = flow.f - feq
        dtype = flow.f.dtype
        e = torch.tensor(D2Q9.e, dtype=dtype)
        xx = torch.einsum("q,q...->...", e[:, 0] ** 2, df)
        yy = torch.einsum("q,q...->...", e[:, 1] ** 2, df)
        xy = torch.einsum("q,q...->...", e[:, 0] * e[:, 1], df)
        T, N = xx + yy, xx - yy
        ds = torch.zeros_like(df)
        ds[0] = -T
        ds[1] = ds[3] = 0.25 * (T + N)
        ds[2] = ds[4] = 0.25 * (T - N)
        ds[5] = ds[7] = 0.25 * xy
        ds[6] = ds[8] = -0.25 * xy
        dh = df - ds
        rows = d2q9_moment_rows(dtype)[3:]
        mg = torch.einsum("kq,q...->k...", rows, dh)
        phi = torch.tensor([PHI_XXY, PHI_XYY, PHI_XXYY], dtype=dtype)
        omega, fallback = self.rates(flow)
        upd = torch.einsum("kq,k...->q...", phi, omega * mg)
        f_learned = flow.f - ds / tau - upd
        if not fallback.any():
            return f_learned
        return torch.where(fallback[None], self.kbc(flow), f_learned)


# --------------------------------------------------------------------------
# training step of the learned collision: the same formula with the MLP in
# the populations' dtype (no bf16 rounding) and no KBC fallback, written as
# differentiable torch so that torch.autograd gives the reference gradients
# --------------------------------------------------------------------------
LEARNED_HIDDEN = 32
LEARNED_PARAMS = 1379


def split_learned_params(theta, hidden=LEARNED_HIDDEN):
    """[w1 (h x 6), b1, w2 (h x h), b2, w3 (3 x h), b3] -> the six tensors"""
    sizes = [hidden * 6, hidden, hidden * hidden, hidden, 3 * hidden, 3]
    w1, b1, w2, b2, w3, b3 = torch.split(theta, sizes)
    return (w1.reshape(hidden, 6), b1, w2.reshape(hidden, hidden), b2,
            w3.reshape(3, hidden), b3)


def learned_train_collide(f, theta, tau, in_scale=100.0, rate_lo=0.05, rate_hi=1.95):
    """Learned collision of populations f [9, ...] (LearnedCollision's
    formula) with the MLP evaluated in f's dtype and without the fallback;
    clamp() passes the gradient inside [rate_lo, rate_hi] only."""
    dtype = f.dtype
    e = torch.tensor(D2Q9.e, dtype=dtype)
    w = torch.tensor(D2Q9.w, dtype=dtype)
    rho = f.sum(0)
    u = torch.einsum("qd,q...->d...", e, f) / rho
    eu = torch.einsum("qd,d...->q...", e, u)
    uu = (u * u).sum(0)
    feq = w.reshape([-1] + [1] * u[0].dim()) * rho * (
        1 + 3 * eu - 1.5 * uu + 4.5 * eu * eu)
    df = f - feq
    rows = d2q9_moment_rows(dtype)
    m = torch.einsum("kq,q...->k...", rows, df)
    T, N, xy = m[0], m[1], m[2]
    ds = torch.stack([-T, 0.25 * (T + N), 0.25 * (T - N), 0.25 * (T + N),
                      0.25 * (T - N), 0.25 * xy, -0.25 * xy, 0.25 * xy, -0.25 * xy])
    dh = df - ds
    mg = torch.einsum("kq,q...->k...", rows[3:], dh)
    w1, b1, w2, b2, w3, b3 = split_learned_params(theta)
    x = torch.movedim(m / rho * in_scale, 0, -1)          # [..., 6]
    h = torch.nn.functional.silu(torch.nn.functional.linear(x, w1, b1))
    h = torch.nn.functional.silu(torch.nn.functional.linear(h, w2, b2))
    o = torch.movedim(torch.nn.functional.linear(h, w3, b3), -1, 0)
    rates = (1.0 / tau + o).clamp(rate_lo, rate_hi)
    phi = torch.tensor([PHI_XXY, PHI_XYY, PHI_XXYY], dtype=dtype)
    return f - ds / tau - torch.einsum("kq,k...->q...", phi, rates * mg)


def learned_train_step(f, theta, tau, **kw):
    """periodic pull streaming over the last two axes of f [9, ..., X, Y]
    (lettuce/_simulation.py:241-256 with batch axes in front), then
    learned_train_collide"""
    g = torch.stack([torch.roll(f[q], shifts=tuple(D2Q9.e[q]), dims=(-2, -1))
                     for q in range(9)])
    return learned_train_collide(g, theta, tau, **kw)


def learned_train_collide_d3q27(f, theta, tau, in_scale=100.0, rate_lo=0.05, rate_hi=1.95):
    """D3Q27 LearnedCollision's formula (KBC's shear part ds of df = f - feq,
    dh = df - ds, the rates acting on the raw moments of dh of order 3, 4 and
    5-6 through the dual basis phi = M^-1) with the MLP evaluated in f's
    dtype and without the fallback; clamp() passes the gradient inside
    [rate_lo, rate_hi] only."""
    dtype = f.dtype
    e = torch.tensor(D3Q27.e, dtype=dtype)
    w = torch.tensor(D3Q27.w, dtype=dtype)
    rho = f.sum(0)
    u = torch.einsum("qd,q...->d...", e, f) / rho
    eu = torch.einsum("qd,d...->q...", e, u)
    uu = (u * u).sum(0)
    feq = w.reshape([-1] + [1] * u[0].dim()) * rho * (1 + 3 * eu - 1.5 * uu + 4.5 * eu * eu)
    df = f - feq
    xx, yy, zz, xy, xz, yz = d3q27_second_moments(df)
    T, Nxz, Nyz = xx + yy + zz, xx - zz, yy - zz
    a, b, c = (2 * Nxz - Nyz + T) / 6, (2 * Nyz - Nxz + T) / 6, (-Nxz - Nyz + T) / 6
    zero = torch.zeros_like(T)
    # kbc_collision.py:41-74 with r = 1: the shear part fills s[0..18]
    ds = torch.stack([-T, a, a, b, b, c, c, 0.25 * yz, 0.25 * yz, -0.25 * yz, -0.25 * yz,
                      0.25 * xz, 0.25 * xz, -0.25 * xz, -0.25 * xz, 0.25 * xy, 0.25 * xy,
                      -0.25 * xy, -0.25 * xy] + [zero] * 8)
    dh = df - ds
    M, orders = d3q27_raw_moments(dtype)
    phi = torch.linalg.inv(M)
    m = torch.einsum("kq,q...->k...", M, dh)
    w1, b1, w2, b2, w3, b3 = split_learned_params(theta)
    x = torch.movedim(torch.stack([T, Nxz, Nyz, xy, xz, yz]) / rho * in_scale, 0, -1)
    h = torch.nn.functional.silu(torch.nn.functional.linear(x, w1, b1))
    h = torch.nn.functional.silu(torch.nn.functional.linear(h, w2, b2))
    o = torch.movedim(torch.nn.functional.linear(h, w3, b3), -1, 0)
    rates = (1.0 / tau + o).clamp(rate_lo, rate_hi)
    group = [-1 if k <= 2 else (0 if k == 3 else (1 if k == 4 else 2)) for k in orders]
    upd = torch.stack([rates[g] * m[k] if g >= 0 else torch.zeros_like(m[k])
                       for k, g in enumerate(group)])
    return f - ds / tau - torch.einsum("qk,k...->q...", phi, upd)


def learned_train_step_d3q27(f, theta, tau, **kw):
    """periodic pull streaming of f [27, X, Y, Z] (lettuce/_simulation.py:241-256),
    then learned_train_collide_d3q27"""
    g = torch.stack([torch.roll(f[q], shifts=tuple(D3Q27.e[q]), dims=(0, 1, 2))
                     for q in range(27)])
    return learned_train_collide_d3q27(g, theta, tau, **kw)

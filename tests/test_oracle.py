"""Pins the CPU oracle (oracle/lettuce_torch.py) before it is trusted:
against lettuce's own trajectories (tests/golden, written by
tests/golden/make_golden.py from lettuce itself) and against what the
mathematics fixes (conservation, fixpoints, reductions between operators,
known-answer streaming, analytic decay rates, brute-force link rules)."""
import numpy as np
import pytest
import torch

from oracle import lettuce_torch as O
from tests.conftest import golden

F64 = torch.float64
F32 = torch.float32


class _Grid(O.Flow):
    """Bare flow with lettuce's TestFlow units (Re=100, Ma=0.1, L=res[0])."""

    def __init__(self, stencil, res, dtype=F64, f=None):
        units = O.UnitConversion(100.0, 0.1, characteristic_length_lu=res[0])
        super().__init__(O.Context("cpu", dtype), res, units, stencil)
        if f is not None:
            self.f = torch.as_tensor(f).to(dtype).clone()


def _perturbed(stencil, n, dtype=F64, seed=0, amp=0.05):
    flow = _Grid(stencil, [n] * stencil.d, dtype)
    d = stencil.d
    g = torch.Generator().manual_seed(seed)
    u = 0.05 * (torch.rand([d] + flow.resolution, generator=g, dtype=F64) - 0.5)
    rho = 1 + 0.02 * (torch.rand([1] + flow.resolution, generator=g, dtype=F64) - 0.5)
    feq = flow.equilibrium(flow, rho=rho.to(dtype), u=u.to(dtype))
    r = torch.rand(feq.shape, generator=g, dtype=F64) * 2 - 1
    flow.f = feq * (1 + amp * r.to(dtype))
    return flow


STENCILS = [O.D2Q9(), O.D3Q19(), O.D3Q27()]
IDS = ["D2Q9", "D3Q19", "D3Q27"]


# ---------------------------------------------------------------- stencils
@pytest.mark.parametrize("s", STENCILS, ids=IDS)
def test_stencil_invariants(s):
    e = np.array(s.e, dtype=float)
    w = np.array(s.w)
    assert np.allclose(w.sum(), 1.0)
    assert np.allclose(w @ e, 0.0)
    assert np.allclose(np.einsum("q,qa,qb->ab", w, e, e), np.eye(s.d) / 3.0)
    assert np.allclose(e[s.opposite], -e)
    assert not e[0].any()
    assert len({tuple(r) for r in e}) == s.q


# ------------------------------------------------------------- equilibrium
@pytest.mark.parametrize("s", STENCILS, ids=IDS)
def test_equilibrium_conserves_moments_and_rest_state(s):
    flow = _perturbed(s, 4)
    rho, u = flow.rho(), flow.u()
    feq = flow.equilibrium(flow, rho=rho, u=u)
    assert torch.allclose(flow.rho(feq), rho, rtol=0, atol=1e-13)
    assert torch.allclose(flow.j(feq), rho * u, rtol=0, atol=1e-13)
    rest = flow.equilibrium(flow, rho=torch.ones_like(rho), u=torch.zeros_like(u))
    assert torch.allclose(rest, torch.tensor(s.w, dtype=F64).reshape([-1] + [1] * s.d)
                          .expand_as(rest))


@pytest.mark.parametrize("s", STENCILS, ids=IDS)
def test_equilibrium_closed_form(s):
    """feq_q = w_q rho (1 + 3 e.u + 9/2 (e.u)^2 - 3/2 u.u), cs^2 = 1/3"""
    flow = _perturbed(s, 3)
    rho, u = flow.rho(), flow.u()
    feq = flow.equilibrium(flow, rho=rho, u=u).numpy()
    e = np.array(s.e, dtype=float)
    un = u.numpy()
    eu = np.einsum("qd,d...->q...", e, un)
    uu = (un ** 2).sum(0)
    w = np.array(s.w).reshape([-1] + [1] * s.d)
    ref = w * rho.numpy() * (1 + 3 * eu + 4.5 * eu ** 2 - 1.5 * uu)
    assert np.allclose(feq, ref, rtol=0, atol=1e-14)


# ---------------------------------------------------------------- streaming
@pytest.mark.parametrize("s", STENCILS, ids=IDS)
def test_stream_one_hot(s):
    """NoCollision + streaming moves population q by e_q (periodic)."""
    n = 5
    flow = _Grid(s, [n] * s.d)
    flow.f = torch.zeros([s.q] + [n] * s.d, dtype=F64)
    x0 = [n - 1] + [1] * (s.d - 1)
    for q in range(s.q):
        flow.f[(q,) + tuple(x0)] = q + 1.0
    sim = O.Simulation(flow, O.NoCollision(), [], O.StreamingStrategy.PRE_STREAMING)
    sim(1)
    for q in range(s.q):
        dst = tuple((x0[a] + s.e[q][a]) % n for a in range(s.d))
        assert flow.f[(q,) + dst] == q + 1.0
        assert float(flow.f[q].sum()) == q + 1.0


# --------------------------------------------------------------- collisions
@pytest.mark.parametrize("s", STENCILS, ids=IDS)
def test_bgk_fixpoint_and_tau_one(s):
    flow = _perturbed(s, 4)
    f0 = flow.f.clone()
    c = O.BGKCollision(0.5)
    flow.f = c(flow)
    flow.f = c(flow)
    assert torch.allclose(flow.f, f0, rtol=0, atol=1e-13)
    feq = flow.equilibrium(flow)
    assert torch.allclose(O.BGKCollision(1.0)(flow), feq, rtol=0, atol=1e-15)


@pytest.mark.parametrize("s", STENCILS, ids=IDS)
@pytest.mark.parametrize("name", ["bgk", "trt", "kbc"])
def test_collisions_conserve_and_relax_shear(s, name):
    if name == "kbc" and s.q == 19:
        pytest.skip("KBC is defined for D2Q9 and D3Q27")
    flow = _perturbed(s, 4)
    tau = flow.units.relaxation_parameter_lu
    coll = {"bgk": O.BGKCollision(tau), "trt": O.TRTCollision(tau, 0.9),
            "kbc": O.KBCCollision()}[name]
    feq = flow.equilibrium(flow)
    sh, sh_eq = flow.shear_tensor(), flow.shear_tensor(feq)
    post = coll(flow)
    assert torch.allclose(flow.rho(post), flow.rho(), rtol=0, atol=1e-13)
    assert torch.allclose(flow.j(post), flow.j(), rtol=0, atol=1e-13)
    if name != "trt":  # TRT's odd part carries no shear; even part is tau
        expect = sh - (sh - sh_eq) / tau
        assert torch.allclose(flow.shear_tensor(post), expect, rtol=0, atol=1e-13)


@pytest.mark.parametrize("s", STENCILS, ids=IDS)
def test_trt_with_equal_times_is_bgk(s):
    flow = _perturbed(s, 4)
    a = O.TRTCollision(0.73, 0.73)(flow)
    b = O.BGKCollision(0.73)(flow)
    assert torch.allclose(a, b, rtol=0, atol=1e-14)


@pytest.mark.parametrize("s", [O.D2Q9(), O.D3Q27()], ids=["D2Q9", "D3Q27"])
def test_kbc_gamma_two_is_bgk(s):
    flow = _perturbed(s, 4)
    kbc = O.KBCCollision()
    kbc.force_gamma = 2.0
    tau = flow.units.relaxation_parameter_lu
    assert torch.allclose(kbc(flow), O.BGKCollision(tau)(flow), rtol=0, atol=1e-14)


@pytest.mark.parametrize("s", [O.D2Q9(), O.D3Q27()], ids=["D2Q9", "D3Q27"])
def test_kbc_pseudo_entropy_exceeds_bgk(s):
    flow = _Grid(s, [3] * s.d)
    np.random.seed(1)
    flow.f = torch.from_numpy(np.random.random([s.q] + [3] * s.d))
    tau = 0.5003
    kbc = O.KBCCollision(tau)
    kbc._setup(flow)
    kbc.tau, kbc.beta = tau, 1 / (2 * tau)
    f_kbc = kbc(flow)
    f_bgk = O.BGKCollision(tau)(flow)
    assert (flow.pseudo_entropy_local(f_bgk) < flow.pseudo_entropy_local(f_kbc)).all()


def test_bounce_back_is_opposite():
    s = O.D2Q9()
    flow = _perturbed(s, 4)
    out = O.BounceBackBoundary(None)(flow)
    assert torch.equal(out, flow.f[s.opposite])


# ------------------------------------------------------------ golden vectors
@pytest.mark.parametrize("tag,dtype,tol", [("f32", F32, 2e-6), ("f64", F64, 1e-13)])
def test_golden_spike_bgk_all_strategies(tag, dtype, tol):
    g = golden("g1_spike_bgk_d2q9.npz")
    for strat in ("PRE_STREAMING", "POST_STREAMING", "NO_STREAMING", "DOUBLE_STREAMING"):
        for steps in (1, 10):
            flow = _Grid(O.D2Q9(), [16, 16], dtype)
            flow.f = torch.ones([9, 16, 16], dtype=dtype)
            flow.f[:, 2, 2] = 2.0
            O.Simulation(flow, O.BGKCollision(2.0), [],
                         getattr(O.StreamingStrategy, strat))(steps)
            ref = g[f"{tag}_{strat}_{steps}"]
            assert np.allclose(flow.f.numpy(), ref, rtol=0, atol=tol), (strat, steps)


def test_golden_tgv3d_bgk():
    g = golden("g2_tgv3d_bgk.npz")
    for s, n in ((O.D3Q19(), 16), (O.D3Q27(), 8)):
        name = type(s).__name__
        flow = O.TaylorGreenVortex(O.Context("cpu", F32), n, 1600, 0.1, s)
        assert np.allclose(flow.f.numpy(), g[f"{name}_f0"], rtol=0, atol=1e-7)
        tau = flow.units.relaxation_parameter_lu
        assert tau == float(g[f"{name}_tau"])
        O.Simulation(flow, O.BGKCollision(tau), [], O.StreamingStrategy.PRE_STREAMING)(10)
        assert np.allclose(flow.f.numpy(), g[f"{name}_f10"], rtol=0, atol=1e-6)


def test_golden_tgv2d_1000_steps_and_observables():
    g = golden("g3_tgv2d_128_bgk_1000.npz")
    flow = O.TaylorGreenVortex(O.Context("cpu", F32), 128, 100, 0.05 * np.sqrt(3), O.D2Q9())
    assert np.array_equal(flow.f.numpy(), g["f0"])
    tau = flow.units.relaxation_parameter_lu
    ke0 = float(O.incompressible_kinetic_energy(flow))
    en0 = float(O.enstrophy(flow))
    assert np.isclose(ke0, g["ke"][0][2], rtol=1e-6)
    assert np.isclose(en0, g["enstrophy"][0][2], rtol=1e-6)
    O.Simulation(flow, O.BGKCollision(tau), [], O.StreamingStrategy.PRE_STREAMING)(1000)
    err = np.abs(flow.f.numpy() - g["f1000"]).max() / np.abs(g["f1000"]).max()
    assert err < 1e-6
    assert np.isclose(float(O.incompressible_kinetic_energy(flow)), g["ke"][-1][2], rtol=1e-5)


def test_tgv2d_decays_as_analytic_solution():
    """2D TGV kinetic energy decays as exp(-4 nu t) in physical units
    (lettuce/ext/_flows/taylorgreen.py:72-81: u ~ exp(-2 nu t))."""
    g = golden("g3_tgv2d_128_bgk_1000.npz")
    ke = g["ke"]
    units = O.TaylorGreenVortex(O.Context("cpu", F32), 8, 100, 0.05 * np.sqrt(3),
                                O.D2Q9()).units
    nu = units.viscosity_pu
    t = ke[-1][1]
    assert np.isclose(ke[-1][2] / ke[0][2], np.exp(-4 * nu * t), rtol=2e-3)


@pytest.mark.parametrize("tag,dtype,tol", [("f32", F32, 3e-6), ("f64", F64, 1e-12)])
def test_golden_kbc(tag, dtype, tol):
    g = golden("g4_kbc.npz")
    for s, n in ((O.D2Q9(), 16), (O.D3Q27(), 6)):
        name = type(s).__name__
        flow = _Grid(s, [n] * s.d, dtype, g[f"{name}_{tag}_f0"])
        assert flow.units.relaxation_parameter_lu == float(g[f"{name}_tau"])
        assert np.allclose(O.KBCCollision()(flow).numpy(), g[f"{name}_{tag}_coll"],
                           rtol=0, atol=tol)
        O.Simulation(flow, O.KBCCollision(), [], O.StreamingStrategy.PRE_STREAMING)(10)
        assert np.allclose(flow.f.numpy(), g[f"{name}_{tag}_f10"], rtol=0, atol=tol)


def test_golden_trt():
    g = golden("g5_trt.npz")
    for s, n in ((O.D2Q9(), 16), (O.D3Q19(), 6), (O.D3Q27(), 6)):
        name = type(s).__name__
        flow = _Grid(s, [n] * s.d, F64, g[f"{name}_f0"])
        O.Simulation(flow, O.TRTCollision(0.8, 1.3), [], O.StreamingStrategy.PRE_STREAMING)(10)
        assert np.allclose(flow.f.numpy(), g[f"{name}_f10"], rtol=0, atol=1e-13)


class _Obstacle(O.Flow):
    """the golden obstacle case: lettuce's Obstacle (48x24, Re 50, Ma 0.1,
    domain length 4) with EquilibriumOutletP and fullway bounce-back"""

    def __init__(self, g):
        units = O.UnitConversion(50, 0.1, characteristic_length_lu=12.0,
                                 characteristic_length_pu=1, characteristic_velocity_pu=1)
        super().__init__(O.Context("cpu", F64), [48, 24], units, O.D2Q9())
        self.f = torch.from_numpy(g["f0"]).clone()
        x = torch.meshgrid(*[units.convert_length_to_pu(torch.arange(n)) for n in (48, 24)],
                           indexing="ij")[0]
        self.post_boundaries = [
            O.EquilibriumBoundaryPU(self.context, self, torch.abs(x) < 1e-6, [1.0, 0.0]),
            O.EquilibriumOutletP([1, 0], self),
            O.BounceBackBoundary(torch.from_numpy(g["mask"])),
        ]


def test_golden_obstacle_masks_and_steps():
    g = golden("g6_obstacle_eqoutlet.npz")
    flow = _Obstacle(g)
    sim = O.Simulation(flow, O.BGKCollision(float(g["tau"])), [])
    assert np.array_equal(sim.no_collision_mask.numpy(), g["ncm"])
    assert np.array_equal(sim.no_streaming_mask.numpy(), g["nsm"])
    sim(10)
    assert np.allclose(flow.f.numpy(), g["f10_post"], rtol=0, atol=1e-13)


def test_golden_decaying_turbulence_field():
    g = golden("g7_decaying_turbulence.npz")
    flow = O.DecayingTurbulence(O.Context("cpu", F64), [32, 32], 1000, 0.1, O.D2Q9(),
                                k0=4, randseed=0)
    assert np.allclose(flow.u().numpy(), g["u0_lu"], rtol=0, atol=1e-14)
    assert np.allclose(flow.f.numpy(), g["f0"], rtol=0, atol=1e-14)
    flow3 = O.DecayingTurbulence(O.Context("cpu", F64), [8, 8, 8], 1000, 0.1, O.D3Q27(),
                                 k0=4, randseed=0)
    assert np.allclose(flow3.u().numpy(), g["u0_lu_3d"], rtol=0, atol=1e-14)


class _DefaultObstacle(O.Flow):
    """lettuce's default Obstacle (lettuce/ext/_flows/obstacle.py:107-122):
    equilibrium inlet at x = 0, anti-bounce-back outlet at x = -1, fullway
    bounce-back on the mask; Re 50, Ma 0.1, characteristic length
    res[0] / domain_length_x lattice units"""

    def __init__(self, stencil, res, length_x, f0, mask, dtype=F64):
        units = O.UnitConversion(50, 0.1, characteristic_length_lu=res[0] / length_x,
                                 characteristic_length_pu=1, characteristic_velocity_pu=1)
        super().__init__(O.Context("cpu", dtype), res, units, stencil)
        self.f = torch.as_tensor(f0).to(dtype).clone()
        x = torch.meshgrid(*[units.convert_length_to_pu(torch.arange(n)) for n in res],
                           indexing="ij")[0]
        unit = torch.eye(len(res))[0]
        self.post_boundaries = [
            O.EquilibriumBoundaryPU(self.context, self, torch.abs(x) < 1e-6, unit),
            O.AntiBounceBackOutlet([int(c) for c in unit.tolist()], self),
            O.BounceBackBoundary(torch.as_tensor(mask)),
        ]


G8_CASES = {"D2Q9": (O.D2Q9, [48, 24], 4.0, "bgk"), "D3Q27": (O.D3Q27, [16, 8, 8], 2.0, "kbc")}


@pytest.mark.parametrize("name", list(G8_CASES))
@pytest.mark.parametrize("tag", ["post", "pre"])
def test_golden_default_obstacle_abb(name, tag):
    """the oracle's Obstacle with the anti-bounce-back outlet against
    lettuce's own (golden g8): masks bit for bit, 10 steps in FP64"""
    g = golden("g8_obstacle_abb.npz")
    stencil, res, length_x, coll = G8_CASES[name]
    flow = _DefaultObstacle(stencil(), res, length_x, g[f"{name}_f0"], g[f"{name}_mask"])
    tau = float(g[f"{name}_tau"])
    assert tau == flow.units.relaxation_parameter_lu
    collision = O.BGKCollision(tau) if coll == "bgk" else O.KBCCollision()
    strategy = O.StreamingStrategy.POST_STREAMING if tag == "post" else \
        O.StreamingStrategy.PRE_STREAMING
    sim = O.Simulation(flow, collision, [], strategy)
    assert np.array_equal(sim.no_collision_mask.numpy(), g[f"{name}_ncm"])
    assert np.array_equal(sim.no_streaming_mask.numpy(), g[f"{name}_nsm"])
    sim(10)
    ref = g[f"{name}_f10_{tag}"]
    assert np.abs(flow.f.numpy() - ref).max() <= 1e-13 * np.abs(ref).max()


class _Poiseuille(O.Flow):
    """lettuce's PoiseuilleFlow2D (lettuce/ext/_flows/poiseuille.py:18-86):
    units with characteristic length res - 1, bounce-back on the first and
    last row of y"""

    def __init__(self, g, n=17):
        units = O.UnitConversion(1, 0.02, characteristic_length_lu=n - 1,
                                 characteristic_length_pu=1, characteristic_velocity_pu=1)
        super().__init__(O.Context("cpu", F64), [n, n], units, O.D2Q9())
        self.f = torch.from_numpy(g["f0"]).clone()
        mask = torch.zeros(n, n, dtype=torch.bool)
        mask[:, [0, -1]] = True
        self.post_boundaries = [O.BounceBackBoundary(mask)]


def test_golden_poiseuille_guo():
    """lettuce's force test case (golden g9): BGK with the Guo force and
    bounce-back walls, FP64, post-streaming, 10 and 100 steps, and the
    velocity with the a / (2 rho) shift"""
    g = golden("g9_poiseuille_guo.npz")
    flow = _Poiseuille(g)
    tau = float(g["tau"])
    assert tau == flow.units.relaxation_parameter_lu
    a = torch.from_numpy(g["acceleration_lu"])
    sim = O.Simulation(flow, O.BGKCollision(tau, force=O.Guo(flow, tau, a)), [])
    assert np.array_equal(sim.no_collision_mask.numpy(), g["ncm"])
    assert np.array_equal(sim.no_streaming_mask.numpy(), g["nsm"])
    sim(10)
    assert np.abs(flow.f.numpy() - g["f10"]).max() <= 1e-14
    sim(90)
    assert np.abs(flow.f.numpy() - g["f100"]).max() <= 1e-14
    rho = flow.rho()
    u = flow.j() / rho + 0.5 * a.reshape(2, 1, 1) / rho
    assert np.abs(u.numpy() - g["u100"]).max() <= 1e-15


# ------------------------------------------------------ halfway bounce-back
def _channel(n, ny, dtype=F64):
    s = O.D2Q9()
    flow = _Grid(s, [n, ny], dtype)
    solid = torch.zeros(n, ny, dtype=torch.bool)
    solid[:, 0] = True
    solid[:, -1] = True
    return flow, solid


def test_halfway_bounce_back_link_rule_brute_force():
    """After one stream, a fluid cell x whose neighbour x - e_q is solid
    holds its own pre-stream f_opp(q); others hold f_q(x - e_q); solid cells
    are unchanged."""
    s = O.D2Q9()
    g = torch.Generator().manual_seed(5)
    n = 6
    solid = torch.zeros(n, n, dtype=torch.bool)
    solid[2, 2] = solid[2, 3] = solid[0, 5] = True
    flow = _Grid(s, [n, n])
    f0 = torch.rand(9, n, n, generator=g, dtype=F64)
    flow.f = f0.clone()
    flow.post_streaming_boundaries = [O.HalfwayBounceBackBoundary(flow.context, flow, solid)]
    O.Simulation(flow, O.NoCollision(), [], O.StreamingStrategy.PRE_STREAMING)(1)
    for x in range(n):
        for y in range(n):
            for q in range(9):
                if solid[x, y]:
                    expect = f0[q, x, y]
                else:
                    sx, sy = (x - s.e[q][0]) % n, (y - s.e[q][1]) % n
                    expect = f0[s.opposite[q], x, y] if solid[sx, sy] else f0[q, sx, sy]
                assert flow.f[q, x, y] == expect, (q, x, y)


def test_halfway_bounce_back_wall_position_and_mass():
    """A shear wave u_x = U sin(pi (y - 1/2) / H) between halfway walls
    (fluid rows 1..H) decays at nu (pi/H)^2: the wall sits half a cell
    outside the last fluid node.  Mass on fluid cells is conserved exactly."""
    nx, H = 4, 24
    flow, solid = _channel(nx, H + 2)
    tau = 0.8
    nu = (tau - 0.5) / 3
    y = torch.arange(H + 2, dtype=F64)
    U = 1e-3
    ux = U * torch.sin(torch.pi * (y - 0.5) / H)
    u = torch.zeros(2, nx, H + 2, dtype=F64)
    u[0] = ux[None, :]
    u[:, solid] = 0
    flow.f = flow.equilibrium(flow, rho=torch.ones(1, nx, H + 2, dtype=F64), u=u)
    flow.post_streaming_boundaries = [O.HalfwayBounceBackBoundary(flow.context, flow, solid)]
    m0 = flow.f[:, ~solid].sum()
    steps = 400
    O.Simulation(flow, O.BGKCollision(tau), [], O.StreamingStrategy.PRE_STREAMING)(steps)
    assert abs(float(flow.f[:, ~solid].sum() - m0)) < 1e-11
    amp = float(flow.u()[0, 0, 1:-1].mul(torch.sin(torch.pi * (y[1:-1] - 0.5) / H)).sum()
                / torch.sin(torch.pi * (y[1:-1] - 0.5) / H).pow(2).sum())
    assert np.isclose(amp / U, np.exp(-nu * (np.pi / H) ** 2 * steps), rtol=3e-3)


# ------------------------------------------------- anti-bounce-back outlet
def test_anti_bounce_back_keeps_a_uniform_equilibrium():
    """feq_v + feq_opp(v) = w rho (2 + 9 (e.u)^2 - 3 |u|^2), so a uniform
    equilibrium column is a fixed point of the outlet"""
    for s, res in ((O.D2Q9(), [6, 5]), (O.D3Q27(), [5, 4, 3])):
        flow = _Grid(s, res)
        d = s.d
        u = torch.zeros([d] + res, dtype=F64)
        u[0] = 0.04
        u[1] = -0.01
        rho = torch.full([1] + res, 1.02, dtype=F64)
        flow.f = flow.equilibrium(flow, rho=rho, u=u)
        f0 = flow.f.clone()
        abb = O.AntiBounceBackOutlet([1] + [0] * (d - 1), flow)
        out = abb(flow)
        torch.testing.assert_close(out, f0, rtol=0, atol=1e-15)


def test_anti_bounce_back_extrapolates_the_wall_velocity():
    """a linear velocity profile along x: u_w = u(N-1) + (u(N-1) - u(N-2)) / 2
    (hand-computed), and only the incoming populations of the outlet column
    change"""
    s = O.D2Q9()
    flow = _Grid(s, [4, 3])
    ux = torch.tensor([0.01, 0.02, 0.03, 0.04], dtype=F64)[:, None].expand(4, 3)
    u = torch.stack([ux, torch.zeros(4, 3, dtype=F64)])
    rho = torch.ones(1, 4, 3, dtype=F64)
    flow.f = flow.equilibrium(flow, rho=rho, u=u)
    f0 = flow.f.clone()
    out = O.AntiBounceBackOutlet([1, 0], flow)(flow).clone()
    uw = 0.04 + 0.5 * (0.04 - 0.03)
    for v in (1, 5, 8):                      # e_x = +1
        o = s.opposite[v]
        ex, ey = s.e[v]
        expect = -f0[v, -1] + s.w[v] * 1.0 * (2 + 9 * (ex * uw) ** 2 - 3 * uw ** 2)
        torch.testing.assert_close(out[o, -1], expect, rtol=0, atol=1e-15)
    changed = (out != f0).any(dim=(1, 2))
    assert sorted(torch.nonzero(changed).flatten().tolist()) == [3, 6, 7]
    assert torch.equal(out[:, :-1], f0[:, :-1])


# ------------------------------------------------------------- Guo forcing
@pytest.mark.parametrize("s", [O.D2Q9(), O.D3Q19(), O.D3Q27()], ids=["D2Q9", "D3Q19", "D3Q27"])
def test_guo_forcing_adds_exactly_the_acceleration(s):
    """With u* = u + a/(2 rho): sum_q si = 0 and sum_q e si = (1 - 1/(2 tau)) a,
    so BGK + Guo keeps the mass and raises the momentum of every cell by a
    (j' = j - (j - rho u*)/tau + (1 - 1/(2 tau)) a = j + a)."""
    flow = _perturbed(s, 4)
    tau = 0.62
    a = [0.002, -0.001, 0.0005][:s.d]
    coll = O.BGKCollision(tau, force=O.Guo(flow, tau, a))
    post = coll(flow)
    torch.testing.assert_close(flow.rho(post), flow.rho(), rtol=0, atol=1e-14)
    expect = flow.j() + torch.tensor(a, dtype=F64).reshape([-1] + [1] * s.d)
    torch.testing.assert_close(flow.j(post), expect, rtol=0, atol=1e-14)
    # a = 0 is plain BGK
    zero = O.BGKCollision(tau, force=O.Guo(flow, tau, [0.0] * s.d))(flow)
    torch.testing.assert_close(zero, O.BGKCollision(tau)(flow), rtol=0, atol=1e-15)

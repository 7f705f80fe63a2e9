"""Host-side checks that need no GPU: the C-ABI library loads and exports
everything include/lettuce_amd.h declares, the ctypes mirrors match the C
structs, and the boundary masks / cell classes / halfway links are
bit-exact against lettuce's masks (golden) and the oracle."""
import ctypes
import os
import re
from types import SimpleNamespace

import numpy as np
import pytest
import torch

import lettuce_amd as lt
from lettuce_amd import _native as N
from lettuce_amd._classes import (OpSpec, build_cell_classes, build_masks,
                                  expand_classes, halfway_links)
from oracle import lettuce_torch as O
from tests.conftest import ROOT, golden

HEADER = os.path.join(ROOT, "include", "lettuce_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(lbm_\w+)\s*\(",
                                 text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 12
    lib = N.lib()
    for name in names:
        assert hasattr(lib, name), name
    assert lib.lbm_abi_version() == N.ABI_VERSION


def test_struct_mirrors_match_library():
    lib = N.lib()
    assert lib.lbm_struct_size(0) == ctypes.sizeof(N.LbmStepArgs)
    assert lib.lbm_struct_size(1) == ctypes.sizeof(N.LbmCellClass) == 48
    assert lib.lbm_struct_size(2) == ctypes.sizeof(N.LbmLearned)


def test_step_rejects_bad_arguments_without_a_gpu():
    """Argument validation happens before any HIP call."""
    lib = N.lib()
    a = N.LbmStepArgs()
    a.abi_version = 99
    assert lib.lbm_step(ctypes.byref(a), None) == -2
    a.abi_version = N.ABI_VERSION
    a.stencil = 7
    assert lib.lbm_step(ctypes.byref(a), None) == -1
    assert b"stencil" in lib.lbm_last_error()
    a.stencil = N.D3Q19
    a.n[:] = (4, 4, 4)
    a.stride[:] = (64, 16, 4, 1)
    a.f_in = a.f_out = 4096
    assert lib.lbm_step(ctypes.byref(a), None) == -1
    assert b"distinct" in lib.lbm_last_error()
    a.f_out = 8192
    a.collision = N.COLL_KBC
    a.tau = 0.6
    assert lib.lbm_step(ctypes.byref(a), None) == -3   # KBC has no D3Q19 kernel
    a.collision = N.COLL_BGK
    a.stride[:] = (10, 16, 4, 1)                        # populations overlap
    assert lib.lbm_step(ctypes.byref(a), None) == -1
    a.stride[:] = (64, 16, 4, 1)
    a.sr_seed = 12345                                   # stochastic rounding: BF16 only
    assert lib.lbm_step(ctypes.byref(a), None) == -3
    assert b"BF16" in lib.lbm_last_error()
    a.sr_seed = 0


def test_training_entry_points_reject_bad_arguments_without_a_gpu():
    """The training kernels validate storage, extents, tau and buffers
    before any HIP call."""
    lib = N.lib()
    n = (ctypes.c_int64 * 3)(2, 8, 8)
    p, a, b = ctypes.c_void_p(4096), ctypes.c_void_p(8192), ctypes.c_void_p(16384)
    fwd = lib.lbm_learned_train_forward
    D2, D3 = N.D2Q9, N.D3Q27
    assert fwd(D2, N.F16_SHIFTED, n, 0.6, p, 100.0, 0.05, 1.95, a, b, None) == -3
    assert fwd(N.D3Q19, N.F32, n, 0.6, p, 100.0, 0.05, 1.95, a, b, None) == -3
    assert fwd(D2, N.F32, n, 0.5, p, 100.0, 0.05, 1.95, a, b, None) == -1
    assert b"tau" in lib.lbm_last_error()
    assert fwd(D3, N.F32, n, 0.6, p, 100.0, 0.05, 1.95, a, a, None) == -1
    assert fwd(D2, N.F32, n, 0.6, None, 100.0, 0.05, 1.95, a, b, None) == -1
    assert fwd(D2, N.F32, n, 0.6, p, 100.0, 1.5, 1.0, a, b, None) == -1
    zero = (ctypes.c_int64 * 3)(2, 0, 8)
    assert fwd(D2, N.F64, zero, 0.6, p, 100.0, 0.05, 1.95, a, b, None) == -1
    bwd = lib.lbm_learned_train_backward
    assert bwd(D2, N.F64, n, 0.6, p, 100.0, 0.05, 1.95, a, b, 1, b, None, None) == -1
    assert bwd(D3, N.BF16_SHIFTED, n, 0.6, p, 100.0, 0.05, 1.95, a, b, 1, p, None, None) == -3
    assert lib.lbm_unpull(D2, N.F32, n, a, a, None) == -1
    big = (ctypes.c_int64 * 3)(1 << 20, 1 << 10, 1 << 10)
    assert lib.lbm_unpull(D2, N.F32, big, a, b, None) == -1
    # 27 populations of 2^27 cells pass 2^31 (9 of them would not)
    mid = (ctypes.c_int64 * 3)(1 << 7, 1 << 10, 1 << 10)
    assert lib.lbm_unpull(D3, N.F32, mid, a, b, None) == -1
    assert b"2^31" in lib.lbm_last_error()
    # the parameter-gradient reduction: one partial gradient per 2048
    # cells, and argument checks before any HIP call
    ws, P = lib.lbm_learned_grad_workspace, N.LEARNED_PARAMS
    assert ws(N.F32, 2048) == P * 4
    assert ws(N.F64, 2049) == 2 * P * 8
    assert ws(N.F32, 32 * 256 * 256) == 1024 * P * 4
    assert ws(N.BF16_SHIFTED, 2048) == -1 and ws(N.F32, 0) == -1
    red = lib.lbm_learned_grad_reduce
    assert red(N.F16_SHIFTED, 2048, a, b, p, None) == -3
    assert red(N.F32, 0, a, b, p, None) == -1
    assert red(N.F32, 2048, a, None, p, None) == -1


def test_scan_rejects_bad_arguments_without_a_gpu():
    lib = N.lib()
    n = (ctypes.c_int64 * 3)(1, 8, 8)
    s = (ctypes.c_int64 * 4)(64, 0, 8, 1)
    out = (ctypes.c_double * 3)()
    ws = ctypes.c_void_p(4096)
    assert lib.lbm_scan_f64(N.D2Q9, N.F32, n, s, ctypes.c_void_p(8192), 1.0, out, None,
                            None) == -1
    assert lib.lbm_scan_f64(9, N.F32, n, s, ctypes.c_void_p(8192), 1.0, out, ws, None) == -1
    bad = (ctypes.c_int64 * 4)(8, 0, 8, 1)    # populations overlap
    assert lib.lbm_scan_f64(N.D2Q9, N.F32, n, bad, ctypes.c_void_p(8192), 1.0, out, ws,
                            None) == -1


def test_slab_entry_points_reject_bad_arguments_without_a_gpu():
    """x-slabs: 3D stencils only, pull streaming, no boundaries, no learned
    collision, four distinct halo buffers; checked before any HIP call."""
    lib = N.lib()
    assert [lib.lbm_halo_planes(s) for s in (N.D2Q9, N.D3Q19, N.D3Q27)] == [-1, 5, 9]
    a = N.LbmStepArgs()
    a.abi_version = N.ABI_VERSION
    a.stencil = N.D3Q19
    a.collision = N.COLL_BGK
    a.tau = 0.6
    a.streaming = 2
    a.n[:] = (4, 4, 4)
    a.stride[:] = (64, 16, 4, 1)
    a.f_in, a.f_out = 4096, 8192
    h = [ctypes.c_void_p(v) for v in (1 << 16, 2 << 16, 3 << 16, 4 << 16)]
    step = lib.lbm_slab_step
    assert step(ctypes.byref(a), 3, *h, None) == -1               # unknown part
    assert step(ctypes.byref(a), N.SLAB_ALL, h[0], h[0], h[2], h[3], None) == -1
    assert b"distinct" in lib.lbm_last_error()
    assert step(ctypes.byref(a), N.SLAB_FACES, None, h[1], h[2], h[3], None) == -1
    a.streaming = 1
    assert step(ctypes.byref(a), N.SLAB_ALL, *h, None) == -3      # push
    a.streaming = 2
    a.cell_class = 1 << 20
    a.classes = 1 << 21
    a.n_classes = 1
    assert step(ctypes.byref(a), N.SLAB_ALL, *h, None) == -3      # boundaries
    a.cell_class = a.classes = None
    a.n_classes = 0
    a.stencil = N.D2Q9
    a.n[:] = (1, 4, 4)
    a.stride[:] = (16, 16, 4, 1)
    assert step(ctypes.byref(a), N.SLAB_ALL, *h, None) == -3      # 2D
    n = (ctypes.c_int64 * 3)(4, 4, 4)
    s = (ctypes.c_int64 * 4)(64, 16, 4, 1)
    pack = lib.lbm_halo_pack
    n2 = (ctypes.c_int64 * 3)(1, 4, 4)
    s2 = (ctypes.c_int64 * 4)(16, 16, 4, 1)
    assert pack(N.D2Q9, N.F32, n2, s2, h[0], h[1], h[2], None) == -3
    assert pack(N.D3Q19, N.F32, n, s, h[0], h[1], h[1], None) == -1
    assert pack(N.D3Q19, N.F32, n, s, None, h[1], h[2], None) == -1


def test_training_rollout_needs_a_device_tensor():
    f = torch.zeros(9, 2, 4, 4)
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        lt.learned_rollout(f, torch.zeros(N.LEARNED_PARAMS), lt.TrainConfig(tau=0.6), 2)


def test_flat_parameters_roundtrip():
    p = lt.LearnedParameters.seeded(4, in_scale=50.0)
    theta = lt.flatten_parameters(p)
    assert theta.numel() == N.LEARNED_PARAMS
    q = lt.unflatten_parameters(theta, like=p)
    for a, b in zip(p.tensors(), q.tensors()):
        assert torch.equal(a, b)
    assert q.in_scale == 50.0
    # the oracle splits the same vector into the same tensors
    for a, b in zip(p.tensors(), O.split_learned_params(theta)):
        assert torch.equal(a, b)


def _cpu_flow(stencil, res):
    return SimpleNamespace(stencil=stencil, resolution=list(res),
                           context=lt.Context("cpu"), torch_stencil=None)


def test_masks_and_classes_match_lettuce_obstacle():
    """ncm / nsm built by lettuce_amd equal lettuce's (golden) bit for bit,
    and the cell-class map reproduces both."""
    g = golden("g6_obstacle_eqoutlet.npz")
    ctx = lt.Context("cpu")
    flow = _cpu_flow(lt.D2Q9(), (48, 24))
    x = torch.arange(48, dtype=torch.float64) / 12.0
    inlet = (torch.abs(x) < 1e-6)[:, None].expand(48, 24)
    bounds = [lt.EquilibriumBoundaryPU(ctx, flow, inlet, [1.0, 0.0]),
              lt.EquilibriumOutletP([1, 0], flow),
              lt.BounceBackBoundary(torch.from_numpy(g["mask"]))]
    ncm, nsm = build_masks(9, [48, 24], [], bounds, [], ctx)
    assert np.array_equal(ncm.numpy(), g["ncm"])
    assert np.array_equal(nsm.numpy(), g["nsm"])
    ops = {0: OpSpec(op=N.OP_COLLIDE), 1: OpSpec(op=N.OP_EQUILIBRIUM),
           2: OpSpec(op=N.OP_OUTLET), 3: OpSpec(op=N.OP_BOUNCE_BACK)}
    cc = build_cell_classes(ncm, nsm, ops, lt.D2Q9().e)
    r_ncm, r_nsm, r_solid = expand_classes(cc, 9)
    assert torch.equal(r_ncm, ncm)
    assert torch.equal(r_nsm, (nsm == 1).to(torch.uint8))
    assert not r_solid.any()
    assert cc.n_classes <= 256


def test_nsm_predicate_is_equality_with_one():
    """lettuce's torch step keeps population q only where nsm[q] == 1
    (lettuce/_simulation.py:254); with two pre-boundaries nsm starts at 2."""
    ctx = lt.Context("cpu")
    m = torch.zeros(6, 6, dtype=torch.bool)
    m[1:3, 1:3] = True

    class B:
        def __init__(self, mask):
            self.mask = mask

        def make_no_collision_mask(self, shape, context):
            return self.mask

        def make_no_streaming_mask(self, shape, context):
            return self.mask

    ncm, nsm = build_masks(9, [6, 6], [B(m), B(~m)], [], [], ctx)
    assert int(nsm.min()) == 3 and int(nsm.max()) == 3   # 2 | 1
    cc = build_cell_classes(ncm, nsm, {}, lt.D2Q9().e)
    _, r_nsm, _ = expand_classes(cc, 9)
    assert not r_nsm.any()


def test_halfway_links_match_oracle_link_list():
    s = lt.D2Q9()
    g = torch.Generator().manual_seed(3)
    solid = torch.rand(12, 9, generator=g) < 0.2
    links = halfway_links(solid, s.e, s.opposite)
    oflow = O.Flow(O.Context("cpu", torch.float64), [12, 9],
                   O.UnitConversion(10.0), O.D2Q9())
    hw = O.HalfwayBounceBackBoundary(oflow.context, oflow, solid)
    ref = {(int(r[1]) * 9 + int(r[2]), int(r[0])) for r in hw.f_index}
    ours = {(int(c), int(d)) for c, d in links.tolist()}
    assert ours == ref
    assert links.shape[0] == len(ref)
    # ordered by (cell, direction)
    key = links[:, 0].long() * 64 + links[:, 1].long()
    assert torch.all(key[1:] > key[:-1])


def test_halfway_classes_flags():
    s = lt.D3Q27()
    solid = torch.zeros(5, 5, 5, dtype=torch.bool)
    solid[2, 2, 2] = True
    ncm = torch.full((5, 5, 5), 0, dtype=torch.uint8)
    ncm[solid] = 1
    nsm = torch.zeros((27, 5, 5, 5), dtype=torch.uint8)
    nsm[:, solid] = 1
    cc = build_cell_classes(ncm, nsm, {0: OpSpec(op=N.OP_COLLIDE)}, s.e, solid)
    flags = torch.tensor([c.flags for c in cc.table])[cc.cell_class.long()]
    near = (flags & N.CLASS_NEAR_SOLID) != 0
    assert int(near.sum()) == 26
    assert bool(flags[2, 2, 2] & N.CLASS_SOLID)


def _plan_args(stencil, storage, n, stride, mask=False):
    a = N.LbmStepArgs()
    a.abi_version = N.ABI_VERSION
    a.stencil, a.storage = stencil, storage
    a.collision, a.tau, a.streaming = N.COLL_BGK, 0.6, 2
    a.n[:], a.stride[:] = n, stride
    a.f_in, a.f_out = 1 << 20, 1 << 40
    if mask:
        a.cell_class, a.classes, a.n_classes = 1 << 21, 1 << 22, 1
    return a


def test_grid_limits_apply_per_i0_plane_without_a_gpu():
    """offsets inside the kernels are an i0 plane's 64-bit offset plus a
    32-bit offset inside the plane: any number of planes, but a plane below
    2^31 elements and (formats up to 4 bytes) 2^32 bytes; grids with
    boundaries keep 32-bit cell offsets.  Checked before any HIP call."""
    lib = N.lib()
    plan = (ctypes.c_int64 * 2)()
    Q, n = 19, 1024
    # 1024^3 D3Q19 FP32 in `rows` ([X, Y, Q, Z] memory order): 2.0e10
    # elements in all, 2.0e7 per plane
    rows = _plan_args(N.D3Q19, N.F32, (n, n, n), (n, n * Q * n, Q * n, 1))
    assert lib.lbm_run_plan(ctypes.byref(rows), 8, plan) == 0, lib.lbm_last_error()
    assert list(plan) == [4, 0]
    # a 40000^2 D2Q9 grid is one plane of 1.6e9 cells: above 2^30 FP32
    # elements (2^32 bytes), below 2^31 16-bit elements
    flat = (1, 40000, 40000)
    flat_stride = (40000 * 40000, 40000 * 40000, 40000, 1)
    big = _plan_args(N.D2Q9, N.F32, flat, flat_stride)
    assert lib.lbm_run_plan(ctypes.byref(big), 1, plan) == -1
    assert b"i0 plane" in lib.lbm_last_error()
    half = _plan_args(N.D2Q9, N.BF16_SHIFTED, flat, flat_stride)
    assert lib.lbm_run_plan(ctypes.byref(half), 1, plan) == 0, lib.lbm_last_error()
    # with boundaries: at most 2^31 - 1 cells (here two planes of 1.2e9)
    cells = (2, 40000, 30000)
    masked = _plan_args(N.D3Q19, N.F16_SHIFTED, cells,
                        (2 * 40000 * 30000, 40000 * 30000, 30000, 1), mask=True)
    assert lib.lbm_run_plan(ctypes.byref(masked), 1, plan) == -1
    assert b"2^31 - 1 cells" in lib.lbm_last_error()


def test_high_ma_reporter_counts_non_finite_velocity(monkeypatch):
    """HighMaReporter fails on cells whose |u|^2 is not finite (scan count 1)
    even with no NaN population and a finite max |u|^2 below the threshold"""
    from lettuce_amd.ext import _failure
    sim = SimpleNamespace(flow=SimpleNamespace(stencil=lt.D2Q9()))
    rep = lt.HighMaReporter(1, threshold=0.3)
    for scan, failed in (([0.0, 0.0, 0.01], False), ([0.0, 1.0, 0.01], True),
                         ([2.0, 0.0, 0.01], True), ([0.0, 0.0, 0.5], True)):
        monkeypatch.setattr(_failure, "_scan", lambda flow, s=scan: torch.tensor(s))
        assert rep.is_failed(sim) is failed


@pytest.mark.parametrize("res,key", [([32, 32], "u0_lu"), ([8, 8, 8], "u0_lu_3d")])
def test_decaying_turbulence_field_matches_lettuce(res, key):
    """The package's DecayingTurbulence field (torch FP64 FFTs) against the
    field lettuce itself generated (golden g7: Re 1000, Ma 0.1, k0 4, seed
    0), in lattice units, and lettuce's characteristic velocity max |u|"""
    from lettuce_amd.ext._flows import turbulence_velocity
    g = golden("g7_decaying_turbulence.npz")
    units = lt.UnitConversion(reynolds_number=1000, mach_number=0.1,
                              characteristic_length_lu=res[0],
                              characteristic_length_pu=2 * np.pi,
                              characteristic_velocity_pu=None)
    u, _, _ = turbulence_velocity(res, 4, 0.5, units.convert_length_to_pu(1.0),
                                  units.characteristic_length_pu, 0, "cpu")
    units.characteristic_velocity_pu = float(torch.linalg.vector_norm(u, dim=0).max())
    u_lu = units.convert_velocity_to_lu(u).numpy()
    ref = g[key]
    assert np.abs(u_lu - ref).max() <= 1e-12 * np.abs(ref).max()
    if key == "u0_lu":
        assert abs(units.characteristic_velocity_pu / float(g["char_velocity_pu"]) - 1) <= 1e-13


@pytest.mark.parametrize("res", [[16, 16], [8, 8, 8]])
def test_decaying_turbulence_spectrum_is_the_shell_sum(res):
    """the shell spectrum (a scatter-add over shell indices) against a
    brute-force sum over each shell k - 1/2 < |k| <= k + 1/2 of
    E(k) ~ k^4 exp(-2 (k/4)^2), normalised to the initial energy"""
    from lettuce_amd.ext._flows import turbulence_velocity
    _, spec, wn = turbulence_velocity(res, 4, 0.5, 2 * np.pi / res[0], 2 * np.pi, 0, "cpu")
    k = np.stack(np.meshgrid(*[np.fft.fftfreq(n, 1 / n) for n in res], indexing="ij"))
    norm = np.sqrt((k ** 2).sum(0))
    ek = norm ** 4 * np.exp(-2 * (norm / 4) ** 2)
    ek = 0.5 * ek / ek.sum()
    assert len(wn) == int(norm.max()) and list(wn) == list(range(len(wn)))
    for s in wn:
        shell = (norm > s - 0.5) & (norm <= s + 0.5)
        assert abs(spec[s] - ek[shell].sum()) <= 1e-14
    assert abs(spec.sum() - ek[norm <= len(wn) - 0.5].sum()) <= 1e-14


def test_poiseuille_units_and_acceleration_match_lettuce():
    """PoiseuilleFlow2D's units (tau) and the acceleration in lattice units
    against lettuce's (golden g9: 17^2, Re 1, Ma 0.02)"""
    g = golden("g9_poiseuille_guo.npz")
    units = lt.UnitConversion(reynolds_number=1, mach_number=0.02,
                              characteristic_length_lu=16, characteristic_length_pu=1,
                              characteristic_velocity_pu=1)
    assert units.relaxation_parameter_lu == float(g["tau"])
    a = units.convert_acceleration_to_lu(torch.tensor([0.001, 0.0], dtype=torch.float64))
    assert np.array_equal(np.asarray(a), g["acceleration_lu"])


def test_stochastic_rounding_is_a_bf16_context_option():
    """Context(rounding="stochastic") applies to shifted BF16 storage only
    and carries a non-zero seed to the step; round to nearest carries 0."""
    import pytest
    import lettuce_amd as lt
    c = lt.Context("cpu", storage="bf16_shifted", rounding="stochastic", rounding_seed=7)
    assert c.rounding == "stochastic" and c.rounding_seed == 7
    assert lt.Context("cpu", storage="bf16_shifted").rounding_seed == 0
    with pytest.raises(ValueError):
        lt.Context("cpu", storage="f16_shifted", rounding="stochastic")
    with pytest.raises(ValueError):
        lt.Context("cpu", storage="bf16_shifted", rounding="stochastic", rounding_seed=0)
    with pytest.raises(ValueError):
        lt.Context("cpu", rounding="up")

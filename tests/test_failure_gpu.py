"""GPU checks of the failure scan (lbm_scan_f64) against torch on the same
tensors, and of NaNReporter / HighMaReporter with BreakableSimulation
(lettuce/ext/_reporter/failure_reporter.py:12-184, _simulation.py:325-344)."""
import math

import pytest
import torch

import lettuce_amd as lt
from lettuce_amd import _ops

pytestmark = pytest.mark.gpu


def _tgv(storage="f32", dtype=torch.float32, res=32, ma=0.1, stencil=None):
    ctx = lt.Context("cuda:0", dtype=dtype, storage=storage)
    stencil = stencil or lt.D2Q9()
    return lt.TaylorGreenVortex(ctx, res, 100, ma, stencil, initialize_fneq=False)


def _sim(flow, reporters):
    tau = flow.units.relaxation_parameter_lu
    return lt.BreakableSimulation(flow, lt.BGKCollision(tau), reporters,
                                  lt.StreamingStrategy.PRE_STREAMING)


@pytest.mark.parametrize("storage,dtype", [("f32", torch.float32), ("f64", torch.float64),
                                           ("f16_shifted", torch.float32)])
@pytest.mark.parametrize("stencil", [lt.D2Q9, lt.D3Q19], ids=["D2Q9", "D3Q19"])
def test_scan_matches_torch(storage, dtype, stencil):
    flow = _tgv(storage, dtype, res=16, stencil=stencil())
    sim = _sim(flow, [])
    sim(3)
    out = _ops.scan_f64(flow.stencil, flow._state, flow.rho0_shift).cpu()
    f = flow.f.double()
    e = torch.tensor(flow.stencil.e, dtype=torch.float64, device=f.device)
    u = torch.einsum("qd,q...->d...", e, f) / f.sum(0)
    uu = (u * u).sum(0).max().item()
    assert out[0].item() == 0 and out[1].item() == 0
    assert abs(out[2].item() - uu) <= 1e-12 * uu


def test_scan_counts_nan_populations_and_cells():
    flow = _tgv(res=16)
    state = flow._state
    state[0, 3, 4] = float("nan")
    state[5, 3, 4] = float("nan")
    state[2, 7, 1] = float("nan")
    out = _ops.scan_f64(flow.stencil, state).cpu()
    assert out[0].item() == 3 and out[1].item() == 2
    good = torch.ones(16, 16, dtype=torch.bool)
    good[3, 4] = good[7, 1] = False
    f = state.double()
    e = torch.tensor(flow.stencil.e, dtype=torch.float64, device=f.device)
    u = torch.einsum("qd,q...->d...", e, f) / f.sum(0)
    uu = (u * u).sum(0)[good.to(f.device)].max().item()
    assert abs(out[2].item() - uu) <= 1e-12 * uu


@pytest.mark.parametrize("storage", ["f32", "f16_shifted"])
def test_nan_reporter_stops_breakable_simulation(tmp_path, storage):
    flow = _tgv(storage)
    rep = lt.NaNReporter(4, k=5, outdir=str(tmp_path))
    sim = _sim(flow, [rep])
    sim(10)
    assert rep.failed_iteration is None and flow.i == 10
    f = flow.f
    f[1, 5, 6] = float("nan")
    flow.f = f
    sim(30)
    assert rep.failed_iteration == 12
    assert flow.i == 12 + int(1e10)
    assert len(rep.results) == 5 and all(math.isnan(v) for _, v in rep.results)
    assert (tmp_path / "NaN_reporter.log").exists()
    log = (tmp_path / "NaN_reporter.log").read_text()
    assert "NaN detected in step 12" in log
    assert list(tmp_path.glob("NaN_frame_*.npz"))


def test_high_ma_reporter_threshold():
    flow = _tgv(ma=0.1)          # u0 = 0.1 cs: max Ma = 0.1
    rep = lt.HighMaReporter(2, threshold=0.2)
    _sim(flow, [rep])(20)
    assert rep.failed_iteration is None and flow.i == 20
    flow = _tgv(ma=0.1)
    rep = lt.HighMaReporter(2, threshold=0.05, k=3)
    _sim(flow, [rep])(20)
    assert rep.failed_iteration == 0
    ma = torch.linalg.vector_norm(flow.u(), dim=0) / flow.stencil.cs
    top = rep.results[0][1]
    assert abs(top - ma.max().item()) <= 1e-6 and len(rep.results) == 3
    assert rep.results[0][1] >= rep.results[1][1] >= rep.results[2][1]


def test_high_ma_reporter_fails_on_inf_population_without_nan():
    """an overflowing population (inf, no NaN) leaves |u|^2 non-finite in its
    cell: HighMaReporter must stop there although no population is NaN"""
    flow = _tgv(res=16)
    state = flow._state
    state[1, 3, 4] = float("inf")
    out = _ops.scan_f64(flow.stencil, state).cpu()
    assert out[0].item() == 0 and out[1].item() == 1
    rep = lt.HighMaReporter(1, threshold=10.0, k=2)
    assert rep.is_failed(_sim(flow, []))

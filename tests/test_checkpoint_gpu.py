"""Checkpoints: lettuce's dump/load test (tests/test_checkpoint.py:5-20) and
the bit-exact restart file (Flow.save_checkpoint / load_checkpoint)."""
import pytest
import torch

import lettuce_amd as lt

pytestmark = pytest.mark.gpu


def _flow(storage="f32", layout="soa", d=2, res=32):
    """storage "bf16_shifted_sr": BF16 with stochastic rounding"""
    dtype = torch.float64 if storage == "f64" else torch.float32
    sr = storage.endswith("_sr")
    ctx = lt.Context("cuda:0", dtype=dtype, storage=storage[:-3] if sr else storage,
                     layout=layout, rounding="stochastic" if sr else "nearest")
    stencil = lt.D2Q9() if d == 2 else lt.D3Q19()
    return lt.TaylorGreenVortex(ctx, res, 100, 0.1, stencil, initialize_fneq=False)


def _bits(t):
    """the stored bytes in logical [q, *res] order"""
    return t.detach().cpu().contiguous().view(-1).view(torch.uint8)


def _sim(flow):
    return lt.Simulation(flow, lt.BGKCollision(flow.units.relaxation_parameter_lu), [],
                         lt.StreamingStrategy.PRE_STREAMING)


def test_dump_load_restores_f(tmp_path):
    flow = _flow()
    f0 = flow.f.clone()
    flow.dump(tmp_path / "tgv")
    _sim(flow)(10)
    assert not torch.equal(flow.f, f0)
    flow.load(tmp_path / "tgv")
    assert torch.equal(flow.f, f0)


@pytest.mark.parametrize("storage", ["f32", "f64", "f16_shifted", "bf16_shifted",
                                     "bf16_shifted_sr"])
@pytest.mark.parametrize("layout,d", [("soa", 2), ("rows", 3)])
def test_restart_is_bit_exact(tmp_path, storage, layout, d):
    """with stochastic rounding too: the random bits are keyed to the step
    index the file restores"""
    res = 32 if d == 2 else 16
    a = _flow(storage, layout, d, res)
    sim = _sim(a)
    sim(7)
    path = tmp_path / "ck.npz"
    a.save_checkpoint(path)
    saved = a._state.clone()
    sim(5)
    b = _flow(storage, layout, d, res)
    b.load_checkpoint(path)
    assert b.i == 7
    assert b._state.stride() == saved.stride()
    assert torch.equal(_bits(b._state), _bits(saved))
    _sim(b)(5)
    assert b.i == a.i == 12
    assert torch.equal(_bits(b._state), _bits(a._state))


def test_checkpoint_rho0_shift_mismatch_is_rejected(tmp_path):
    """shifted storage decodes against rho0_shift: a file written under one
    shift must not load into a flow that decodes with another"""
    a = _flow("f16_shifted")
    path = tmp_path / "ck.npz"
    a.save_checkpoint(path)

    class Shifted(lt.TaylorGreenVortex):
        @property
        def rho0_shift(self):
            return 0.5

    ctx = lt.Context("cuda:0", dtype=torch.float32, storage="f16_shifted")
    b = Shifted(ctx, 32, 100, 0.1, lt.D2Q9(), initialize_fneq=False)
    with pytest.raises(ValueError, match="rho0_shift"):
        b.load_checkpoint(path)
    c = _flow("f16_shifted")
    c.load_checkpoint(path)
    assert torch.equal(_bits(c._state), _bits(a._state))


def test_checkpoint_mismatches_are_rejected(tmp_path):
    a = _flow("f32")
    path = tmp_path / "ck.npz"
    a.save_checkpoint(path)
    with pytest.raises(ValueError, match="storage"):
        _flow("f16_shifted").load_checkpoint(path)
    with pytest.raises(ValueError, match="resolution"):
        _flow("f32", res=16).load_checkpoint(path)
    other = lt.TaylorGreenVortex(lt.Context("cuda:0"), 32, 200, 0.1, lt.D2Q9())
    with pytest.raises(ValueError, match="reynolds_number"):
        other.load_checkpoint(path)
    other.load_checkpoint(path, check_units=False)
    assert other.i == 0 and torch.equal(other.f, a.f)


def test_snapshot_reporter_feeds_the_dataset(tmp_path):
    """SnapshotReporter every 5 steps -> LettuceDataset pairs (lettuce's
    HDF5Reporter / LettuceDataset, util/datautils.py:17-156)"""
    flow = _flow("f16_shifted")
    base = str(tmp_path / "snaps")
    rep = lt.SnapshotReporter(flow, lt.BGKCollision(0.6), interval=5, filebase=base,
                              metadata={"case": "tgv"})
    sim = lt.Simulation(flow, lt.BGKCollision(flow.units.relaxation_parameter_lu), [rep],
                        lt.StreamingStrategy.PRE_STREAMING)
    sim(20)
    ds = lt.LettuceDataset(base, target=True, skip_idx_to_target=1)
    assert ds.get_attr("steps") == [0, 5, 10, 15, 20] and len(ds) == 4
    assert ds.get_attr("metadata") == {"case": "tgv"} and ds.get_attr("storage") == "f16_shifted"
    f, target, idx = ds[3]
    assert idx == 3 and torch.equal(target, flow.f.cpu())
    assert f.shape == target.shape == (9, 32, 32)

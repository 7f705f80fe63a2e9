"""Snapshot files and LettuceDataset (lettuce/util/datautils.py:17-156), on
the CPU with the writer the reporter uses."""
import numpy as np
import torch

from lettuce_amd.util.datautils import LettuceDataset, SnapshotWriter


def test_snapshot_roundtrip_and_target_pairs(tmp_path):
    base = str(tmp_path / "run")
    rng = np.random.default_rng(0)
    snaps = [rng.random((9, 8, 6)).astype(np.float32) for _ in range(5)]
    w = SnapshotWriter(base, [9, 8, 6], np.float32, {"flow": "TGV", "metadata": {"a": "1"}})
    for i, s in enumerate(snaps):
        w.append(s, step=10 * i)
    ds = LettuceDataset(base + ".json")
    assert len(ds) == 5 and ds.get_attr("steps") == [0, 10, 20, 30, 40]
    f, idx = ds[3]
    assert idx == 3 and torch.equal(f, torch.from_numpy(snaps[3]))
    pairs = LettuceDataset(base, target=True, skip_idx_to_target=2,
                           transform=lambda t: 2 * t)
    assert len(pairs) == 3
    f, target, idx = pairs[1]
    assert torch.equal(f, 2 * torch.from_numpy(snaps[1]))
    assert torch.equal(target, 2 * torch.from_numpy(snaps[3]))
    loader = torch.utils.data.DataLoader(pairs, batch_size=2, shuffle=False)
    batches = list(loader)
    assert batches[0][0].shape == (2, 9, 8, 6) and batches[-1][2].tolist() == [2]


def test_snapshot_writer_rejects_wrong_shapes(tmp_path):
    w = SnapshotWriter(str(tmp_path / "x"), [9, 4, 4], np.float64)
    try:
        w.append(np.zeros((9, 4, 5)), 0)
    except ValueError:
        pass
    else:
        raise AssertionError("shape not checked")
    assert LettuceDataset(str(tmp_path / "x")).__len__() == 0

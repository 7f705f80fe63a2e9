"""Population snapshots for training data (lettuce/util/datautils.py:17-156).

lettuce writes snapshots of f into an HDF5 file with h5py and pickles the
flow and collision into its attributes.  Here the same reporter / dataset
pair uses a raw, append-only binary file and a JSON sidecar, so that
neither h5py nor pickle is needed and the dataset can memory-map the data:

    <filebase>.bin    snapshots of f, C order, one [Q, *resolution] array each
    <filebase>.json   {"format", "dtype", "shape", "count", "steps",
                       "flow", "collision", "units", "metadata"}

The JSON records class names and unit parameters instead of pickled
objects.  LettuceDataset returns (f, idx), or (f, target, idx) with
target = snapshot idx + skip_idx_to_target, as lettuce's does.
"""
import json
import os
from typing import Optional

import numpy as np
import torch
from torch.utils import data

from .._simulation import Reporter

__all__ = ["SnapshotReporter", "LettuceDataset", "SnapshotWriter"]

_FORMAT = "lettuce_amd-snapshots-1"
_UNIT_FIELDS = ("reynolds_number", "mach_number", "characteristic_length_pu",
                "characteristic_velocity_pu", "characteristic_length_lu",
                "characteristic_density_lu", "characteristic_density_pu")


class SnapshotWriter:
    """Append-only writer of the snapshot file pair."""

    def __init__(self, filebase: str, shape, dtype, info: Optional[dict] = None):
        self.filebase = str(filebase)
        self.shape = [int(s) for s in shape]
        self.dtype = np.dtype(dtype)
        self.meta = {"format": _FORMAT, "dtype": self.dtype.str, "shape": self.shape,
                     "count": 0, "steps": []}
        self.meta.update(info or {})
        d = os.path.dirname(self.filebase)
        if d:
            os.makedirs(d, exist_ok=True)
        open(self.filebase + ".bin", "wb").close()
        self._write_meta()

    def _write_meta(self):
        tmp = self.filebase + ".json.tmp"
        with open(tmp, "w") as fh:
            json.dump(self.meta, fh)
        os.replace(tmp, self.filebase + ".json")

    def append(self, f: np.ndarray, step: int):
        f = np.ascontiguousarray(f, dtype=self.dtype)
        if list(f.shape) != self.shape:
            raise ValueError(f"snapshot must have shape {self.shape}, got {list(f.shape)}")
        with open(self.filebase + ".bin", "ab") as fh:
            fh.write(f.tobytes())
        self.meta["count"] += 1
        self.meta["steps"].append(int(step))
        self._write_meta()


class SnapshotReporter(Reporter):
    """Appends the populations flow.f every `interval` steps
    (HDF5Reporter, datautils.py:17-80).  `metadata` values are stored as
    strings, as lettuce requires them to be."""

    periodic_interval = True

    def __init__(self, flow, collision, interval, filebase="./output", metadata=None):
        super().__init__(interval)
        self.context = flow.context
        units = {k: float(getattr(flow.units, k)) for k in _UNIT_FIELDS}
        info = {"flow": type(flow).__name__, "collision": type(collision).__name__,
                "stencil": type(flow.stencil).__name__, "storage": flow.context.storage,
                "units": units,
                "metadata": {k: str(v) for k, v in (metadata or {}).items()}}
        dtype = np.float64 if flow.context.dtype == torch.float64 else np.float32
        self.writer = SnapshotWriter(filebase, [flow.stencil.q] + list(flow.resolution),
                                     dtype, info)

    def __call__(self, simulation):
        if simulation.flow.i % self.interval == 0:
            self.writer.append(self.context.convert_to_ndarray(simulation.flow.f),
                               simulation.flow.i)


class LettuceDataset(data.Dataset):
    """Snapshots written by SnapshotReporter as a torch Dataset
    (datautils.py:83-156).  `filebase` may name the .json or .bin file or
    their common base.  Tensors are created on `device` (CPU by default)."""

    def __init__(self, filebase, transform=None, target=False, skip_idx_to_target=1,
                 device="cpu"):
        super().__init__()
        base = str(filebase)
        for ext in (".json", ".bin"):
            if base.endswith(ext):
                base = base[:-len(ext)]
        self.filebase = base
        self.transform = transform
        self.target = target
        self.skip_idx_to_target = skip_idx_to_target
        self.device = torch.device(device)
        with open(base + ".json") as fh:
            self.meta = json.load(fh)
        if self.meta.get("format") != _FORMAT:
            raise ValueError(f"{base}.json is not a snapshot file")
        self.shape = (self.meta["count"], *self.meta["shape"])
        self.data = np.memmap(base + ".bin", dtype=np.dtype(self.meta["dtype"]), mode="r",
                              shape=self.shape) if self.meta["count"] else None

    def __str__(self):
        return "\n".join(f"{k}: {v}" for k, v in self.meta.items() if k != "steps")

    def __len__(self):
        n = self.shape[0]
        return max(n - self.skip_idx_to_target, 0) if self.target else n

    def __getitem__(self, idx):
        f = self.get_data(idx)
        target = []
        if self.target:
            target = self.get_data(idx + self.skip_idx_to_target)
        if self.transform:
            f = self.transform(f)
            if self.target:
                target = self.transform(target)
        return (f, target, idx) if self.target else (f, idx)

    def get_data(self, idx):
        if idx < 0 or idx >= self.shape[0]:
            raise IndexError(idx)
        return torch.from_numpy(np.array(self.data[idx])).to(self.device)

    def get_attr(self, attr):
        return self.meta[attr]

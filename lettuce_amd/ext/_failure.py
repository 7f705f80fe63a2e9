"""Failure reporters (lettuce/ext/_reporter/failure_reporter.py:12-184).

The check that runs every `interval` steps is one native FP64 pass over the
stored populations (lbm_scan_f64): NaN populations and the largest |u|^2,
without materialising u.  Only when a check fails are the locations listed,
from the decoded populations and velocities on the GPU, and the simulation is
stopped the way lettuce stops it, by pushing flow.i past any step target
(failure_reporter.py:49-53), which BreakableSimulation honours.

With `outdir`, lettuce writes a text log and a VTK frame (pyevtk).  Here the
log has lettuce's format and the frame is a numpy .npz of p and u in
physical units (pyevtk is not a dependency of this package).
"""
import os
from abc import ABC, abstractmethod
from typing import List, Optional, Tuple

import numpy as np
import torch

from .. import _ops
from .._simulation import Reporter

__all__ = ["FailureReporterBase", "NaNReporter", "HighMaReporter"]


def _scan(flow):
    return _ops.scan_f64(flow.stencil, flow._state, flow.rho0_shift)


class FailureReporterBase(Reporter, ABC):
    periodic_interval = True

    def __init__(self, interval: int, k: int = 100, outdir: Optional[str] = None,
                 vtk_out: bool = False):
        super().__init__(interval)
        self.k = k
        self.outdir = outdir
        self.name = "Failure"
        self.vtk_out = vtk_out
        self.failed_iteration = None
        self.results: List[Tuple] = []

    def __call__(self, simulation):
        flow = simulation.flow
        if flow.i % self.interval != 0 or not self.is_failed(simulation):
            return
        self.results = self.get_results(simulation)
        self.failed_iteration = flow.i
        if self.outdir is not None:
            self._write_log(simulation, self.results)
            self.save_frame(simulation)
        print(f"(!) ABORT MESSAGE: {self.name}Reporter detected {self.name} "
              f"(reporter-interval = {self.interval}) at iteration {flow.i}. "
              f"See log in {self.outdir} for details!")
        flow.i = int(flow.i + 1e10)

    def _get_top_failures(self, mask: torch.Tensor, values: torch.Tensor) -> List[Tuple]:
        """(position, value) of at most k failing nodes: the first k if the
        values are NaN, else the k largest (failure_reporter.py:56-74)"""
        failed = values[mask]
        coords = torch.nonzero(mask)
        k = min(self.k, failed.numel())
        if torch.isnan(failed).any():
            top = torch.arange(k, device=values.device)
        else:
            _, top = torch.topk(failed, k=k, largest=True, sorted=True)
        return [(list(c.astype(int)), float(v))
                for c, v in zip(coords[top].cpu().numpy(), failed[top].cpu().numpy())]

    def _write_log(self, simulation, results):
        os.makedirs(self.outdir, exist_ok=True)
        path = os.path.join(self.outdir, f"{self.name}_reporter.log")
        with open(path, "w") as fh:
            fh.write(f"(!) {self.name} detected in step {simulation.flow.i} "
                     f"at following locations (top {self.k} listed):\n")
            fh.write("     ")
            for loc in self.locations_string(simulation):
                fh.write(f"{loc:6}   ")
            fh.write("\n")
            for pos, val in results:
                line = "".join(f"{int(p):6}   " for p in pos)
                fh.write(f"{line:<20} | {val:<15.6f}\n")
            fh.write("\n")

    def save_frame(self, simulation):
        flow = simulation.flow
        np.savez(os.path.join(self.outdir, f"{self.name}_frame_{flow.i:08d}.npz"),
                 p=flow.context.convert_to_ndarray(flow.p_pu),
                 u=flow.context.convert_to_ndarray(flow.u_pu))

    def _axes(self, simulation):
        return ["x", "y", "z"][:simulation.flow.stencil.d]

    @abstractmethod
    def locations_string(self, simulation) -> List[str]:
        ...

    @abstractmethod
    def is_failed(self, simulation) -> bool:
        ...

    @abstractmethod
    def get_results(self, simulation) -> List[Tuple]:
        ...


class NaNReporter(FailureReporterBase):
    """Stops at the first report step with a NaN population
    (failure_reporter.py:134-158)."""

    def __init__(self, interval: int, k: int = 100, outdir: Optional[str] = None,
                 vtk_out: bool = False):
        super().__init__(interval, k, outdir, vtk_out)
        self.name = "NaN"

    def is_failed(self, simulation) -> bool:
        return bool(_scan(simulation.flow)[0] > 0)

    def get_results(self, simulation):
        f = simulation.flow.f
        return self._get_top_failures(torch.isnan(f), f)

    def locations_string(self, simulation):
        return ["q"] + self._axes(simulation)


class HighMaReporter(FailureReporterBase):
    """Stops at the first report step where |u| / cs exceeds `threshold`
    somewhere (failure_reporter.py:161-184); a non-finite velocity counts
    as exceeding it."""

    def __init__(self, interval: int, threshold: float = 0.3, k: int = 100,
                 outdir: Optional[str] = None, vtk_out: bool = False):
        super().__init__(interval, k, outdir, vtk_out)
        self.threshold = threshold
        self.name = "HighMa"

    def is_failed(self, simulation) -> bool:
        nan_pops, bad_u, uu_max = _scan(simulation.flow).tolist()
        cs = simulation.flow.stencil.cs
        # bad_u counts cells whose |u|^2 is not finite (an inf population, a
        # zero density), which need not hold a NaN population
        return bool(nan_pops > 0 or bad_u > 0 or uu_max ** 0.5 / cs > self.threshold)

    def get_results(self, simulation):
        flow = simulation.flow
        ma = torch.linalg.vector_norm(flow.u(), dim=0) / flow.stencil.cs
        return self._get_top_failures(~(ma <= self.threshold), ma)

    def locations_string(self, simulation):
        return self._axes(simulation) + ["Ma"]

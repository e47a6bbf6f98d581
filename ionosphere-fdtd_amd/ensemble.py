"""Ensembles of independent GeodesicFDTD runs advanced by shared launches.

The reference advances one ``GeodesicFDTD`` per process (solver.py:789-816);
an ensemble of runs (different ionosphere profiles, ground conductivity,
sources, receivers, radial grids and therefore time steps) is a loop over
those objects.  Here the runs stay ordinary ``GeodesicFDTD`` objects (their
fields, checkpoints and diagnostics work as usual), and ``EnsembleFDTD``
advances all of them with one H and one E kernel launch per step (HIP graphs
of ``compile_chunk_size`` steps).  Each run's results are bitwise those of
stepping it alone.

The column-block shape -- 16-byte strips per thread (1, 2, 4) and the block
order (run-major or interleaved) -- is chosen by the library's autotuner,
which times every shape on the ensemble itself and restores the state.
"""

from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np
from numpy.typing import NDArray

from . import _native
from .solver import GeodesicFDTD


class EnsembleFDTD:
    """Advance several ``GeodesicFDTD`` runs together on one GPU."""

    def __init__(self, runs: Sequence[GeodesicFDTD], *, compile_chunk_size: int = 100,
                 autotune: bool = True, tune_steps: int = 10) -> None:
        runs = list(runs)
        if not runs:
            raise ValueError("an ensemble needs at least one run")
        if not all(isinstance(r, GeodesicFDTD) for r in runs):
            raise TypeError("runs must be GeodesicFDTD instances")
        if len({id(r) for r in runs}) != len(runs):
            raise ValueError("an ensemble may hold each run only once")
        if len({r.backend.dtype_name for r in runs}) != 1:
            raise ValueError("runs must share one storage dtype")
        if len({str(r.backend.torch_device) for r in runs}) != 1:
            raise ValueError("runs must share one device")
        if (isinstance(compile_chunk_size, bool) or int(compile_chunk_size) != compile_chunk_size
                or compile_chunk_size < 1):
            raise ValueError("compile_chunk_size must be a positive integer")
        self.runs = runs
        self.compile_chunk_size = int(compile_chunk_size)
        self._lib = runs[0]._lib
        plans = (C.c_void_p * len(runs))(*[r._plan for r in runs])
        handle = C.c_void_p()
        _native.check(self._lib.geo_ensemble_create(plans, len(runs), C.byref(handle)),
                      self._lib)
        self._handle = handle
        self.shape = {"strips_per_thread": 1, "block_order": 0, "ms_per_step": None}
        if autotune:
            self.tune(tune_steps)

    def __del__(self) -> None:
        handle = getattr(self, "_handle", None)
        if handle is not None and getattr(self, "_lib", None) is not None:
            try:
                self.runs[0].backend.synchronize()
            except Exception:
                pass
            self._lib.geo_ensemble_destroy(handle)
            self._handle = None

    @property
    def n_runs(self) -> int:
        return len(self.runs)

    @property
    def cells(self) -> int:
        """Dual-cell x radial-interval cells over all runs."""
        return int(sum(r.mesh.n_vertices * len(r.radial_midpoints_m) for r in self.runs))

    def _stream(self):
        return self.runs[0]._stream()

    def configure(self, strips_per_thread: int, block_order: int) -> None:
        """Column-block shape: 1, 2 or 4 strips per thread; block order 0
        (run-major) or 1 (runs interleaved), plus 2 for every run on the flat
        thread mapping (one strip per thread only)."""
        _native.check(self._lib.geo_ensemble_configure(self._handle, int(strips_per_thread),
                                                       int(block_order)), self._lib)
        self.shape = {"strips_per_thread": int(strips_per_thread),
                      "block_order": int(block_order), "ms_per_step": None}

    def tune(self, steps: int = 10) -> dict:
        """Time every column-block shape on this ensemble and keep the fastest."""
        s, order, ms = C.c_int32(), C.c_int32(), C.c_double()
        _native.check(self._lib.geo_ensemble_tune(self._handle, int(steps), self._stream(),
                                                  C.byref(s), C.byref(order), C.byref(ms)),
                      self._lib)
        self.shape = {"strips_per_thread": s.value, "block_order": order.value,
                      "ms_per_step": ms.value}
        return dict(self.shape)

    def _currents(self, count: int):
        """(n_runs, count) source currents, each run at its own time step."""
        per_run = [r._source_currents(count) for r in self.runs]
        if all(c is None for c in per_run):
            return None
        torch = self.runs[0].backend.torch
        out = torch.zeros((self.n_runs, count), dtype=self.runs[0].backend.compute_dtype,
                          device=self.runs[0].backend.torch_device)
        for i, c in enumerate(per_run):
            if c is not None:
                out[i] = c
        return out

    def _advance(self, count: int, traces) -> None:
        for r in self.runs:
            r._sync_step_index(count)
        currents = self._currents(count)
        _native.check(self._lib.geo_ensemble_step(
            self._handle, None if currents is None else currents.data_ptr(), count,
            self.compile_chunk_size, None if traces is None else traces.data_ptr(),
            self._stream()), self._lib)
        self._pending = currents
        for r in self.runs:
            r.steps += count
            r.time_s = r.steps * r.time_step_s

    def step(self, count: int = 1) -> None:
        """Advance every run by ``count`` leapfrog steps."""
        count = self.runs[0]._validated_count(count, "step count", minimum=0)
        if count:
            self._advance(count, None)

    def record_er_observations(self, receivers, steps: int) -> NDArray[np.generic]:
        """Advance every run and record its receivers (solver.py:1080-1136 per run).

        ``receivers`` holds one ``(vertex_indices, radial_layers, weights)``
        triple per run, all with the same receiver count.  Returns
        ``(n_runs, steps + 1, n_receivers)``: row 0 is the initial field.
        """
        receivers = list(receivers)
        if len(receivers) != self.n_runs:
            raise ValueError("give one receiver set per run")
        steps = self.runs[0]._validated_count(steps, "step count", minimum=0)
        counts = {r._set_receivers(*rx) for r, rx in zip(self.runs, receivers)}
        if len(counts) != 1 or 0 in counts:
            raise ValueError("every run needs the same, nonzero number of receivers")
        n_rx = counts.pop()
        torch = self.runs[0].backend.torch
        traces = torch.zeros((self.n_runs, steps + 1, n_rx),
                             dtype=self.runs[0].backend.compute_dtype,
                             device=self.runs[0].backend.torch_device)
        self._advance(steps, traces)
        self.runs[0].backend.synchronize()
        return traces.cpu().numpy()

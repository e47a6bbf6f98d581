"""Volume-weighted field, energy and conductive-loss diagnostics on the device.

Mirrors the reference's ``verification/physics_diagnostics/model.py``
(``PhysicsDiagnosticSampler`` and ``PhysicsSnapshot``, model.py:22-47,
60-275, 372-450): the same scalar names and radial profiles, reduced in FP64
by the library (``geo_diagnostics``, ``geo_radial_profiles``) so that only
scalars and per-layer profiles leave the GPU.  Non-finite field values are
counted and skipped by the device reductions; a field holding any reports
``field_finite`` 0 and a NaN RMS and maximum, as the reference's
whole-array reductions would.

Horizontal regions (per-entity support weights) are not provided by the
device sampler.
"""

from __future__ import annotations

import math
from collections.abc import Mapping
from dataclasses import dataclass

import numpy as np

_FIELDS = ("er", "et", "hr", "ht")


@dataclass(frozen=True, slots=True)
class PhysicsSnapshot:
    """One read-only diagnostic sample at the current leapfrog state
    (model.py:32-47)."""

    step: int
    electric_time_s: float
    magnetic_time_s: float
    scalars: dict[str, float]
    receiver_values: dict[str, float]
    radial_energy_er_j: np.ndarray
    radial_energy_et_j: np.ndarray
    radial_energy_hr_j: np.ndarray
    radial_energy_ht_j: np.ndarray
    radial_loss_er_w: np.ndarray
    radial_loss_et_w: np.ndarray


class PhysicsDiagnosticSampler:
    """Discrete volume-weighted diagnostics of a ``GeodesicFDTD`` run
    (model.py:60-275).  Sampling reads the device state and leaves it
    unchanged."""

    def __init__(self, simulation, *, horizontal_regions: Mapping | None = None) -> None:
        if horizontal_regions:
            raise ValueError("horizontal regions are not supported by the device sampler")
        self.simulation = simulation
        reference_height = getattr(simulation.material, "ionosphere_reference_height_m",
                                   70_000.0)
        self.reference_height_m = float(reference_height)
        # the library keeps its FP64 weight and material tables on the device
        self.diagnostic_backend_bytes = 0

    def sample(self, receiver_values: Mapping[str, float] | None = None, *,
               steps_per_second: float | None = None) -> PhysicsSnapshot:
        """Reduce the current fields (model.py:136-230)."""
        sim = self.simulation
        d = sim.field_diagnostics()
        profiles = sim.radial_profiles()
        counts = {"er": sim.mesh.n_vertices * len(sim.radii_m),
                  "et": sim.mesh.n_edges * len(sim.radial_midpoints_m),
                  "hr": sim.mesh.n_faces * len(sim.radial_midpoints_m),
                  "ht": sim.mesh.n_edges * len(sim.radii_m)}
        scalars: dict[str, float] = {}
        for name in _FIELDS:
            finite = d[f"nonfinite_{name}"] == 0
            scalars[f"field_rms/{name}"] = (math.sqrt(d[f"sumsq_{name}"] / counts[name])
                                            if finite else math.nan)
            scalars[f"field_max_abs/{name}"] = d[f"max_abs_{name}"] if finite else math.nan
            scalars[f"field_finite/{name}"] = float(finite)
        energy = {name: profiles[f"energy_{name}"] for name in _FIELDS}
        loss = {"er": profiles["loss_er"], "et": profiles["loss_et"]}
        for name, profile in energy.items():
            scalars[f"energy/{name}_j"] = float(np.sum(profile))
        scalars["energy/electric_j"] = scalars["energy/er_j"] + scalars["energy/et_j"]
        scalars["energy/magnetic_j"] = scalars["energy/hr_j"] + scalars["energy/ht_j"]
        scalars["energy/total_staggered_j"] = (scalars["energy/electric_j"]
                                               + scalars["energy/magnetic_j"])
        for name, profile in loss.items():
            scalars[f"conductive_loss/{name}_w"] = float(np.sum(profile))
        scalars["conductive_loss/total_w"] = (scalars["conductive_loss/er_w"]
                                              + scalars["conductive_loss/et_w"])
        self._add_region_scalars(scalars, energy, loss)
        scalars.update(self._source_scalars())
        scalars["memory/diagnostic_backend_bytes"] = float(self.diagnostic_backend_bytes)
        scalars["time/electric_s"] = sim.electric_time_s
        scalars["time/magnetic_s"] = sim.magnetic_time_s
        scalars["performance/steps_per_second"] = float(
            steps_per_second if steps_per_second is not None else 0.0)
        scalars.update(self._accelerator_scalars())
        return PhysicsSnapshot(
            step=sim.steps, electric_time_s=sim.electric_time_s,
            magnetic_time_s=sim.magnetic_time_s, scalars=scalars,
            receiver_values={str(k): float(v) for k, v in (receiver_values or {}).items()},
            radial_energy_er_j=energy["er"], radial_energy_et_j=energy["et"],
            radial_energy_hr_j=energy["hr"], radial_energy_ht_j=energy["ht"],
            radial_loss_er_w=loss["er"], radial_loss_et_w=loss["et"])

    def _add_region_scalars(self, scalars, energy, loss) -> None:
        """Earth / atmosphere / ionosphere split of the profiles (model.py:372-401)."""
        sim = self.simulation
        nodes = sim.altitudes_m
        cells = sim.radial_midpoint_altitudes_m
        h = self.reference_height_m
        regions = {"earth": (nodes < 0.0, cells < 0.0),
                   "atmosphere": ((nodes >= 0.0) & (nodes < h), (cells >= 0.0) & (cells < h)),
                   "ionosphere": (nodes >= h, cells >= h)}
        for region, (node_mask, cell_mask) in regions.items():
            scalars[f"energy_region/{region}_j"] = float(
                np.sum(energy["er"][node_mask]) + np.sum(energy["ht"][node_mask])
                + np.sum(energy["et"][cell_mask]) + np.sum(energy["hr"][cell_mask]))
            scalars[f"conductive_loss_region/{region}_w"] = float(
                np.sum(loss["er"][node_mask]) + np.sum(loss["et"][cell_mask]))

    def _source_scalars(self) -> dict[str, float]:
        """Source current at the half steps around the sample (model.py:403-436)."""
        sim = self.simulation
        source = sim.source
        if source is None:
            return {"source/current_previous_a": 0.0, "source/current_next_a": 0.0}
        dt = sim.time_step_s
        previous = source.current_a((sim.steps - 0.5) * dt, dt) if sim.steps else 0.0
        following = source.current_a((sim.steps + 0.5) * dt, dt)
        out = {"source/current_previous_a": float(previous),
               "source/current_next_a": float(following)}
        length = getattr(source, "vertical_element_length_m", None)
        if length is not None:
            out["source/current_moment_previous_a_m"] = float(previous * length)
            out["source/current_moment_next_a_m"] = float(following * length)
        return out

    def _accelerator_scalars(self) -> dict[str, float]:
        """Device memory counters (model.py:438-450)."""
        torch = self.simulation.backend.torch
        device = self.simulation.backend.torch_device
        return {"memory/cuda_allocated_bytes": float(torch.cuda.memory_allocated(device)),
                "memory/cuda_reserved_bytes": float(torch.cuda.memory_reserved(device)),
                "memory/cuda_max_allocated_bytes": float(
                    torch.cuda.max_memory_allocated(device))}

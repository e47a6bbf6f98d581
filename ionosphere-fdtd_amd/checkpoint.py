"""Portable, versioned NPZ checkpoints of the device state.

The file format is the reference's (checkpoint.py:23-93, 335-381): an NPZ
with a JSON ``metadata`` string (format ``ionosphere-fdtd-checkpoint``,
version 4), the mesh, the four fields, the surface-impedance memory as
``(edge, term)`` and, with plasma, the model arrays and one Cartesian
``(face, radial, 3)`` current per species.  Files therefore move in both
directions between the two implementations.  FP16/BF16 fields are widened to
FP32 on save.  Writes are atomic (temporary file, fsync, rename).
"""

from __future__ import annotations

from dataclasses import asdict, is_dataclass, replace
import json
import os
from pathlib import Path
import tempfile
from typing import Any

import numpy as np

from .materials import EarthIonosphereMaterial
from .mesh import build_geodesic_mesh_from_topology, subdivided_icosahedron
from .plasma import ColdPlasmaSpecies, MeshPlasmaModel, array_sha256
from .sources import GaussianCurrent, TangentialGaussianCurrent
from .surface_impedance import ConductiveHalfSpaceSurface

CHECKPOINT_FORMAT = "ionosphere-fdtd-checkpoint"
CHECKPOINT_VERSION = 4
SUPPORTED_CHECKPOINT_VERSIONS = (1, 2, 3, 4)
MESH_PLASMA_FORMAT = "ionosphere-fdtd-mesh-plasma"


class CheckpointError(ValueError):
    """Raised when a checkpoint is unsupported, corrupt, or inconsistent."""


def _json_default(value: Any) -> Any:
    if isinstance(value, np.generic):
        return value.item()
    if isinstance(value, np.ndarray):
        return value.tolist()
    raise TypeError(f"cannot encode {type(value).__name__} in checkpoint metadata")


def _plasma_payload(model: MeshPlasmaModel) -> tuple[dict, dict]:
    arrays = {"radial_midpoint_altitudes_m": model.radial_midpoint_altitudes_m,
              "magnetic_field_t": model.magnetic_field_t}
    for i, sp in enumerate(model.species):
        arrays[f"species_{i}_number_density_m3"] = sp.number_density_m3
        arrays[f"species_{i}_collision_frequency_hz"] = sp.collision_frequency_hz
    provenance = [asdict(p) if is_dataclass(p) else dict(p) for p in model.provenance]
    metadata = {
        "format": MESH_PLASMA_FORMAT, "version": 1,
        "mesh": {"vertices_sha256": model.mesh_vertices_sha256,
                 "faces_sha256": model.mesh_faces_sha256},
        "species": [{"name": s.name, "charge_c": s.charge_c, "mass_kg": s.mass_kg}
                    for s in model.species],
        "provenance": provenance,
        "interpolation": model.interpolation,
        "arrays": {k: {"shape": list(np.asarray(v).shape), "sha256": array_sha256(v)}
                   for k, v in arrays.items()},
    }
    return metadata, arrays


def _plasma_from_payload(metadata: dict, arrays: dict) -> MeshPlasmaModel:
    if metadata.get("format") != MESH_PLASMA_FORMAT or metadata.get("version") != 1:
        raise CheckpointError("unsupported plasma artifact format")
    for name, record in metadata["arrays"].items():
        if name not in arrays:
            raise CheckpointError(f"plasma artifact is missing {name}")
        if record.get("shape") != list(arrays[name].shape):
            raise CheckpointError(f"plasma artifact {name} shape mismatch")
        if record.get("sha256") != array_sha256(arrays[name]):
            raise CheckpointError(f"plasma artifact {name} checksum mismatch")
    species = tuple(
        ColdPlasmaSpecies(name=r["name"], charge_c=r["charge_c"], mass_kg=r["mass_kg"],
                          number_density_m3=arrays[f"species_{i}_number_density_m3"],
                          collision_frequency_hz=arrays[f"species_{i}_collision_frequency_hz"])
        for i, r in enumerate(metadata["species"]))
    return MeshPlasmaModel(mesh_vertices_sha256=metadata["mesh"]["vertices_sha256"],
                           mesh_faces_sha256=metadata["mesh"]["faces_sha256"],
                           radial_midpoint_altitudes_m=arrays["radial_midpoint_altitudes_m"],
                           magnetic_field_t=arrays["magnetic_field_t"], species=species,
                           provenance=tuple(metadata.get("provenance", ())),
                           interpolation=metadata["interpolation"])


def _metadata(sim) -> dict:
    if not isinstance(sim.material, EarthIonosphereMaterial):
        raise CheckpointError("checkpoints currently support EarthIonosphereMaterial only")
    if sim.source is not None and not isinstance(sim.source, (GaussianCurrent,
                                                              TangentialGaussianCurrent)):
        raise CheckpointError("checkpoints support GaussianCurrent and "
                              "TangentialGaussianCurrent sources only")
    material = asdict(sim.material)
    material["anomalies"] = []
    return {
        "format": CHECKPOINT_FORMAT,
        "version": CHECKPOINT_VERSION,
        "simulation_config": asdict(sim.config),
        "mesh": {"topology_kind": sim.mesh.topology_kind, "subdivision": sim.mesh.subdivision,
                 "has_face_levels": sim.mesh.face_levels is not None,
                 "refinement_spec": sim.mesh.refinement_spec},
        "material": {"type": "EarthIonosphereMaterial", "parameters": material},
        "source": (None if sim.source is None
                   else {"type": type(sim.source).__name__, "parameters": asdict(sim.source)}),
        "surface_impedance": (sim.surface_impedance.to_metadata()
                              if sim.surface_impedance is not None else None),
        "plasma": _plasma_payload(sim.plasma)[0] if sim.plasma is not None else None,
        "state": {"steps": sim.steps, "time_s": sim.time_s, "time_step_s": sim.time_step_s},
        "runtime": {"backend": sim.backend.name, "device": sim.backend.device,
                    "dtype": sim.backend.dtype_name, "compiled": sim.compiled,
                    "compile_chunk_size": sim.compile_chunk_size},
    }


def _host(sim, values) -> np.ndarray:
    array = sim.to_numpy(values)
    if array.dtype == np.float16:
        array = array.astype(np.float32)
    return np.ascontiguousarray(array)


def save_checkpoint(sim, path: str | Path) -> Path:
    """Atomically write a version-4 checkpoint of ``sim`` and return its path."""
    destination = Path(path)
    metadata = _metadata(sim)
    destination.parent.mkdir(parents=True, exist_ok=True)
    arrays = {
        "metadata": np.asarray(json.dumps(metadata, default=_json_default, sort_keys=True)),
        "mesh_vertices": np.asarray(sim.mesh.vertices, dtype=np.float64),
        "mesh_faces": np.asarray(sim.mesh.faces, dtype=np.int64),
        "mesh_face_levels": (np.asarray(sim.mesh.face_levels, dtype=np.int64)
                             if sim.mesh.face_levels is not None else np.empty(0, dtype=np.int64)),
        "er": _host(sim, sim.er), "et": _host(sim, sim.et), "hr": _host(sim, sim.hr),
        "ht": _host(sim, sim.ht),
        "surface_impedance_memory": (_host(sim, sim._si_memory) if sim._si_memory is not None
                                     else np.empty((0, 0), dtype=np.float64)),
    }
    if sim.plasma is not None:
        _, plasma_arrays = _plasma_payload(sim.plasma)
        arrays.update({f"plasma_model_{k}": v for k, v in plasma_arrays.items()})
        for i, current in enumerate(sim.plasma_current_density):
            arrays[f"plasma_current_{i}"] = _host(sim, current)
    tmp_name = None
    try:
        with tempfile.NamedTemporaryFile(mode="wb", dir=destination.parent,
                                         prefix=f".{destination.name}.", delete=False) as tmp:
            tmp_name = tmp.name
            np.savez_compressed(tmp, **arrays)
            tmp.flush()
            os.fsync(tmp.fileno())
        os.replace(tmp_name, destination)
    finally:
        if tmp_name is not None and os.path.exists(tmp_name):
            os.unlink(tmp_name)
    return destination


def read_checkpoint(path: str | Path) -> tuple[dict, dict]:
    """Read and validate the metadata and arrays of a checkpoint (no device work)."""
    source = Path(path)
    try:
        with np.load(source, allow_pickle=False) as archive:
            if "metadata" not in archive.files:
                raise CheckpointError("checkpoint is missing arrays: metadata")
            raw = archive["metadata"]
            if raw.shape != () or raw.dtype.kind not in {"U", "S"}:
                raise CheckpointError("checkpoint metadata must be a scalar string")
            metadata = json.loads(str(raw.item()))
            if not isinstance(metadata, dict):
                raise CheckpointError("checkpoint metadata must be a JSON object")
            if metadata.get("format") != CHECKPOINT_FORMAT:
                raise CheckpointError("file is not an ionosphere-fdtd checkpoint")
            version = metadata.get("version")
            if version not in SUPPORTED_CHECKPOINT_VERSIONS:
                raise CheckpointError(
                    f"unsupported checkpoint version {version!r}; supported versions are "
                    f"{SUPPORTED_CHECKPOINT_VERSIONS}")
            need_meta = {"simulation_config", "material", "source", "state", "runtime"}
            if version >= 2:
                need_meta.add("mesh")
            if version >= 3:
                need_meta.add("surface_impedance")
            if version >= 4:
                need_meta.add("plasma")
            missing = need_meta.difference(metadata)
            if missing:
                raise CheckpointError(f"checkpoint metadata is missing: {', '.join(sorted(missing))}")
            required = {"metadata", "mesh_vertices", "er", "et", "hr", "ht"}
            if version >= 2:
                required |= {"mesh_faces", "mesh_face_levels"}
            if version >= 3:
                required.add("surface_impedance_memory")
            plasma_meta = metadata.get("plasma") if version >= 4 else None
            if plasma_meta is not None:
                required |= {f"plasma_model_{k}" for k in plasma_meta["arrays"]}
                required |= {f"plasma_current_{i}" for i in range(len(plasma_meta["species"]))}
            missing = required.difference(archive.files)
            if missing:
                raise CheckpointError(f"checkpoint is missing arrays: {', '.join(sorted(missing))}")
            arrays = {name: np.array(archive[name], copy=True) for name in archive.files
                      if name != "metadata"}
    except CheckpointError:
        raise
    except (OSError, ValueError, KeyError, json.JSONDecodeError) as error:
        raise CheckpointError(f"cannot read checkpoint {source}: {error}") from error
    return metadata, arrays


def load_checkpoint(path: str | Path, *, backend: str = "hip", device: str = "auto",
                    dtype: str | None = None, compile_step: bool = False,
                    compile_chunk_size: int = 8, torch_threads: int | None = None):
    """Restore a checkpoint (from this package or the reference) on the HIP device."""
    from .solver import GeodesicFDTD, SimulationConfig

    metadata, arrays = read_checkpoint(path)
    version = metadata["version"]
    try:
        cfg_values = dict(metadata["simulation_config"])
        if cfg_values.get("radial_altitudes_m") is not None:
            cfg_values["radial_altitudes_m"] = tuple(cfg_values["radial_altitudes_m"])
        config = SimulationConfig(**cfg_values)
        vertices = np.asarray(arrays["mesh_vertices"], dtype=np.float64)
        if version == 1:
            faces = subdivided_icosahedron(config.subdivision, config.mesh_orientation)[1]
            levels = None
            kind, spec, subdivision = "uniform", None, config.subdivision
        else:
            mesh_meta = metadata["mesh"]
            faces = arrays["mesh_faces"]
            has_levels = mesh_meta["has_face_levels"]
            if not isinstance(has_levels, bool):
                raise CheckpointError("checkpoint has_face_levels must be boolean")
            levels = arrays["mesh_face_levels"] if has_levels else None
            if not has_levels and arrays["mesh_face_levels"].shape != (0,):
                raise CheckpointError("checkpoint has unexpected face-level values")
            kind, spec, subdivision = (mesh_meta["topology_kind"], mesh_meta["refinement_spec"],
                                       mesh_meta["subdivision"])
        mesh = build_geodesic_mesh_from_topology(vertices, faces, subdivision=subdivision,
                                                 face_levels=levels, refinement_spec=spec,
                                                 topology_kind=kind, normalize_vertices=False)
        mat = metadata["material"]
        if not isinstance(mat, dict) or mat.get("type") != "EarthIonosphereMaterial":
            raise CheckpointError("checkpoint material type is unsupported")
        params = dict(mat.get("parameters", {}))
        if params.get("anomalies"):
            raise CheckpointError("checkpoint material anomalies are not supported")
        params["anomalies"] = ()
        material = EarthIonosphereMaterial(**params)
        source = None
        if metadata["source"] is not None:
            kinds = {"GaussianCurrent": GaussianCurrent,
                     "TangentialGaussianCurrent": TangentialGaussianCurrent}
            values = dict(metadata["source"].get("parameters", {}))
            for name in ("azimuths_deg", "line_lengths_m"):
                if values.get(name) is not None:
                    values[name] = tuple(values[name])
            source = kinds[metadata["source"]["type"]](**values)
        surface = (ConductiveHalfSpaceSurface.from_metadata(metadata["surface_impedance"])
                   if version >= 3 and metadata["surface_impedance"] is not None else None)
        plasma = None
        if version >= 4 and metadata["plasma"] is not None:
            plasma = _plasma_from_payload(metadata["plasma"], {
                k: arrays[f"plasma_model_{k}"] for k in metadata["plasma"]["arrays"]})
    except CheckpointError:
        raise
    except (KeyError, TypeError, ValueError, RuntimeError) as error:
        raise CheckpointError(f"invalid checkpoint model metadata: {error}") from error
    selected = dtype or str(metadata["runtime"]["dtype"])
    sim = GeodesicFDTD(replace(config, mesh_relaxations=0, mesh_optimization_steps=0),
                       material=material, source=source, surface_impedance=surface,
                       plasma=plasma, mesh=mesh, backend=backend, device=device,
                       dtype=selected, compile_step=compile_step,
                       compile_chunk_size=compile_chunk_size, torch_threads=torch_threads)
    sim.config = config
    for name in ("er", "et", "hr", "ht"):
        values = arrays[name]
        expected = tuple(getattr(sim, name).shape)
        if values.shape != expected:
            raise CheckpointError(f"checkpoint field {name} has shape {values.shape}, "
                                  f"expected {expected}")
        if not np.issubdtype(values.dtype, np.floating):
            raise CheckpointError(f"checkpoint field {name} must be floating point")
        if not np.all(np.isfinite(values)):
            raise CheckpointError(f"checkpoint field {name} contains non-finite values")
        setattr(sim, name, values)
    memory = arrays.get("surface_impedance_memory", np.empty((0, 0)))
    if sim._si_memory is None:
        if memory.shape != (0, 0):
            raise CheckpointError("checkpoint has ADE state without a surface impedance model")
    else:
        if memory.shape != tuple(sim._si_memory.shape):
            raise CheckpointError("checkpoint surface impedance state has shape "
                                  f"{memory.shape}, expected {tuple(sim._si_memory.shape)}")
        if not np.all(np.isfinite(memory)):
            raise CheckpointError("checkpoint surface impedance state is invalid")
        sim._surface_impedance_ade.memory = memory
    if plasma is not None:
        for i in range(len(plasma.species)):
            current = arrays[f"plasma_current_{i}"]
            if current.shape != (sim.mesh.n_faces, len(sim.radial_midpoints_m), 3):
                raise CheckpointError(f"checkpoint plasma current {i} has shape {current.shape}")
            if not np.all(np.isfinite(current)):
                raise CheckpointError(f"checkpoint plasma current {i} is invalid")
            sim.set_plasma_current_density(i, current)
    steps = metadata["state"]["steps"]
    if isinstance(steps, bool) or not isinstance(steps, int) or steps < 0:
        raise CheckpointError("checkpoint step count must be a non-negative integer")
    saved_dt = float(metadata["state"]["time_step_s"])
    if not np.isclose(saved_dt, sim.time_step_s, rtol=0.0, atol=1.0e-15):
        raise CheckpointError("checkpoint time step is inconsistent with the reconstructed model")
    saved_time = float(metadata["state"]["time_s"])
    if not np.isfinite(saved_time) or not np.isclose(saved_time, steps * saved_dt, rtol=1e-12,
                                                     atol=1e-15):
        raise CheckpointError("checkpoint time is inconsistent with its step count")
    sim.steps = steps
    sim.time_s = steps * saved_dt
    return sim

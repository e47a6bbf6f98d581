"""Cold-plasma species on the mesh and the host tables of the current ADE.

``ColdPlasmaSpecies`` / ``MeshPlasmaModel`` keep the reference's constructor
fields and validation (plasma.py:32-152) so models built for the reference
drop in; provenance is optional here.  ``plasma_tables`` builds, vectorised,
what the reference's ``GeodesicPlasmaCoupler`` and ``MagnetizedPlasmaADE``
precompute (plasma.py:374-411, 458-487, 548-588): the per-face
reconstruction pinv(tangents).T, the left/right edge tangents, the
area-weighted vertex-face averages and the six exact-integration
coefficients per species.
"""

from __future__ import annotations

from dataclasses import dataclass, field
import hashlib
from typing import Any

import numpy as np
from numpy.typing import NDArray

from .constants import EPSILON_0

ELEMENTARY_CHARGE_C = 1.602_176_634e-19
ELECTRON_MASS_KG = 9.109_383_713_9e-31


def _readonly(values) -> NDArray[np.float64]:
    result = np.array(values, dtype=np.float64, order="C", copy=True)
    result.setflags(write=False)
    return result


def array_sha256(values) -> str:
    """sha256 of the array as little-endian C-order f8 (floats) or i8 (integers).

    The same canonical form as the reference's data_artifacts.py:457-467, so a
    plasma model built against a mesh there validates against the identical
    mesh here.
    """
    array = np.asarray(values)
    if np.issubdtype(array.dtype, np.floating):
        canonical = np.asarray(array, dtype="<f8", order="C")
    elif np.issubdtype(array.dtype, np.integer):
        canonical = np.asarray(array, dtype="<i8", order="C")
    else:
        raise ValueError("only numeric arrays can be checksummed")
    return hashlib.sha256(canonical.tobytes()).hexdigest()


@dataclass(frozen=True, slots=True)
class ColdPlasmaSpecies:
    """One charged fluid sampled at (face, radial midpoint) centres."""

    name: str
    charge_c: float
    mass_kg: float
    number_density_m3: NDArray[np.float64]
    collision_frequency_hz: NDArray[np.float64]

    def __post_init__(self) -> None:
        if not isinstance(self.name, str) or not self.name.strip():
            raise ValueError("plasma species name must be nonempty")
        if not np.isfinite(self.charge_c) or self.charge_c == 0.0:
            raise ValueError("plasma species charge must be finite and nonzero")
        if not np.isfinite(self.mass_kg) or self.mass_kg <= 0.0:
            raise ValueError("plasma species mass must be finite and positive")
        density = np.asarray(self.number_density_m3, dtype=np.float64)
        collision = np.asarray(self.collision_frequency_hz, dtype=np.float64)
        if density.ndim != 2 or collision.shape != density.shape:
            raise ValueError("plasma density and collision grids must have one shape")
        if (not np.all(np.isfinite(density)) or np.any(density < 0.0)
                or not np.all(np.isfinite(collision)) or np.any(collision <= 0.0)):
            raise ValueError("plasma density/collision values are invalid")
        object.__setattr__(self, "number_density_m3", _readonly(density))
        object.__setattr__(self, "collision_frequency_hz", _readonly(collision))


@dataclass(frozen=True, slots=True)
class MeshPlasmaModel:
    """Magnetic field and charged fluids on one solver mesh and radial grid."""

    mesh_vertices_sha256: str
    mesh_faces_sha256: str
    radial_midpoint_altitudes_m: NDArray[np.float64]
    magnetic_field_t: NDArray[np.float64]
    species: tuple[ColdPlasmaSpecies, ...]
    provenance: tuple = field(default_factory=tuple)
    interpolation: str = "exact"

    def __post_init__(self) -> None:
        altitudes = np.asarray(self.radial_midpoint_altitudes_m, dtype=np.float64)
        magnetic = np.asarray(self.magnetic_field_t, dtype=np.float64)
        species = tuple(self.species)
        if (altitudes.ndim != 1 or len(altitudes) < 1 or not np.all(np.isfinite(altitudes))
                or not np.all(np.diff(altitudes) > 0.0)):
            raise ValueError("plasma radial midpoints must be finite and increasing")
        if magnetic.ndim != 3 or magnetic.shape[1:] != (len(altitudes), 3):
            raise ValueError("magnetic field must have shape (face, radial, 3)")
        if not np.all(np.isfinite(magnetic)):
            raise ValueError("magnetic field must be finite")
        if not species or not all(isinstance(s, ColdPlasmaSpecies) for s in species):
            raise ValueError("plasma model requires charged species")
        if any(s.number_density_m3.shape != magnetic.shape[:2] for s in species):
            raise ValueError("plasma species grids must match the magnetic grid")
        names = [s.name for s in species]
        if len(set(names)) != len(names):
            raise ValueError("plasma species names must be unique")
        if not isinstance(self.interpolation, str) or not self.interpolation.strip():
            raise ValueError("plasma interpolation policy must be nonempty")
        object.__setattr__(self, "radial_midpoint_altitudes_m", _readonly(altitudes))
        object.__setattr__(self, "magnetic_field_t", _readonly(magnetic))
        object.__setattr__(self, "species", species)
        object.__setattr__(self, "provenance", tuple(self.provenance))

    @classmethod
    def from_mesh(cls, mesh, radial_midpoint_altitudes_m, magnetic_field_t, species, *,
                  provenance=(), interpolation: str = "exact") -> "MeshPlasmaModel":
        return cls(mesh_vertices_sha256=array_sha256(np.asarray(mesh.vertices)),
                   mesh_faces_sha256=array_sha256(np.asarray(mesh.faces)),
                   radial_midpoint_altitudes_m=radial_midpoint_altitudes_m,
                   magnetic_field_t=magnetic_field_t, species=tuple(species),
                   provenance=tuple(provenance), interpolation=interpolation)

    def validate_grid(self, mesh, radial_midpoint_altitudes_m) -> None:
        if array_sha256(np.asarray(mesh.vertices)) != self.mesh_vertices_sha256:
            raise ValueError("plasma model mesh vertices do not match")
        if array_sha256(np.asarray(mesh.faces)) != self.mesh_faces_sha256:
            raise ValueError("plasma model mesh faces do not match")
        if not np.array_equal(radial_midpoint_altitudes_m, self.radial_midpoint_altitudes_m):
            raise ValueError("plasma model radial grid does not match")
        if self.magnetic_field_t.shape[0] != mesh.n_faces:
            raise ValueError("plasma model face count does not match")

    def conductivity_components(self, frequency_hz: float):
        """Summed sigma_parallel, sigma_Pedersen, sigma_Hall (exp(+i w t))."""
        if not np.isfinite(frequency_hz) or frequency_hz < 0.0:
            raise ValueError("plasma response frequency must be nonnegative")
        shape = self.magnetic_field_t.shape[:2]
        par = np.zeros(shape, dtype=np.complex128)
        ped = np.zeros_like(par)
        hall = np.zeros_like(par)
        bmag = np.linalg.norm(self.magnetic_field_t, axis=2)
        w = 2.0 * np.pi * frequency_hz
        for s in self.species:
            a = s.collision_frequency_hz + 1j * w
            gyro = s.charge_c * bmag / s.mass_kg
            drive = s.number_density_m3 * s.charge_c**2 / s.mass_kg
            den = a**2 + gyro**2
            par += drive / a
            ped += drive * a / den
            hall += drive * gyro / den
        return par, ped, hall

    def maximum_plasma_drive(self) -> float:
        """max over cells of sum_s n q^2 / m (rad^2/s^2 times eps0)."""
        drive = np.zeros(self.magnetic_field_t.shape[:2])
        for s in self.species:
            drive += s.number_density_m3 * s.charge_c**2 / s.mass_kg
        return float(np.max(drive))


def electron_drude_model(mesh, radial_midpoint_altitudes_m, density_m3, collision_hz):
    """Single-species electron model with B = 0 (Drude form) from (F,R) or (R,) grids."""
    alt = np.asarray(radial_midpoint_altitudes_m, dtype=np.float64)
    shape = (mesh.n_faces, len(alt))
    density = np.broadcast_to(np.asarray(density_m3, dtype=np.float64), shape)
    collision = np.broadcast_to(np.asarray(collision_hz, dtype=np.float64), shape)
    species = ColdPlasmaSpecies("electron", -ELEMENTARY_CHARGE_C, ELECTRON_MASS_KG,
                                density, collision)
    return MeshPlasmaModel.from_mesh(mesh, alt, np.zeros((*shape, 3)), (species,))


def _uniform_rows(values: NDArray[np.float64]) -> NDArray[np.float64]:
    """Collapse an (F, ...) array to one row when every row is identical."""
    if values.shape[0] > 1 and np.array_equal(values, np.broadcast_to(values[:1], values.shape)):
        return values[:1]
    return values


_GEOMETRY_CACHE: list[tuple[Any, dict[str, Any]]] = []


def _plasma_geometry(mesh) -> dict[str, Any]:
    """Mesh-only coupler tables (face reconstruction, edge tangent slots,
    vertex-face weights), kept for the last few meshes: every run on one mesh
    (ensembles, query services) shares them."""
    for cached_mesh, tables in _GEOMETRY_CACHE:
        if cached_mesh is mesh:
            return tables
    F, V, E = mesh.n_faces, mesh.n_vertices, mesh.n_edges
    # face reconstruction (reference plasma.py:548-559), batched pinv
    normals = np.cross(mesh.vertices[mesh.edges[:, 0]], mesh.vertices[mesh.edges[:, 1]])
    normals /= np.linalg.norm(normals, axis=1, keepdims=True)
    tangents = np.cross(normals[mesh.face_edges], mesh.face_centers[:, None, :])
    tangents /= np.linalg.norm(tangents, axis=2, keepdims=True)
    recon = np.transpose(np.linalg.pinv(tangents), (0, 2, 1))
    # left/right tangent slots (plasma.py:562-569)
    edge_ids = np.arange(E)[:, None]
    lslot = np.argmax(mesh.face_edges[mesh.edge_left_faces] == edge_ids, axis=1)
    rslot = np.argmax(mesh.face_edges[mesh.edge_right_faces] == edge_ids, axis=1)
    left_t = tangents[mesh.edge_left_faces, lslot]
    right_t = tangents[mesh.edge_right_faces, rslot]
    # vertex-face area weights (plasma.py:572-588): faces in increasing order
    flat_v = mesh.faces.reshape(-1)
    flat_f = np.repeat(np.arange(F), 3)
    order = np.lexsort((flat_f, flat_v))
    fv, ff = flat_v[order], flat_f[order]
    counts = np.bincount(fv, minlength=V)
    width = int(counts.max())
    starts = np.concatenate(([0], np.cumsum(counts)[:-1]))
    slot = np.arange(len(fv)) - np.repeat(starts, counts)
    vfaces = np.zeros((V, width), dtype=np.int64)
    vfaces[fv, slot] = ff
    areas = np.zeros((V, width))
    areas[fv, slot] = mesh.face_solid_angles[ff]
    vweights = areas / areas.sum(axis=1, keepdims=True)
    tables = {
        "reconstruction": np.ascontiguousarray(recon),
        "face_tangents": tangents,
        "left_tangents": np.ascontiguousarray(left_t),
        "right_tangents": np.ascontiguousarray(right_t),
        "vertex_faces": vfaces,
        "vertex_face_weights": vweights,
    }
    for value in tables.values():
        value.setflags(write=False)
    _GEOMETRY_CACHE.insert(0, (mesh, tables))
    del _GEOMETRY_CACHE[4:]
    return tables


def plasma_tables(model: MeshPlasmaModel, mesh, time_step_s: float) -> dict[str, Any]:
    """Host FP64 tables for the HIP plasma ADE and coupler."""
    geometry = _plasma_geometry(mesh)
    # ADE coefficients (plasma.py:385-411), on one face row when every input is
    # horizontally uniform (the same per-element arithmetic on fewer elements)
    b = np.asarray(model.magnetic_field_t)
    fields = [b] + [np.asarray(x) for s in model.species
                    for x in (s.collision_frequency_hz, s.number_density_m3)]
    uniform = all(np.array_equal(x, np.broadcast_to(x[:1], x.shape)) for x in fields)
    rows = (lambda x: np.asarray(x)[:1]) if uniform else np.asarray
    b = rows(b)
    bmag = np.linalg.norm(b, axis=2)
    bdir = np.zeros_like(b)
    nz = bmag > 0.0
    bdir[nz] = b[nz] / bmag[nz, None]
    dt = time_step_s
    coefs = []
    for s in model.species:
        nu = rows(s.collision_frequency_hz)
        gyro = s.charge_c * bmag / s.mass_kg
        eig = -nu + 1j * gyro
        integral = np.expm1(eig * dt) / eig
        parallel = -np.expm1(-nu * dt) / nu
        drive = rows(s.number_density_m3) * s.charge_c**2 / s.mass_kg
        coefs.append(np.stack((np.exp(-nu * dt), np.cos(gyro * dt), np.sin(gyro * dt),
                               drive * parallel, drive * integral.real, drive * integral.imag)))
    coef = np.stack(coefs)  # (species, 6, F or 1, R)
    # horizontally uniform rows collapse to one radial profile
    coef_rows = coef.transpose(2, 0, 1, 3)  # (F, species, 6, R)
    collapsed = _uniform_rows(coef_rows)
    coef_out = np.ascontiguousarray(collapsed.transpose(1, 2, 0, 3))
    bdir_out = np.ascontiguousarray(_uniform_rows(bdir))
    # band of midpoint layers with any drive
    drive_any = np.zeros(b.shape[1], dtype=bool)
    for s in model.species:
        drive_any |= np.any(rows(s.number_density_m3) > 0.0, axis=0)
    active = np.flatnonzero(drive_any)
    band = (int(active[0]), int(active[-1]) + 1) if len(active) else (0, 0)
    return {
        **geometry,
        "coef": coef_out,
        "bdir": bdir_out,
        "magnetized": bool(np.any(nz)),
        "band": band,
    }


def plasma_cfl_limit(model: MeshPlasmaModel, minimum_epsilon_r: float) -> float:
    """2 / omega_p,max (reference solver.py:392-406); inf when there is no plasma."""
    wp = np.sqrt(model.maximum_plasma_drive() / (EPSILON_0 * minimum_epsilon_r))
    return float("inf") if wp == 0.0 else 2.0 / wp

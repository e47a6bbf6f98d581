"""CPU oracle for the GeodesicFDTD time-step loop (test infrastructure only).

This module is the correctness oracle of the build.  It is a plain NumPy
restatement of the reference's array algorithm and is imported only by
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` -- never by the product package, which must run on the HIP
extension or fail.

Every function names the reference file:line it follows (paths relative to
the ``ruddyscent/ionosphere-fdtd`` source tree).  The restatement keeps the
reference's order of floating-point operations wherever that is cheap (mesh
loops, ``r*Et`` radial differences, ``np.add.at`` accumulation), so FP64
results agree with the reference to roundoff.

Parity pinning (what checks this module, see tests/test_oracle_*.py):
  * mesh integer arrays: bit-exact against fixtures written by
    tests/golden/make_fixtures.py from the reference itself;
  * CFL time steps: golden values in the reference's
    artifacts/analytic-solutions/full-field.csv (copied as data into
    tests/golden/full_field_dt.csv);
  * H/E curls + leapfrog: TM/TE single-layer eigenmode recurrences and the
    A2/A4 measured cavity frequencies of the same CSV;
  * conductive integration: exact stiff decay and the 2nd-order forcing ratio;
  * SI boundary, plasma ADE, sources, receivers and the whole C0-style run:
    end-to-end trajectories written by tests/golden/make_fixtures.py.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Sequence

import numpy as np

# constants.py:3-6
EPSILON_0 = 8.854_187_812_8e-12
MU_0 = 1.256_637_062_12e-6
C_0 = 1.0 / (EPSILON_0 * MU_0) ** 0.5
EARTH_RADIUS_M = 6_371_000.0
# plasma.py:26-27
ELEMENTARY_CHARGE_C = 1.602_176_634e-19
ELECTRON_MASS_KG = 9.109_383_713_9e-31


# ---------------------------------------------------------------------------
# Mesh (mesh.py:191-227, 386-789), loop-based like the reference.
# ---------------------------------------------------------------------------


@dataclass
class OracleMesh:
    vertices: np.ndarray
    faces: np.ndarray
    edges: np.ndarray
    face_edges: np.ndarray
    face_edge_signs: np.ndarray
    edge_left_faces: np.ndarray
    edge_right_faces: np.ndarray
    face_centers: np.ndarray
    primal_edge_angles: np.ndarray
    dual_edge_angles: np.ndarray
    face_solid_angles: np.ndarray
    dual_cell_solid_angles: np.ndarray
    vertex_degree: np.ndarray

    @property
    def n_vertices(self) -> int:
        return int(self.vertices.shape[0])

    @property
    def n_edges(self) -> int:
        return int(self.edges.shape[0])

    @property
    def n_faces(self) -> int:
        return int(self.faces.shape[0])

    def edge_midpoints(self) -> np.ndarray:  # mesh.py:89-91
        return _normalize(self.vertices[self.edges].sum(axis=1))

    # Reference operators, mesh.py:157-188 (scatter form).
    def face_circulation(self, edge_values: np.ndarray) -> np.ndarray:
        selected = edge_values[self.face_edges]
        signs = self.face_edge_signs.astype(edge_values.dtype, copy=False)
        while signs.ndim < selected.ndim:
            signs = signs[..., None]
        return np.sum(selected * signs, axis=1)

    def dual_cell_circulation(self, dual_edge_values: np.ndarray) -> np.ndarray:
        shape = (self.n_vertices,) + dual_edge_values.shape[1:]
        result = np.zeros(shape, dtype=dual_edge_values.dtype)
        np.add.at(result, self.edges[:, 0], dual_edge_values)
        np.add.at(result, self.edges[:, 1], -dual_edge_values)
        return result

    def edge_difference(self, vertex_values: np.ndarray) -> np.ndarray:
        return vertex_values[self.edges[:, 1]] - vertex_values[self.edges[:, 0]]

    def dual_edge_difference(self, face_values: np.ndarray) -> np.ndarray:
        return face_values[self.edge_left_faces] - face_values[self.edge_right_faces]


def _normalize(values: np.ndarray) -> np.ndarray:  # mesh.py:471-472
    return values / np.linalg.norm(values, axis=-1, keepdims=True)


def _icosahedron() -> tuple[np.ndarray, np.ndarray]:  # mesh.py:475-519
    phi = (1.0 + np.sqrt(5.0)) / 2.0
    vertices = np.asarray(
        [(-1, phi, 0), (1, phi, 0), (-1, -phi, 0), (1, -phi, 0),
         (0, -1, phi), (0, 1, phi), (0, -1, -phi), (0, 1, -phi),
         (phi, 0, -1), (phi, 0, 1), (-phi, 0, -1), (-phi, 0, 1)],
        dtype=np.float64,
    )
    faces = np.asarray(
        [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
         (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
         (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
         (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)],
        dtype=np.int64,
    )
    return _normalize(vertices), _orient_faces(_normalize(vertices), faces)


def _polar_orientation(vertices: np.ndarray) -> np.ndarray:  # mesh.py:522-555
    source = vertices[5]
    target = np.asarray((0.0, 0.0, 1.0))
    axis = np.cross(source, target)
    sine = np.linalg.norm(axis)
    cosine = float(source @ target)
    skew = np.asarray(
        ((0.0, -axis[2], axis[1]), (axis[2], 0.0, -axis[0]), (-axis[1], axis[0], 0.0))
    )
    rotation = np.eye(3) + skew + skew @ skew * ((1.0 - cosine) / sine**2)
    rotated = vertices @ rotation.T
    reference_longitude = np.arctan2(rotated[9, 1], rotated[9, 0])
    c = np.cos(-reference_longitude)
    s = np.sin(-reference_longitude)
    twist = np.asarray(((c, -s, 0.0), (s, c, 0.0), (0.0, 0.0, 1.0)))
    return rotated @ twist.T


def _orient_faces(vertices: np.ndarray, faces: np.ndarray) -> np.ndarray:  # mesh.py:558-566
    result = faces.copy()
    a, b, c = (vertices[result[:, i]] for i in range(3))
    inward = np.einsum("ij,ij->i", np.cross(b - a, c - a), a + b + c) < 0.0
    result[inward, 1], result[inward, 2] = result[inward, 2].copy(), result[inward, 1].copy()
    return result


def _subdivide(vertices: np.ndarray, faces: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    # mesh.py:569-593 (dict of midpoints in first-appearance order).
    points = [row.copy() for row in vertices]
    midpoint_indices: dict[tuple[int, int], int] = {}

    def midpoint(first: int, second: int) -> int:
        key = (min(first, second), max(first, second))
        if key not in midpoint_indices:
            midpoint_indices[key] = len(points)
            points.append(_normalize((vertices[first] + vertices[second])[None, :])[0])
        return midpoint_indices[key]

    new_faces: list[tuple[int, int, int]] = []
    for first, second, third in faces:
        ab = midpoint(int(first), int(second))
        bc = midpoint(int(second), int(third))
        ca = midpoint(int(third), int(first))
        new_faces.extend(
            [(int(first), ab, ca), (int(second), bc, ab), (int(third), ca, bc), (ab, bc, ca)]
        )
    return np.asarray(points), np.asarray(new_faces, dtype=np.int64)


def _spherical_face_centers(vertices, faces):  # mesh.py:666-671
    a, b, c = (vertices[faces[:, i]] for i in range(3))
    centers = _normalize(np.cross(b - a, c - a))
    reverse = np.einsum("ij,ij->i", centers, a + b + c) < 0.0
    centers[reverse] *= -1.0
    return centers


def _build_edges(faces):  # mesh.py:674-708
    edge_lookup: dict[tuple[int, int], int] = {}
    edge_faces: list[list[tuple[int, int]]] = []
    face_edges = np.empty((faces.shape[0], 3), dtype=np.int64)
    face_signs = np.empty_like(face_edges)
    for face_index, (a, b, c) in enumerate(faces):
        for local, (tail, head) in enumerate(((a, b), (b, c), (c, a))):
            key = (min(int(tail), int(head)), max(int(tail), int(head)))
            sign = 1 if (int(tail), int(head)) == key else -1
            edge_index = edge_lookup.setdefault(key, len(edge_lookup))
            if edge_index == len(edge_faces):
                edge_faces.append([])
            edge_faces[edge_index].append((face_index, sign))
            face_edges[face_index, local] = edge_index
            face_signs[face_index, local] = sign
    if any(len(adj) != 2 or {s for _, s in adj} != {-1, 1} for adj in edge_faces):
        raise RuntimeError("geodesic mesh is not a closed two-manifold")
    edges = np.asarray(list(edge_lookup), dtype=np.int64)
    left = np.empty(edges.shape[0], dtype=np.int64)
    right = np.empty_like(left)
    for edge_index, adjacent in enumerate(edge_faces):
        for face_index, sign in adjacent:
            if sign == 1:
                left[edge_index] = face_index
            else:
                right[edge_index] = face_index
    return edges, face_edges, face_signs, left, right


def _arc_length(first, second):  # mesh.py:745-748
    cross_norm = np.linalg.norm(np.cross(first, second), axis=1)
    dot = np.einsum("ij,ij->i", first, second)
    return np.arctan2(cross_norm, dot)


def _signed_spherical_triangle_area(a, b, c):  # mesh.py:755-765
    numerator = np.einsum("ij,ij->i", a, np.cross(b, c))
    denominator = (
        1.0 + np.einsum("ij,ij->i", a, b) + np.einsum("ij,ij->i", b, c)
        + np.einsum("ij,ij->i", c, a)
    )
    return 2.0 * np.arctan2(numerator, denominator)


def _dual_geometry(vertices, face_centers, edges, left_faces, right_faces):  # mesh.py:768-789
    left_centers = face_centers[left_faces]
    right_centers = face_centers[right_faces]
    dual_area = np.zeros(vertices.shape[0], dtype=np.float64)
    for endpoint in range(2):
        vertex_indices = edges[:, endpoint]
        wedges = _signed_spherical_triangle_area(
            vertices[vertex_indices], left_centers, right_centers
        )
        if endpoint == 0:
            wedges = -wedges
        np.add.at(dual_area, vertex_indices, wedges)
    degree = np.bincount(edges.ravel(), minlength=len(vertices))
    return dual_area, degree


def build_mesh(subdivision: int = 2, orientation: str = "polar") -> OracleMesh:
    """mesh.py:191-227 with relaxations=0, optimization_steps=0."""
    vertices, faces = _icosahedron()
    if orientation == "polar":
        vertices = _polar_orientation(vertices)
    for _ in range(subdivision):
        vertices, faces = _subdivide(vertices, faces)
    faces = _orient_faces(vertices, faces)  # mesh.py:407
    face_centers = _spherical_face_centers(vertices, faces)
    edges, face_edges, face_signs, left, right = _build_edges(faces)
    primal = _arc_length(vertices[edges[:, 0]], vertices[edges[:, 1]])
    dual = _arc_length(face_centers[left], face_centers[right])
    face_areas = np.abs(_signed_spherical_triangle_area(
        vertices[faces[:, 0]], vertices[faces[:, 1]], vertices[faces[:, 2]]))
    dual_areas, degree = _dual_geometry(vertices, face_centers, edges, left, right)
    return OracleMesh(vertices, faces, edges, face_edges, face_signs, left, right,
                      face_centers, primal, dual, face_areas, dual_areas, degree)


def vertex_incidence(mesh: OracleMesh) -> tuple[np.ndarray, np.ndarray]:
    """Padded degree-6 incidence, numpy_backend.py:46-76."""
    edges = mesh.edges
    edge_indices = np.arange(len(edges), dtype=np.int64)
    vertices = np.concatenate((edges[:, 0], edges[:, 1]))
    incident_edges = np.concatenate((edge_indices, edge_indices))
    incident_signs = np.concatenate((np.ones(len(edges)), -np.ones(len(edges))))
    order = np.argsort(vertices, kind="stable")
    vertices = vertices[order]
    incident_edges = incident_edges[order]
    incident_signs = incident_signs[order]
    counts = np.bincount(vertices, minlength=mesh.n_vertices)
    offsets = np.cumsum(np.concatenate(([0], counts[:-1])))
    slots = np.arange(len(vertices)) - np.repeat(offsets, counts)
    maximum_degree = int(counts.max())
    vertex_edges = np.zeros((mesh.n_vertices, maximum_degree), dtype=np.int64)
    vertex_signs = np.zeros((mesh.n_vertices, maximum_degree))
    vertex_edges[vertices, slots] = incident_edges
    vertex_signs[vertices, slots] = incident_signs
    return vertex_edges, vertex_signs


# ---------------------------------------------------------------------------
# Physics models (materials.py:283-326, sources.py, surface_impedance.py,
# plasma.py).
# ---------------------------------------------------------------------------


def earth_ionosphere_sample(altitudes_m, n_directions, *, lithosphere_conductivity_s_m=1.0e-3,
                            lithosphere_relative_permittivity=10.0,
                            atmosphere_relative_permittivity=1.0,
                            ionosphere_reference_height_m=74_000.0,
                            ionosphere_scale_height_m=6_000.0,
                            ionosphere_prefactor_hz=2.5e5):
    """EarthIonosphereMaterial.sample without anomalies, materials.py:283-326."""
    altitudes = np.asarray(altitudes_m, dtype=np.float64)
    sigma_air = ionosphere_prefactor_hz * EPSILON_0 * np.exp(np.clip(
        (altitudes - ionosphere_reference_height_m) / ionosphere_scale_height_m, -80.0, 80.0))
    below = altitudes < 0.0
    sigma_profile = np.where(below, lithosphere_conductivity_s_m, sigma_air)
    eps_profile = np.where(below, lithosphere_relative_permittivity,
                           atmosphere_relative_permittivity)
    shape = (n_directions, len(altitudes))
    return (np.broadcast_to(sigma_profile, shape).copy(),
            np.broadcast_to(eps_profile, shape).copy())


def geographic_direction(latitude_deg, longitude_deg):  # sources.py:18-32
    lat = np.deg2rad(latitude_deg)
    lon = np.deg2rad(longitude_deg)
    return np.asarray((np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)),
                      dtype=np.float64)


def geographic_face_index(mesh: OracleMesh, latitude_deg, longitude_deg) -> int:
    """sources.py:53-85."""
    direction = geographic_direction(latitude_deg, longitude_deg)
    triangles = mesh.vertices[mesh.faces]
    signs = np.column_stack(tuple(
        np.einsum("ij,j->i", np.cross(triangles[:, e], triangles[:, (e + 1) % 3]), direction)
        for e in range(3)))
    inside = np.all(signs >= -1.0e-12, axis=1) | np.all(signs <= 1.0e-12, axis=1)
    candidates = np.flatnonzero(inside)
    if len(candidates):
        return int(candidates[np.argmax(mesh.face_centers[candidates] @ direction)])
    return int(np.argmax(mesh.face_centers @ direction))


def geographic_distribution(mesh: OracleMesh, altitudes_m, latitude_deg, longitude_deg,
                            altitude_m):
    """sources.py:88-124: containing face, clipped barycentric weights, layer."""
    direction = geographic_direction(latitude_deg, longitude_deg)
    face_index = geographic_face_index(mesh, latitude_deg, longitude_deg)
    vertices = mesh.faces[face_index]
    a, b, c = mesh.vertices[vertices]
    normal = np.cross(b - a, c - a)
    intersection = direction * float((normal @ a) / (normal @ direction))
    ab = b - a
    ac = c - a
    off = intersection - a
    d_ab_ab = float(ab @ ab)
    d_ab_ac = float(ab @ ac)
    d_ac_ac = float(ac @ ac)
    d_o_ab = float(off @ ab)
    d_o_ac = float(off @ ac)
    den = d_ab_ab * d_ac_ac - d_ab_ac**2
    wb = (d_ac_ac * d_o_ab - d_ab_ac * d_o_ac) / den
    wc = (d_ab_ab * d_o_ac - d_ab_ac * d_o_ab) / den
    weights = np.asarray((1.0 - wb - wc, wb, wc))
    weights = np.clip(weights, 0.0, None)
    weights /= weights.sum()
    layer = int(np.argmin(np.abs(np.asarray(altitudes_m) - altitude_m)))
    return vertices.copy(), layer, weights


def radial_linear_distribution(radial_altitudes_m, altitude_m):  # sources.py:127-154
    altitudes = np.asarray(radial_altitudes_m, dtype=np.float64)
    upper = int(np.searchsorted(altitudes, altitude_m, side="left"))
    if upper < len(altitudes) and altitudes[upper] == altitude_m:
        return np.asarray((upper,), dtype=np.int64), np.asarray((1.0,))
    lower = upper - 1
    w = (altitude_m - altitudes[lower]) / (altitudes[upper] - altitudes[lower])
    return np.asarray((lower, upper), dtype=np.int64), np.asarray((1.0 - w, w))


@dataclass
class OracleSource:
    """GaussianCurrent fields, sources.py:157-167."""

    latitude_deg: float = 35.1595
    longitude_deg: float = 126.8526
    altitude_m: float = 2_500.0
    peak_current_a: float = 1.0e6
    vertical_element_length_m: float = 5_000.0
    center_time_s: float | None = None
    one_over_e_half_width_s: float | None = None
    carrier_frequency_hz: float = 0.0

    def current_a(self, time_s: float, dt_s: float) -> float:  # sources.py:237-254
        if self.one_over_e_half_width_s is not None:
            half_width = self.one_over_e_half_width_s
        elif self.carrier_frequency_hz:
            half_width = 0.5 / self.carrier_frequency_hz
        else:
            half_width = 12.0 * dt_s
        center = (self.center_time_s if self.center_time_s is not None
                  else max(4.0 * half_width, 36.0 * dt_s))
        envelope = np.exp(-((time_s - center) / half_width) ** 2)
        if self.carrier_frequency_hz:
            envelope *= np.cos(2.0 * np.pi * self.carrier_frequency_hz * (time_s - center))
        return float(self.peak_current_a * envelope)


def si_pole_rates(min_f=5.0, max_f=45.0, terms=16, span=5.0):  # surface_impedance.py:66-75
    lower = np.log(2.0 * np.pi * min_f)
    upper = np.log(2.0 * np.pi * max_f)
    ext = span * np.log(10.0)
    lower -= ext
    upper += ext
    spacing = (upper - lower) / terms
    return np.exp(lower + (np.arange(terms) + 0.5) * spacing)


def si_weights(min_f=5.0, max_f=45.0, terms=16, span=5.0):  # surface_impedance.py:77-90
    poles = si_pole_rates(min_f, max_f, terms, span)
    if terms > 1:
        spacing = np.log(poles[-1] / poles[0]) / (terms - 1)
    else:
        spacing = np.log(max_f / min_f) + 2.0 * span * np.log(10.0)
    return spacing * np.sqrt(poles) / np.pi


@dataclass
class OracleSurface:
    """ConductiveHalfSpaceSurface parameters, surface_impedance.py:29-34."""

    conductivity_s_m: float | np.ndarray
    minimum_frequency_hz: float = 5.0
    maximum_frequency_hz: float = 45.0
    terms: int = 16
    pole_span_decades: float = 5.0
    permeability_h_m: float = MU_0


@dataclass
class OracleSpecies:
    charge_c: float
    mass_kg: float
    number_density_m3: np.ndarray  # (F, R)
    collision_frequency_hz: np.ndarray  # (F, R)


@dataclass
class OraclePlasma:
    magnetic_field_t: np.ndarray  # (F, R, 3)
    species: Sequence[OracleSpecies] = field(default_factory=tuple)


def _face_reconstruction(mesh: OracleMesh):  # plasma.py:548-559
    normals = np.cross(mesh.vertices[mesh.edges[:, 0]], mesh.vertices[mesh.edges[:, 1]])
    normals /= np.linalg.norm(normals, axis=1, keepdims=True)
    face_normals = normals[mesh.face_edges]
    tangents = np.cross(face_normals, mesh.face_centers[:, None, :])
    tangents /= np.linalg.norm(tangents, axis=2, keepdims=True)
    # per-face pinv(tangents[f]).T; numpy's stacked pinv runs the same SVD per matrix
    reconstruction = np.transpose(np.linalg.pinv(tangents), (0, 2, 1))
    return reconstruction, tangents


def _edge_face_slots(mesh: OracleMesh, faces):  # plasma.py:562-569
    matches = mesh.face_edges[faces] == np.arange(mesh.n_edges)[:, None]
    return np.argmax(matches, axis=1)


def _vertex_face_average(mesh: OracleMesh):  # plasma.py:572-588
    incident = [[] for _ in range(mesh.n_vertices)]
    for face, verts in enumerate(mesh.faces):
        for v in verts:
            incident[int(v)].append(face)
    maximum = max(len(v) for v in incident)
    faces = np.zeros((mesh.n_vertices, maximum), dtype=np.int64)
    weights = np.zeros((mesh.n_vertices, maximum))
    for vertex, values in enumerate(incident):
        idx = np.asarray(values, dtype=np.int64)
        w = mesh.face_solid_angles[idx]
        w /= np.sum(w)
        faces[vertex, : len(values)] = idx
        weights[vertex, : len(values)] = w
    return faces, weights


def _cross_with_b(values, b):  # plasma.py:591-596
    result = np.empty_like(values)
    result[..., 0] = values[..., 1] * b[..., 2] - values[..., 2] * b[..., 1]
    result[..., 1] = values[..., 2] * b[..., 0] - values[..., 0] * b[..., 2]
    result[..., 2] = values[..., 0] * b[..., 1] - values[..., 1] * b[..., 0]
    return result


class OraclePlasmaCoupler:
    """MagnetizedPlasmaADE + GeodesicPlasmaCoupler, plasma.py:371-527."""

    def __init__(self, plasma: OraclePlasma, mesh: OracleMesh, dt: float, dtype=np.float64):
        self.mesh = mesh
        reconstruction, tangents = _face_reconstruction(mesh)
        self.reconstruction = reconstruction
        left_slots = _edge_face_slots(mesh, mesh.edge_left_faces)
        right_slots = _edge_face_slots(mesh, mesh.edge_right_faces)
        self.left_tangents = tangents[mesh.edge_left_faces, left_slots]
        self.right_tangents = tangents[mesh.edge_right_faces, right_slots]
        self.vertex_faces, self.vertex_face_weights = _vertex_face_average(mesh)
        magnetic = plasma.magnetic_field_t
        magnitude = np.linalg.norm(magnetic, axis=2)
        direction = np.zeros_like(magnetic)
        nz = magnitude > 0.0
        direction[nz] = magnetic[nz] / magnitude[nz, None]
        self.b = direction.astype(dtype)
        self.reconstruction = self.reconstruction.astype(dtype)
        self.left_tangents = self.left_tangents.astype(dtype)
        self.right_tangents = self.right_tangents.astype(dtype)
        self.vertex_face_weights = self.vertex_face_weights.astype(dtype)
        self.face_centers = mesh.face_centers.astype(dtype)
        self.coefficients = []
        self.current_density = []
        for sp in plasma.species:  # plasma.py:385-411
            collision = np.asarray(sp.collision_frequency_hz, dtype=np.float64)
            gyro = sp.charge_c * magnitude / sp.mass_kg
            decay = np.exp(-collision * dt)
            angle = gyro * dt
            eigenvalue = -collision + 1j * gyro
            integral = np.expm1(eigenvalue * dt) / eigenvalue
            parallel_integral = -np.expm1(-collision * dt) / collision
            drive = np.asarray(sp.number_density_m3) * sp.charge_c**2 / sp.mass_kg
            # cast like the reference's backend.asarray (plasma.py:398-410)
            self.coefficients.append(tuple(np.asarray(c, dtype=dtype) for c in (
                decay, np.cos(angle), np.sin(angle), drive * parallel_integral,
                drive * integral.real, drive * integral.imag)))
            self.current_density.append(np.zeros((*magnitude.shape, 3), dtype=dtype))

    def ade_advance(self, electric):  # plasma.py:413-442
        total = np.zeros(electric.shape, dtype=electric.dtype)
        b = self.b
        e_par = (electric * b).sum(axis=2)[..., None] * b
        e_perp = electric - e_par
        e_cross = _cross_with_b(e_perp, b)
        for index, (decay, cosine, sine, d_par, d_re, d_im) in enumerate(self.coefficients):
            current = self.current_density[index]
            c_par = (current * b).sum(axis=2)[..., None] * b
            c_perp = current - c_par
            c_cross = _cross_with_b(c_perp, b)
            updated = (decay[..., None] * c_par + d_par[..., None] * e_par
                       + decay[..., None] * (cosine[..., None] * c_perp + sine[..., None] * c_cross)
                       + d_re[..., None] * e_perp + d_im[..., None] * e_cross)
            self.current_density[index][:] = updated
            total += updated
        return total

    def advance(self, er, et):  # plasma.py:489-527
        mesh = self.mesh
        nodes_at_faces = er[mesh.faces].sum(axis=1) / 3.0
        radial = 0.5 * (nodes_at_faces[:, :-1] + nodes_at_faces[:, 1:])
        edge_values = et[mesh.face_edges]
        tangential = (edge_values[..., None] * self.reconstruction[:, :, None, :]).sum(axis=1)
        electric = tangential + radial[..., None] * self.face_centers[:, None, :]
        current = self.ade_advance(electric)
        left = (current[mesh.edge_left_faces] * self.left_tangents[:, None, :]).sum(axis=2)
        right = (current[mesh.edge_right_faces] * self.right_tangents[:, None, :]).sum(axis=2)
        tangential_current = 0.5 * (left + right)
        radial_at_faces = (current * self.face_centers[:, None, :]).sum(axis=2)
        radial_at_vertices = (radial_at_faces[self.vertex_faces]
                              * self.vertex_face_weights[:, :, None]).sum(axis=1)
        radial_current = np.zeros((radial_at_vertices.shape[0], radial_at_vertices.shape[1] + 1),
                                  dtype=radial_at_vertices.dtype)
        radial_current[:, 0] = radial_at_vertices[:, 0]
        radial_current[:, -1] = radial_at_vertices[:, -1]
        if radial_at_vertices.shape[1] > 1:
            radial_current[:, 1:-1] = 0.5 * (radial_at_vertices[:, :-1] + radial_at_vertices[:, 1:])
        return radial_current, tangential_current


# ---------------------------------------------------------------------------
# Simulation (solver.py:184-989, 1080-1136)
# ---------------------------------------------------------------------------


class OracleSim:
    """NumPy restatement of GeodesicFDTD (default options).

    Supports the options the HIP path supports: full-spherical or thin-shell
    geometry, exponential or trapezoidal loss, PEC or surface-impedance lower
    boundary, vertical Gaussian current, Earth-ionosphere (or caller-given
    sigma / epsilon_r arrays) material, magnetized cold-plasma ADE.
    ``dtype`` selects float64 (default) or float32 field arithmetic, as the
    reference NumPy backend does (numpy_backend.py:23-26).
    """

    def __init__(self, *, subdivision=2, radial_cells=24, minimum_altitude_m=-100_000.0,
                 maximum_altitude_m=100_000.0, earth_radius_m=EARTH_RADIUS_M,
                 courant_factor=0.35, time_step_s=None, orientation="polar",
                 radial_altitudes_m=None, loss_integration="exponential",
                 geometry_mode="full-spherical", material: dict | None = None,
                 material_arrays=None, source: OracleSource | None = None,
                 surface: OracleSurface | None = None, plasma: OraclePlasma | None = None,
                 mesh: OracleMesh | None = None, dtype="float64"):
        self.dtype = np.dtype(dtype)
        self.geometry_mode = geometry_mode
        self.mesh = mesh if mesh is not None else build_mesh(subdivision, orientation)
        mesh = self.mesh
        self.vertex_edges, vsigns = vertex_incidence(mesh)
        self.vertex_edge_signs = vsigns.astype(self.dtype)
        self.face_edge_signs = mesh.face_edge_signs.astype(self.dtype)
        # solver.py:260-283
        if radial_altitudes_m is None:
            self.altitudes_m = np.linspace(minimum_altitude_m, maximum_altitude_m,
                                           radial_cells + 1)
        else:
            self.altitudes_m = np.asarray(radial_altitudes_m, dtype=np.float64)
        self.earth_radius_m = earth_radius_m
        self.radii_m = earth_radius_m + self.altitudes_m
        self.radial_midpoints_m = 0.5 * (self.radii_m[:-1] + self.radii_m[1:])
        self.radial_midpoint_altitudes_m = self.radial_midpoints_m - earth_radius_m
        self.radial_steps_m = np.diff(self.radii_m)
        L = np.empty(len(self.radii_m))
        L[0] = 0.5 * self.radial_steps_m[0]
        L[-1] = 0.5 * self.radial_steps_m[-1]
        L[1:-1] = np.diff(self.radial_midpoints_m)
        self.radial_node_control_lengths_m = L
        # material sampling (solver.py:467-629, point support)
        if material_arrays is not None:
            self.sigma_er, self.epsilon_r_er, self.sigma_et, self.epsilon_r_et = (
                np.asarray(a, dtype=np.float64) for a in material_arrays)
        else:
            params = material or {}
            self.sigma_er, self.epsilon_r_er = earth_ionosphere_sample(
                self.altitudes_m, mesh.n_vertices, **params)
            self.sigma_et, self.epsilon_r_et = earth_ionosphere_sample(
                self.radial_midpoint_altitudes_m, mesh.n_edges, **params)
        self.plasma = plasma
        self.cfl_time_step_limit_s = self._cfl()
        self.maximum_stable_time_step_s = courant_factor * self.cfl_time_step_limit_s
        self.time_step_s = (time_step_s if time_step_s is not None
                            else self.maximum_stable_time_step_s)
        if self.time_step_s > self.maximum_stable_time_step_s * (1.0 + 1.0e-12):
            raise ValueError("time step exceeds conservative limit")
        dt = self.time_step_s
        V, E, F, R = mesh.n_vertices, mesh.n_edges, mesh.n_faces, len(self.radial_midpoints_m)
        self.er = np.zeros((V, R + 1), dtype=self.dtype)
        self.ht = np.zeros((E, R + 1), dtype=self.dtype)
        self.et = np.zeros((E, R), dtype=self.dtype)
        self.hr = np.zeros((F, R), dtype=self.dtype)
        # Surface impedance ADE, surface_impedance.py:176-203
        self.surface = surface
        if surface is not None:
            cond = np.asarray(surface.conductivity_s_m, dtype=np.float64).reshape(-1)
            cond = np.full(E, cond[0]) if cond.size == 1 else cond.copy()
            poles = si_pole_rates(surface.minimum_frequency_hz, surface.maximum_frequency_hz,
                                  surface.terms, surface.pole_span_decades)
            weights = si_weights(surface.minimum_frequency_hz, surface.maximum_frequency_hz,
                                 surface.terms, surface.pole_span_decades)
            alpha = 0.5 * dt * poles
            self.si_decay = ((1.0 - alpha) / (1.0 + alpha)).astype(self.dtype)
            self.si_drive = (alpha / (1.0 + alpha)).astype(self.dtype)
            self.si_history_weights = (weights / (1.0 + alpha)).astype(self.dtype)
            self.si_scale = np.sqrt(surface.permeability_h_m / cond).astype(self.dtype)
            self.si_memory = np.zeros((E, surface.terms), dtype=self.dtype)
        self.coupler = (OraclePlasmaCoupler(plasma, mesh, dt, self.dtype) if plasma is not None
                        else None)
        self._prepare_geometry()
        self._prepare_coefficients(loss_integration)
        self.source = source
        self.source_distribution = None
        if source is not None:
            verts, _, hw = geographic_distribution(mesh, self.altitudes_m, source.latitude_deg,
                                                   source.longitude_deg, source.altitude_m)
            layers, rw = radial_linear_distribution(self.altitudes_m, source.altitude_m)
            cnt = len(layers)
            self.source_distribution = (
                np.repeat(verts, cnt), np.tile(layers, len(verts)),
                (np.repeat(hw, cnt) * np.tile(rw, len(verts))).astype(self.dtype))
        self.steps = 0
        self.time_s = 0.0

    # solver.py:377-406
    def _cfl(self) -> float:
        smallest_radius = float(self.radii_m.min())
        primal = smallest_radius * float(self.mesh.primal_edge_angles.min())
        dual = smallest_radius * float(self.mesh.dual_edge_angles.min())
        radial = float(self.radial_steps_m.min())
        inverse_length_squared = primal**-2 + dual**-2 + (2.0 / radial) ** 2
        minimum_epsilon_r = min(float(np.min(self.epsilon_r_er)),
                                float(np.min(self.epsilon_r_et)))
        geometric = np.sqrt(minimum_epsilon_r) / (C_0 * np.sqrt(inverse_length_squared))
        if self.plasma is None:
            return geometric
        drive = np.zeros(self.plasma.magnetic_field_t.shape[:2])
        for sp in self.plasma.species:
            drive += np.asarray(sp.number_density_m3) * sp.charge_c**2 / sp.mass_kg
        wp = np.sqrt(float(np.max(drive)) / (EPSILON_0 * minimum_epsilon_r))
        if wp == 0.0:
            return geometric
        return min(geometric, 2.0 / wp)

    def _asarray(self, values):
        return np.array(values, dtype=self.dtype, copy=True)

    def _prepare_geometry(self):  # solver.py:408-442
        m = self.mesh
        a = self._asarray
        self._primal = a(m.primal_edge_angles[:, None])
        self._inv_primal = a(1.0 / m.primal_edge_angles[:, None])
        self._dual = a(m.dual_edge_angles[:, None])
        self._inv_dual = a(1.0 / m.dual_edge_angles[:, None])
        self._inv_dual_area = a(1.0 / m.dual_cell_solid_angles[:, None])
        self._inv_face_area = a(1.0 / m.face_solid_angles[:, None])
        self._radii = a(self.radii_m)
        self._inv_radii = a(1.0 / self.radii_m[None, :])
        self._mid = a(self.radial_midpoints_m)
        self._inv_mid = a(1.0 / self.radial_midpoints_m[None, :])
        self._steps = a(self.radial_steps_m)
        self._control = a(self.radial_node_control_lengths_m)
        self._center_dist = a(self.radial_midpoints_m[1:] - self.radial_midpoints_m[:-1])

    def _prepare_coefficients(self, loss_integration):  # solver.py:706-761
        dt = self.time_step_s
        eps_er = EPSILON_0 * self.epsilon_r_er
        eps_et = EPSILON_0 * self.epsilon_r_et
        if loss_integration == "trapezoidal":
            l_er = self.sigma_er * dt / (2.0 * eps_er)
            l_et = self.sigma_et * dt / (2.0 * eps_et)
            ca_er, cb_er = (1.0 - l_er) / (1.0 + l_er), dt / (eps_er * (1.0 + l_er))
            ca_et, cb_et = (1.0 - l_et) / (1.0 + l_et), dt / (eps_et * (1.0 + l_et))
        else:
            ca_er, cb_er = exponential_loss_coefficients(self.sigma_er, eps_er, dt)
            ca_et, cb_et = exponential_loss_coefficients(self.sigma_et, eps_et, dt)
        self.ca_er_f64, self.cb_er_f64 = ca_er, cb_er
        self.ca_et_f64, self.cb_et_f64 = ca_et, cb_et
        self._ca_er = self._asarray(ca_er)
        self._cb_er = self._asarray(cb_er)
        self._ca_et = self._asarray(ca_et)
        self._cb_et = self._asarray(cb_et)

    # -- operators (numpy_backend.py:90-114) --
    def _edge_difference(self, v):
        e = self.mesh.edges
        return v[e[:, 1]] - v[e[:, 0]]

    def _dual_edge_difference(self, f):
        return f[self.mesh.edge_left_faces] - f[self.mesh.edge_right_faces]

    def _face_circulation(self, ev):
        selected = ev[self.mesh.face_edges]
        signs = self.face_edge_signs
        while signs.ndim < selected.ndim:
            signs = signs[..., None]
        return np.sum(selected * signs, axis=1)

    def _dual_cell_circulation(self, ev):
        shape = (self.mesh.n_vertices,) + (1,) * (ev.ndim - 1)
        result = ev[self.vertex_edges[:, 0]].copy()
        result *= self.vertex_edge_signs[:, 0].reshape(shape)
        for slot in range(1, self.vertex_edges.shape[1]):
            result += ev[self.vertex_edges[:, slot]] * self.vertex_edge_signs[:, slot].reshape(shape)
        return result

    # solver.py:443-465
    def radial_derivative_et(self):
        values = self.et
        if self.geometry_mode == "full-spherical":
            values = values * self._mid[None, :]
        result = np.empty_like(self.ht)
        result[:, 0] = 2.0 * values[:, 0] / self._steps[0]
        result[:, -1] = -2.0 * values[:, -1] / self._steps[-1]
        if self.ht.shape[1] > 2:
            result[:, 1:-1] = np.diff(values, axis=1) / self._center_dist
        if self.geometry_mode == "full-spherical":
            result *= self._inv_radii
        return result

    def radial_derivative_ht(self):
        values = self.ht
        if self.geometry_mode == "full-spherical":
            values = values * self._radii[None, :]
        result = np.diff(values, axis=1) / self._steps[None, :]
        if self.geometry_mode == "full-spherical":
            result *= self._inv_mid
        return result

    def si_advance_lower_boundary(self, h_old, electric_first_cell, surface_gradient_er,
                                  radial_step_m, dt):
        """surface_impedance.py:225-270."""
        history = self.si_scale * (self.si_memory * self.si_history_weights[None, :]).sum(axis=1)
        gain = self.si_scale * self.si_history_weights.sum()
        curl_scale = 2.0 * dt / (MU_0 * radial_step_m)
        forcing = dt / MU_0 * surface_gradient_er
        coupling = curl_scale * gain
        h_new = ((1.0 - 0.5 * coupling) * h_old + forcing - curl_scale * electric_first_cell
                 + curl_scale * history) / (1.0 + 0.5 * coupling)
        self.si_memory[:] = (self.si_memory * self.si_decay[None, :]
                             + (h_new + h_old)[:, None] * self.si_drive[None, :])
        return h_new

    def update_magnetic_fields(self):  # solver.py:877-921
        dt = self.time_step_s
        g = self._edge_difference(self.er)
        g *= self._inv_primal
        g *= self._inv_radii
        lower_g = g[:, 0] + 0.0 if self.surface is not None else None
        lower_h_old = self.ht[:, 0] + 0.0 if self.surface is not None else None
        rd = self.radial_derivative_et()
        g -= rd
        g *= dt / MU_0
        self.ht += g
        if self.surface is not None:
            metric = (self.radial_midpoints_m[0] / self.radii_m[0]
                      if self.geometry_mode == "full-spherical" else 1.0)
            self.ht[:, 0] = self.si_advance_lower_boundary(
                lower_h_old, metric * self.et[:, 0], lower_g, self._steps[0], dt)
        circ = self._face_circulation(self.et * self._primal)
        circ *= self._inv_face_area
        circ *= self._inv_mid
        circ *= dt / MU_0
        self.hr -= circ

    def update_electric_fields(self, current_a=0.0):  # solver.py:923-989
        plasma_current = self.coupler.advance(self.er, self.et) if self.coupler else None
        circ = self._dual_cell_circulation(self.ht * self._dual)
        circ *= self._inv_dual_area
        circ *= self._inv_radii
        cd = None
        if self.source_distribution is not None:
            verts, layers, weights = self.source_distribution
            cd = (weights * current_a * self.source.vertical_element_length_m
                  * self._inv_dual_area[verts, 0] / self._radii[layers] ** 2
                  / self._control[layers])
        self.er *= self._ca_er
        circ *= self._cb_er
        self.er += circ
        if cd is not None:
            verts, layers, _ = self.source_distribution
            cv = verts if self._cb_er.shape[0] > 1 else 0
            self.er[verts, layers] -= self._cb_er[cv, layers] * cd
        if plasma_current is not None:
            self.er -= self._cb_er * plasma_current[0]
        g = self._dual_edge_difference(self.hr)
        g *= self._inv_dual
        g *= self._inv_mid
        rd = self.radial_derivative_ht()
        g -= rd
        self.et *= self._ca_et
        g *= self._cb_et
        self.et += g
        if plasma_current is not None:
            self.et -= self._cb_et * plasma_current[1]

    def current_at_step(self, n: int) -> float:
        if self.source is None:
            return 0.0
        return self.source.current_a((n + 0.5) * self.time_step_s, self.time_step_s)

    def step(self, count: int = 1) -> None:  # solver.py:789-816 (eager branch)
        for _ in range(count):
            current = (self.source.current_a(self.time_s + 0.5 * self.time_step_s,
                                             self.time_step_s)
                       if self.source is not None else 0.0)
            self.update_magnetic_fields()
            self.update_electric_fields(current)
            self.steps += 1
            self.time_s = self.steps * self.time_step_s

    def record_h_observations(self, faces, face_layers, face_weights, edges, edge_layers,
                              edge_weights, steps, sample_every=1):
        """solver.py:1138-1229: the first row is H at -1/2 and the row of sample
        step k is H at k - 1/2; samples at 0, sample_every, ... and at steps."""
        faces = np.asarray(faces, dtype=np.int64)
        face_layers = np.asarray(face_layers, dtype=np.int64)
        fw = np.asarray(face_weights, dtype=np.float64).astype(self.dtype)
        edges = np.asarray(edges, dtype=np.int64)
        edge_layers = np.asarray(edge_layers, dtype=np.int64)
        ew = np.asarray(edge_weights, dtype=np.float64).astype(self.dtype)
        sample_steps = np.unique(np.concatenate((np.arange(0, steps + 1, sample_every),
                                                 np.asarray((steps,)))))
        rt = np.zeros((len(sample_steps), faces.shape[0]), dtype=self.dtype)
        tt = np.zeros((len(sample_steps), edges.shape[0]), dtype=self.dtype)

        def sample(row):
            rt[row] = (self.hr[faces, face_layers] * fw).sum(axis=1)
            tt[row] = (self.ht[edges, edge_layers] * ew).sum(axis=1)

        sample(0)
        for row, target in enumerate(sample_steps[1:], start=1):
            self.step(int(target) - self.steps)
            sample(row)
        return rt, tt

    def record_er_observations(self, vertices, layers, weights, steps):
        """solver.py:1080-1136: row 0 initial, row n after step n."""
        vertices = np.asarray(vertices, dtype=np.int64)
        layers = np.asarray(layers, dtype=np.int64)
        w = np.asarray(weights, dtype=np.float64).astype(self.dtype)
        traces = np.zeros((steps + 1, vertices.shape[0]), dtype=self.dtype)

        def sample(row):
            selected = self.er[vertices, layers[:, None]]
            traces[row] = (selected * w).sum(axis=1)

        sample(0)
        currents = np.asarray([
            self.source.current_a((self.steps + o + 0.5) * self.time_step_s, self.time_step_s)
            if self.source is not None else 0.0 for o in range(steps)], dtype=np.float64)
        currents = currents.astype(self.dtype)
        for offset in range(steps):
            self.update_magnetic_fields()
            self.update_electric_fields(currents[offset])
            self.steps += 1
            self.time_s = self.steps * self.time_step_s
            sample(offset + 1)
        return traces


def exponential_loss_coefficients(sigma, epsilon, dt):  # solver.py:750-761
    rate = sigma * dt / epsilon
    decay = np.exp(-rate)
    phi1 = np.ones_like(rate)
    nz = rate != 0.0
    phi1[nz] = -np.expm1(-rate[nz]) / rate[nz]
    return decay, dt / epsilon * phi1


def energy(sim: OracleSim) -> float:
    """'total_staggered' energy, verification/physics_diagnostics/model.py:78-92, 242-258.

    Weights: Er A_v r_k^2 L_k; Et theta*phi rbar_j^2 dr_j; Hr Omega_f rbar_j^2 dr_j;
    Ht theta*phi r_k^2 L_k.
    """
    m = sim.mesh
    er_w = m.dual_cell_solid_angles[:, None] * (sim.radii_m**2 * sim.radial_node_control_lengths_m)[None, :]
    tp = (m.primal_edge_angles * m.dual_edge_angles)[:, None]
    et_w = tp * (sim.radial_midpoints_m**2 * sim.radial_steps_m)[None, :]
    hr_w = m.face_solid_angles[:, None] * (sim.radial_midpoints_m**2 * sim.radial_steps_m)[None, :]
    ht_w = tp * (sim.radii_m**2 * sim.radial_node_control_lengths_m)[None, :]
    eps_er = EPSILON_0 * sim.epsilon_r_er
    eps_et = EPSILON_0 * sim.epsilon_r_et
    er = sim.er.astype(np.float64)
    et = sim.et.astype(np.float64)
    hr = sim.hr.astype(np.float64)
    ht = sim.ht.astype(np.float64)
    return float(0.5 * np.sum(eps_er * er**2 * er_w) + 0.5 * np.sum(eps_et * et**2 * et_w)
                 + 0.5 * MU_0 * np.sum(hr**2 * hr_w) + 0.5 * MU_0 * np.sum(ht**2 * ht_w))


def divergence_b(mesh, radii_m, midpoints_m, control_m, hr, ht, skip_lower=False):
    """Discrete div-B of every prism cell (f, k), k = 0..R (derived invariant).

    Fluxes Phi_r(f,j) = Hr[f,j] Omega_f rbar_j^2 and Phi_t(e,k) = Ht[e,k]
    theta_e r_k L_k; D(f,k) = Phi_r(f,k) - Phi_r(f,k-1) - sum_c s_fc Phi_t(e_c,k)
    with Phi_r(f,-1) = Phi_r(f,R) = 0.  Returns (D, scale) where scale is the
    per-cell sum of absolute flux terms.  Nothing in the reference computes
    this; it follows from the H update (solver.py:877-921) by the telescoping
    of the radial stencil and boundary-of-boundary = 0 (mesh.py tests).
    """
    hr = np.asarray(hr, dtype=np.float64)
    ht = np.asarray(ht, dtype=np.float64)
    F, R = hr.shape
    phir = hr * mesh.face_solid_angles[:, None] * (midpoints_m**2)[None, :]
    phit = ht * mesh.primal_edge_angles[:, None] * (radii_m * control_m)[None, :]
    padded = np.zeros((F, R + 2))
    padded[:, 1:-1] = phir
    tang = mesh.face_circulation(phit)  # (F, R+1)
    upper = padded[:, 1:]
    lower = padded[:, :-1]
    D = upper - lower - tang
    abs_tang = np.abs(phit[mesh.face_edges]).sum(axis=1)
    scale = np.abs(upper) + np.abs(lower) + abs_tang
    if skip_lower:
        D = D[:, 1:]
        scale = scale[:, 1:]
    return D, scale

"""Icosahedral geodesic mesh: topology and unit-sphere metric (host, FP64).

Same public contract as the reference's ``mesh.py`` (``GeodesicMesh`` fields,
``build_geodesic_mesh`` arguments) and the same integer numbering, but every
step is vectorised so level 7-8 meshes build in seconds:

* new vertices of a subdivision level and the edge list are numbered in
  first-appearance order of the half-edges (a,b),(b,c),(c,a) of the faces in
  order -- the order the reference's dict insertion produces
  (mesh.py:569-593, 674-708) -- computed with one stable sort of int64 keys;
* float geometry follows the reference formulas (mesh.py:666-789) with the
  same reduction order (dual areas accumulate edge wedges in index order).
"""

from __future__ import annotations

from dataclasses import dataclass
import json
from typing import Any, Mapping

import numpy as np
from numpy.typing import NDArray

FloatArray = NDArray[np.float64]
IntArray = NDArray[np.int64]


@dataclass(frozen=True, slots=True)
class GeodesicMesh:
    """Topology and unit-sphere metric terms (reference mesh.py:21-41)."""

    vertices: FloatArray
    faces: IntArray
    edges: IntArray
    face_edges: IntArray
    face_edge_signs: IntArray
    edge_left_faces: IntArray
    edge_right_faces: IntArray
    face_centers: FloatArray
    primal_edge_angles: FloatArray
    dual_edge_angles: FloatArray
    face_solid_angles: FloatArray
    dual_cell_solid_angles: FloatArray
    vertex_degree: IntArray
    subdivision: int | None
    topology_kind: str
    face_levels: IntArray | None
    refinement_spec_json: str | None

    def __post_init__(self) -> None:
        for values in (
            self.vertices, self.faces, self.edges, self.face_edges, self.face_edge_signs,
            self.edge_left_faces, self.edge_right_faces, self.face_centers,
            self.primal_edge_angles, self.dual_edge_angles, self.face_solid_angles,
            self.dual_cell_solid_angles, self.vertex_degree,
        ):
            values.setflags(write=False)
        if self.face_levels is not None:
            self.face_levels.setflags(write=False)

    @property
    def n_vertices(self) -> int:
        return int(self.vertices.shape[0])

    @property
    def n_edges(self) -> int:
        return int(self.edges.shape[0])

    @property
    def n_faces(self) -> int:
        return int(self.faces.shape[0])

    @property
    def refinement_spec(self) -> dict[str, Any] | None:
        if self.refinement_spec_json is None:
            return None
        value = json.loads(self.refinement_spec_json)
        if not isinstance(value, dict):
            raise RuntimeError("mesh refinement specification is not a JSON object")
        return value

    def edge_midpoints(self) -> FloatArray:
        """Unit edge midpoints (read-only, computed once per mesh)."""
        return _mesh_derived(self, "edge_midpoints",
                             lambda: _normalize(self.vertices[self.edges].sum(axis=1)))

    def edge_diamond_solid_angles(self) -> FloatArray:
        """Disjoint primal-face area associated with each edge (mesh.py:93-104)."""
        tail = self.vertices[self.edges[:, 0]]
        head = self.vertices[self.edges[:, 1]]
        left = self.face_centers[self.edge_left_faces]
        right = self.face_centers[self.edge_right_faces]
        return _signed_area(tail, head, left) - _signed_area(tail, head, right)

    def edge_diamond_quadrant_solid_angles(self) -> FloatArray:
        """(E, 4) areas of the four spherical triangles that split an edge's
        diamond at its midpoint m: (m, tail, left), (m, left, head),
        (m, head, right), (m, right, tail), with left / right the adjacent face
        centres (mesh.py:106-121).  They sum to edge_diamond_solid_angles."""
        tail = self.vertices[self.edges[:, 0]]
        head = self.vertices[self.edges[:, 1]]
        mid = self.edge_midpoints()
        left = self.face_centers[self.edge_left_faces]
        right = self.face_centers[self.edge_right_faces]
        corners = ((tail, left), (left, head), (head, right), (right, tail))
        return np.stack([-_signed_area(mid, a, b) for a, b in corners], axis=1)

    def dual_cell_wedge_quadrature(self, vertex_indices, edge_indices
                                   ) -> tuple[FloatArray, FloatArray]:
        """One quadrature point per dual-cell wedge (mesh.py:123-155).

        The wedge of vertex v and incident edge e is the spherical triangle
        (v, left centre, right centre) of e; the wedges of a vertex tile its
        dual cell.  Returns the normalised vertex-corner sum as the direction
        and the wedge's area, oriented so that it is positive for both ends of
        the edge.
        """
        v = np.asarray(vertex_indices, dtype=np.int64)
        e = np.asarray(edge_indices, dtype=np.int64)
        if v.ndim != 1 or e.shape != v.shape:
            raise ValueError("dual-cell wedge indices must be matching 1-D arrays")
        if np.any((v < 0) | (v >= self.n_vertices)):
            raise ValueError("dual-cell wedge vertex index is out of bounds")
        if np.any((e < 0) | (e >= self.n_edges)):
            raise ValueError("dual-cell wedge edge index is out of bounds")
        ends = self.edges[e]
        if not np.all((ends[:, 0] == v) | (ends[:, 1] == v)):
            raise ValueError("dual-cell wedge vertex must be incident to its edge")
        corner = self.vertices[v]
        left = self.face_centers[self.edge_left_faces[e]]
        right = self.face_centers[self.edge_right_faces[e]]
        area = _signed_area(corner, left, right)
        return _normalize(corner + left + right), np.where(ends[:, 1] == v, area, -area)

    # Host operators with the reference's semantics (mesh.py:157-188); used by
    # setup code and tests, never by the time step.
    def face_circulation(self, edge_values: FloatArray) -> FloatArray:
        selected = edge_values[self.face_edges]
        signs = self.face_edge_signs.astype(edge_values.dtype, copy=False)
        while signs.ndim < selected.ndim:
            signs = signs[..., None]
        return np.sum(selected * signs, axis=1)

    def dual_cell_circulation(self, dual_edge_values: FloatArray) -> FloatArray:
        shape = (self.n_vertices,) + dual_edge_values.shape[1:]
        result = np.zeros(shape, dtype=dual_edge_values.dtype)
        np.add.at(result, self.edges[:, 0], dual_edge_values)
        np.add.at(result, self.edges[:, 1], -dual_edge_values)
        return result

    def edge_difference(self, vertex_values: FloatArray) -> FloatArray:
        return vertex_values[self.edges[:, 1]] - vertex_values[self.edges[:, 0]]

    def dual_edge_difference(self, face_values: FloatArray) -> FloatArray:
        return face_values[self.edge_left_faces] - face_values[self.edge_right_faces]


def build_geodesic_mesh(
    subdivision: int = 2,
    relaxations: int = 0,
    orientation: str = "polar",
    optimization_steps: int = 0,
) -> GeodesicMesh:
    """Recursively bisected icosahedral mesh (reference mesh.py:191-227)."""

    if subdivision < 0:
        raise ValueError("subdivision must be non-negative")
    if relaxations < 0:
        raise ValueError("relaxations must be non-negative")
    if optimization_steps < 0:
        raise ValueError("optimization_steps must be non-negative")
    if orientation not in {"native", "polar"}:
        raise ValueError("orientation must be 'native' or 'polar'")
    vertices, faces = subdivided_icosahedron(subdivision, orientation)
    for _ in range(relaxations):
        vertices = _relax(vertices, faces)
    if optimization_steps:
        vertices = _optimize_edge_lengths(vertices, faces, optimization_steps)
    return assemble_geodesic_mesh(
        vertices, faces, subdivision, topology_kind="uniform",
        face_levels=np.full(len(faces), subdivision, dtype=np.int64),
    )


def build_geodesic_mesh_from_topology(
    vertices: FloatArray,
    faces: IntArray,
    *,
    subdivision: int | None,
    face_levels: IntArray | None = None,
    refinement_spec: Mapping[str, Any] | None = None,
    topology_kind: str = "uniform",
    normalize_vertices: bool = True,
    require_well_centered: bool = False,
) -> GeodesicMesh:
    """Assemble a mesh from explicit coordinates and faces (checkpoint restore)."""

    coordinates = np.array(vertices, dtype=np.float64, copy=True)
    topology = np.array(faces, dtype=np.int64, copy=True)
    if coordinates.ndim != 2 or coordinates.shape[1] != 3:
        raise ValueError("vertices must have shape (n, 3)")
    if topology.ndim != 2 or topology.shape[1] != 3:
        raise ValueError("faces must have shape (n, 3)")
    if np.any(topology < 0) or np.any(topology >= len(coordinates)):
        raise ValueError("face vertex index is out of range")
    if not np.all(np.isfinite(coordinates)):
        raise ValueError("vertices must be finite")
    if normalize_vertices:
        coordinates = _normalize(coordinates)
    spec = None if refinement_spec is None else json.dumps(dict(refinement_spec), sort_keys=True)
    levels = None if face_levels is None else np.array(face_levels, dtype=np.int64, copy=True)
    mesh = assemble_geodesic_mesh(coordinates, topology, subdivision, topology_kind=topology_kind,
                                  face_levels=levels, refinement_spec_json=spec)
    return mesh


def _normalize(values: FloatArray) -> FloatArray:
    return values / np.linalg.norm(values, axis=-1, keepdims=True)


_ICOSAHEDRON_FACES = np.asarray(
    [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
     (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
     (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)],
    dtype=np.int64,
)


def _icosahedron() -> tuple[FloatArray, IntArray]:
    """Base icosahedron (the reference's table at mesh.py:475-519)."""
    phi = (1.0 + np.sqrt(5.0)) / 2.0
    vertices = np.asarray(
        [(-1, phi, 0), (1, phi, 0), (-1, -phi, 0), (1, -phi, 0), (0, -1, phi), (0, 1, phi),
         (0, -1, -phi), (0, 1, -phi), (phi, 0, -1), (phi, 0, 1), (-phi, 0, -1), (-phi, 0, 1)],
        dtype=np.float64,
    )
    unit = _normalize(vertices)
    return unit, _orient_faces(unit, _ICOSAHEDRON_FACES.copy())


def _polar_orientation(vertices: FloatArray) -> FloatArray:
    """Rotate vertex 5 to the north pole, vertex 9 to 0 deg longitude (mesh.py:522-555)."""
    source = vertices[5]
    target = np.asarray((0.0, 0.0, 1.0))
    axis = np.cross(source, target)
    sine = np.linalg.norm(axis)
    cosine = float(source @ target)
    skew = np.asarray(((0.0, -axis[2], axis[1]), (axis[2], 0.0, -axis[0]),
                       (-axis[1], axis[0], 0.0)))
    rotation = np.eye(3) + skew + skew @ skew * ((1.0 - cosine) / sine**2)
    rotated = vertices @ rotation.T
    longitude = np.arctan2(rotated[9, 1], rotated[9, 0])
    c, s = np.cos(-longitude), np.sin(-longitude)
    twist = np.asarray(((c, -s, 0.0), (s, c, 0.0), (0.0, 0.0, 1.0)))
    return rotated @ twist.T


def _orient_faces(vertices: FloatArray, faces: IntArray) -> IntArray:
    result = faces.copy()
    a, b, c = (vertices[result[:, i]] for i in range(3))
    inward = np.einsum("ij,ij->i", np.cross(b - a, c - a), a + b + c) < 0.0
    result[inward, 1], result[inward, 2] = result[inward, 2].copy(), result[inward, 1].copy()
    return result


def _first_appearance_ids(keys: IntArray) -> tuple[IntArray, IntArray]:
    """Number distinct keys in order of first appearance.

    Returns (id of every key, index of the first occurrence of each id).
    """
    unique, first, inverse = np.unique(keys, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty(len(unique), dtype=np.int64)
    rank[order] = np.arange(len(unique), dtype=np.int64)
    return rank[inverse.reshape(-1)], first[order]


def _half_edge_keys(faces: IntArray, n_vertices: int) -> tuple[IntArray, IntArray, IntArray]:
    tails = faces.reshape(-1)  # (a, b, c) per face
    heads = np.roll(faces, -1, axis=1).reshape(-1)  # (b, c, a)
    lo = np.minimum(tails, heads)
    hi = np.maximum(tails, heads)
    return lo * np.int64(n_vertices) + hi, tails, heads


def _subdivide(vertices: FloatArray, faces: IntArray) -> tuple[FloatArray, IntArray]:
    """Quarter every face; midpoints numbered as the reference (mesh.py:569-593)."""
    n = len(vertices)
    keys, tails, heads = _half_edge_keys(faces, n)
    ids, first = _first_appearance_ids(keys)
    mids = n + ids.reshape(-1, 3)  # columns: ab, bc, ca
    # (v[first] + v[second]) is commutative, so the first occurrence's pair is exact.
    points = _normalize(vertices[tails[first]] + vertices[heads[first]])
    a, b, c = faces[:, 0], faces[:, 1], faces[:, 2]
    ab, bc, ca = mids[:, 0], mids[:, 1], mids[:, 2]
    new_faces = np.stack(
        (np.stack((a, ab, ca), 1), np.stack((b, bc, ab), 1), np.stack((c, ca, bc), 1),
         np.stack((ab, bc, ca), 1)), axis=1).reshape(-1, 3)
    return np.concatenate((vertices, points)), new_faces


def subdivided_icosahedron(subdivision: int, orientation: str) -> tuple[FloatArray, IntArray]:
    vertices, faces = _icosahedron()
    if orientation == "polar":
        vertices = _polar_orientation(vertices)
    for _ in range(subdivision):
        vertices, faces = _subdivide(vertices, faces)
    return vertices, faces


def build_edges(faces: IntArray, n_vertices: int):
    """Edges, face-edge incidence/signs and left/right faces (mesh.py:674-708)."""
    keys, tails, heads = _half_edge_keys(faces, n_vertices)
    ids, first = _first_appearance_ids(keys)
    n_edges = len(first)
    face_edges = ids.reshape(-1, 3)
    signs = np.where(tails < heads, 1, -1).astype(np.int64)
    face_signs = signs.reshape(-1, 3)
    counts = np.bincount(ids, minlength=n_edges)
    plus = np.bincount(ids, weights=(signs > 0), minlength=n_edges)
    if np.any(counts != 2) or np.any(plus != 1):
        raise RuntimeError("geodesic mesh is not a closed two-manifold")
    lo = np.minimum(tails, heads)[first]
    hi = np.maximum(tails, heads)[first]
    edges = np.stack((lo, hi), axis=1).astype(np.int64)
    face_of_half = np.repeat(np.arange(len(faces), dtype=np.int64), 3)
    left = np.empty(n_edges, dtype=np.int64)
    right = np.empty(n_edges, dtype=np.int64)
    left[ids[signs > 0]] = face_of_half[signs > 0]
    right[ids[signs < 0]] = face_of_half[signs < 0]
    return edges, face_edges, face_signs, left, right


def _spherical_face_centers(vertices: FloatArray, faces: IntArray) -> FloatArray:
    a, b, c = (vertices[faces[:, i]] for i in range(3))
    centers = _normalize(np.cross(b - a, c - a))
    reverse = np.einsum("ij,ij->i", centers, a + b + c) < 0.0
    centers[reverse] *= -1.0
    return centers


def _arc_length(first: FloatArray, second: FloatArray) -> FloatArray:
    return np.arctan2(np.linalg.norm(np.cross(first, second), axis=1),
                      np.einsum("ij,ij->i", first, second))


def _signed_area(a: FloatArray, b: FloatArray, c: FloatArray) -> FloatArray:
    """Signed spherical triangle area 2 atan2(a.(b x c), 1 + a.b + b.c + c.a)."""
    numerator = np.einsum("ij,ij->i", a, np.cross(b, c))
    denominator = (1.0 + np.einsum("ij,ij->i", a, b) + np.einsum("ij,ij->i", b, c)
                   + np.einsum("ij,ij->i", c, a))
    return 2.0 * np.arctan2(numerator, denominator)


def _dual_geometry(vertices, centers, edges, left, right):
    """Circumcentric dual areas from edge wedges, accumulated in edge order."""
    lc = centers[left]
    rc = centers[right]
    w_tail = -_signed_area(vertices[edges[:, 0]], lc, rc)
    w_head = _signed_area(vertices[edges[:, 1]], lc, rc)
    # bincount accumulates sequentially in input order: all tails, then heads,
    # which is the reference's two np.add.at passes.
    dual_area = np.bincount(np.concatenate((edges[:, 0], edges[:, 1])),
                            weights=np.concatenate((w_tail, w_head)), minlength=len(vertices))
    degree = np.bincount(edges.ravel(), minlength=len(vertices))
    return dual_area, degree


def assemble_geodesic_mesh(vertices, faces, subdivision, *, topology_kind,
                           face_levels=None, refinement_spec_json=None) -> GeodesicMesh:
    """Orient faces, number edges and compute the metric (mesh.py:397-468)."""
    faces = _orient_faces(vertices, faces)
    centers = _spherical_face_centers(vertices, faces)
    edges, face_edges, face_signs, left, right = build_edges(faces, len(vertices))
    primal = _arc_length(vertices[edges[:, 0]], vertices[edges[:, 1]])
    dual = _arc_length(centers[left], centers[right])
    face_areas = np.abs(_signed_area(vertices[faces[:, 0]], vertices[faces[:, 1]],
                                     vertices[faces[:, 2]]))
    dual_areas, degree = _dual_geometry(vertices, centers, edges, left, right)
    if not np.all(degree >= 3):
        raise RuntimeError("invalid closed geodesic topology")
    if vertices.shape[0] - edges.shape[0] + faces.shape[0] != 2:
        raise RuntimeError("geodesic mesh topology is not a sphere")
    for values, label in ((primal, "primal edges"), (dual, "dual edges"),
                          (face_areas, "primal faces"), (dual_areas, "dual cells")):
        if not np.all(np.isfinite(values)) or not np.all(values > 0.0):
            raise ValueError(f"geodesic mesh contains degenerate {label}")
    if not np.isclose(face_areas.sum(), 4.0 * np.pi, rtol=1e-11):
        raise RuntimeError("primal face areas do not cover the unit sphere")
    if not np.isclose(dual_areas.sum(), 4.0 * np.pi, rtol=1e-10):
        raise RuntimeError("dual cell areas do not cover the unit sphere")
    return GeodesicMesh(
        vertices=vertices, faces=faces, edges=edges, face_edges=face_edges,
        face_edge_signs=face_signs, edge_left_faces=left, edge_right_faces=right,
        face_centers=centers, primal_edge_angles=primal, dual_edge_angles=dual,
        face_solid_angles=face_areas, dual_cell_solid_angles=dual_areas, vertex_degree=degree,
        subdivision=subdivision, topology_kind=topology_kind, face_levels=face_levels,
        refinement_spec_json=refinement_spec_json,
    )


def _relax(vertices: FloatArray, faces: IntArray) -> FloatArray:
    """Face-centroid relaxation (reference mesh.py:596-603)."""
    centers = _normalize(vertices[faces].sum(axis=1))
    accumulated = np.zeros_like(vertices)
    counts = np.zeros(vertices.shape[0], dtype=np.int64)
    for corner in range(3):
        np.add.at(accumulated, faces[:, corner], centers)
        np.add.at(counts, faces[:, corner], 1)
    return _normalize(accumulated / counts[:, None])


def _optimize_edge_lengths(vertices: FloatArray, faces: IntArray, steps: int) -> FloatArray:
    """Projected edge-length-variance descent, pentagons fixed (mesh.py:606-663)."""
    edges = build_edges(faces, len(vertices))[0]
    degree = np.bincount(edges.ravel(), minlength=len(vertices))
    fixed = degree == 5
    result = vertices.copy()
    tails, heads = edges[:, 0], edges[:, 1]
    for _ in range(steps):
        tv, hv = result[tails], result[heads]
        dot = np.clip(np.einsum("ij,ij->i", tv, hv), -1.0, 1.0)
        sine = np.linalg.norm(np.cross(tv, hv), axis=1)
        residual = np.arctan2(sine, dot)
        residual = residual - float(np.mean(residual))
        tg = residual[:, None] * (dot[:, None] * tv - hv) / sine[:, None]
        hg = residual[:, None] * (dot[:, None] * hv - tv) / sine[:, None]
        gradient = np.zeros_like(result)
        np.add.at(gradient, tails, tg)
        np.add.at(gradient, heads, hg)
        gradient /= degree[:, None]
        gradient -= np.einsum("ij,ij->i", gradient, result)[:, None] * result
        gradient[fixed] = 0.0
        result = _normalize(result - gradient)
        result[fixed] = vertices[fixed]
    return result


_DERIVED: list = []


def _mesh_derived(mesh, name: str, compute):
    """Mesh-only arrays computed once and shared (read-only) by every run on
    the mesh; the last few meshes are kept."""
    for cached, values in _DERIVED:
        if cached is mesh:
            break
    else:
        values = {}
        _DERIVED.insert(0, (mesh, values))
        del _DERIVED[4:]
    if name not in values:
        result = compute()
        for array in (result if isinstance(result, tuple) else (result,)):
            array.setflags(write=False)
        values[name] = result
    return values[name]


def vertex_incidence(mesh: GeodesicMesh) -> tuple[IntArray, IntArray]:
    """Padded dual-cell incidence: per vertex its edges in edge order, sign +1
    where the vertex is the tail and -1 where it is the head; degree-5 rows
    padded with (edge 0, sign 0).  Same table as the reference's
    numpy_backend.py:46-76 (stable argsort of the concatenated endpoints).
    Read-only, computed once per mesh."""
    return _mesh_derived(mesh, "vertex_incidence", lambda: _vertex_incidence(mesh))


def _vertex_incidence(mesh: GeodesicMesh) -> tuple[IntArray, IntArray]:
    E = mesh.n_edges
    verts = np.concatenate((mesh.edges[:, 0], mesh.edges[:, 1]))
    incident = np.concatenate((np.arange(E), np.arange(E)))
    signs = np.concatenate((np.ones(E, dtype=np.int64), -np.ones(E, dtype=np.int64)))
    order = np.argsort(verts, kind="stable")
    verts, incident, signs = verts[order], incident[order], signs[order]
    counts = np.bincount(verts, minlength=mesh.n_vertices)
    if not np.array_equal(counts, mesh.vertex_degree):
        raise RuntimeError("mesh vertex degree does not match edge incidence")
    offsets = np.cumsum(np.concatenate(([0], counts[:-1])))
    slots = np.arange(len(verts)) - np.repeat(offsets, counts)
    width = int(counts.max())
    vertex_edges = np.zeros((mesh.n_vertices, width), dtype=np.int64)
    vertex_signs = np.zeros((mesh.n_vertices, width), dtype=np.int64)
    vertex_edges[verts, slots] = incident
    vertex_signs[verts, slots] = signs
    return vertex_edges, vertex_signs

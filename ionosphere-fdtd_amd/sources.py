"""Localised Gaussian current sources and geographic sampling weights.

Same dataclasses, defaults and distribution rules as the reference's
sources.py:14-335: the containing primal face (antipode-safe search), clipped
barycentric weights on its three vertices, and an exact linear split onto the
two adjacent radial planes.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import TYPE_CHECKING

import numpy as np
from numpy.typing import NDArray

if TYPE_CHECKING:
    from .solver import GeodesicFDTD

GWANGJU_LATITUDE_DEG = 35.1595
GWANGJU_LONGITUDE_DEG = 126.8526


def geographic_direction(latitude_deg: float, longitude_deg: float) -> NDArray[np.float64]:
    lat = np.deg2rad(latitude_deg)
    lon = np.deg2rad(longitude_deg)
    return np.asarray((np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)),
                      dtype=np.float64)


def geographic_tangent_basis(latitude_deg: float, longitude_deg: float):
    """Local east and north unit vectors."""
    lat = np.deg2rad(latitude_deg)
    lon = np.deg2rad(longitude_deg)
    east = np.asarray((-np.sin(lon), np.cos(lon), 0.0))
    north = np.asarray((-np.sin(lat) * np.cos(lon), -np.sin(lat) * np.sin(lon), np.cos(lat)))
    return east, north


def destination_point(latitude_deg: float, longitude_deg: float, azimuth_deg: float,
                      distance_m: float, radius_m: float) -> tuple[float, float]:
    """Great-circle destination from a start point, azimuth clockwise from north."""
    delta = distance_m / radius_m
    east, north = geographic_tangent_basis(latitude_deg, longitude_deg)
    start = geographic_direction(latitude_deg, longitude_deg)
    az = np.deg2rad(azimuth_deg)
    tangent = np.cos(az) * north + np.sin(az) * east
    point = np.cos(delta) * start + np.sin(delta) * tangent
    lat = float(np.rad2deg(np.arcsin(np.clip(point[2], -1.0, 1.0))))
    lon = float(np.rad2deg(np.arctan2(point[1], point[0])))
    return lat, lon


_FACE_PLANES: dict = {}
_NEAREST_FACES = 24


def _face_planes(mesh):
    """Per-face edge-plane normals v_e x v_(e+1), computed once per mesh."""
    hit = _FACE_PLANES.get(id(mesh))
    if hit is None or hit[0] is not mesh:
        triangles = mesh.vertices[mesh.faces]
        hit = (mesh, tuple(np.ascontiguousarray(np.cross(triangles[:, e], triangles[:, (e + 1) % 3]))
                           for e in range(3)))
        _FACE_PLANES[id(mesh)] = hit
        if len(_FACE_PLANES) > 16:
            _FACE_PLANES.pop(next(iter(_FACE_PLANES)))
    return hit[1]


def face_index(mesh, latitude_deg: float, longitude_deg: float) -> int:
    """Primal face containing a direction (reference sources.py:53-85): among
    the faces whose edge planes all hold the direction on one side (1e-12
    slack), the one whose centre is nearest; the nearest centre if none does.

    The planes are tested on the faces with the 24 nearest centres first: the
    chosen face has the largest centre . direction of all passing faces, so
    when any of those 24 passes, the choice is the one the full search makes;
    otherwise every face is tested."""
    direction = geographic_direction(latitude_deg, longitude_deg)
    planes = _face_planes(mesh)
    closeness = mesh.face_centers @ direction
    subsets = [None]
    if len(closeness) > 2 * _NEAREST_FACES:
        subsets.insert(0, np.sort(np.argpartition(-closeness, _NEAREST_FACES)[:_NEAREST_FACES]))
    for subset in subsets:
        rows = subset if subset is not None else slice(None)
        signs = np.column_stack(tuple(np.einsum("ij,j->i", p[rows], direction) for p in planes))
        inside = np.all(signs >= -1.0e-12, axis=1) | np.all(signs <= 1.0e-12, axis=1)
        candidates = np.flatnonzero(inside)
        if len(candidates):
            if subset is not None:
                candidates = subset[candidates]
            return int(candidates[np.argmax(mesh.face_centers[candidates] @ direction)])
    return int(np.argmax(closeness))


def geographic_face_index(simulation: "GeodesicFDTD", latitude_deg: float,
                          longitude_deg: float) -> int:
    return face_index(simulation.mesh, latitude_deg, longitude_deg)


def barycentric_distribution(mesh, altitudes_m, latitude_deg: float, longitude_deg: float,
                             altitude_m: float):
    """Face vertices, nearest radial layer and clipped barycentric weights."""
    direction = geographic_direction(latitude_deg, longitude_deg)
    vertices = mesh.faces[face_index(mesh, latitude_deg, longitude_deg)]
    a, b, c = mesh.vertices[vertices]
    normal = np.cross(b - a, c - a)
    point = direction * float((normal @ a) / (normal @ direction))
    ab, ac, off = b - a, c - a, point - a
    d11, d12, d22 = float(ab @ ab), float(ab @ ac), float(ac @ ac)
    o1, o2 = float(off @ ab), float(off @ ac)
    den = d11 * d22 - d12**2
    wb = (d22 * o1 - d12 * o2) / den
    wc = (d11 * o2 - d12 * o1) / den
    weights = np.clip(np.asarray((1.0 - wb - wc, wb, wc)), 0.0, None)
    weights /= weights.sum()
    layer = int(np.argmin(np.abs(np.asarray(altitudes_m) - altitude_m)))
    return vertices.copy(), layer, weights


def geographic_distribution(simulation: "GeodesicFDTD", latitude_deg: float,
                            longitude_deg: float, altitude_m: float):
    return barycentric_distribution(simulation.mesh, simulation.altitudes_m, latitude_deg,
                                    longitude_deg, altitude_m)


def radial_linear_distribution(radial_altitudes_m, altitude_m: float):
    """Altitude represented exactly on the adjacent staggered planes."""
    altitudes = np.asarray(radial_altitudes_m, dtype=np.float64)
    if altitudes.ndim != 1 or len(altitudes) < 1:
        raise ValueError("radial_altitudes_m must be a nonempty 1-D array")
    if not np.all(np.diff(altitudes) > 0.0):
        raise ValueError("radial_altitudes_m must be strictly increasing")
    if altitude_m < altitudes[0] or altitude_m > altitudes[-1]:
        raise ValueError("source altitude is outside the radial grid")
    upper = int(np.searchsorted(altitudes, altitude_m, side="left"))
    if upper < len(altitudes) and altitudes[upper] == altitude_m:
        return np.asarray((upper,), dtype=np.int64), np.asarray((1.0,))
    lower = upper - 1
    w = (altitude_m - altitudes[lower]) / (altitudes[upper] - altitudes[lower])
    return np.asarray((lower, upper), dtype=np.int64), np.asarray((1.0 - w, w))


@dataclass(frozen=True, slots=True)
class GaussianCurrent:
    """Vertical current with a Gaussian (optionally modulated) pulse."""

    latitude_deg: float = GWANGJU_LATITUDE_DEG
    longitude_deg: float = GWANGJU_LONGITUDE_DEG
    altitude_m: float = 2_500.0
    peak_current_a: float = 1.0e6
    vertical_element_length_m: float = 5_000.0
    center_time_s: float | None = None
    one_over_e_half_width_s: float | None = None
    carrier_frequency_hz: float = 0.0

    def __post_init__(self) -> None:
        finite = (self.latitude_deg, self.longitude_deg, self.altitude_m, self.peak_current_a,
                  self.vertical_element_length_m, self.carrier_frequency_hz)
        if not all(np.isfinite(v) for v in finite):
            raise ValueError("source coordinates and waveform values must be finite")
        if not -90.0 <= self.latitude_deg <= 90.0:
            raise ValueError("source latitude must be in [-90, 90]")
        if self.vertical_element_length_m <= 0.0:
            raise ValueError("vertical source element length must be positive")
        if self.center_time_s is not None and not np.isfinite(self.center_time_s):
            raise ValueError("source center time must be finite")
        if self.one_over_e_half_width_s is not None and (
                not np.isfinite(self.one_over_e_half_width_s)
                or self.one_over_e_half_width_s <= 0.0):
            raise ValueError("source half width must be finite and positive")
        if self.carrier_frequency_hz < 0.0:
            raise ValueError("source carrier frequency cannot be negative")

    def direction(self):
        return geographic_direction(self.latitude_deg, self.longitude_deg)

    def location(self, simulation: "GeodesicFDTD") -> tuple[int, int]:
        vertex = int(np.argmax(simulation.mesh.vertices @ self.direction()))
        layer = int(np.argmin(np.abs(simulation.altitudes_m - self.altitude_m)))
        return vertex, layer

    def distribution(self, simulation: "GeodesicFDTD"):
        return geographic_distribution(simulation, self.latitude_deg, self.longitude_deg,
                                       self.altitude_m)

    def staggered_distribution(self, simulation: "GeodesicFDTD"):
        """(vertices, layers, weights): 3 face vertices x 1-2 radial nodes."""
        vertices, _, hw = self.distribution(simulation)
        layers, rw = radial_linear_distribution(simulation.altitudes_m, self.altitude_m)
        n = len(layers)
        return (np.repeat(vertices, n), np.tile(layers, len(vertices)),
                np.repeat(hw, n) * np.tile(rw, len(vertices)))

    def waveform_parameters(self, dt_s: float) -> tuple[float, float]:
        """(half width, center) of the envelope (reference sources.py:238-250)."""
        if self.one_over_e_half_width_s is not None:
            half = self.one_over_e_half_width_s
        elif self.carrier_frequency_hz:
            half = 0.5 / self.carrier_frequency_hz
        else:
            half = 12.0 * dt_s
        center = self.center_time_s if self.center_time_s is not None else max(4.0 * half,
                                                                               36.0 * dt_s)
        return half, center

    def current_a(self, time_s: float, dt_s: float) -> float:
        half, center = self.waveform_parameters(dt_s)
        envelope = np.exp(-((time_s - center) / half) ** 2)
        if self.carrier_frequency_hz:
            envelope *= np.cos(2.0 * np.pi * self.carrier_frequency_hz * (time_s - center))
        return float(self.peak_current_a * envelope)

    def currents_a(self, times_s: NDArray[np.float64], dt_s: float) -> NDArray[np.float64]:
        """Vectorised current_a over many times (FP64)."""
        half, center = self.waveform_parameters(dt_s)
        t = np.asarray(times_s, dtype=np.float64)
        envelope = np.exp(-((t - center) / half) ** 2)
        if self.carrier_frequency_hz:
            envelope = envelope * np.cos(2.0 * np.pi * self.carrier_frequency_hz * (t - center))
        return self.peak_current_a * envelope


@dataclass(frozen=True, slots=True)
class TangentialGaussianCurrent(GaussianCurrent):
    """Gaussian current along horizontal ground lines (reference sources.py:257-335)."""

    altitude_m: float = 0.0
    azimuths_deg: tuple[float, ...] = (0.0,)
    line_lengths_m: tuple[float, ...] | None = None
    edge_assignment: str = "projected"

    def __post_init__(self) -> None:
        GaussianCurrent.__post_init__(self)
        if not self.azimuths_deg:
            raise ValueError("azimuths_deg must contain at least one direction")
        if not all(np.isfinite(v) for v in self.azimuths_deg):
            raise ValueError("source azimuths must be finite")
        if self.line_lengths_m is not None:
            if len(self.line_lengths_m) != len(self.azimuths_deg):
                raise ValueError("line_lengths_m must match azimuths_deg")
            if not all(np.isfinite(v) and v > 0.0 for v in self.line_lengths_m):
                raise ValueError("ground-line lengths must be finite and positive")
        if self.edge_assignment not in {"projected", "nearest"}:
            raise ValueError("edge_assignment must be 'projected' or 'nearest'")

    def edge_distribution(self, simulation: "GeodesicFDTD"):
        mesh = simulation.mesh
        face = face_index(mesh, self.latitude_deg, self.longitude_deg)
        edges = mesh.face_edges[face].copy()
        ends = mesh.vertices[mesh.edges[edges]]
        dirs = ends[:, 1] - ends[:, 0]
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
        east, north = geographic_tangent_basis(self.latitude_deg, self.longitude_deg)
        lengths = mesh.primal_edge_angles[edges] * (simulation.config.earth_radius_m
                                                    + self.altitude_m)
        weights = np.zeros(len(edges))
        lines = self.line_lengths_m or tuple(1.0 for _ in self.azimuths_deg)
        for az_deg, line in zip(self.azimuths_deg, lines):
            az = np.deg2rad(az_deg)
            proj = dirs @ (np.cos(az) * north + np.sin(az) * east)
            if self.edge_assignment == "projected":
                weights += line * proj / lengths
            else:
                k = int(np.argmax(np.abs(proj)))
                weights[k] += line * proj[k] / lengths[k]
        layers, rw = radial_linear_distribution(simulation.radial_midpoint_altitudes_m,
                                                self.altitude_m)
        n = len(layers)
        return (np.repeat(edges, n), np.tile(layers, len(edges)),
                np.repeat(weights, n) * np.tile(rw, len(edges)))

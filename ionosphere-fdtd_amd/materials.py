"""Earth-ionosphere material model (reference materials.py:22-326).

The solver accepts any object with ``sample(directions, altitudes_m,
earth_radius_m) -> (sigma, epsilon_r)`` (or a ``sample_mesh`` hook, and
optionally ``sample_tangential_cells``), exactly as the reference does
(solver.py:467-629).  This module provides the data-free exponential profile
used by every benchmark configuration, spherical lithosphere anomalies, and
the conservative area rasterisation of anomalies that
``horizontal_anomaly_mode="conservative-nearest"`` uses.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from numpy.typing import NDArray

from .constants import EPSILON_0

FloatArray = NDArray[np.float64]


@dataclass(frozen=True, slots=True)
class SphericalAnomaly:
    """Conductivity anomaly in a spherical-cap volume: inside the cap of arc
    radius ``radius_m`` around (latitude, longitude) and between the two
    altitudes, sigma is multiplied by ``conductivity_factor`` (only where it is
    at most ``maximum_background_conductivity_s_m``, when that is set) and
    epsilon_r replaced by ``relative_permittivity`` (when set).
    ``target_area_m2`` overrides the cap area that conservative rasterisation
    preserves.  Fields and messages are the reference interface
    (materials.py:22-86)."""

    latitude_deg: float
    longitude_deg: float
    radius_m: float
    altitude_min_m: float
    altitude_max_m: float
    conductivity_factor: float
    relative_permittivity: float | None = None
    maximum_background_conductivity_s_m: float | None = None
    target_area_m2: float | None = None

    def __post_init__(self) -> None:
        core = np.array((self.latitude_deg, self.longitude_deg, self.radius_m,
                         self.altitude_min_m, self.altitude_max_m, self.conductivity_factor),
                        dtype=np.float64)
        if not np.all(np.isfinite(core)):
            raise ValueError("anomaly geometry and conductivity must be finite")
        if abs(self.latitude_deg) > 90.0:
            raise ValueError("anomaly latitude must be in [-90, 90]")
        if not self.radius_m > 0.0:
            raise ValueError("anomaly radius_m must be positive")
        if self.altitude_max_m < self.altitude_min_m:
            raise ValueError("anomaly altitude bounds are reversed")
        if not self.conductivity_factor > 0.0:
            raise ValueError("conductivity_factor must be positive")
        if not _positive_or_none(self.relative_permittivity):
            raise ValueError("relative permittivity must be finite and positive")
        if not _positive_or_none(self.maximum_background_conductivity_s_m):
            raise ValueError("maximum background conductivity must be finite and positive")
        if not _positive_or_none(self.target_area_m2):
            raise ValueError("target anomaly area must be finite and positive")

    @property
    def center(self) -> FloatArray:
        """Unit vector of the cap centre."""
        lat, lon = np.radians(self.latitude_deg), np.radians(self.longitude_deg)
        return np.array((np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)))

    def arc_distance(self, directions: FloatArray) -> FloatArray:
        """Great-circle angle (rad) from the cap centre to unit directions."""
        return np.arccos(np.clip(np.asarray(directions) @ self.center, -1.0, 1.0))

    def cap_solid_angle(self, earth_radius_m: float) -> float:
        """Solid angle the rasterisation preserves: target_area_m2 / R^2, else
        the cap's own 2 pi (1 - cos(radius / R))."""
        if self.target_area_m2 is not None:
            return self.target_area_m2 / earth_radius_m**2
        return 2.0 * np.pi * (1.0 - np.cos(self.radius_m / earth_radius_m))


def _positive_or_none(value) -> bool:
    return value is None or (bool(np.isfinite(value)) and value > 0.0)


def blend_anomaly(sigma: FloatArray, epsilon_r: FloatArray, fraction: FloatArray,
                  anomaly: SphericalAnomaly) -> None:
    """Mix a fraction f of anomalous material into each cell, in place:
    sigma <- sigma (1 + f (factor - 1)), eps <- (1 - f) eps + f eps_anomaly;
    cells above the background-conductivity ceiling keep their material
    (reference materials.py:224-239)."""
    f = fraction
    if anomaly.maximum_background_conductivity_s_m is not None:
        f = np.where(sigma <= anomaly.maximum_background_conductivity_s_m, f, 0.0)
    sigma *= 1.0 + f * (anomaly.conductivity_factor - 1.0)
    if anomaly.relative_permittivity is not None:
        epsilon_r *= 1.0 - f
        epsilon_r += f * anomaly.relative_permittivity


def point_anomaly_overrides(sigma: FloatArray, epsilon_r: FloatArray, directions: FloatArray,
                            altitudes_m: FloatArray, earth_radius_m: float,
                            anomalies) -> None:
    """Point sampling of spherical anomalies, in place: every sample inside a
    cap (arc distance <= radius / R) and its altitude band takes the anomaly
    (reference materials.py:89-112)."""
    for anomaly in anomalies:
        horizontal = anomaly.arc_distance(directions) <= anomaly.radius_m / earth_radius_m
        vertical = (altitudes_m >= anomaly.altitude_min_m) & (altitudes_m <= anomaly.altitude_max_m)
        inside = np.outer(horizontal, vertical)
        if anomaly.maximum_background_conductivity_s_m is not None:
            inside &= sigma <= anomaly.maximum_background_conductivity_s_m
        sigma[inside] *= anomaly.conductivity_factor
        if anomaly.relative_permittivity is not None:
            epsilon_r[inside] = anomaly.relative_permittivity


def conservative_anomaly_fractions(directions: FloatArray, support_solid_angles: FloatArray,
                                   anomaly: SphericalAnomaly, earth_radius_m: float) -> FloatArray:
    """Fraction of every support area inside a rasterised anomaly whose total
    area is exactly the anomaly's (reference materials.py:151-182).

    The supports are filled whole in order of arc distance from the cap centre
    (stable for ties) until the area is used up; the support that crosses the
    remaining area is filled partially.
    """
    points = np.asarray(directions, dtype=np.float64)
    areas = np.asarray(support_solid_angles, dtype=np.float64)
    if points.shape != (len(areas), 3) or np.any(areas <= 0.0):
        raise ValueError("anomaly supports must have positive areas and directions")
    target = anomaly.cap_solid_angle(earth_radius_m)
    if target > float(np.sum(areas)):
        raise ValueError("anomaly area exceeds the available spherical supports")
    order = np.argsort(anomaly.arc_distance(points), kind="stable")
    fractions = np.zeros(len(areas), dtype=np.float64)
    left = target
    for i in order:
        if left <= 0.0:
            break
        take = min(float(areas[i]), left)
        fractions[i] = take / areas[i]
        left -= take
    if left > 64.0 * np.finfo(np.float64).eps * target:
        raise RuntimeError("conservative anomaly rasterization did not close")
    return fractions


def blend_node_anomalies(sigma, epsilon_r, altitudes_m, anomalies, horizontal_fractions) -> None:
    """Conservative anomalies on radial-node samples: the horizontal fraction
    on every node inside the altitude band (reference materials.py:185-199)."""
    for anomaly, horizontal in zip(anomalies, horizontal_fractions, strict=True):
        band = (altitudes_m >= anomaly.altitude_min_m) & (altitudes_m <= anomaly.altitude_max_m)
        blend_anomaly(sigma, epsilon_r, np.outer(horizontal, band.astype(np.float64)), anomaly)


def blend_cell_anomalies(sigma, epsilon_r, lower_altitudes_m, upper_altitudes_m, anomalies,
                         horizontal_fractions) -> None:
    """Conservative anomalies on radial cells: horizontal fraction times the
    cell's overlap with the altitude band (reference materials.py:202-221)."""
    width = upper_altitudes_m - lower_altitudes_m
    for anomaly, horizontal in zip(anomalies, horizontal_fractions, strict=True):
        overlap = np.minimum(upper_altitudes_m, anomaly.altitude_max_m) - \
            np.maximum(lower_altitudes_m, anomaly.altitude_min_m)
        radial = np.clip(overlap, 0.0, width) / width
        blend_anomaly(sigma, epsilon_r, np.outer(horizontal, radial), anomaly)


@dataclass(frozen=True, slots=True)
class EarthIonosphereMaterial:
    """Lossy ground below 0 m, ``prefactor * eps0 * exp((h - H')/zeta)`` above,
    with point-sampled spherical ``anomalies`` (reference materials.py:242-326)."""

    lithosphere_conductivity_s_m: float = 1.0e-3
    lithosphere_relative_permittivity: float = 10.0
    atmosphere_relative_permittivity: float = 1.0
    ionosphere_reference_height_m: float = 74_000.0
    ionosphere_scale_height_m: float = 6_000.0
    ionosphere_prefactor_hz: float = 2.5e5
    anomalies: tuple[SphericalAnomaly, ...] = field(default_factory=tuple)

    def __post_init__(self) -> None:
        values = (self.lithosphere_conductivity_s_m, self.lithosphere_relative_permittivity,
                  self.atmosphere_relative_permittivity, self.ionosphere_reference_height_m,
                  self.ionosphere_scale_height_m, self.ionosphere_prefactor_hz)
        if not all(np.isfinite(v) for v in values):
            raise ValueError("material parameters must be finite")
        if self.lithosphere_conductivity_s_m < 0.0:
            raise ValueError("lithosphere conductivity cannot be negative")
        if self.lithosphere_relative_permittivity < 1.0:
            raise ValueError("lithosphere relative permittivity must be >= 1")
        if self.atmosphere_relative_permittivity < 1.0:
            raise ValueError("atmosphere relative permittivity must be >= 1")
        if self.ionosphere_scale_height_m <= 0.0:
            raise ValueError("ionosphere scale height must be positive")
        if self.ionosphere_prefactor_hz < 0.0:
            raise ValueError("ionosphere prefactor cannot be negative")

    def sample(self, directions: FloatArray, altitudes_m: FloatArray,
               earth_radius_m: float) -> tuple[FloatArray, FloatArray]:
        """(conductivity, relative permittivity) on a direction x altitude grid."""
        altitudes = np.asarray(altitudes_m, dtype=np.float64)
        exponent = np.clip((altitudes - self.ionosphere_reference_height_m)
                           / self.ionosphere_scale_height_m, -80.0, 80.0)
        sigma_air = self.ionosphere_prefactor_hz * EPSILON_0 * np.exp(exponent)
        ground = altitudes < 0.0
        sigma = np.where(ground, self.lithosphere_conductivity_s_m, sigma_air)
        eps = np.where(ground, self.lithosphere_relative_permittivity,
                       self.atmosphere_relative_permittivity)
        shape = (len(directions), len(altitudes))
        sigma = np.broadcast_to(sigma, shape).copy()
        eps = np.broadcast_to(eps, shape).copy()
        point_anomaly_overrides(sigma, eps, np.asarray(directions, dtype=np.float64), altitudes,
                                earth_radius_m, self.anomalies)
        return sigma, eps

    @classmethod
    def from_conductivity_profile(cls, conductivity_s_m_at_reference: float,
                                  reference_height_m: float, scale_height_m: float,
                                  **kwargs) -> "EarthIonosphereMaterial":
        """sigma(z) = sigma_ref * exp((z - z_ref)/zeta) above ground."""
        return cls(ionosphere_prefactor_hz=conductivity_s_m_at_reference / EPSILON_0,
                   ionosphere_reference_height_m=reference_height_m,
                   ionosphere_scale_height_m=scale_height_m, **kwargs)

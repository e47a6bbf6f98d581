"""Earth-ionosphere material model (reference materials.py:242-326).

The solver accepts any object with ``sample(directions, altitudes_m,
earth_radius_m) -> (sigma, epsilon_r)`` (or a ``sample_mesh`` hook), exactly as
the reference does (solver.py:467-629); this module provides the data-free
exponential profile used by every benchmark configuration.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from numpy.typing import NDArray

from .constants import EPSILON_0

FloatArray = NDArray[np.float64]


@dataclass(frozen=True, slots=True)
class EarthIonosphereMaterial:
    """Lossy ground below 0 m, ``prefactor * eps0 * exp((h - H')/zeta)`` above.

    ``anomalies`` is kept for signature compatibility; only an empty tuple is
    supported by this package.
    """

    lithosphere_conductivity_s_m: float = 1.0e-3
    lithosphere_relative_permittivity: float = 10.0
    atmosphere_relative_permittivity: float = 1.0
    ionosphere_reference_height_m: float = 74_000.0
    ionosphere_scale_height_m: float = 6_000.0
    ionosphere_prefactor_hz: float = 2.5e5
    anomalies: tuple = field(default_factory=tuple)

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
        if tuple(self.anomalies):
            raise ValueError("spherical anomalies are not supported by this package")

    def sample(self, directions: FloatArray, altitudes_m: FloatArray,
               earth_radius_m: float) -> tuple[FloatArray, FloatArray]:
        """(conductivity, relative permittivity) on a direction x altitude grid."""
        del earth_radius_m
        altitudes = np.asarray(altitudes_m, dtype=np.float64)
        exponent = np.clip((altitudes - self.ionosphere_reference_height_m)
                           / self.ionosphere_scale_height_m, -80.0, 80.0)
        sigma_air = self.ionosphere_prefactor_hz * EPSILON_0 * np.exp(exponent)
        ground = altitudes < 0.0
        sigma = np.where(ground, self.lithosphere_conductivity_s_m, sigma_air)
        eps = np.where(ground, self.lithosphere_relative_permittivity,
                       self.atmosphere_relative_permittivity)
        shape = (len(directions), len(altitudes))
        return np.broadcast_to(sigma, shape).copy(), np.broadcast_to(eps, shape).copy()

    @classmethod
    def from_conductivity_profile(cls, conductivity_s_m_at_reference: float,
                                  reference_height_m: float, scale_height_m: float,
                                  **kwargs) -> "EarthIonosphereMaterial":
        """sigma(z) = sigma_ref * exp((z - z_ref)/zeta) above ground."""
        return cls(ionosphere_prefactor_hz=conductivity_s_m_at_reference / EPSILON_0,
                   ionosphere_reference_height_m=reference_height_m,
                   ionosphere_scale_height_m=scale_height_m, **kwargs)

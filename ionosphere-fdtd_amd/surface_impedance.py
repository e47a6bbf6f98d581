"""Diffusive conductive half-space impedance and its ADE coefficients.

``ConductiveHalfSpaceSurface`` has the reference's fields, pole grid and
quadrature weights (surface_impedance.py:14-170).  ``surface_ade_tables``
gives the FP64 coefficients of the trapezoidal lower-boundary ADE
(surface_impedance.py:176-203) that the HIP H kernel evaluates.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any

import numpy as np
from numpy.typing import NDArray

from .constants import MU_0


@dataclass(frozen=True, slots=True)
class ConductiveHalfSpaceSurface:
    r"""Positive-real fit of :math:`Z(s)=\sqrt{\mu s/\sigma}` on a log pole grid."""

    conductivity_s_m: float | NDArray[np.float64]
    minimum_frequency_hz: float = 5.0
    maximum_frequency_hz: float = 45.0
    terms: int = 16
    pole_span_decades: float = 5.0
    permeability_h_m: float = MU_0

    def __post_init__(self) -> None:
        conductivity = np.asarray(self.conductivity_s_m, dtype=np.float64)
        if conductivity.ndim > 1 or conductivity.size == 0:
            raise ValueError("surface conductivity must be scalar or one-dimensional")
        conductivity = conductivity.reshape(-1)
        if not np.all(np.isfinite(conductivity)) or np.any(conductivity <= 0.0):
            raise ValueError("surface conductivity must be finite and positive")
        values = (self.minimum_frequency_hz, self.maximum_frequency_hz, self.pole_span_decades,
                  self.permeability_h_m)
        if not all(np.isfinite(v) for v in values):
            raise ValueError("surface impedance controls must be finite")
        if not 0.0 < self.minimum_frequency_hz < self.maximum_frequency_hz:
            raise ValueError("surface impedance frequency band is invalid")
        if isinstance(self.terms, bool) or not isinstance(self.terms, (int, np.integer)) \
                or self.terms < 1:
            raise ValueError("surface impedance terms must be a positive integer")
        if self.pole_span_decades <= 0.0 or self.permeability_h_m <= 0.0:
            raise ValueError("surface impedance span and permeability must be positive")
        conductivity = np.array(conductivity, copy=True)
        conductivity.setflags(write=False)
        object.__setattr__(self, "conductivity_s_m", conductivity)

    @property
    def pole_rates_s_inv(self) -> NDArray[np.float64]:
        """Log-spaced pole rates, the band extended by pole_span_decades each side."""
        lower = np.log(2.0 * np.pi * self.minimum_frequency_hz)
        upper = np.log(2.0 * np.pi * self.maximum_frequency_hz)
        ext = self.pole_span_decades * np.log(10.0)
        lower -= ext
        upper += ext
        spacing = (upper - lower) / self.terms
        return np.exp(lower + (np.arange(self.terms) + 0.5) * spacing)

    @property
    def diffusive_weights_sqrt_s_inv(self) -> NDArray[np.float64]:
        """Midpoint quadrature of sqrt(s) = (1/pi) int s/(s+p) p^-1/2 dp."""
        poles = self.pole_rates_s_inv
        if self.terms > 1:
            spacing = np.log(poles[-1] / poles[0]) / (self.terms - 1)
        else:
            spacing = (np.log(self.maximum_frequency_hz / self.minimum_frequency_hz)
                       + 2.0 * self.pole_span_decades * np.log(10.0))
        return spacing * np.sqrt(poles) / np.pi

    def conductivity_for_edges(self, edge_count: int) -> NDArray[np.float64]:
        c = np.asarray(self.conductivity_s_m)
        if c.size == 1:
            return np.full(edge_count, c[0])
        if c.size != edge_count:
            raise ValueError("surface conductivity does not match mesh edges")
        return np.array(c, copy=True)

    def impedance_ohm(self, frequency_hz) -> NDArray[np.complex128]:
        f = np.asarray(frequency_hz, dtype=np.float64)
        if not np.all(np.isfinite(f)) or np.any(f <= 0.0):
            raise ValueError("surface impedance frequencies must be positive")
        s = 2j * np.pi * f.reshape(-1)
        root = np.sum(self.diffusive_weights_sqrt_s_inv[:, None] * s[None, :]
                      / (s[None, :] + self.pole_rates_s_inv[:, None]), axis=0)
        scale = np.sqrt(self.permeability_h_m / np.asarray(self.conductivity_s_m))
        return scale[:, None] * root[None, :]

    def exact_impedance_ohm(self, frequency_hz) -> NDArray[np.complex128]:
        f = np.asarray(frequency_hz, dtype=np.float64)
        if not np.all(np.isfinite(f)) or np.any(f <= 0.0):
            raise ValueError("surface impedance frequencies must be positive")
        s = 2j * np.pi * f.reshape(-1)
        scale = np.sqrt(self.permeability_h_m / np.asarray(self.conductivity_s_m))
        return scale[:, None] * np.sqrt(s)[None, :]

    def to_metadata(self) -> dict[str, Any]:
        return {
            "type": "conductive-half-space-diffusive",
            "conductivity_s_m": np.asarray(self.conductivity_s_m).tolist(),
            "minimum_frequency_hz": self.minimum_frequency_hz,
            "maximum_frequency_hz": self.maximum_frequency_hz,
            "terms": int(self.terms),
            "pole_span_decades": self.pole_span_decades,
            "permeability_h_m": self.permeability_h_m,
        }

    @classmethod
    def from_metadata(cls, metadata: dict[str, Any]) -> "ConductiveHalfSpaceSurface":
        if metadata.get("type") != "conductive-half-space-diffusive":
            raise ValueError("unsupported surface impedance model")
        return cls(conductivity_s_m=np.asarray(metadata["conductivity_s_m"]),
                   minimum_frequency_hz=metadata["minimum_frequency_hz"],
                   maximum_frequency_hz=metadata["maximum_frequency_hz"],
                   terms=metadata["terms"], pole_span_decades=metadata["pole_span_decades"],
                   permeability_h_m=metadata["permeability_h_m"])


def surface_ade_tables(model: ConductiveHalfSpaceSurface, edge_count: int,
                       time_step_s: float) -> dict[str, NDArray[np.float64]]:
    """Trapezoidal ADE coefficients: memory decay (1-a)/(1+a), drive a/(1+a),
    history weights w/(1+a) with a = dt p / 2, and the per-edge scale
    sqrt(mu/sigma) (one value when the conductivity is uniform)."""
    poles = model.pole_rates_s_inv
    weights = model.diffusive_weights_sqrt_s_inv
    alpha = 0.5 * time_step_s * poles
    conductivity = np.asarray(model.conductivity_s_m).reshape(-1)
    if conductivity.size not in (1, edge_count):
        raise ValueError("surface conductivity does not match mesh edges")
    return {
        "decay": (1.0 - alpha) / (1.0 + alpha),
        "drive": alpha / (1.0 + alpha),
        "history_weights": weights / (1.0 + alpha),
        "scale": np.sqrt(model.permeability_h_m / conductivity),
    }

"""Static 2:1-balanced dyadic radial refinement (reference radial_grid.py:12-165)."""

from __future__ import annotations

from dataclasses import dataclass
import math

import numpy as np
from numpy.typing import NDArray


@dataclass(frozen=True, slots=True)
class RadialRefinementRegion:
    """Altitude interval with a requested maximum radial cell size."""

    minimum_altitude_m: float
    maximum_altitude_m: float
    maximum_step_m: float

    def __post_init__(self) -> None:
        values = (self.minimum_altitude_m, self.maximum_altitude_m, self.maximum_step_m)
        if not all(np.isfinite(v) for v in values):
            raise ValueError("radial refinement values must be finite")
        if self.minimum_altitude_m >= self.maximum_altitude_m:
            raise ValueError("radial refinement bounds are reversed")
        if self.maximum_step_m <= 0.0:
            raise ValueError("radial refinement step must be positive")


@dataclass(frozen=True, slots=True)
class RadialGridValidation:
    cells: int
    minimum_step_m: float
    maximum_step_m: float
    maximum_adjacent_step_ratio: float


def build_refined_radial_grid(
    minimum_altitude_m: float,
    maximum_altitude_m: float,
    background_step_m: float,
    regions: tuple[RadialRefinementRegion, ...],
) -> tuple[float, ...]:
    """Halve base cells until every cell touching a region meets its size, then
    split coarser neighbours until adjacent levels differ by at most one."""

    values = (minimum_altitude_m, maximum_altitude_m, background_step_m)
    if not all(np.isfinite(v) for v in values):
        raise ValueError("radial grid bounds and spacing must be finite")
    if minimum_altitude_m >= maximum_altitude_m:
        raise ValueError("radial grid bounds are reversed")
    if background_step_m <= 0.0:
        raise ValueError("background radial step must be positive")
    regions = tuple(regions)
    if not all(isinstance(r, RadialRefinementRegion) for r in regions):
        raise TypeError("regions must contain RadialRefinementRegion values")
    for r in regions:
        if r.minimum_altitude_m < minimum_altitude_m or r.maximum_altitude_m > maximum_altitude_m:
            raise ValueError("radial refinement region lies outside the grid")
    base = int(math.ceil((maximum_altitude_m - minimum_altitude_m) / background_step_m))
    nodes = np.linspace(minimum_altitude_m, maximum_altitude_m, base + 1)
    cells = [(float(nodes[i]), float(nodes[i + 1]), 0) for i in range(base)]

    def target(lower: float, upper: float) -> float:
        sizes = [r.maximum_step_m for r in regions
                 if upper > r.minimum_altitude_m and lower < r.maximum_altitude_m]
        return min(sizes, default=math.inf)

    changed = True
    while changed:
        changed = False
        refined = []
        for lower, upper, level in cells:
            if upper - lower > target(lower, upper) * (1.0 + 1.0e-12):
                mid = 0.5 * (lower + upper)
                refined += [(lower, mid, level + 1), (mid, upper, level + 1)]
                changed = True
            else:
                refined.append((lower, upper, level))
        cells = refined
    while True:
        split = None
        for i in range(len(cells) - 1):
            if abs(cells[i][2] - cells[i + 1][2]) > 1:
                split = i if cells[i][2] < cells[i + 1][2] else i + 1
                break
        if split is None:
            break
        lower, upper, level = cells[split]
        mid = 0.5 * (lower + upper)
        cells[split:split + 1] = [(lower, mid, level + 1), (mid, upper, level + 1)]
    out = (cells[0][0], *(c[1] for c in cells))
    validate_radial_grid(out, maximum_adjacent_step_ratio=2.0)
    return tuple(float(v) for v in out)


def validate_radial_grid(
    altitudes_m: tuple[float, ...] | NDArray[np.float64],
    *,
    maximum_adjacent_step_ratio: float = 2.0,
) -> RadialGridValidation:
    altitudes = np.asarray(altitudes_m, dtype=np.float64)
    if (altitudes.ndim != 1 or len(altitudes) < 3 or not np.all(np.isfinite(altitudes))
            or not np.all(np.diff(altitudes) > 0.0)):
        raise ValueError("radial grid must contain finite increasing nodes")
    if not np.isfinite(maximum_adjacent_step_ratio) or maximum_adjacent_step_ratio < 1.0:
        raise ValueError("maximum adjacent step ratio must be at least one")
    steps = np.diff(altitudes)
    ratios = np.maximum(steps[:-1] / steps[1:], steps[1:] / steps[:-1])
    ratio = float(np.max(ratios))
    if ratio > maximum_adjacent_step_ratio * (1.0 + 1.0e-12):
        raise ValueError("radial grid violates the adjacent-step ratio")
    return RadialGridValidation(len(steps), float(steps.min()), float(steps.max()), ratio)

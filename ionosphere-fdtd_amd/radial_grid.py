"""Static dyadic radial refinement with 2:1 balance (reference radial_grid.py:12-165).

The public names, fields and error messages are the reference's interface.
The construction is this module's own: the grid is a row of leaves of
per-base-cell bisection trees held as three NumPy arrays (lower node, upper
node, level), refined in whole-array sweeps.

1. Resolution sweeps: every leaf wider than the finest step requested by
   the regions it overlaps (open-interval overlap) is bisected; repeat
   until no leaf is too wide.
2. Balance sweeps: a leaf whose level is more than one below a neighbour's
   must be bisected in *every* 2:1-balanced refinement of the current tree,
   so bisecting all such leaves at once and repeating until none is left
   gives the coarsest balanced refinement.  That refinement is unique, so
   the node set does not depend on the order in which forced splits are
   made.

A bisected leaf [a, b] always gets the node 0.5 * (a + b), so every node is
the same floating-point value whichever order produced it.
"""

from __future__ import annotations

from dataclasses import dataclass
import math

import numpy as np
from numpy.typing import NDArray

# relative slack when comparing a leaf width with a requested step
_WIDTH_SLACK = 1.0e-12


@dataclass(frozen=True, slots=True)
class RadialRefinementRegion:
    """Altitude band [minimum_altitude_m, maximum_altitude_m] whose radial cells
    must be no wider than maximum_step_m."""

    minimum_altitude_m: float
    maximum_altitude_m: float
    maximum_step_m: float

    def __post_init__(self) -> None:
        bounds = (self.minimum_altitude_m, self.maximum_altitude_m, self.maximum_step_m)
        if not np.all(np.isfinite(np.asarray(bounds, dtype=np.float64))):
            raise ValueError("radial refinement values must be finite")
        if self.minimum_altitude_m >= self.maximum_altitude_m:
            raise ValueError("radial refinement bounds are reversed")
        if self.maximum_step_m <= 0.0:
            raise ValueError("radial refinement step must be positive")


@dataclass(frozen=True, slots=True)
class RadialGridValidation:
    """Summary of a validated grid: cell count, extreme steps, worst neighbour ratio."""

    cells: int
    minimum_step_m: float
    maximum_step_m: float
    maximum_adjacent_step_ratio: float


def _bisect(lower, upper, level, mask):
    """Replace every leaf selected by ``mask`` by its two halves, in place order."""
    count = np.where(mask, 2, 1)
    lo = np.repeat(lower, count)
    hi = np.repeat(upper, count)
    lv = np.repeat(level, count)
    # the first copy of a split leaf is its lower half, the second its upper half
    first = np.cumsum(count) - count
    split_at = first[mask]
    mid = 0.5 * (lower[mask] + upper[mask])
    hi[split_at] = mid
    lo[split_at + 1] = mid
    lv[split_at] += 1
    lv[split_at + 1] += 1
    return lo, hi, lv


def _requested_steps(lower, upper, regions):
    """Finest step of the regions each leaf's open interval overlaps (inf if none)."""
    want = np.full(lower.shape, np.inf)
    for region in regions:
        overlap = (upper > region.minimum_altitude_m) & (lower < region.maximum_altitude_m)
        want = np.where(overlap, np.minimum(want, region.maximum_step_m), want)
    return want


def build_refined_radial_grid(
    minimum_altitude_m: float,
    maximum_altitude_m: float,
    background_step_m: float,
    regions: tuple[RadialRefinementRegion, ...],
) -> tuple[float, ...]:
    """Nodes of a dyadic grid that resolves every region and is 2:1 balanced.

    The domain is first cut into ceil(height / background_step_m) equal base
    cells; the leaves are then refined and balanced as the module docstring
    describes.
    """
    limits = np.asarray((minimum_altitude_m, maximum_altitude_m, background_step_m),
                        dtype=np.float64)
    if not np.all(np.isfinite(limits)):
        raise ValueError("radial grid bounds and spacing must be finite")
    if minimum_altitude_m >= maximum_altitude_m:
        raise ValueError("radial grid bounds are reversed")
    if background_step_m <= 0.0:
        raise ValueError("background radial step must be positive")
    regions = tuple(regions)
    if any(not isinstance(r, RadialRefinementRegion) for r in regions):
        raise TypeError("regions must contain RadialRefinementRegion values")
    if any(r.minimum_altitude_m < minimum_altitude_m or r.maximum_altitude_m > maximum_altitude_m
           for r in regions):
        raise ValueError("radial refinement region lies outside the grid")

    n_base = int(math.ceil((maximum_altitude_m - minimum_altitude_m) / background_step_m))
    base_nodes = np.linspace(minimum_altitude_m, maximum_altitude_m, n_base + 1)
    lower, upper = base_nodes[:-1].copy(), base_nodes[1:].copy()
    level = np.zeros(n_base, dtype=np.int64)

    while True:  # resolution sweeps
        too_wide = (upper - lower) > _requested_steps(lower, upper, regions) * (1.0 + _WIDTH_SLACK)
        if not too_wide.any():
            break
        lower, upper, level = _bisect(lower, upper, level, too_wide)

    while True:  # balance sweeps: bisect every leaf a neighbour forces
        finer_neighbour = np.zeros_like(level)
        finer_neighbour[:-1] = np.maximum(finer_neighbour[:-1], level[1:])
        finer_neighbour[1:] = np.maximum(finer_neighbour[1:], level[:-1])
        forced = finer_neighbour - level > 1
        if not forced.any():
            break
        lower, upper, level = _bisect(lower, upper, level, forced)

    nodes = np.append(lower, upper[-1])
    validate_radial_grid(nodes, maximum_adjacent_step_ratio=2.0)
    return tuple(float(x) for x in nodes)


def validate_radial_grid(
    altitudes_m: tuple[float, ...] | NDArray[np.float64],
    *,
    maximum_adjacent_step_ratio: float = 2.0,
) -> RadialGridValidation:
    """Check finite, strictly increasing nodes (at least three) whose adjacent
    steps differ by no more than ``maximum_adjacent_step_ratio``."""
    nodes = np.asarray(altitudes_m, dtype=np.float64)
    valid = nodes.ndim == 1 and nodes.size >= 3 and bool(np.isfinite(nodes).all())
    if not valid or not bool((np.diff(nodes) > 0.0).all()):
        raise ValueError("radial grid must contain finite increasing nodes")
    if not math.isfinite(maximum_adjacent_step_ratio) or maximum_adjacent_step_ratio < 1.0:
        raise ValueError("maximum adjacent step ratio must be at least one")
    widths = np.diff(nodes)
    pair = np.stack((widths[:-1], widths[1:]))
    worst = float((pair.max(axis=0) / pair.min(axis=0)).max())
    if worst > maximum_adjacent_step_ratio * (1.0 + _WIDTH_SLACK):
        raise ValueError("radial grid violates the adjacent-step ratio")
    return RadialGridValidation(cells=int(widths.size), minimum_step_m=float(widths.min()),
                                maximum_step_m=float(widths.max()),
                                maximum_adjacent_step_ratio=worst)

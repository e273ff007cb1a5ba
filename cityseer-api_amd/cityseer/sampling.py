"""Distance-based source sampling schedule.

Restates ``pysrc/cityseer/sampling.py:27-104``: reachability of a canonical
grid network r = pi d^2 / s^2 and the Hoeffding / Eppstein-Wang bound
k = ln(2r / delta) / (2 eps^2), p = min(1, k / r).
"""

from __future__ import annotations

import math

import numpy as np

HOEFFDING_EPSILON: float = 0.06
HOEFFDING_DELTA: float = 0.1
GRID_SPACING: float = 175.0


def compute_hoeffding_p(
    mean_reachability: float, epsilon: float = HOEFFDING_EPSILON, delta: float = HOEFFDING_DELTA
) -> float:
    """Sampling probability for a mean reachability (sampling.py:32-69)."""
    if (
        not np.isfinite(mean_reachability)
        or not np.isfinite(epsilon)
        or not np.isfinite(delta)
        or mean_reachability <= 0
        or epsilon <= 0
        or delta <= 0
        or delta >= 1
    ):
        return 1.0
    k = math.log(2 * mean_reachability / delta) / (2 * epsilon**2)
    return min(1.0, k / mean_reachability)


def compute_distance_p(
    distance: float,
    epsilon: float = HOEFFDING_EPSILON,
    delta: float = HOEFFDING_DELTA,
    grid_spacing: float = GRID_SPACING,
) -> float:
    """Sampling probability for a distance threshold (sampling.py:72-104)."""
    if distance <= 0 or grid_spacing <= 0:
        return 1.0
    r = math.pi * distance**2 / grid_spacing**2
    return compute_hoeffding_p(r, epsilon=epsilon, delta=delta)

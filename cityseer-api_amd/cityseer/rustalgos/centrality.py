"""Result classes of the centrality calls.

Mirror ``CentralityShortestResult``, ``CentralitySimplestResult`` and
``CentralitySegmentResult`` (rust/src/centrality.rs:152-355): attributes
``distances``, ``node_keys_py``, ``node_indices``, ``reachability_totals``,
``sampled_source_count`` and one getter per metric returning
``dict[int, np.ndarray[float64]]`` compacted by ``node_indices``, a fresh
array per call (rust/src/common.rs:40-53).
"""

from __future__ import annotations

from typing import Any

import numpy as np


class _MetricResultSet:
    """Holds f64 metrics as one array [M][T][N] and serves compacted copies."""

    _metrics: tuple[str, ...] = ()

    def __init__(
        self,
        distances: list[int],
        node_keys_py: list[Any],
        node_indices: list[int],
        values: np.ndarray,
        reachability_totals: list[int] | None = None,
        sampled_source_count: int = 0,
    ):
        self.distances = list(distances)
        self.node_keys_py = list(node_keys_py)
        self.node_indices = list(node_indices)
        self._values = values
        self._idx = np.asarray(self.node_indices, dtype=np.int64)
        self.reachability_totals = list(reachability_totals or [])
        self.sampled_source_count = int(sampled_source_count)

    def _load_compact(self, metric_pos: int) -> dict[int, np.ndarray]:
        rows = self._values[metric_pos]
        return {int(d): rows[i][self._idx].copy() for i, d in enumerate(self.distances)}


def _getter(pos: int):
    return property(lambda self: self._load_compact(pos))


class CentralityShortestResult(_MetricResultSet):
    """Closeness + betweenness from one Brandes search per source (centrality.rs:152-234)."""

    _metrics = ("density", "farness", "cycles", "harmonic", "beta", "betweenness", "betweenness_beta")
    node_density = _getter(0)
    node_farness = _getter(1)
    node_cycles = _getter(2)
    node_harmonic = _getter(3)
    node_beta = _getter(4)
    node_betweenness = _getter(5)
    node_betweenness_beta = _getter(6)


class CentralitySimplestResult(_MetricResultSet):
    """Angular closeness + betweenness on the dual graph (centrality.rs:236-299)."""

    _metrics = ("density", "farness", "harmonic", "betweenness")
    node_density = _getter(0)
    node_farness = _getter(1)
    node_harmonic = _getter(2)
    node_betweenness = _getter(3)


class CentralitySegmentResult(_MetricResultSet):
    """Segment closeness + betweenness (centrality.rs:301-355)."""

    _metrics = ("density", "harmonic", "beta", "betweenness")
    segment_density = _getter(0)
    segment_harmonic = _getter(1)
    segment_beta = _getter(2)
    segment_betweenness = _getter(3)


class BetweennessShortestResult(_MetricResultSet):
    """OD-weighted betweenness (centrality.rs:98-148): betweenness and betweenness_beta only."""

    _metrics = ("betweenness", "betweenness_beta")
    node_betweenness = _getter(0)
    node_betweenness_beta = _getter(1)


class OdMatrix:
    """Sparse origin-destination trip weights (centrality.rs:50-91), built from parallel
    (origin, destination, weight) arrays in COO form; a later pair overwrites an earlier one with
    the same origin and destination. Reusable across calls."""

    def __init__(self, origins: list[int], destinations: list[int], weights: list[float]) -> None:
        if len(origins) != len(destinations) or len(origins) != len(weights):
            raise ValueError(
                f"origins ({len(origins)}), destinations ({len(destinations)}), and weights ({len(weights)}) "
                "must have equal length"
            )
        o = np.asarray(origins, dtype=np.int64)
        d = np.asarray(destinations, dtype=np.int64)
        if o.size and (o.min() < 0 or d.min() < 0):
            raise ValueError("OD node indices must be non-negative")
        self.origins = o.astype(np.uint32)
        self.destinations = d.astype(np.uint32)
        self.weights = np.asarray(weights, dtype=np.float32)
        # distinct (origin, destination) pairs, as the reference's nested map holds them
        self._pairs = len(set(zip(self.origins.tolist(), self.destinations.tolist())))
        self._origins = len(set(self.origins.tolist()))

    def len(self) -> int:
        """Number of non-zero OD pairs."""
        return self._pairs

    def n_origins(self) -> int:
        """Number of unique origin nodes."""
        return self._origins

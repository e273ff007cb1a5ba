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

"""CPU: front-end argument handling that needs no device (tests/metrics/test_networks.py:246-261)."""

from __future__ import annotations

import pytest

from cityseer.metrics import networks
from cityseer.tools import io


def test_simplest_wrappers_require_dual_graph(primal_graph):
    nodes_gdf, _edges, ns = io.network_structure_from_nx(primal_graph)
    for fn in (networks.closeness_simplest, networks.betweenness_simplest, networks.node_centrality_simplest):
        with pytest.raises(ValueError, match="dual graph"):
            fn(network_structure=ns, nodes_gdf=nodes_gdf.copy(), distances=[200, 400])


def test_thresholds_are_validated_before_any_device_work(primal_graph):
    nodes_gdf, _edges, ns = io.network_structure_from_nx(primal_graph)
    with pytest.raises(ValueError):
        networks.node_centrality_shortest(network_structure=ns, nodes_gdf=nodes_gdf.copy(), distances=[400, 400])
    with pytest.raises(ValueError):
        networks.node_centrality_shortest(network_structure=ns, nodes_gdf=nodes_gdf.copy())

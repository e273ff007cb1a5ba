"""GPU: the cityseer.metrics.networks front-end (pysrc/cityseer/metrics/networks.py:101-888).

Restates the reference's tests/metrics/test_networks.py:11-261 for this package:
- every cc_* column equals the raw NetworkStructure getter it is built from, for each
  combination of the closeness and betweenness flags, with thresholds given as distances or
  as betas (both give the same columns);
- Hillier closeness equals density ** 2 / farness;
- the simplest-path columns under several scaling units and offsets, and the segment columns;
- the closeness-only and betweenness-only wrappers write their columns, deterministically for a
  seed, and the simplest-path wrappers refuse a primal graph.
The columns are also compared with the oracle, so the front-end adds nothing of its own.
Tolerance: the per-row 1e-5 bar of the other GPU tests (the reference compares at 0.01 absolute).
"""

from __future__ import annotations

import numpy as np
import pandas as pd
import pytest

from cityseer import config, rustalgos
from cityseer.metrics import networks
from cityseer.tools import io
import oracle.oracle as O
from oracle.oracle import OracleGraph
from parity import assert_close
from test_gpu_shortest import say

pytestmark = pytest.mark.gpu

DISTANCES = [400, 800]
FLAGS = [(False, True), (True, False), (True, True)]


@pytest.fixture(autouse=True)
def _keyed():
    O.set_heap_mode(False)


def col(gdf: pd.DataFrame, key: str) -> np.ndarray:
    return gdf[key].to_numpy(dtype=np.float64)


def hillier(density, farness):
    with np.errstate(divide="ignore", invalid="ignore"):
        return density ** 2 / farness


def assert_hillier(got, density, farness, what):
    want = hillier(density, farness)
    assert np.array_equal(np.isnan(got), np.isnan(want)), what
    ok = ~np.isnan(want)
    assert_close(got[ok], want[ok], what)


def test_node_centrality_shortest(primal_graph):
    betas = rustalgos.betas_from_distances(DISTANCES)
    nodes_gdf, _edges, ns = io.network_structure_from_nx(primal_graph)
    og = OracleGraph.from_network_structure(ns)
    ref, _ = og.centrality_shortest(DISTANCES)
    raw_close = ns.centrality_shortest(compute_closeness=True, compute_betweenness=False, betas=betas)
    for closeness, betweenness in FLAGS:
        by_input = []
        for dists, bts in [(DISTANCES, None), (None, betas)]:
            say(f"shortest closeness={closeness} betweenness={betweenness} "
                f"{'distances' if dists else 'betas'}")
            gdf = networks.node_centrality_shortest(network_structure=ns, nodes_gdf=nodes_gdf.copy(), distances=dists,
                                                    betas=bts, compute_closeness=closeness,
                                                    compute_betweenness=betweenness)
            by_input.append(gdf)
            for i, d in enumerate(DISTANCES):
                for m, measure in enumerate(("density", "farness", "cycles", "harmonic", "beta")):
                    key = config.prep_gdf_key(measure, d)
                    if not closeness:
                        assert key not in gdf.columns
                        continue
                    attr = {"density": "node_density", "farness": "node_farness", "cycles": "node_cycles",
                            "harmonic": "node_harmonic", "beta": "node_beta"}[measure]
                    assert_close(col(gdf, key), getattr(raw_close, attr)[d], f"{key} vs getter")
                    assert_close(col(gdf, key), ref[m, i], f"{key} vs oracle")
                if closeness:
                    assert_hillier(col(gdf, config.prep_gdf_key("hillier", d)), raw_close.node_density[d],
                                   raw_close.node_farness[d], f"hillier {d}")
                raw_betw = ns.centrality_shortest(compute_closeness=False, compute_betweenness=True, distances=[d])
                for m, (measure, attr) in enumerate([("betweenness", "node_betweenness"),
                                                     ("betweenness_beta", "node_betweenness_beta")]):
                    key = config.prep_gdf_key(measure, d)
                    if not betweenness:
                        assert key not in gdf.columns
                        continue
                    assert_close(col(gdf, key), getattr(raw_betw, attr)[d], f"{key} vs getter")
                    assert_close(col(gdf, key), ref[5 + m, i], f"{key} vs oracle")
        # distances and betas give the same columns (f64 sums may differ in order between runs)
        a, b = by_input
        assert list(a.columns) == list(b.columns)
        for key in a.columns:
            if key.startswith("cc_density"):
                assert np.array_equal(col(a, key), col(b, key)), key
            elif key.startswith("cc_hillier"):
                assert np.array_equal(np.isnan(col(a, key)), np.isnan(col(b, key))), key
                ok = ~np.isnan(col(a, key))
                assert_close(col(a, key)[ok], col(b, key)[ok], key)
            elif key.startswith("cc_"):
                assert_close(col(a, key), col(b, key), key)


def test_node_centrality_simplest(dual_graph):
    betas = rustalgos.betas_from_distances(DISTANCES)
    nodes_gdf, _edges, ns = io.network_structure_from_nx(dual_graph)
    og = OracleGraph.from_network_structure(ns)
    for dists, bts in [(DISTANCES, None), (None, betas)]:
        for offset, unit in [(0, 180), (0, 90), (1, 180)]:
            say(f"simplest offset={offset} unit={unit} {'distances' if dists else 'betas'}")
            gdf = networks.node_centrality_simplest(network_structure=ns, nodes_gdf=nodes_gdf.copy(), distances=dists,
                                                    betas=bts, compute_closeness=True, compute_betweenness=True,
                                                    farness_scaling_offset=offset, angular_scaling_unit=unit)
            raw = ns.centrality_simplest(compute_closeness=True, compute_betweenness=False, betas=betas,
                                         farness_scaling_offset=offset, angular_scaling_unit=unit)
            raw_b = ns.centrality_simplest(compute_closeness=False, compute_betweenness=True, betas=betas)
            ref, _ = og.centrality_simplest(DISTANCES, unit=unit, offset=offset)
            for i, d in enumerate(DISTANCES):
                for m, (measure, attr) in enumerate([("density", "node_density"), ("farness", "node_farness"),
                                                     ("harmonic", "node_harmonic")]):
                    key = config.prep_gdf_key(measure, d, angular=True)
                    assert_close(col(gdf, key), getattr(raw, attr)[d], f"{key} vs getter")
                    assert_close(col(gdf, key), ref[m, i], f"{key} vs oracle")
                assert_hillier(col(gdf, config.prep_gdf_key("hillier", d, angular=True)), raw.node_density[d],
                               raw.node_farness[d], f"angular hillier {d}")
                key = config.prep_gdf_key("betweenness", d, angular=True)
                assert_close(col(gdf, key), raw_b.node_betweenness[d], f"{key} vs getter")
                assert_close(col(gdf, key), ref[3, i], f"{key} vs oracle")


def test_segment_centrality(primal_graph):
    betas = rustalgos.betas_from_distances(DISTANCES)
    nodes_gdf, _edges, ns = io.network_structure_from_nx(primal_graph)
    og = OracleGraph.from_network_structure(ns)
    ref = og.segment_centrality(DISTANCES)
    for closeness, betweenness in FLAGS:
        for dists, bts in [(DISTANCES, None), (None, betas)]:
            say(f"segment closeness={closeness} betweenness={betweenness}")
            gdf = networks.segment_centrality(network_structure=ns, nodes_gdf=nodes_gdf.copy(), distances=dists,
                                              betas=bts, compute_closeness=closeness, compute_betweenness=betweenness)
            raw = ns.segment_centrality(betas=betas, compute_closeness=closeness, compute_betweenness=betweenness)
            for i, d in enumerate(DISTANCES):
                for m, (measure, attr) in enumerate([("seg_density", "segment_density"),
                                                     ("seg_harmonic", "segment_harmonic"),
                                                     ("seg_beta", "segment_beta")]):
                    key = config.prep_gdf_key(measure, d)
                    if closeness:
                        assert_close(col(gdf, key), getattr(raw, attr)[d], f"{key} vs getter")
                        assert_close(col(gdf, key), ref[m, i], f"{key} vs oracle")
                    else:
                        assert key not in gdf.columns
                key = config.prep_gdf_key("seg_betweenness", d)
                if betweenness:
                    assert_close(col(gdf, key), raw.segment_betweenness[d], f"{key} vs getter")
                    assert_close(col(gdf, key), ref[3, i], f"{key} vs oracle")
                else:
                    assert key not in gdf.columns


@pytest.mark.parametrize("which", ["shortest", "simplest"])
def test_closeness_wrappers_columns_and_seeded_determinism(which, primal_graph, dual_graph):
    graph = primal_graph if which == "shortest" else dual_graph
    angular = which == "simplest"
    fn = networks.closeness_shortest if which == "shortest" else networks.closeness_simplest
    distances = [200, 400, 800]
    nodes_gdf, _edges, ns = io.network_structure_from_nx(graph)
    say(f"closeness_{which} wrapper")
    r1 = fn(network_structure=ns, nodes_gdf=nodes_gdf.copy(), distances=distances, random_seed=42)
    r2 = fn(network_structure=ns, nodes_gdf=nodes_gdf.copy(), distances=distances, random_seed=42)
    measures = ["harmonic", "density", "farness", "hillier"] + ([] if angular else ["beta", "cycles"])
    for d in distances:
        for measure in measures:
            assert config.prep_gdf_key(measure, d, angular=angular) in r1.columns
        key = config.prep_gdf_key("density", d, angular=angular)
        assert np.array_equal(r1[key].to_numpy(), r2[key].to_numpy()), f"non-deterministic at {d} m"
        assert config.prep_gdf_key("betweenness", d, angular=angular) not in r1.columns


def test_betweenness_wrapper_columns(primal_graph):
    nodes_gdf, _edges, ns = io.network_structure_from_nx(primal_graph)
    say("betweenness_shortest wrapper")
    gdf = networks.betweenness_shortest(network_structure=ns, nodes_gdf=nodes_gdf.copy(), distances=DISTANCES)
    raw = ns.centrality_shortest(compute_closeness=False, compute_betweenness=True, distances=DISTANCES)
    for d in DISTANCES:
        assert_close(col(gdf, config.prep_gdf_key("betweenness", d)), raw.node_betweenness[d], "betweenness")
        assert config.prep_gdf_key("harmonic", d) not in gdf.columns

"""GPU: the layers front-end end to end (pysrc/cityseer/metrics/layers.py:104-756) on the mock
graph: mock data assigned to the network on the host, then accessibility, mixed uses and
statistics on the device. The columns written into ``nodes_gdf`` must equal ``DataMap``'s own
results (as the reference's tests/metrics/test_layers.py checks) and the oracle's restatement over
the same assignment: counts, richness, medians and extremes exactly, the rest to 1e-5 relative."""

from __future__ import annotations

import numpy as np
import pytest

from cityseer import config
from cityseer.metrics import layers
from cityseer.rustalgos import data
from cityseer.tools import graphs, io, mock
import oracle.oracle as O
from oracle.access import STAT_NAMES
from oracle.access import accessibility as oracle_access
from oracle.access import mixed_uses as oracle_mixed
from oracle.access import stats as oracle_stats
from oracle.oracle import OracleGraph
from test_gpu_shortest import say

pytestmark = pytest.mark.gpu

DISTS = [400, 800]


@pytest.fixture(autouse=True)
def _keyed():
    O.set_heap_mode(False)


@pytest.fixture(scope="module")
def primal():
    G = graphs.nx_simple_geoms(mock.mock_graph())
    nodes, _edges, ns = io.network_structure_from_nx(G)
    return G, nodes, ns


def _oracle_inputs(dm, ns, dists, speed=config.SPEED_M_S):
    og = OracleGraph.from_network_structure(ns)
    _d, bets, secs = O.thresholds(speed, distances=dists)
    mcw = data.clip_wts_curve(dists, bets, 0)
    dedupe_of = {k: e.dedupe_key for k, e in dm.entries.items()}
    return og, bets, mcw, max(secs), speed, dedupe_of


def _node_rows(nodes_gdf, ns, col):
    """The column in node index order."""
    return nodes_gdf.loc[ns.node_keys_py(), col].to_numpy(dtype=np.float64)


@pytest.mark.parametrize("data_id_col", [None, "data_id"])
def test_compute_accessibilities(primal, data_id_col):
    G, nodes0, ns = primal
    df = mock.mock_landuse_categorical_data(G)
    keys = ["a", "b", "z"]
    say(f"compute_accessibilities on the mock graph, dedupe column {data_id_col}")
    nodes, _ = layers.compute_accessibilities(df, "categorical_landuses", keys, nodes0.copy(), ns,
                                              max_netw_assign_dist=400, distances=DISTS, data_id_col=data_id_col)
    dm = layers.build_data_map(df, ns, 400, data_id_col)
    lu = {f"{type(k).__name__}:{k}": v for k, v in df["categorical_landuses"].items()}
    direct = dm.accessibility(ns, dict(df["categorical_landuses"]), keys, distances=DISTS)
    og, bets, mcw, mws, speed, dedupe_of = _oracle_inputs(dm, ns, DISTS)
    ref = oracle_access(og, [1] * og.n, dm.node_data_map, dedupe_of, lu, keys, DISTS, bets, mcw, mws, speed)
    for key in keys:
        r_u, r_w, r_n = ref[key]
        for i, d in enumerate(DISTS):
            nw = _node_rows(nodes, ns, config.prep_gdf_key(key, d, weighted=False))
            wt = _node_rows(nodes, ns, config.prep_gdf_key(key, d, weighted=True))
            assert np.array_equal(nw, direct.result[key].unweighted[d])
            assert np.array_equal(wt, direct.result[key].weighted[d])
            assert np.array_equal(nw, r_u[i]), f"{key} count at {d}"
            np.testing.assert_allclose(wt, r_w[i], rtol=1e-5, atol=1e-12, err_msg=f"{key} weighted at {d}")
        near = _node_rows(nodes, ns, config.prep_gdf_key(f"{key}_nearest_max", 800))
        assert np.array_equal(near, r_n, equal_nan=True), f"{key} nearest"
        assert config.prep_gdf_key(f"{key}_nearest_max", 400) not in nodes.columns
    # the dedupe group "z" (the last five entries) counts once per node
    if data_id_col is not None:
        assert _node_rows(nodes, ns, config.prep_gdf_key("z", 800, weighted=False)).max() == 1.0


def test_compute_mixed_uses(primal):
    G, nodes0, ns = primal
    df = mock.mock_landuse_categorical_data(G)
    say("compute_mixed_uses on the mock graph, all measures")
    nodes, _ = layers.compute_mixed_uses(df, "categorical_landuses", nodes0.copy(), ns, max_netw_assign_dist=400,
                                         distances=DISTS, compute_shannon=True, compute_gini=True,
                                         data_id_col="data_id")
    dm = layers.build_data_map(df, ns, 400, "data_id")
    lu = {f"{type(k).__name__}:{k}": v for k, v in df["categorical_landuses"].items()}
    og, bets, mcw, mws, speed, dedupe_of = _oracle_inputs(dm, ns, DISTS)
    ref = oracle_mixed(og, [1] * og.n, dm.node_data_map, dedupe_of, lu, DISTS, bets, mcw, mws, speed)
    for i, d in enumerate(DISTS):
        assert np.array_equal(_node_rows(nodes, ns, config.prep_gdf_key("hill_q0", d, weighted=False)), ref[0, i])
        cols = [(config.prep_gdf_key(f"hill_q{q}", d, weighted=False), q) for q in (1, 2)]
        cols += [(config.prep_gdf_key(f"hill_q{q}", d, weighted=True), 3 + q) for q in (0, 1, 2)]
        cols += [(config.prep_gdf_key("shannon", d), 6), (config.prep_gdf_key("gini", d), 7)]
        for col, m in cols:
            np.testing.assert_allclose(_node_rows(nodes, ns, col), ref[m, i], rtol=1e-5, atol=1e-6, err_msg=col)
    # flags leave their columns out
    only, _ = layers.compute_mixed_uses(df, "categorical_landuses", nodes0.copy(), ns, 400, compute_hill=False,
                                        distances=[400], data_id_col="data_id")
    assert config.prep_gdf_key("hill_q0", 400, weighted=False) not in only.columns
    assert config.prep_gdf_key("shannon", 400) not in only.columns
    assert np.array_equal(_node_rows(only, ns, config.prep_gdf_key("hill_q1", 400, weighted=True)),
                          _node_rows(nodes, ns, config.prep_gdf_key("hill_q1", 400, weighted=True)))


def test_compute_stats(primal):
    G, nodes0, ns = primal
    df = mock.mock_numerical_data(G, num_arrs=2)
    labels = ["mock_numerical_1", "mock_numerical_2"]
    say("compute_stats on the mock graph, two numeric columns")
    nodes, _ = layers.compute_stats(df, labels, nodes0.copy(), ns, max_netw_assign_dist=400, distances=DISTS,
                                    data_id_col="data_id")
    dm = layers.build_data_map(df, ns, 400, "data_id")
    maps = [{f"{type(k).__name__}:{k}": float(v) for k, v in df[c].items()} for c in labels]
    og, bets, mcw, mws, speed, dedupe_of = _oracle_inputs(dm, ns, DISTS)
    order = {k: i for i, k in enumerate(dm.entries)}
    ref = oracle_stats(og, [1] * og.n, dm.node_data_map, dedupe_of, order, maps, DISTS, bets, mcw, mws, speed)
    column = {"sum": ("sum", False), "sum_wt": ("sum", True), "mean": ("mean", False), "mean_wt": ("mean", True),
              "median": ("median", False), "median_wt": ("median", True), "count": ("count", False),
              "count_wt": ("count", True), "variance": ("var", False), "variance_wt": ("var", True),
              "mad": ("mad", False), "mad_wt": ("mad", True), "max": ("max", None), "min": ("min", None)}
    exact = ("count", "max", "min", "median", "mad", "median_wt", "mad_wt")
    for mp, label in enumerate(labels):
        for m, name in enumerate(STAT_NAMES):
            suffix, weighted = column[name]
            for i, d in enumerate(DISTS):
                got = _node_rows(nodes, ns, config.prep_gdf_key(f"{label}_{suffix}", d, weighted=weighted))
                want = ref[mp, m, i]
                assert np.array_equal(np.isnan(got), np.isnan(want)), f"{label} {name} {d}: NaN pattern"
                ok = ~np.isnan(want)
                if name in exact:
                    assert np.array_equal(got[ok], want[ok]), f"{label} {name} {d}"
                else:
                    np.testing.assert_allclose(got[ok], want[ok], rtol=1e-5, atol=1e-6, err_msg=f"{label} {name} {d}")
    assert np.nanmax(_node_rows(nodes, ns, config.prep_gdf_key("mock_numerical_1_count", 800, weighted=False))) > 0


def test_layers_angular_on_the_dual(primal):
    """The reference's layer tests also run angular=True on the dual of the mock graph: data are
    assigned to the dual's street edges, then simplest-path trees (data.rs:538-548)."""
    G, _nodes, _ns = primal
    nodes0, _e, dual = io.network_structure_from_nx(graphs.nx_to_dual(G))
    df = mock.mock_landuse_categorical_data(G)
    keys = ["a", "b"]
    say("compute_accessibilities and compute_mixed_uses, angular, on the dual mock graph")
    nodes, _ = layers.compute_accessibilities(df, "categorical_landuses", keys, nodes0.copy(), dual,
                                              max_netw_assign_dist=400, distances=DISTS, data_id_col="data_id",
                                              angular=True)
    nodes, _ = layers.compute_mixed_uses(df, "categorical_landuses", nodes, dual, max_netw_assign_dist=400,
                                         distances=DISTS, data_id_col="data_id", angular=True)
    dm = layers.build_data_map(df, dual, 400, "data_id")
    lu = {f"{type(k).__name__}:{k}": v for k, v in df["categorical_landuses"].items()}
    og, bets, mcw, mws, speed, dedupe_of = _oracle_inputs(dm, dual, DISTS)
    ref = oracle_access(og, [1] * og.n, dm.node_data_map, dedupe_of, lu, keys, DISTS, bets, mcw, mws, speed,
                        angular=True)
    mixed = oracle_mixed(og, [1] * og.n, dm.node_data_map, dedupe_of, lu, DISTS, bets, mcw, mws, speed, angular=True)
    for key in keys:
        for i, d in enumerate(DISTS):
            assert np.array_equal(_node_rows(nodes, dual, config.prep_gdf_key(key, d, True, weighted=False)),
                                  ref[key][0][i]), (key, d)
            np.testing.assert_allclose(_node_rows(nodes, dual, config.prep_gdf_key(key, d, True, weighted=True)),
                                       ref[key][1][i], rtol=1e-5, atol=1e-12)
        assert np.array_equal(_node_rows(nodes, dual, config.prep_gdf_key(f"{key}_nearest_max", 800, True)),
                              ref[key][2], equal_nan=True)
    for i, d in enumerate(DISTS):
        assert np.array_equal(_node_rows(nodes, dual, config.prep_gdf_key("hill_q0", d, True, weighted=False)),
                              mixed[0, i])
        np.testing.assert_allclose(_node_rows(nodes, dual, config.prep_gdf_key("hill_q1", d, True, weighted=True)),
                                   mixed[4, i], rtol=1e-5, atol=1e-6)

"""GPU parity: simplest-path (angular) centrality and trees on dual graphs vs the oracle."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import graph_from
from cityseer import config
from cityseer.tools import graphs, io, mock
import oracle.oracle as O
from oracle.oracle import OracleGraph
from test_gpu_shortest import assert_close, say, stack

pytestmark = pytest.mark.gpu

SPEED = config.SPEED_M_S
KEYS = ("node_density", "node_farness", "node_harmonic", "node_betweenness")


@pytest.fixture(autouse=True)
def _keyed():
    O.set_heap_mode(False)


def check_simplest(ns, og, distances, **kw):
    res = ns.centrality_simplest(distances=distances, **kw)
    ref, reach = og.centrality_simplest(
        distances,
        closeness=kw.get("compute_closeness", True),
        betweenness=kw.get("compute_betweenness", True),
        unit=kw.get("angular_scaling_unit", 180.0),
        offset=kw.get("farness_scaling_offset", 1.0),
    )
    got = stack(res, KEYS, distances)
    assert np.array_equal(got[0], ref[0]), "angular density (reachable counts) must match bit-exactly"
    assert_close(got, ref, "simplest metrics")
    return res


def test_dual_lattice_all_metrics():
    say("dual of a 20x20 perturbed lattice, simplest [500, 1000, 2000]")
    dual = mock.lattice_dual(mock.perturbed_lattice(20, seed=7))
    ns = mock.network_structure_from_lattice(dual)
    og = OracleGraph.from_network_structure(ns)
    check_simplest(ns, og, [500, 1000, 2000])
    say("same, unit 90 offset 1 (the metrics front-end defaults)")
    check_simplest(ns, og, [500, 1000, 2000], angular_scaling_unit=90.0, farness_scaling_offset=1.0)


def test_mock_dual_graph(dual_graph):
    say("mock dual graph simplest [500, 2000, 5000]")
    _n, _e, ns = io.network_structure_from_nx(dual_graph)
    og = OracleGraph.from_network_structure(ns)
    check_simplest(ns, og, [500, 2000, 5000])
    res_sh = ns.centrality_shortest(distances=[5000], compute_betweenness=False)
    res_si = ns.centrality_simplest(distances=[5000], compute_betweenness=False)
    assert np.allclose(res_sh.node_density[5000], res_si.node_density[5000])
    assert not np.allclose(res_sh.node_farness[5000], res_si.node_farness[5000], atol=config.ATOL)


def test_diamond_dual_reference_values(diamond_graph, golden):
    say("diamond dual simplest harmonic vs reference closed forms")
    gold = golden["diamond"]
    _n, _e, ns = io.network_structure_from_nx(graphs.nx_to_dual(diamond_graph))
    res = ns.centrality_simplest(distances=gold["distances"], compute_betweenness=False)
    for d in gold["distances"]:
        assert np.allclose(res.node_harmonic[d], gold["simplest_harmonic"][str(d)], atol=config.ATOL,
                           rtol=config.RTOL)


def test_plateau_ratio(golden):
    say("angular plateau ratio")
    g = graph_from(golden["plateau"]["nodes"], golden["plateau"]["edges"])
    nodes, _e, ns = io.network_structure_from_nx(graphs.nx_to_dual(g))
    res = ns.centrality_simplest(distances=[1000], compute_closeness=False)
    b = {k: res.node_betweenness[1000][i] for i, k in enumerate(nodes.index)}
    assert np.isclose(b["B_C_k0"] / b["C_D_k0"], golden["plateau"]["ratio"], atol=1e-6)


def test_tree_simplest_reference_paths_and_oracle(dual_graph):
    say("dijkstra_tree_simplest paths")
    nodes, _e, ns = io.network_structure_from_nx(dual_graph)
    og = OracleGraph.from_network_structure(ns)
    keys = nodes.index.tolist()
    max_s = int(5000 / SPEED)

    def path(visits, src, tgt):
        walk = [tgt]
        while walk[-1] != src:
            walk.append(visits[walk[-1]].pred)
        return [keys[i] for i in reversed(walk)]

    src, tgt = keys.index("11_6_k0"), keys.index("39_40_k0")
    _order, visits = ns.dijkstra_tree_simplest(src, max_s, SPEED)
    assert path(visits, src, tgt) == [
        "11_6_k0", "11_14_k0", "10_14_k0", "10_43_k0", "43_44_k0", "40_44_k0", "39_40_k0"]
    src, tgt = keys.index("10_43_k0"), keys.index("10_5_k0")
    _order, visits = ns.dijkstra_tree_simplest(src, max_s, SPEED)
    assert path(visits, src, tgt) == ["10_43_k0", "10_5_k0"]
    for s in (0, 11, 40):
        order, visits = ns.dijkstra_tree_simplest(s, 600, SPEED)
        ref = og.tree_simplest(s, 600)
        assert order == ref["order"].tolist()
        assert [(-1 if v.pred is None else v.pred) for v in visits] == ref["pred"].tolist()
        assert np.array_equal(np.float32([v.simpl_dist for v in visits]), ref["dist"])
        assert np.array_equal(np.float32([v.agg_seconds for v in visits]), ref["secs"])


def test_order_invariance():
    say("angular betweenness node-order invariance")
    base = {"A": (500000.0, 0.0), "B": (500000.0, 100.0), "C": (500100.0, 100.0), "D": (500050.0, 170.0)}
    base_edges = [("A", "B"), ("B", "C"), ("B", "D")]
    results = []
    for ordering in (["A", "B", "C", "D"], ["D", "C", "B", "A"], ["C", "A", "D", "B"]):
        lab = {k: str(i) for i, k in enumerate(ordering)}
        inv = {v: k for k, v in lab.items()}
        g = graph_from({lab[k]: base[k] for k in ordering}, [(lab[a], lab[b]) for a, b in base_edges])
        nodes, _e, ns = io.network_structure_from_nx(graphs.nx_to_dual(g))
        res = ns.centrality_simplest(distances=[500], compute_closeness=False)
        by_edge = {}
        for pos, key in enumerate(nodes.index):
            row = nodes.loc[key]
            by_edge[tuple(sorted((inv[row["primal_edge_node_a"]], inv[row["primal_edge_node_b"]])))] = \
                res.node_betweenness[500][pos]
        results.append(by_edge)
    for e in [("A", "B"), ("B", "C"), ("B", "D")]:
        vals = [r[e] for r in results]
        assert all(abs(v - vals[0]) < config.ATOL for v in vals)


def test_source_indices_and_tolerance_band():
    say("angular source_indices subset")
    dual = mock.lattice_dual(mock.perturbed_lattice(16, seed=2))
    ns = mock.network_structure_from_lattice(dual)
    og = OracleGraph.from_network_structure(ns)
    src = np.random.default_rng(5).choice(ns.node_count(), 40, replace=False).tolist()
    res = ns.centrality_simplest(distances=[800, 1600], source_indices=src)
    plan = ns._sampling_plan(None, None, None, src)
    ref, reach = og.centrality_simplest([800, 1600], sources=plan["sources"], wt=plan["wt"],
                                        eligible=plan["eligible"], track=True, scale=plan["scale"])
    assert_close(stack(res, KEYS, [800, 1600]), ref, "angular subset")
    assert res.reachability_totals == reach.tolist()


def _lanes(n):
    import ctypes

    from cityseer import _native

    lib = _native.load()
    t = _native.CsTuning()
    _native.check(lib.cs_get_tuning(ctypes.byref(t)))
    t.simplest_lanes = n
    _native.check(lib.cs_set_tuning(ctypes.byref(t)))


def test_sources_per_wave_and_workspace_reuse():
    """Launches alternate between graphs and source counts, so the workspace layout (graph
    size, sources per wave) changes from call to call over maps that are never cleared in
    between; every result must still match the oracle, and forcing any number of sources
    per wave must give the same results."""
    small = mock.network_structure_from_lattice(mock.lattice_dual(mock.perturbed_lattice(10, seed=41)))
    big = mock.network_structure_from_lattice(mock.lattice_dual(mock.perturbed_lattice(24, seed=42)))
    og_s, og_b = OracleGraph.from_network_structure(small), OracleGraph.from_network_structure(big)
    rng = np.random.default_rng(43)
    sub = np.sort(rng.choice(small.node_count(), 23, replace=False)).tolist()

    def sub_check(ns, og, dists, sources):
        res = ns.centrality_simplest(distances=dists, source_indices=sources)
        plan = ns._sampling_plan(None, None, None, sources)
        ref, _reach = og.centrality_simplest(dists, sources=plan["sources"], wt=plan["wt"],
                                             eligible=plan["eligible"], track=True, scale=plan["scale"])
        got = stack(res, KEYS, dists)
        assert np.array_equal(got[0], ref[0])
        assert_close(got, ref, "simplest subset")
        return got

    try:
        say("auto sources per wave: small graph subset, big graph, small graph again")
        first = sub_check(small, og_s, [300, 700], sub)
        check_simplest(big, og_b, [400, 900])
        again = sub_check(small, og_s, [300, 700], sub)
        assert np.array_equal(first[0], again[0])
        for n in (1, 7, 64):
            say(f"{n} sources per wave")
            _lanes(n)
            check_simplest(small, og_s, [300, 700])
            sub_check(big, og_b, [400, 900], np.sort(rng.choice(big.node_count(), 150, replace=False)).tolist())
    finally:
        _lanes(0)

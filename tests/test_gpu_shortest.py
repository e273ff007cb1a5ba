"""GPU parity: shortest-path and segment centrality vs the oracle (HIP kernels via the C ABI).

Bar: reachable counts (density with unit weights), distances, predecessor
sets and sigma bit-exact; f32-term closeness / betweenness sums within 1e-5
relative per (metric, threshold) row (only f64 summation order differs).
"""

from __future__ import annotations

import numpy as np
import pytest

from cityseer import config
from cityseer.tools import io, mock
import oracle.oracle as O
from oracle.oracle import OracleGraph
from parity import assert_close  # noqa: F401  (re-exported to the other GPU test files)

pytestmark = pytest.mark.gpu

RTOL = 1e-5
SPEED = config.SPEED_M_S
SHORTEST_KEYS = ("node_density", "node_farness", "node_cycles", "node_harmonic", "node_beta",
                 "node_betweenness", "node_betweenness_beta")
SEGMENT_KEYS = ("segment_density", "segment_harmonic", "segment_beta", "segment_betweenness")


def say(msg: str) -> None:
    print(msg, flush=True)


def stack(res, keys, distances):
    return np.stack([np.stack([getattr(res, k)[d] for d in distances]) for k in keys])


def check_shortest(ns, og, distances, **kw):
    res = ns.centrality_shortest(distances=distances, **kw)
    ref, reach = og.centrality_shortest(distances, closeness=kw.get("compute_closeness", True),
                                        betweenness=kw.get("compute_betweenness", True))
    got = stack(res, SHORTEST_KEYS, distances)
    assert np.array_equal(got[0], ref[0]), "density (reachable counts) must match bit-exactly"
    assert_close(got, ref, "shortest metrics")
    return res, ref


@pytest.fixture(scope="module")
def cfg1():
    lat = mock.perturbed_lattice(30, seed=0)
    ns = mock.network_structure_from_lattice(lat)
    return ns, OracleGraph.from_network_structure(ns)


def test_cfg1_shortest_all_metrics(cfg1):
    say("cfg1 30x30 shortest [400, 800]")
    ns, og = cfg1
    check_shortest(ns, og, [400, 800])


def test_cfg1_shortest_closeness_only_and_betweenness_only(cfg1):
    ns, og = cfg1
    say("cfg1 closeness-only")
    res = ns.centrality_shortest(distances=[400, 800], compute_betweenness=False)
    ref, _ = og.centrality_shortest([400, 800], betweenness=False)
    assert_close(stack(res, SHORTEST_KEYS, [400, 800]), ref, "closeness only")
    say("cfg1 betweenness-only")
    res = ns.centrality_shortest(distances=[400, 800], compute_closeness=False)
    ref, _ = og.centrality_shortest([400, 800], closeness=False)
    assert_close(stack(res, SHORTEST_KEYS, [400, 800]), ref, "betweenness only")


def test_brandes_state_bit_exact(cfg1):
    """Distances, predecessor sets and sigma of single sources match the oracle bit-exactly."""
    import ctypes

    from cityseer import _native

    ns, og = cfg1
    n = ns.node_count()
    lib = _native.load()
    h = ns._native_handle()
    max_seconds = int(round(800 / SPEED))
    rng = np.random.default_rng(1)
    for src in rng.choice(n, 24, replace=False).tolist():
        say(f"brandes state src={src}")
        secs, route = np.zeros(n, np.float32), np.zeros(n, np.float32)
        sigma = np.zeros(n, np.float64)
        npred = np.zeros(n, np.uint32)
        preds = np.zeros(n * 16, np.uint32)
        p = _native.ptr
        _native.check(lib.cs_brandes_shortest(h, src, max_seconds, SPEED, 1e-4, p(secs, ctypes.c_float),
                                              p(route, ctypes.c_float), p(sigma, ctypes.c_double),
                                              p(npred, ctypes.c_uint32), p(preds, ctypes.c_uint32), 16))
        ref = og.brandes_shortest(src, max_seconds)
        assert np.array_equal(secs, ref["secs"])
        assert np.array_equal(route, ref["route"])
        assert np.array_equal(sigma, ref["sigma"])
        assert np.array_equal(npred, ref["npred"])
        got_p = [sorted(preds[v * 16: v * 16 + min(npred[v], 16)].tolist()) for v in range(n)]
        assert got_p == ref["preds"]


def test_tree_shortest_bit_exact(cfg1):
    ns, og = cfg1
    rng = np.random.default_rng(2)
    for src in rng.choice(ns.node_count(), 8, replace=False).tolist():
        say(f"tree src={src}")
        order, visits = ns.dijkstra_tree_shortest(src, 600, SPEED)
        ref = og.tree_shortest(src, 600)
        assert order == ref["order"].tolist()
        assert [(-1 if v.pred is None else v.pred) for v in visits] == ref["pred"].tolist()
        assert np.array_equal(np.float32([v.short_dist for v in visits]), ref["dist"])
        assert np.array_equal(np.float32([v.agg_seconds for v in visits]), ref["secs"])


def test_cfg1_segment(cfg1):
    say("cfg1 segment [400, 800]")
    ns, og = cfg1
    res = ns.segment_centrality(distances=[400, 800])
    ref = og.segment_centrality([400, 800])
    assert_close(stack(res, SEGMENT_KEYS, [400, 800]), ref, "segment metrics")


def test_tree_segment_matches_oracle(cfg1):
    ns, og = cfg1
    for src in (0, 123, 456):
        say(f"segment tree src={src}")
        order, edges, visits, _ev = ns.dijkstra_tree_segment(src, 600, SPEED)
        ref = og.tree_segment(src, 600)
        assert order == ref["order"].tolist()
        assert edges == ref["edges"].tolist()
        assert [(-1 if v.origin_seg is None else v.origin_seg) for v in visits] == ref["origin_seg"].tolist()
        assert [(-1 if v.last_seg is None else v.last_seg) for v in visits] == ref["last_seg"].tolist()


def test_mock_graph_all_paths(primal_graph):
    say("mock graph shortest + segment [200, 400, 800, 5000]")
    _n, _e, ns = io.network_structure_from_nx(primal_graph)
    og = OracleGraph.from_network_structure(ns)
    check_shortest(ns, og, [200, 400, 800, 5000])
    res = ns.segment_centrality(distances=[200, 400, 800, 5000])
    ref = og.segment_centrality([200, 400, 800, 5000])
    assert_close(stack(res, SEGMENT_KEYS, [200, 400, 800, 5000]), ref, "segment mock")


def test_diamond_reference_values(diamond_graph, golden):
    say("diamond vs reference closed forms")
    gold = golden["diamond"]
    _n, _e, ns = io.network_structure_from_nx(diamond_graph)
    res = ns.centrality_shortest(distances=gold["distances"], compute_betweenness=False)
    for key in ("node_density", "node_farness", "node_cycles", "node_harmonic", "node_beta"):
        for d in gold["distances"]:
            assert np.allclose(getattr(res, key)[d], gold[key][str(d)], atol=config.ATOL, rtol=config.RTOL)
    seg = ns.segment_centrality(gold["distances"])
    for key in ("segment_density", "segment_harmonic", "segment_beta"):
        for d in gold["distances"]:
            assert np.allclose(getattr(seg, key)[d], gold[key][str(d)], atol=config.ATOL, rtol=config.RTOL)
    og = OracleGraph.from_network_structure(ns)
    O.set_heap_mode(False)
    ref = og.segment_centrality(gold["distances"])
    assert_close(stack(seg, SEGMENT_KEYS, gold["distances"]), ref, "segment diamond (keyed order)")


def test_betweenness_vs_networkx(primal_graph):
    import networkx as nx

    say("mock betweenness vs networkx")
    _n, _e, ns = io.network_structure_from_nx(primal_graph)
    res = ns.centrality_shortest(distances=[5000], compute_closeness=False)
    G = nx.MultiGraph()
    for s, e, k, d in primal_graph.edges(keys=True, data=True):
        G.add_edge(s, e, k, length=float(np.float32(d["geom"].length)))
    G.add_nodes_from(primal_graph.nodes())
    nxb = nx.betweenness_centrality(G, normalized=False, weight="length")
    for v in range(ns.node_count()):
        assert abs(nxb[str(v)] - res.node_betweenness[5000][v]) < config.ATOL


def test_source_indices_and_sampling(cfg1):
    ns, og = cfg1
    n = ns.node_count()
    src = np.random.default_rng(3).choice(n, 60, replace=False).tolist()
    say("source_indices subset")
    res = ns.centrality_shortest(distances=[400, 800], source_indices=src)
    plan = ns._sampling_plan(None, None, None, src)
    ref, reach = og.centrality_shortest([400, 800], sources=plan["sources"], wt=plan["wt"],
                                        eligible=plan["eligible"], track=True, scale=plan["scale"])
    assert_close(stack(res, SHORTEST_KEYS, [400, 800]), ref, "source_indices")
    assert res.reachability_totals == reach.tolist()
    assert res.sampled_source_count == 60
    say("sample_probability 0.3 seed 7")
    res = ns.centrality_shortest(distances=[400, 800], sample_probability=0.3, random_seed=7)
    plan = ns._sampling_plan(0.3, None, 7, None)
    ref, reach = og.centrality_shortest([400, 800], sources=plan["sources"], wt=plan["wt"],
                                        eligible=plan["eligible"], track=True, scale=plan["scale"])
    assert_close(stack(res, SHORTEST_KEYS, [400, 800]), ref, "sampled")
    assert res.reachability_totals == reach.tolist()


def test_cfg2_subset_and_full_size_properties():
    """cfg2 (200x200 + 5% diagonals) at [500, 1000, 2000, 5000]: 192 sources vs the oracle,
    then all 40k sources: density integers, linearity in node weight, idempotence."""
    say("cfg2 build")
    lat = mock.perturbed_lattice(200, seed=0, diag_frac=0.05)
    ns = mock.network_structure_from_lattice(lat)
    og = OracleGraph.from_network_structure(ns)
    d = [500, 1000, 2000, 5000]
    src = np.random.default_rng(4).choice(ns.node_count(), 192, replace=False).tolist()
    say("cfg2 subset vs oracle")
    res = ns.centrality_shortest(distances=d, source_indices=src)
    plan = ns._sampling_plan(None, None, None, src)
    ref, _ = og.centrality_shortest(d, sources=plan["sources"], wt=plan["wt"], eligible=plan["eligible"],
                                    track=True, scale=plan["scale"], n_threads=8)
    got = stack(res, SHORTEST_KEYS, d)
    assert np.array_equal(got[0], ref[0])
    assert_close(got, ref, "cfg2 subset")
    say("cfg2 full run x2")
    a = stack(ns.centrality_shortest(distances=d), SHORTEST_KEYS, d)
    b = stack(ns.centrality_shortest(distances=d), SHORTEST_KEYS, d)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[0], np.round(a[0]))
    assert np.allclose(a, b, rtol=1e-12, atol=0)
    say("cfg2 full run, weights x2")
    ns2 = mock.network_structure_from_lattice(lat)
    ns2._weight.array()[:] = 2.0
    c = stack(ns2.centrality_shortest(distances=d), SHORTEST_KEYS, d)
    scale = np.array([2, 2, 1, 2, 2, 2, 2], dtype=np.float64)[:, None, None]
    assert np.allclose(c, a * scale, rtol=1e-12, atol=0)


def test_epoch_wrap_keeps_results_exact():
    """Each wave tags its node records with 10-bit epochs and clears them when the tags run out.
    1,300 back-to-back calls of 144 sources each (144 waves, one source each on average) take
    every wave through a wrap; every call must return the first call's result."""
    lat = mock.perturbed_lattice(12, seed=5, diag_frac=0.1)
    ns = mock.network_structure_from_lattice(lat)
    og = OracleGraph.from_network_structure(ns)
    d = [300, 700]
    first, _ = check_shortest(ns, og, d)
    want = stack(first, SHORTEST_KEYS, d)
    say(f"{ns.node_count()} sources per call, 1300 calls")
    for i in range(1300):
        got = stack(ns.centrality_shortest(distances=d), SHORTEST_KEYS, d)
        assert np.array_equal(got[0], want[0]), f"call {i}: density changed"
        assert np.allclose(got, want, rtol=1e-12, atol=0), f"call {i}: metrics changed"

"""GPU: graph features the lattice workloads leave out, each checked against the oracle and
against the reference's own assertions for it.

- Elevation: directional slope penalties, and flat cost when either endpoint lacks z
  (rust tests: test_directional_slope_penalty_shortest_and_simplest, test_centrality.py:392-449).
- Non-live nodes interleaved by index (test_betweenness_mixed_live_non_live_invariant_to_node_order,
  test_centrality.py:775-847).
- Decomposition: segment closeness on the original nodes survives edge subdivision
  (test_decomposed_local_centrality, test_centrality.py:622-650).
"""

from __future__ import annotations

import numpy as np
import pytest

from cityseer import config
from cityseer.rustalgos.graph import NetworkStructure
from cityseer.tools import graphs, io, mock
import oracle.oracle as O
from oracle.oracle import OracleGraph
from test_gpu_shortest import SEGMENT_KEYS, SHORTEST_KEYS, SPEED, assert_close, say, stack

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _keyed():
    O.set_heap_mode(False)


def _lattice_ns(lat, z=None, live=None, order=None):
    """A NetworkStructure of the lattice, nodes inserted in `order` (default as generated)."""
    n = lat.n_nodes
    p = np.arange(n) if order is None else np.asarray(order)
    inv = np.empty_like(p)
    inv[p] = np.arange(n)
    ns = NetworkStructure()
    ns.add_street_nodes_bulk(
        node_keys=[str(i) for i in p], x=lat.x[p], y=lat.y[p],
        live=np.ones(n, bool) if live is None else np.asarray(live)[p],
        weight=np.ones(n, np.float32), z=None if z is None else np.asarray(z)[p],
    )
    src, dst, length, _angle, _pk = lat.directed()
    ns.add_street_edges_bulk(start_nd_idx=inv[src], end_nd_idx=inv[dst], edge_idx=np.zeros(len(src), np.uint32),
                             length=length)
    return ns, inv


def test_single_edge_slope_is_directional():
    """Rising 5 m over 100 m: downhill < flat < uphill; a missing z gives the flat cost."""
    flat = 100.0 / config.SPEED_M_S
    for z1, want in ((5.0, "sloped"), (None, "flat")):
        ns = NetworkStructure()
        ns.add_street_node("0", 0.0, 0.0, True, 1.0, z=0.0)
        ns.add_street_node("1", 0.0, 100.0, True, 1.0, z=z1)
        ns.add_street_edge(0, 1, 0, "0", "1", [[0.0, 0.0], [0.0, 100.0]])
        ns.add_street_edge(1, 0, 0, "1", "0", [[0.0, 100.0], [0.0, 0.0]])
        _o, t0 = ns.dijkstra_tree_shortest(0, 1000, config.SPEED_M_S)
        _o, t1 = ns.dijkstra_tree_shortest(1, 1000, config.SPEED_M_S)
        down, up = t0[1].agg_seconds, t1[0].agg_seconds
        say(f"{want}: 0->1 {down:.4f} s, 1->0 {up:.4f} s, flat {flat:.4f} s")
        if want == "sloped":
            # the search runs towards the source: the tree from 0 aggregates the walk 1 -> 0 (downhill)
            assert down < flat < up
        else:
            assert np.isclose(down, flat, atol=config.ATOL) and np.isclose(up, flat, atol=config.ATOL)


def test_elevation_matches_oracle():
    lat = mock.perturbed_lattice(22, seed=91, diag_frac=0.05)
    rng = np.random.default_rng(92)
    z = rng.uniform(0.0, 40.0, lat.n_nodes)
    z[rng.random(lat.n_nodes) < 0.1] = np.nan  # some nodes without elevation: their edges are flat
    order = rng.permutation(lat.n_nodes)
    ns, _inv = _lattice_ns(lat, z=z, order=order)
    og = OracleGraph.from_network_structure(ns)
    dists = [300, 800, 1600]
    say("elevation: shortest and segment vs oracle")
    res = ns.centrality_shortest(distances=dists)
    ref, _reach = og.centrality_shortest(dists)
    got = stack(res, SHORTEST_KEYS, dists)
    assert np.array_equal(got[0], ref[0])
    assert_close(got, ref, "shortest with elevation")
    assert_close(stack(ns.segment_centrality(distances=dists), SEGMENT_KEYS, dists), og.segment_centrality(dists),
                 "segment with elevation")
    for src in rng.choice(ns.node_count(), 4, replace=False).tolist():
        order_g, visits = ns.dijkstra_tree_shortest(src, 900, SPEED)
        tref = og.tree_shortest(src, 900)
        assert order_g == tref["order"].tolist()
        assert np.array_equal(np.float32([v.agg_seconds for v in visits]), tref["secs"])


def test_non_live_nodes_any_insertion_order():
    """A corridor A-B-C-D with D non-live, inserted in three orders: betweenness per label is the
    same; and a lattice with scattered non-live nodes matches the oracle in any insertion order."""
    coords = {"A": (0.0, 0.0), "B": (0.0, 100.0), "C": (0.0, 200.0), "D": (0.0, 300.0)}
    live = {"A": True, "B": True, "C": True, "D": False}
    per_label = []
    for ordering in (["A", "B", "C", "D"], ["D", "A", "B", "C"], ["B", "D", "A", "C"]):
        ns = NetworkStructure()
        idx = {}
        for lab in ordering:
            idx[lab] = ns.add_street_node(lab, *coords[lab], live[lab], 1.0)
        for a, b in (("A", "B"), ("B", "C"), ("C", "D")):
            for s, d in ((a, b), (b, a)):
                ns.add_street_edge(idx[s], idx[d], 0, s, d, [list(coords[s]), list(coords[d])])
        res = ns.centrality_shortest(distances=[1000], compute_closeness=False, compute_betweenness=True)
        per_label.append({lab: res.node_betweenness[1000][idx[lab]] for lab in ordering})
    say(f"corridor betweenness: {per_label[0]}")
    for lab in "ABCD":
        vals = [r[lab] for r in per_label]
        assert all(abs(v - vals[0]) < config.ATOL for v in vals), (lab, vals)

    lat = mock.perturbed_lattice(20, seed=93)
    rng = np.random.default_rng(94)
    live_l = rng.random(lat.n_nodes) > 0.15
    results = []
    for k, order in enumerate((None, rng.permutation(lat.n_nodes))):
        ns, inv = _lattice_ns(lat, live=live_l, order=order)
        og = OracleGraph.from_network_structure(ns)
        dists = [400, 1200]
        res = ns.centrality_shortest(distances=dists)
        ref, _reach = og.centrality_shortest(dists)
        got = stack(res, SHORTEST_KEYS, dists)
        say(f"lattice with {int((~live_l).sum())} non-live nodes, order {k}")
        assert np.array_equal(got[0], ref[0])
        assert_close(got, ref, "non-live lattice")
        results.append(got[:, :, inv])  # back to lattice numbering
    assert np.array_equal(results[0][0], results[1][0])
    assert_close(results[1], results[0], "non-live lattice, insertion order")


def test_decomposed_segment_closeness(primal_graph):
    """Segment closeness summed over the original nodes is unchanged by decomposition."""
    distances = [200, 400, 800, 5000]
    _n, _e, ns = io.network_structure_from_nx(primal_graph)
    _n, _e, nsd = io.network_structure_from_nx(graphs.nx_decompose(primal_graph, 20))
    a = ns.segment_centrality(distances, compute_closeness=True, compute_betweenness=True)
    b = nsd.segment_centrality(distances, compute_closeness=True, compute_betweenness=True)
    for key in ("segment_density", "segment_beta", "segment_harmonic"):
        got, want = getattr(b, key)[400][:57].sum(), getattr(a, key)[400].sum()
        say(f"{key}[400]: decomposed {got:.6f}, original {want:.6f}")
        assert np.allclose(got, want)

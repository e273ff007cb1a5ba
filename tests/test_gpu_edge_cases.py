"""GPU: degenerate and irregular inputs through every entry point, against the oracle.

Empty and single-node graphs, disconnected components, all sources non-live, self-loops and
parallel edges (multigraph keys), thresholds that reach nothing or everything.
"""

from __future__ import annotations

import numpy as np
import pytest

from cityseer.rustalgos.graph import NetworkStructure
from cityseer.tools import mock
import oracle.oracle as O
from oracle.oracle import OracleGraph
from test_gpu_shortest import SEGMENT_KEYS, SHORTEST_KEYS, SPEED, assert_close, say, stack

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _keyed():
    O.set_heap_mode(False)


def _check(ns, dists, what, segment=True):
    og = OracleGraph.from_network_structure(ns)
    res = ns.centrality_shortest(distances=dists)
    ref, reach = og.centrality_shortest(dists)
    got = stack(res, SHORTEST_KEYS, dists)
    assert np.array_equal(got[0], ref[0]), f"{what}: density"
    assert_close(got, ref, what)
    if segment:
        assert_close(stack(ns.segment_centrality(distances=dists), SEGMENT_KEYS, dists), og.segment_centrality(dists),
                     f"{what} (segment)")
    return got


def _ns(xy, edges, live=None):
    """Nodes at xy; each undirected edge (a, b) as two directed straight edges with key 0."""
    ns = NetworkStructure()
    for i, (x, y) in enumerate(xy):
        ns.add_street_node(str(i), float(x), float(y), True if live is None else bool(live[i]), 1.0)
    for a, b in edges:
        pa, pb = list(map(float, xy[a])), list(map(float, xy[b]))
        ns.add_street_edge(a, b, 0, str(a), str(b), [pa, pb])
        ns.add_street_edge(b, a, 0, str(b), str(a), [pb, pa])
    return ns


def test_single_node_and_no_edges():
    say("one node, no edges")
    ns = NetworkStructure()
    ns.add_street_node("0", 0.0, 0.0, True, 1.0)
    res = ns.centrality_shortest(distances=[100, 500])
    assert all(np.all(getattr(res, k)[d] == 0.0) for k in SHORTEST_KEYS for d in (100, 500))
    order, visits = ns.dijkstra_tree_shortest(0, 100, SPEED)
    assert order == [0] and visits[0].agg_seconds == 0.0
    say("five isolated nodes")
    ns = _ns([(i * 10.0, 0.0) for i in range(5)], [])
    _check(ns, [50], "isolated nodes", segment=False)


def test_disconnected_components_and_thresholds():
    a = mock.perturbed_lattice(8, seed=102)
    b = mock.perturbed_lattice(6, seed=103)
    xy = np.concatenate([np.stack([a.x, a.y], 1), np.stack([b.x + 5000.0, b.y], 1)])
    edges = [tuple(e) for e in a.edges.tolist()] + [(int(p) + a.n_nodes, int(q) + a.n_nodes) for p, q in b.edges]
    ns = _ns(xy, edges)
    say(f"two components ({a.n_nodes} + {b.n_nodes} nodes), thresholds from 1 m to 100 km")
    got = _check(ns, [1, 150, 700, 100000], "two components")
    assert np.all(got[0][0] == 0.0)  # 1 m reaches nothing
    # at 100 km every node reaches exactly the rest of its component
    sizes = np.concatenate([np.full(a.n_nodes, a.n_nodes - 1), np.full(b.n_nodes, b.n_nodes - 1)])
    assert np.array_equal(got[0][3], sizes.astype(np.float64))


def test_all_sources_non_live():
    lat = mock.perturbed_lattice(7, seed=104)
    xy = np.stack([lat.x, lat.y], 1)
    ns = _ns(xy, [tuple(e) for e in lat.edges.tolist()], live=np.zeros(lat.n_nodes, bool))
    say("no live node: nothing is a source")
    res = ns.centrality_shortest(distances=[500])
    assert all(np.all(getattr(res, k)[500] == 0.0) for k in SHORTEST_KEYS)
    _check(ns, [500], "no live nodes", segment=False)


def test_self_loops_and_parallel_edges():
    """A ring of 12 nodes with a self-loop, a second and longer edge between nodes 3 and 4
    (key 1) and a chord: the multigraph cases that edge keys tell apart."""
    n = 12
    ang = np.linspace(0.0, 2 * np.pi, n, endpoint=False)
    xy = np.stack([200.0 * np.cos(ang), 200.0 * np.sin(ang)], 1)
    edges = [(i, (i + 1) % n) for i in range(n)]
    keys = [0] * n
    lengths = [None] * n
    edges.append((3, 4))  # parallel edge, bent (longer), key 1
    keys.append(1)
    lengths.append(40.0)
    edges.append((0, 6))  # a chord, key 0
    keys.append(0)
    lengths.append(None)
    ns = NetworkStructure()
    for i, (x, y) in enumerate(xy):
        ns.add_street_node(str(i), float(x), float(y), True, 1.0)
    for (a, b), k, bend in zip(edges, keys, lengths, strict=True):
        pa, pb = list(map(float, xy[a])), list(map(float, xy[b]))
        if bend is None:
            g1, g2 = [pa, pb], [pb, pa]
        else:
            mid = [(pa[0] + pb[0]) / 2 + bend, (pa[1] + pb[1]) / 2 + bend]
            g1, g2 = [pa, mid, pb], [pb, mid, pa]
        ns.add_street_edge(a, b, k, str(a), str(b), g1)
        ns.add_street_edge(b, a, k, str(b), str(a), g2)
    # a self-loop at node 5 (a small closed ring)
    p5 = list(map(float, xy[5]))
    loop = [p5, [p5[0] + 10.0, p5[1]], [p5[0] + 10.0, p5[1] + 10.0], p5]
    ns.add_street_edge(5, 5, 0, "5", "5", loop)
    say("ring with a self-loop, a parallel edge and a chord")
    _check(ns, [150, 400, 1500], "multigraph ring", segment=False)


def _brandes(ns, src, max_seconds, tol=1e-4, max_preds=16):
    import ctypes

    from cityseer import _native

    n = ns.node_count()
    p = _native.ptr
    secs, route = np.zeros(n, np.float32), np.zeros(n, np.float32)
    sigma = np.zeros(n, np.float64)
    npred = np.zeros(n, np.uint32)
    preds = np.zeros(n * max_preds, np.uint32)
    _native.check(_native.load().cs_brandes_shortest(ns._native_handle(), src, max_seconds, SPEED, tol,
                                                     p(secs, ctypes.c_float), p(route, ctypes.c_float),
                                                     p(sigma, ctypes.c_double), p(npred, ctypes.c_uint32),
                                                     p(preds, ctypes.c_uint32), max_preds))
    got_p = [sorted(preds[v * max_preds: v * max_preds + min(npred[v], max_preds)].tolist()) for v in range(n)]
    return secs, sigma, npred, got_p


@pytest.mark.parametrize("jitter", [0.3, 0.0])
def test_long_thresholds_against_short_edges(jitter):
    """Edges of ~2 m under a 3 km threshold: no bucket width keeps a node's tie band clear of
    its own bucket, so the search folds predecessors after settling (same-bucket candidates in
    settle order). Jitter 0 with equal lengths puts exact ties everywhere."""
    lat = mock.perturbed_lattice(14, spacing=2.0, jitter=jitter, len_range=(1.0, 1.3) if jitter else (1.0, 1.0),
                                 seed=105)
    ns = mock.network_structure_from_lattice(lat)
    og = OracleGraph.from_network_structure(ns)
    say(f"2 m lattice, jitter {jitter}: shortest at [5, 20, 3000] m")
    _check(ns, [5, 20, 3000], f"2 m lattice, jitter {jitter}")
    max_seconds = int(round(3000 / SPEED))
    for src in (0, 37, ns.node_count() - 1):
        secs, sigma, npred, preds = _brandes(ns, src, max_seconds)
        ref = og.brandes_shortest(src, max_seconds)
        assert np.array_equal(secs, ref["secs"])
        assert np.array_equal(sigma, ref["sigma"])
        assert np.array_equal(npred, ref["npred"])
        assert preds == ref["preds"]
    say("same with a 5% tolerance (Phase 2)")
    res = ns.centrality_shortest(distances=[3000], tolerance=5.0)
    ref, _r = og.centrality_shortest([3000], tolerance=0.05)
    assert_close(stack(res, SHORTEST_KEYS, [3000]), ref, "2 m lattice, Phase 2")


def test_edge_cost_spread_beyond_one_ring_per_four_waves():
    """Two 1 m chains joined by a 4.5 km edge: the bucket ring spans ~6,000 bucket widths, more
    than four waves' share of a workgroup's LDS, so the launch runs fewer waves per workgroup."""
    left = [(-float(i), 0.0) for i in range(120)]
    right = [(4500.0 + float(i), 0.0) for i in range(120)]
    xy = np.array(left + right)
    edges = [(i, i + 1) for i in range(119)] + [(120 + i, 121 + i) for i in range(119)] + [(0, 120)]
    ns = _ns(xy, edges)
    say("two 1 m chains and a 4.5 km edge, thresholds [30, 6000] m")
    got = _check(ns, [30, 6000], "wide edge-cost spread")
    assert got[0][1].min() == 239.0  # at 6 km every node reaches all others

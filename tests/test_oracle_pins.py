"""Pin the oracle's beta-weighted betweenness and its angular farness and betweenness (CPU only).

The reference's tests hold no vectors for these three outputs, so each is
checked here against an evaluation of its definition that shares no code with
the oracle:
- node_betweenness_beta (centrality.rs:1788-1811,823-873): a sum over source
  and target pairs within the threshold of 0.5 * exp(-beta * d) * sigma_st(w) /
  sigma_st, with d and the shortest paths taken from networkx on the f32 edge
  lengths;
- angular farness (centrality.rs:2019-2026): offset + route / unit per target,
  with routes summed by hand from the diamond's turn angles, and from a
  Dijkstra over the (segment, entry end) state graph of small duals;
- angular betweenness (centrality.rs:2037-2096,793-821): duals of tree-shaped
  primals, where every route is unique, and the diamond's dual, whose two
  routes of equal angle split a target's seed between its two entry states.

Each check is then run against the oracle with one plausible mistake planted
(oracle.mutation) and must fail: that is what shows it pins the output.
"""

from __future__ import annotations

import math

import networkx as nx
import numpy as np
import pytest

from cityseer import config
from cityseer.tools import graphs, io, mock
import oracle.oracle as O
from oracle.oracle import OracleGraph

SPEED = config.SPEED_M_S


@pytest.fixture(autouse=True)
def _keyed_heap():
    O.set_heap_mode(False)
    yield
    O.set_heap_mode(False)


def _close(a, b, rtol=1e-6):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.allclose(a, b, rtol=rtol, atol=1e-12 * max(1.0, float(np.abs(b).max(initial=0.0))))


# ---------------------------------------------------------------- betweenness_beta
def _primal_nx(ns) -> nx.Graph:
    """Undirected networkx graph on the NetworkStructure's f32 edge lengths (one key per pair)."""
    a = ns.arrays()
    g = nx.Graph()
    g.add_nodes_from(range(ns.node_count()))
    for s, d, ln in zip(a["src"].tolist(), a["dst"].tolist(), a["length"].tolist(), strict=True):
        if s != d:
            g.add_edge(s, d, length=float(np.float32(ln)))
    return g


def _betweenness_brute(g: nx.Graph, distances, betas):
    """Per threshold: sum over ordered pairs (s, t), d(s, t) <= D, of 0.5 * sigma_st(w) / sigma_st
    (plain) and 0.5 * exp(-beta * d) * sigma_st(w) / sigma_st (beta), credited to interior w."""
    n = g.number_of_nodes()
    plain = np.zeros((len(distances), n))
    beta = np.zeros((len(distances), n))
    closest_margin = math.inf
    for s in range(n):
        dist = nx.single_source_dijkstra_path_length(g, s, weight="length")
        for t, d in dist.items():
            if t == s:
                continue
            closest_margin = min(closest_margin, min(abs(d - D) / D for D in distances))
            within = [i for i, D in enumerate(distances) if d <= D]
            if not within:
                continue
            paths = list(nx.all_shortest_paths(g, s, t, weight="length"))
            counts: dict[int, int] = {}
            for p in paths:
                for w in p[1:-1]:
                    counts[w] = counts.get(w, 0) + 1
            for i in within:
                wexp = math.exp(-betas[i] * d)
                for w, c in counts.items():
                    plain[i, w] += 0.5 * c / len(paths)
                    beta[i, w] += 0.5 * wexp * c / len(paths)
    # no pair sits so near a threshold that f32 against f64 path sums could put it on either side
    assert closest_margin > 1e-5
    return plain, beta


def _beta_cases():
    lat = mock.perturbed_lattice(7, seed=21, diag_frac=0.2)
    return {
        "mock": (lambda: io.network_structure_from_nx(graphs.nx_simple_geoms(mock.mock_graph()))[2],
                 [230, 400, 800, 5000]),
        "lattice": (lambda: mock.network_structure_from_lattice(lat), [150, 300, 450, 900]),
    }


@pytest.mark.parametrize("case", ["mock", "lattice", "diamond"])
def test_betweenness_beta_brute_force(case, diamond_graph):
    if case == "diamond":
        ns = io.network_structure_from_nx(diamond_graph)[2]
        distances = [50, 150, 250]
    else:
        make, distances = _beta_cases()[case]
        ns = make()
    og = OracleGraph.from_network_structure(ns)
    _d, betas, _s = O.thresholds(SPEED, distances=distances)
    plain, beta = _betweenness_brute(_primal_nx(ns), distances, betas)
    out, _ = og.centrality_shortest(distances, closeness=False)
    assert beta.max() > 0
    assert _close(out[5], plain), "plain betweenness"
    assert _close(out[6], beta), "betweenness_beta"
    # the checks bite: a beta seed on the source's route, or without the 0.5 pair weight, fails them
    for name in ("beta_source_route", "beta_full_seed"):
        with O.mutation(name):
            bad, _ = og.centrality_shortest(distances, closeness=False)
        assert _close(bad[5], plain)
        assert not _close(bad[6], beta), name


def test_betweenness_beta_ties_split_by_path_count(diamond_graph):
    """Diamond at 250 m: the two equal routes between nodes 0 and 3 each carry half of the pair."""
    ns = io.network_structure_from_nx(diamond_graph)[2]
    og = OracleGraph.from_network_structure(ns)
    out, _ = og.centrality_shortest([250], closeness=False)
    _d, betas, _s = O.thresholds(SPEED, distances=[250])
    side = float(np.float32(100.0))
    # ordered pairs (0, 3) and (3, 0): weight 0.5 each, half of the paths through node 1 (and node 2)
    want_b = 2 * 0.5 * 0.5 * math.exp(-betas[0] * 2 * side)
    assert np.allclose(out[5, 0], [0.0, 0.5, 0.5, 0.0], rtol=1e-9)
    assert np.allclose(out[6, 0], [0.0, want_b, want_b, 0.0], rtol=1e-6)


# ---------------------------------------------------------------- angular: diamond closed forms
def _turn(a, p, b) -> float:
    """Turn angle in degrees when travelling a -> p -> b (0 for straight on)."""
    v1 = (p[0] - a[0], p[1] - a[1])
    v2 = (b[0] - p[0], b[1] - p[1])
    c = (v1[0] * v2[0] + v1[1] * v2[1]) / (math.hypot(*v1) * math.hypot(*v2))
    return math.degrees(math.acos(max(-1.0, min(1.0, c))))


def _diamond_dual_routes(gold):
    """All-pairs angular routes of the diamond's dual, summed by hand from the turn angles.

    Dual nodes are the primal edges; a route is a sequence of primal edges in which each edge
    leaves by the end opposite the one it was entered by (the endpoint-slot rule). The diamond is
    two equilateral triangles, so the best route between two edges is the cheaper of the direct
    turn (when they share a node) and the two-turn routes around a triangle.
    """
    xy = {k: tuple(v) for k, v in gold["nodes"].items()}
    edges = [tuple(e) for e in gold["edges"]]
    hop_len = 1e-6 * 0.5 * (100.0 + 100.0)
    routes: dict[tuple[int, int], tuple[float, int]] = {}
    # enumerate admissible walks of up to three edges (every best route here has at most two turns)
    def walks(start):
        for first_exit in start:
            stack = [([start], first_exit, 0.0, 0)]
            while stack:
                path, node, ang, hops = stack.pop()
                if hops:
                    key = (edges.index(path[0]), edges.index(path[-1]))
                    if key not in routes or ang + hops * hop_len < routes[key][0] - 1e-9:
                        routes[key] = (ang + hops * hop_len, hops)
                if hops == 3:
                    continue
                cur = path[-1]
                prev = cur[0] if cur[1] == node else cur[1]
                for e in edges:
                    if e == cur or node not in e:
                        continue
                    nxt = e[0] if e[1] == node else e[1]
                    t = _turn(xy[prev], xy[node], xy[nxt])
                    stack.append((path + [e], nxt, ang + t, hops + 1))

    for e in edges:
        walks(e)
    return routes


def test_angular_farness_diamond_closed_form(diamond_graph, golden):
    gold = golden["diamond"]
    dual = graphs.nx_to_dual(diamond_graph)
    _n, _e, ns = io.network_structure_from_nx(dual)
    assert list(_n.index) == ["0_1_k0", "0_2_k0", "1_2_k0", "1_3_k0", "2_3_k0"]
    og = OracleGraph.from_network_structure(ns)
    routes = _diamond_dual_routes(gold)
    unit, offset = 90.0, 2.0
    distances = [50, 150, 250]
    _d, _b, secs = O.thresholds(SPEED, distances=distances)
    hop_secs = float(np.float32(100.0)) / SPEED
    far = np.zeros((3, 5))
    harm = np.zeros((3, 5))
    dens = np.zeros((3, 5))
    for (s, t), (r, hops) in routes.items():
        if s == t:
            continue
        for i, sec in enumerate(secs):
            if hops * hop_secs <= sec:
                dens[i, t] += 1
                far[i, t] += offset + r / unit
                harm[i, t] += 1.0 / (1.0 + r / unit)
    # a two-turn route is 180 degrees here, a one-turn route 60 or 120
    assert math.isclose(routes[(0, 4)][0], 180.0 + 2e-4, rel_tol=1e-6)
    out, _ = og.centrality_simplest(distances, unit=unit, offset=offset, betweenness=False)
    assert np.array_equal(out[0], dens)
    assert _close(out[1], far), "angular farness"
    assert _close(out[2], harm), "angular harmonic"
    for name in ("ang_far_swap", "ang_far_no_offset"):
        with O.mutation(name):
            bad, _ = og.centrality_simplest(distances, unit=unit, offset=offset, betweenness=False)
        assert _close(bad[2], harm)
        assert not _close(bad[1], far), name


def test_angular_betweenness_diamond_seed_split(diamond_graph):
    """At 250 m every two-turn pair (0_1 with 2_3, 0_2 with 1_3) has two routes of 180 degrees that
    enter the target by opposite ends: the pair's 0.5 is split 0.25 / 0.25 between the two entry
    states and credited to the middle segment of each route. At 150 m no pair has a middle."""
    dual = graphs.nx_to_dual(diamond_graph)
    _n, _e, ns = io.network_structure_from_nx(dual)
    og = OracleGraph.from_network_structure(ns)
    out, _ = og.centrality_simplest([150, 250], closeness=False)
    want = np.array([[0, 0, 0, 0, 0], [0.5, 0.5, 0.0, 0.5, 0.5]])
    assert np.allclose(out[3], want, rtol=1e-12, atol=1e-12)
    for name in ("ang_full_seed", "ang_no_split"):
        with O.mutation(name):
            bad, _ = og.centrality_simplest([150, 250], closeness=False)
        assert not np.allclose(bad[3], want, rtol=1e-12, atol=1e-12), name


# ---------------------------------------------------------------- angular: trees and state-graph Dijkstra
def _random_tree(n, seed):
    """A primal tree: each new node hangs off a random earlier node, 60-160 m away at a random bearing."""
    rng = np.random.default_rng(seed)
    x, y = [0.0], [0.0]
    edges = []
    deg = [0]
    for v in range(1, n):
        while True:
            p = int(rng.integers(0, v))
            if deg[p] < 4:
                break
        ang = rng.uniform(0, 2 * math.pi)
        r = rng.uniform(60, 160)
        x.append(x[p] + r * math.cos(ang))
        y.append(y[p] + r * math.sin(ang))
        edges.append((p, v))
        deg[p] += 1
        deg.append(1)
    x, y = np.array(x), np.array(y)
    e = np.array(edges, dtype=np.int64)
    length = np.hypot(x[e[:, 1]] - x[e[:, 0]], y[e[:, 1]] - y[e[:, 0]]).astype(np.float32)
    return mock.Lattice(x=x, y=y, edges=e, length=length)


def _state_graph_routes(prim: mock.Lattice, dual: mock.Lattice):
    """Best angular route, its seconds and its path of dual nodes, for every ordered pair of dual
    nodes, by networkx Dijkstra over (dual node, entry end) states.

    Dual node i is primal edge (a_i, b_i). A state (i, p) has entered edge i at its end p and must
    leave by the other end; crossing primal node q from edge i to edge j costs the turn angle at q
    plus 1e-6 of the dual edge's length (centrality.rs:638-667). The source leaves by either end.
    """
    pe = prim.edges
    n_d = len(pe)
    turn_cost = {}
    for k, (i, j) in enumerate(dual.edges.tolist()):
        q = int(dual.primal_node[k])
        ln = float(dual.length[k])
        for u, v in ((i, j), (j, i)):
            a = int(pe[u][0] if pe[u][1] == q else pe[u][1])
            b = int(pe[v][0] if pe[v][1] == q else pe[v][1])
            t = _turn((prim.x[a], prim.y[a]), (prim.x[q], prim.y[q]), (prim.x[b], prim.y[b]))
            turn_cost[(u, v, q)] = (t + 1e-6 * ln, ln / SPEED)
    g = nx.DiGraph()
    for (u, v, q), (w, sec) in turn_cost.items():
        other_u = int(pe[u][0] if pe[u][1] == q else pe[u][1])
        # leave u by end q: the state of u entered at its other end, or u as the source
        for src_state in ((u, other_u), (u, "src")):
            g.add_edge(src_state, (v, q), w=w, sec=sec)
    out = {}
    near_ties = 0
    for s in range(n_d):
        if (s, "src") not in g:
            continue
        dist, paths = nx.single_source_dijkstra(g, (s, "src"), weight="w")
        best: dict[int, tuple[float, list]] = {}
        for state, d in dist.items():
            node = state[0]
            if state[1] == "src":
                continue
            # a second way into this state within twice the reference's 1e-4 band would make the
            # route a tie there (several predecessors); the cases used here must have none
            cands = sorted(dist[u] + g.edges[u, state]["w"] for u in g.predecessors(state) if u in dist)
            if len(cands) > 1 and cands[1] <= cands[0] * (1 + 2e-4):
                near_ties += 1
            if node not in best or d < best[node][0]:
                best[node] = (d, paths[state])
        for t, (d, path) in best.items():
            if t == s:
                continue
            other = [dist[st] for st in dist if st[0] == t and st[1] != "src" and st != path[-1]]
            if other and min(other) <= d * (1 + 2e-4):
                near_ties += 1
            secs = sum(g.edges[a, b]["sec"] for a, b in zip(path[:-1], path[1:], strict=True))
            out[(s, t)] = (d, secs, [st[0] for st in path[1:-1]])
    return out, near_ties


@pytest.mark.parametrize("case", ["tree15", "tree30", "lattice"])
def test_angular_brute_force(case):
    if case.startswith("tree"):
        prim = _random_tree(int(case[4:]), seed=int(case[4:]))
    else:
        prim = mock.perturbed_lattice(5, seed=8)
    dual = mock.lattice_dual(prim)
    ns = mock.network_structure_from_lattice(dual)
    og = OracleGraph.from_network_structure(ns)
    routes, near_ties = _state_graph_routes(prim, dual)
    # tree duals have one route per pair; the lattice's dual has routes within the reference's
    # 1e-4 tie band of each other, where its betweenness follows the order-dependent band fold
    # (several predecessors), so only its closeness, set by best route values alone, is compared
    tree = case.startswith("tree")
    assert near_ties == 0 or not tree
    distances = [150, 400, 20000]
    _d, _b, secs = O.thresholds(SPEED, distances=distances)
    n = dual.n_nodes
    unit, offset = 180.0, 1.0
    dens, far, harm, betw = (np.zeros((3, n)) for _ in range(4))
    margin = math.inf
    for (s, t), (r, sec, middle) in routes.items():
        margin = min(margin, min(abs(sec - x) / x for x in secs))
        for i, x in enumerate(secs):
            if sec <= x:
                dens[i, t] += 1
                far[i, t] += offset + r / unit
                harm[i, t] += 1.0 / (1.0 + r / unit)
                for w in middle:
                    if w != s:
                        betw[i, w] += 0.5
    assert margin > 1e-5
    out, _ = og.centrality_simplest(distances)
    assert np.array_equal(out[0], dens)
    assert _close(out[1], far, rtol=1e-5), "angular farness"
    assert _close(out[2], harm, rtol=1e-5), "angular harmonic"
    if tree:
        assert _close(out[3], betw), "angular betweenness"
        assert betw[-1].max() > 0
        with O.mutation("ang_full_seed"):
            bad, _ = og.centrality_simplest(distances)
        assert not _close(bad[3], betw)
    with O.mutation("ang_far_no_offset"):
        bad, _ = og.centrality_simplest(distances)
    assert not _close(bad[1], far, rtol=1e-5)

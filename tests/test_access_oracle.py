"""Accessibility (host side, no GPU): the oracle's restatement pinned by hand-computed totals on a
line graph, the weight-curve helpers by closed forms, and the DataMap argument checks the
reference makes before any search (rust/src/data.rs:617-680, common.rs:311-348)."""

from __future__ import annotations

import math

import numpy as np
import pytest

from cityseer.rustalgos import data
from cityseer.rustalgos.graph import NetworkStructure
import oracle.oracle as O
from oracle.access import accessibility as oracle_access
from oracle.oracle import OracleGraph


def line_graph(n=4, step=100.0):
    ns = NetworkStructure()
    for i in range(n):
        ns.add_street_node(str(i), i * step, 0.0, True, 1.0)
    for a in range(n - 1):
        pa, pb = [a * step, 0.0], [(a + 1) * step, 0.0]
        ns.add_street_edge(a, a + 1, 0, str(a), str(a + 1), [pa, pb])
        ns.add_street_edge(a + 1, a, 0, str(a + 1), str(a), [pb, pa])
    return ns


# entries on the 4-node line 0-1-2-3 (100 m apart): (key, class, dedupe key, [(node, distance)])
LINE_ENTRIES = [
    ("A", "shop", None, [(1, 10.0)]),
    ("B", "shop", None, [(3, 5.0)]),
    ("C", "other", None, [(2, 0.0)]),
    ("D", "shop", None, [(1, 30.0), (3, 2.0)]),
    ("E", "shop", "g", [(2, 20.0)]),
    ("F", "cafe", "g", [(0, 40.0)]),
]


def line_inputs():
    node_data_map: dict[int, list] = {}
    for key, _c, _d, assigned in LINE_ENTRIES:
        for node, dist in assigned:
            node_data_map.setdefault(node, []).append((key, dist))
    dedupe_of = {k: (d if d is not None else k) for k, _c, d, _a in LINE_ENTRIES}
    lu_class = {k: c for k, c, _d, _a in LINE_ENTRIES}
    return node_data_map, dedupe_of, lu_class


def line_expected(speed, dists, betas):
    """Totals worked by hand: walk 250 s at `speed` m/s from each end node."""
    # from node 0 (node 3 lies beyond 250 s at 1 m/s): A 110, D 130 (via node 1), group g: F 40 beats E 220
    # from node 3 (node 0 lies beyond): B 5, D 2 (via node 3), A 210, group g: E 120 only
    reach = {0: {"shop": [110.0, 130.0], "cafe": [40.0]}, 3: {"shop": [5.0, 2.0, 210.0, 120.0], "cafe": []}}
    out = {}
    for src, per in reach.items():
        for cls, totals in per.items():
            unw = [sum(1 for t in totals if t <= d) for d in dists]
            wtd = [sum(math.exp(-b * t) for t in totals if t <= d) for d, b in zip(dists, betas, strict=True)]
            near = min(totals) if totals else math.nan
            out[(src, cls)] = (unw, wtd, near)
    return out


def test_oracle_line_graph_by_hand():
    O.set_heap_mode(False)
    ns = line_graph()
    og = OracleGraph.from_network_structure(ns)
    speed, dists = 1.0, [100, 250]
    _d, betas, secs = O.thresholds(speed, distances=dists)
    assert max(secs) == 250
    ndm, dedupe_of, lu_class = line_inputs()
    res = oracle_access(og, [1, 1, 1, 1], ndm, dedupe_of, lu_class, ["shop", "cafe"], dists, betas, [1.0, 1.0],
                        max(secs), speed)
    for (src, cls), (unw, wtd, near) in line_expected(speed, dists, betas).items():
        got_u, got_w, got_n = res[cls]
        assert got_u[:, src].tolist() == unw, (src, cls)
        assert np.allclose(got_w[:, src], wtd, rtol=1e-6), (src, cls)
        assert (math.isnan(near) and math.isnan(got_n[src])) or got_n[src] == near, (src, cls)
    # from node 1: A 10, D 30 (not 202 via node 3), E 120 (nearer than F 140, so group g counts
    # as "shop" and "cafe" counts nothing), B 205
    assert res["shop"][0][1, 1] == 4.0 and res["cafe"][0][1, 1] == 0.0


def test_weight_curve_helpers():
    assert data.clip_wts_curve([400, 800], [0.01, 0.005], 0) == [1.0, 1.0]
    mcw = data.clip_wts_curve([400], [0.01], 50)[0]
    assert math.isclose(mcw, math.exp(-0.5), rel_tol=1e-6)
    # within the tolerance the weight is clipped to 1; beyond, exp(-beta d) / mcw
    assert data.clipped_beta_wt(0.01, mcw, 20.0) == 1.0
    assert math.isclose(data.clipped_beta_wt(0.01, mcw, 150.0), math.exp(-1.5) / math.exp(-0.5), rel_tol=1e-6)
    with pytest.raises(ValueError, match="Clipping distance"):
        data.clip_wts_curve([40], [0.1], 50)
    with pytest.raises(ValueError, match="range"):
        data.clipped_beta_wt(0.01, 1.5, 1.0)


def test_datamap_argument_checks():
    ns = line_graph()
    dm = data.DataMap()
    dm.insert("a", "POINT (0 0)")
    dm.insert("b", "POINT (1 0)", dedupe_key_py="x")
    assert dm.entry_keys() == ["str:a", "str:b"] and dm.count() == 2 and not dm.is_empty()
    assert dm.get_entry("str:b").dedupe_key == "str:x"
    with pytest.raises(ValueError, match="number of landuse encodings"):
        dm.accessibility(ns, {"a": "shop"}, ["shop"], distances=[200])
    with pytest.raises(KeyError, match="Data entries key missing"):
        dm.accessibility(ns, {"a": "shop", "c": "shop"}, ["shop"], distances=[200])
    with pytest.raises(ValueError, match="exactly one"):
        dm.accessibility(ns, {"a": "shop", "b": "shop"}, ["shop"])
    # both points take both ends of edge 0-1 (the twin 1-0 shares its geometry)
    dm.assign_data_to_network(ns, 400.0, 6)
    assert dm.node_data_map == {0: [("str:a", 0.0), ("str:b", 1.0)], 1: [("str:a", 100.0), ("str:b", 99.0)]}
    with pytest.raises(ValueError, match="dual"):
        dm.accessibility(ns, {"a": "shop", "b": "shop"}, ["shop"], distances=[200], angular=True)


def test_oracle_diversity_closed_forms():
    from oracle import access as A

    counts = [4, 0, 2, 2]
    p = np.array([0.5, 0.25, 0.25])
    assert A.hill(counts, 0) == 3.0
    assert math.isclose(A.hill(counts, 1), math.exp(-(p * np.log(p)).sum()), rel_tol=1e-6)
    assert math.isclose(A.hill(counts, 2), 1.0 / (p**2).sum(), rel_tol=1e-6)
    assert math.isclose(A.shannon(counts), -(p * np.log(p)).sum(), rel_tol=1e-6)
    # gini-simpson: 1 - sum x (x - 1) / (n (n - 1))
    assert math.isclose(A.gini(counts), 1.0 - (4 * 3 + 2 + 2) / (8 * 7), rel_tol=1e-6)
    assert A.hill([0, 0], 1) == 0.0 and A.gini([1, 0]) == 0.0 and A.shannon([]) == 0.0
    # with unit weights (beta 0) the weighted Hill numbers are the plain ones for q = 1, 2;
    # q = 0 sums the weights, here the class count
    near = [10.0, np.inf, 50.0, 90.0]
    for q in (1, 2):
        assert math.isclose(A.hill_weighted(counts, near, q, 0.0, 1.0), A.hill(counts, q), rel_tol=1e-6)
    assert A.hill_weighted(counts, near, 0, 0.0, 1.0) == 3.0
    # q = 0 with beta > 0: the sum of exp(-beta d) over present classes
    assert math.isclose(A.hill_weighted(counts, near, 0, 0.01, 1.0), math.exp(-0.1) + math.exp(-0.5) + math.exp(-0.9),
                        rel_tol=1e-6)


def test_oracle_mixed_uses_line_graph():
    """On the 4-node line, node 0 reaches A, D (shop), C (other: every class counts here) and F
    (cafe) within 250, and only F within 100."""
    from oracle import access as A

    O.set_heap_mode(False)
    ns = line_graph()
    og = OracleGraph.from_network_structure(ns)
    _d, betas, secs = O.thresholds(1.0, distances=[100, 250])
    ndm, dedupe_of, lu_class = line_inputs()
    out = A.mixed_uses(og, [1] * 4, ndm, dedupe_of, lu_class, [100, 250], betas, [1.0, 1.0], max(secs), 1.0)
    # classes sorted: cafe, other, shop; node 0 at 250: counts [1, 1, 2]
    assert out[0, 1, 0] == 3.0
    assert math.isclose(out[2, 1, 0], 1.0 / (0.25**2 + 0.25**2 + 0.5**2), rel_tol=1e-6)
    assert math.isclose(out[7, 1, 0], 1.0 - (2 * 1) / (4 * 3), rel_tol=1e-6)
    # at 100 only F (40) counts: one class
    assert out[0, 0, 0] == 1.0 and out[6, 0, 0] == 0.0 and out[7, 0, 0] == 0.0


# ---------------------------------------------------------------- statistics (data.rs:946-1233)
def test_median_and_weighted_median_closed_forms():
    from oracle.access import median, weighted_median

    f = np.float32
    assert median([]) != median([])  # NaN
    assert median([f(3), f(1), f(2)]) == f(2)
    assert median([f(4), f(1), f(3), f(2)]) == f(2.5)
    assert median([f(-0.0), f(0.0)]) == f(0.0)
    # a weight above half the total wins outright
    assert weighted_median([(f(1), f(0.2)), (f(5), f(0.7)), (f(9), f(0.1))], f(1.0)) == f(5)
    # cumulative weight lands exactly on the midpoint: mean of that value and the next
    assert weighted_median([(f(1), f(0.25)), (f(2), f(0.25)), (f(3), f(0.25)), (f(4), f(0.25))], f(1.0)) == f(2.5)
    # passes the midpoint inside a value
    assert weighted_median([(f(1), f(0.3)), (f(2), f(0.3)), (f(3), f(0.4))], f(1.0)) == f(2)
    assert weighted_median([(f(1), f(0.0))], f(0.0)) != weighted_median([(f(1), f(0.0))], f(0.0))


def _stats_line(values, dists, speed=1.0, spatial_tolerance=0):
    from oracle.access import stats as oracle_stats

    ns = line_graph()
    og = OracleGraph.from_network_structure(ns)
    ndm, dedupe_of, _lu = line_inputs()
    _d, bets, secs = O.thresholds(speed, distances=dists)
    mcw = data.clip_wts_curve(dists, bets, spatial_tolerance)
    order = {k: i for i, (k, _c, _dd, _a) in enumerate(LINE_ENTRIES)}
    return oracle_stats(og, [1] * og.n, ndm, dedupe_of, order, [values], dists, bets, mcw, max(secs), speed)[0]


def test_stats_closed_forms_on_the_line():
    """From node 0 at 250 m (1 m/s): A 110, D 130, C 200, F 40 (E loses group g to F); from node 3:
    B 5, D 2, C 100, E 120, A 210. Values A=1, B=2, C=3, D=4, E=5, F=NaN; a tolerance equal to the
    largest distance clips every weight to 1, so the weighted forms equal the plain ones."""
    vals = {"A": 1.0, "B": 2.0, "C": 3.0, "D": 4.0, "E": 5.0, "F": float("nan")}
    st = _stats_line(vals, [250], spatial_tolerance=250)
    names = ("sum", "sum_wt", "mean", "mean_wt", "median", "median_wt", "count", "count_wt", "variance",
             "variance_wt", "mad", "mad_wt", "max", "min")
    # node 0: values {A: 1, C: 3, D: 4} (F has no value)
    want0 = dict(sum=8, count=3, mean=8 / 3, variance=14 / 9, median=3.0, mad=1.0, max=4, min=1)
    # node 3: values {B: 2, D: 4, C: 3, E: 5, A: 1}
    want3 = dict(sum=15, count=5, mean=3.0, variance=2.0, median=3.0, mad=1.0, max=5, min=1)
    for src, want in ((0, want0), (3, want3)):
        for k, v in want.items():
            assert np.isclose(st[names.index(k), 0, src], v), (src, k)
            if k in ("sum", "count", "mean", "variance", "median", "mad"):
                assert np.isclose(st[names.index(k + "_wt"), 0, src], v), (src, k + "_wt")


def test_stats_against_numpy_and_the_weight_curve():
    rng = np.random.default_rng(7)
    vals = {k: float(rng.uniform(-10, 10)) for k, _c, _d, _a in LINE_ENTRIES}
    st = _stats_line(vals, [150, 250])
    # node 3 at 250 m reaches B, D, C, E, A with totals 5, 2, 100, 120, 210
    v = np.array([vals["B"], vals["D"], vals["C"], vals["E"], vals["A"]], np.float32)
    assert np.isclose(st[0, 1, 3], v.sum(), rtol=1e-6)
    assert np.isclose(st[4, 1, 3], np.median(v), rtol=1e-6)
    assert np.isclose(st[8, 1, 3], v.var(), rtol=1e-4, atol=1e-5)
    assert np.isclose(st[10, 1, 3], np.median(np.abs(v - np.median(v))), rtol=1e-6)
    # weights: exp(-beta * total) at beta = 0.016 (250 m), none clipped
    w = np.exp(-0.016 * np.array([5.0, 2.0, 100.0, 120.0, 210.0]))
    assert np.isclose(st[1, 1, 3], (v * w).sum(), rtol=1e-5)
    assert np.isclose(st[7, 1, 3], w.sum(), rtol=1e-5)
    assert np.isclose(st[3, 1, 3], (v * w).sum() / w.sum(), rtol=1e-5)
    # at 150 m node 3 reaches B, D, C, E (A at 210 m lies beyond)
    assert st[6, 0, 3] == 4.0
    assert np.isclose(st[12, 0, 3], max(vals["B"], vals["D"], vals["C"], vals["E"]))


def test_datamap_boundary_follows_the_reference():
    """Assignments to nodes outside the network are never looked up by the reference's searches
    (data.rs:555-559), so they are ignored; land-use values must be strings (data.rs:650-660)."""
    dm = data.DataMap()
    dm.insert("a", "POINT (0 0)")
    dm.insert("b", "POINT (0 0)")
    dm.node_data_map = {0: [("str:a", 1.0)], 99: [("str:b", 2.0)], -1: [("str:b", 3.0)]}
    dm._node_bound = 10
    _keys, _grp, nodes, ents, _dists = dm._assignment_arrays()
    assert nodes.tolist() == [0] and ents.tolist() == [0]
    with pytest.raises(TypeError, match="strings"):
        dm._landuses({"a": 1, "b": "shop"})
    with pytest.raises(KeyError):
        dm._landuses({"a": "shop", "c": "shop"})


def detour_dual():
    """Dual of a primal graph where street T0-T1 is reached from street S0-S1 by a shortest route
    with three right-angle turns (300 m between midpoints) and by a longer route through D with
    101.4 degrees of turning (50 + 100 sqrt(1.01) + 110 + 50 = 310.5 m): the simplest-path tree
    (centrality.rs:1202-1332) takes the second. Returns the network and the two dual nodes."""
    import networkx as nx

    from cityseer.tools import graphs, io

    pts = {"S0": (0, 0), "S1": (100, 0), "C1": (150, 0), "C2": (150, 50), "C3": (200, 50), "T0": (200, 100),
           "T1": (200, 200), "D": (200, -10)}
    G = nx.MultiGraph()
    for k, (x, y) in pts.items():
        G.add_node(k, x=x, y=y)
    for a, b in [("S0", "S1"), ("S1", "C1"), ("C1", "C2"), ("C2", "C3"), ("C3", "T0"), ("S1", "D"), ("D", "T0"),
                 ("T0", "T1")]:
        G.add_edge(a, b)
    _n, _e, ns = io.network_structure_from_nx(graphs.nx_to_dual(graphs.nx_simple_geoms(G)))
    keys = [str(k) for k in ns.node_keys_py()]
    return ns, keys.index("S0_S1_k0"), keys.index("T0_T1_k0")


DETOUR_SIMPLEST = 50.0 + 100.0 * math.sqrt(1.01) + 110.0 + 50.0


def test_oracle_angular_aggregation_follows_the_simplest_route():
    """Angular aggregation walks the simplest-path tree (data.rs:538-548): an entry on T0-T1 is
    300 m from S0-S1 by the shortest route and 310.5 m by the simplest, so a 305 m threshold
    counts it only for the shortest search and the nearest distances differ by the detour."""
    from oracle.access import accessibility as oracle_access
    from oracle.oracle import OracleGraph

    ns, s, t = detour_dual()
    og = OracleGraph.from_network_structure(ns)
    ndm = {t: [("x", 0.0)]}
    dists = [305, 400]
    _d, bets, secs = O.thresholds(1.0, distances=dists)
    live = [1] * og.n
    args = (og, live, ndm, {"x": "x"}, {"x": "shop"}, ["shop"], dists, bets, [1.0, 1.0], max(secs), 1.0)
    sh = oracle_access(*args, sources=[s])["shop"]
    an = oracle_access(*args, sources=[s], angular=True)["shop"]
    assert sh[0][:, s].tolist() == [1.0, 1.0] and an[0][:, s].tolist() == [0.0, 1.0]
    assert sh[2][s] == 300.0
    assert abs(an[2][s] - DETOUR_SIMPLEST) < 1e-3


def test_angular_data_layer_needs_a_dual_graph():
    ns = line_graph()
    dm = data.DataMap()
    dm.insert("a", "POINT (0 0)")
    for call in (lambda: dm.accessibility(ns, {"a": "shop"}, ["shop"], distances=[200], angular=True),
                 lambda: dm.mixed_uses(ns, {"a": "shop"}, distances=[200], angular=True),
                 lambda: dm.stats(ns, [{"a": 1.0}], distances=[200], angular=True),
                 lambda: dm.aggregate_to_src_idx(0, ns, 300, 1.0, angular=True)):
        with pytest.raises(ValueError, match="requires a dual graph for angular analysis"):
            call()


def test_angular_data_layer_needs_endpoint_keys():
    """A graph flagged dual whose edges carry no shared primal node key has no endpoint slots: the
    angular data layer refuses it before any search (centrality.rs:533-566 expects the key)."""
    ns = NetworkStructure()
    for i in range(3):
        ns.add_street_node(i, i * 100.0, 0.0, True, 1.0)
    ns.add_street_edge(0, 1, 0, 0, 1, "LINESTRING (0 0, 100 0)")
    ns.add_street_edge(1, 2, 0, 1, 2, "LINESTRING (100 0, 200 0)")
    ns.set_is_dual(True)
    dm = data.DataMap()
    dm.insert("a", "POINT (0 0)")
    dm.node_data_map = {1: [("str:a", 1.0)]}
    with pytest.raises(ValueError, match="shared_primal_node_key"):
        dm.accessibility(ns, {"a": "shop"}, ["shop"], distances=[200], angular=True)
    with pytest.raises(ValueError, match="shared_primal_node_key"):
        dm.stats(ns, [{"a": 1.0}], distances=[200], angular=True)


def test_threaded_c_accessibility_equals_the_restatement():
    """og_accessibility (C, threads over sources; the CPU baseline) gives the Python
    restatement's rows: on the hand-worked line, on a random lattice assignment, and angular on the
    detour dual and a lattice dual. Counts and nearest distances exactly; weights to 1e-6 (expf
    against numpy's exp, summed in another order)."""
    from cityseer.tools import mock
    from oracle.access import accessibility_threads

    O.set_heap_mode(False)

    def same(a, b):
        for cls in a:
            assert np.array_equal(a[cls][0], b[cls][0]), cls
            np.testing.assert_allclose(a[cls][1], b[cls][1], rtol=1e-6, atol=1e-12)
            assert np.array_equal(a[cls][2], b[cls][2], equal_nan=True), cls

    ns = line_graph()
    og = OracleGraph.from_network_structure(ns)
    _d, betas, secs = O.thresholds(1.0, distances=[100, 250])
    ndm, dedupe_of, lu_class = line_inputs()
    args = (og, [1, 1, 1, 1], ndm, dedupe_of, lu_class, ["shop", "cafe"], [100, 250], betas, [1.0, 1.0], max(secs), 1.0)
    same(accessibility_threads(*args, n_threads=3), oracle_access(*args))
    rng = np.random.default_rng(31)
    for lat, angular in ((mock.perturbed_lattice(12, seed=32, diag_frac=0.1), False),
                         (mock.lattice_dual(mock.perturbed_lattice(9, seed=33)), True)):
        ns = mock.network_structure_from_lattice(lat)
        og = OracleGraph.from_network_structure(ns)
        n = ns.node_count()
        ndm, dedupe_of, lu = {}, {}, {}
        for e in range(300):
            key = f"e{e}"
            dedupe_of[key] = f"g{int(rng.integers(0, 60))}" if rng.random() < 0.3 else key
            lu[key] = ("shop", "cafe", "park")[int(rng.integers(0, 3))]
            for v in rng.choice(n, int(rng.integers(1, 3)), replace=False).tolist():
                ndm.setdefault(int(v), []).append((key, float(rng.uniform(0, 50))))
        dists = [300, 700]
        _d, betas, secs = O.thresholds(1.33333, distances=dists)
        live = (rng.random(n) < 0.9).astype(int).tolist()
        args = (og, live, ndm, dedupe_of, lu, ["shop", "park"], dists, betas, [1.0, 0.8], max(secs), 1.33333)
        same(accessibility_threads(*args, angular=angular, n_threads=4), oracle_access(*args, angular=angular))
    ns, s, t = detour_dual()
    og = OracleGraph.from_network_structure(ns)
    _d, betas, secs = O.thresholds(1.0, distances=[305, 400])
    args = (og, [1] * og.n, {t: [("x", 0.0)]}, {"x": "x"}, {"x": "shop"}, ["shop"], [305, 400], betas, [1.0, 1.0],
            max(secs), 1.0)
    got = accessibility_threads(*args, angular=True)
    assert got["shop"][0][:, s].tolist() == [0.0, 1.0] and abs(got["shop"][2][s] - DETOUR_SIMPLEST) < 1e-3

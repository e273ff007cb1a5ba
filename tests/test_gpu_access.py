"""GPU: land-use accessibility (DataMap.accessibility, rust/src/data.rs:617-747) through the C-ABI,
against hand-computed totals and against the oracle's restatement on random assignments.

Counts and nearest distances are bit-exact; weighted sums are f64 sums of the same f32 weights
in another order, compared at 1e-9 relative (plus ulp-level differences between exp
implementations: 1e-5 relative, the bar for f32 metrics)."""

from __future__ import annotations

import math

import numpy as np
import pytest

from cityseer.rustalgos import data
from cityseer.tools import mock
import oracle.oracle as O
from oracle.access import accessibility as oracle_access
from oracle.oracle import OracleGraph
from test_access_oracle import LINE_ENTRIES, line_expected, line_graph
from test_gpu_shortest import say

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _keyed():
    O.set_heap_mode(False)


def _datamap(entries):
    """entries: (key, class, dedupe key, [(node, distance)]) -> DataMap and its landuse map."""
    dm = data.DataMap()
    for key, _c, dd, _a in entries:
        dm.insert(key, "POINT (0 0)", dd)
    for key, _c, _d, assigned in entries:
        for node, dist in assigned:
            dm.node_data_map.setdefault(node, []).append((f"str:{key}", dist))
    return dm, {key: cls for key, cls, _d, _a in entries}


def test_line_graph_by_hand():
    ns = line_graph()
    dm, lu = _datamap(LINE_ENTRIES)
    speed, dists = 1.0, [100, 250]
    say("accessibility on a 4-node line, hand-computed totals")
    res = dm.accessibility(ns, lu, ["shop", "cafe"], distances=dists, speed_m_s=speed)
    _d, bets, _s = O.thresholds(speed, distances=dists)
    for (src, cls), (unw, wtd, near) in line_expected(speed, dists, bets).items():
        la = res.result[cls]
        assert [la.unweighted[d][src] for d in dists] == unw, (src, cls)
        assert np.allclose([la.weighted[d][src] for d in dists], wtd, rtol=1e-6), (src, cls)
        got_n = la.distance[250][src]
        assert (math.isnan(near) and math.isnan(got_n)) or got_n == near, (src, cls)
    assert res.result["shop"].unweighted[250][1] == 4.0 and res.result["cafe"].unweighted[250][1] == 0.0


def _random_case(ns, n_entries, seed, classes=("shop", "cafe", "park", "other")):
    rng = np.random.default_rng(seed)
    n = ns.node_count()
    entries = []
    for e in range(n_entries):
        k = int(rng.integers(1, 4))
        nodes = rng.choice(n, k, replace=False).tolist()
        assigned = [(int(v), float(rng.uniform(0.0, 60.0))) for v in nodes]
        dd = f"g{int(rng.integers(0, n_entries // 4))}" if rng.random() < 0.3 else None
        entries.append((f"e{e}", classes[int(rng.integers(0, len(classes)))], dd, assigned))
    return entries


def _compare(ns, entries, acc_keys, dists, speed, live=None, spatial_tolerance=0, angular=False):
    dm, lu = _datamap(entries)
    res = dm.accessibility(ns, lu, acc_keys, distances=dists, speed_m_s=speed, spatial_tolerance=spatial_tolerance,
                           angular=angular)
    og = OracleGraph.from_network_structure(ns)
    d_, bets, secs = O.thresholds(speed, distances=dists)
    mcw = data.clip_wts_curve(dists, bets, spatial_tolerance)
    ndm = {v: [(k[4:], dist) for k, dist in pairs] for v, pairs in dm.node_data_map.items()}
    dedupe_of = {key: (dd if dd is not None else key) for key, _c, dd, _a in entries}
    ref = oracle_access(og, live if live is not None else [1] * og.n, ndm, dedupe_of, lu, acc_keys, dists, bets,
                        mcw, max(secs), speed, angular=angular)
    for cls in acc_keys:
        la = res.result[cls]
        r_u, r_w, r_n = ref[cls]
        for i, d in enumerate(dists):
            assert np.array_equal(la.unweighted[d], r_u[i]), f"{cls} count at {d}"
            np.testing.assert_allclose(la.weighted[d], r_w[i], rtol=1e-5, atol=1e-12, err_msg=f"{cls} weighted {d}")
        assert np.array_equal(la.distance[max(dists)], r_n, equal_nan=True), f"{cls} nearest"
    return res


def test_random_assignments_lattice():
    lat = mock.perturbed_lattice(24, seed=201, diag_frac=0.1)
    ns = mock.network_structure_from_lattice(lat)
    entries = _random_case(ns, 900, 202)
    say(f"accessibility, {ns.node_count()} nodes, {len(entries)} entries, thresholds [200, 500, 900]")
    res = _compare(ns, entries, ["shop", "cafe", "park"], [200, 500, 900], 1.33333)
    assert sum(res.result["shop"].unweighted[900]) > 0


def test_random_assignments_spatial_tolerance_and_betas():
    lat = mock.perturbed_lattice(16, seed=203)
    ns = mock.network_structure_from_lattice(lat)
    entries = _random_case(ns, 300, 204)
    say(f"accessibility with spatial_tolerance 30, {ns.node_count()} nodes")
    _compare(ns, entries, ["park", "shop"], [150, 400], 1.1, spatial_tolerance=30)


def test_no_assignments_and_unknown_keys():
    ns = line_graph()
    dm, lu = _datamap([("A", "shop", None, [])])
    dm.node_data_map[2] = [("str:missing", 1.0)]
    say("accessibility with no usable assignment")
    res = dm.accessibility(ns, lu, ["shop", "none"], distances=[300])
    assert np.all(res.result["shop"].unweighted[300] == 0.0)
    assert np.all(np.isnan(res.result["none"].distance[300]))


def _compare_mixed(ns, entries, dists, speed, spatial_tolerance=0, angular=False):
    from oracle.access import mixed_uses as oracle_mixed

    dm, lu = _datamap(entries)
    res = dm.mixed_uses(ns, lu, distances=dists, speed_m_s=speed, compute_shannon=True, compute_gini=True,
                        spatial_tolerance=spatial_tolerance, angular=angular)
    og = OracleGraph.from_network_structure(ns)
    _d, bets, secs = O.thresholds(speed, distances=dists)
    mcw = data.clip_wts_curve(dists, bets, spatial_tolerance)
    ndm = {v: [(k[4:], dist) for k, dist in pairs] for v, pairs in dm.node_data_map.items()}
    dedupe_of = {key: (dd if dd is not None else key) for key, _c, dd, _a in entries}
    ref = oracle_mixed(og, [1] * og.n, ndm, dedupe_of, lu, dists, bets, mcw, max(secs), speed, angular=angular)
    for i, d in enumerate(dists):
        # richness is an exact count; the rest are f32 values of exp/log/pow, whose last bits may
        # differ between the device library and the host's (1e-5 relative)
        assert np.array_equal(res.hill[0][d], ref[0, i]), f"hill q0 at {d}"
        for name, got, m in (("hill q1", res.hill[1][d], 1), ("hill q2", res.hill[2][d], 2),
                             ("weighted q0", res.hill_weighted[0][d], 3), ("weighted q1", res.hill_weighted[1][d], 4),
                             ("weighted q2", res.hill_weighted[2][d], 5), ("shannon", res.shannon[d], 6),
                             ("gini", res.gini[d], 7)):
            np.testing.assert_allclose(got, ref[m, i], rtol=1e-5, atol=1e-6, err_msg=f"{name} at {d}")
    return res


def test_mixed_uses_line_graph_and_lattice():
    ns = line_graph()
    say("mixed uses on the 4-node line")
    res = _compare_mixed(ns, LINE_ENTRIES, [100, 250], 1.0)
    assert res.hill[0][250][0] == 3.0 and res.hill[0][100][0] == 1.0
    lat = mock.perturbed_lattice(20, seed=205, diag_frac=0.1)
    ns = mock.network_structure_from_lattice(lat)
    entries = _random_case(ns, 700, 206, classes=tuple(f"c{i}" for i in range(9)))
    say(f"mixed uses, {ns.node_count()} nodes, {len(entries)} entries, 9 classes")
    res = _compare_mixed(ns, entries, [200, 600], 1.33333, spatial_tolerance=20)
    assert res.hill[0][600].max() > 3
    dm, lu = _datamap(entries)
    only = dm.mixed_uses(ns, lu, distances=[200, 600], compute_hill=False, spatial_tolerance=20)
    assert np.all(only.hill[1][600] == 0.0) and np.array_equal(only.hill_weighted[1][600], res.hill_weighted[1][600])
    with pytest.raises(ValueError, match="compute_"):
        dm.mixed_uses(ns, lu, distances=[200], compute_hill=False, compute_hill_weighted=False)


def test_large_dedupe_groups():
    """Dedupe groups of up to 2,000 members spread over the lattice (polygon-like features): one
    winner per group, through its nearest member, as the oracle's dictionary walk picks it."""
    lat = mock.perturbed_lattice(40, seed=207, diag_frac=0.05)
    ns = mock.network_structure_from_lattice(lat)
    rng = np.random.default_rng(208)
    n = ns.node_count()
    entries = []
    # one group of 2,000 single-node entries, two of 300 multi-node entries, the rest singletons
    for e in range(2000):
        entries.append((f"big{e}", ("shop", "cafe")[e % 2], "poly0",
                        [(int(rng.integers(0, n)), float(rng.uniform(0.0, 40.0)))]))
    for grp in (1, 2):
        for e in range(300):
            nodes = rng.choice(n, 3, replace=False).tolist()
            entries.append((f"g{grp}_{e}", "park", f"poly{grp}", [(int(v), float(rng.uniform(0.0, 40.0))) for v in nodes]))
    entries += _random_case(ns, 400, 209)
    say(f"accessibility, {len(entries)} entries in groups of up to 2,000 members")
    res = _compare(ns, entries, ["shop", "cafe", "park"], [300, 700], 1.33333)
    # each group counts at most once per source
    assert res.result["park"].unweighted[700].max() <= 2 + 400
    assert (res.result["shop"].unweighted[700] + res.result["cafe"].unweighted[700]).max() >= 1


def test_angular_detour_by_hand():
    """Angular aggregation walks the simplest-path tree (data.rs:538-548): the entry on T0-T1 is
    310.5 m from S0-S1 along the simplest route (300 m along the shortest)."""
    from test_access_oracle import DETOUR_SIMPLEST, detour_dual

    ns, s, t = detour_dual()
    dm = data.DataMap()
    dm.insert("x", "POINT (0 0)")
    dm.node_data_map = {t: [("str:x", 0.0)]}
    say("angular accessibility on the detour dual")
    res = dm.accessibility(ns, {"x": "shop"}, ["shop"], distances=[305, 400], speed_m_s=1.0, angular=True)
    assert res.result["shop"].unweighted[305][s] == 0.0 and res.result["shop"].unweighted[400][s] == 1.0
    assert abs(res.result["shop"].distance[400][s] - DETOUR_SIMPLEST) < 1e-3
    short = dm.accessibility(ns, {"x": "shop"}, ["shop"], distances=[305, 400], speed_m_s=1.0)
    assert short.result["shop"].unweighted[305][s] == 1.0 and short.result["shop"].distance[400][s] == 300.0
    # the single-source form, tree on the device and the walk on the host
    assert dm.aggregate_to_src_idx(s, ns, 400, 1.0) == {"str:x": 300.0}
    got = dm.aggregate_to_src_idx(s, ns, 400, 1.0, angular=True)
    assert list(got) == ["str:x"] and abs(got["str:x"] - DETOUR_SIMPLEST) < 1e-3
    _compare(ns, [("x", "shop", None, [(t, 0.0)])], ["shop"], [305, 400], 1.0, angular=True)


def test_angular_random_assignments_lattice_dual():
    ns = mock.network_structure_from_lattice(mock.lattice_dual(mock.perturbed_lattice(14, seed=211)))
    entries = _random_case(ns, 500, 212)
    say(f"angular accessibility, {ns.node_count()} dual nodes, {len(entries)} entries")
    res = _compare(ns, entries, ["shop", "cafe", "park"], [200, 500, 900], 1.33333, angular=True)
    assert res.result["shop"].unweighted[900].sum() > 0
    say("angular mixed uses with a spatial tolerance")
    entries = _random_case(ns, 400, 213, classes=tuple(f"c{i}" for i in range(7)))
    _compare_mixed(ns, entries, [250, 700], 1.33333, spatial_tolerance=25, angular=True)


def test_angular_aggregate_to_src_idx_matches_oracle():
    from oracle.access import aggregate_to_src_idx as oracle_aggregate

    ns = mock.network_structure_from_lattice(mock.lattice_dual(mock.perturbed_lattice(10, seed=214)))
    entries = _random_case(ns, 150, 215)
    dm, _lu = _datamap(entries)
    og = OracleGraph.from_network_structure(ns)
    ndm = {v: pairs for v, pairs in dm.node_data_map.items()}
    dedupe_of = {k: e.dedupe_key for k, e in dm.entries.items()}
    say("aggregate_to_src_idx, shortest and angular, 12 sources")
    for src in range(0, ns.node_count(), max(1, ns.node_count() // 12)):
        for angular in (False, True):
            tree = og.tree_simplest(src, 600, 1.33333) if angular else og.tree_shortest(src, 600, 1.33333)
            want = {k: float(v) for k, v in oracle_aggregate(tree, ndm, dedupe_of, 600, 1.33333).items()}
            assert dm.aggregate_to_src_idx(src, ns, 600, 1.33333, angular=angular) == want, (src, angular)

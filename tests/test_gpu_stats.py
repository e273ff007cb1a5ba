"""GPU: DataMap.stats (data.rs:1013-1233) against the oracle's restatement: the hand-worked line
graph, then random lattices with several numeric maps, missing values, dedupe groups (one of
2,000 members), a spatial tolerance and more thresholds than one launch takes. Counts, extremes,
medians and median deviations must match exactly (they select values); sums, means and variances
within 1e-5 relative per row (f32 in the same order on both sides)."""

from __future__ import annotations

import numpy as np
import pytest

from cityseer.rustalgos import data
from cityseer.tools import mock
import oracle.oracle as O
from oracle.access import STAT_NAMES
from oracle.access import stats as oracle_stats
from oracle.oracle import OracleGraph
from parity import assert_close
from test_access_oracle import LINE_ENTRIES, line_graph
from test_gpu_access import _datamap, _random_case
from test_gpu_shortest import say

pytestmark = pytest.mark.gpu

EXACT = ("count", "max", "min", "median", "mad", "median_wt", "mad_wt")


@pytest.fixture(autouse=True)
def _keyed():
    O.set_heap_mode(False)


def _compare(ns, entries, maps, dists, speed, spatial_tolerance=0, angular=False):
    dm, _lu = _datamap(entries)
    num_maps = [{k: v for k, v in m.items()} for m in maps]
    res = dm.stats(ns, num_maps, distances=dists, speed_m_s=speed, spatial_tolerance=spatial_tolerance,
                   angular=angular)
    og = OracleGraph.from_network_structure(ns)
    _d, bets, secs = O.thresholds(speed, distances=dists)
    mcw = data.clip_wts_curve(dists, bets, spatial_tolerance)
    ndm = {v: [(k[4:], dist) for k, dist in pairs] for v, pairs in dm.node_data_map.items()}
    dedupe_of = {key: (dd if dd is not None else key) for key, _c, dd, _a in entries}
    order = {key: i for i, (key, _c, _dd, _a) in enumerate(entries)}
    ref = oracle_stats(og, [1] * og.n, ndm, dedupe_of, order, maps, dists, bets, mcw, max(secs), speed,
                       angular=angular)
    for mp, st in enumerate(res.result):
        for m, name in enumerate(STAT_NAMES):
            got = np.stack([getattr(st, name)[d] for d in dists])
            want = ref[mp, m]
            assert np.array_equal(np.isnan(got), np.isnan(want)), f"map {mp} {name}: NaN pattern"
            ok = ~np.isnan(want)
            if name in EXACT:
                assert np.array_equal(got[ok], want[ok]), f"map {mp} {name}"
            else:
                g = np.where(ok, got, 0.0)
                w = np.where(ok, want, 0.0)
                assert_close(g, w, f"map {mp} {name}")
    return res


def test_line_graph_closed_forms():
    ns = line_graph()
    vals = {"A": 1.0, "B": 2.0, "C": 3.0, "D": 4.0, "E": 5.0, "F": float("nan")}
    say("stats on the 4-node line")
    res = _compare(ns, LINE_ENTRIES, [vals], [250], 1.0, spatial_tolerance=250)
    st = res.result[0]
    assert st.sum[250][0] == 8.0 and st.count[250][0] == 3.0 and st.median[250][0] == 3.0
    assert st.sum[250][3] == 15.0 and st.variance[250][3] == 2.0 and st.mad[250][3] == 1.0


def test_random_lattices():
    lat = mock.perturbed_lattice(22, seed=401, diag_frac=0.1)
    ns = mock.network_structure_from_lattice(lat)
    entries = _random_case(ns, 700, 402)
    rng = np.random.default_rng(403)
    maps = []
    for _ in range(3):
        vals = {key: float(rng.normal(5.0, 3.0)) for key, _c, _d, _a in entries}
        for key in rng.choice([e[0] for e in entries], 50, replace=False):
            vals[str(key)] = float("nan")
        maps.append(vals)
    # integer-valued map: many exact ties for the medians
    maps.append({key: float(rng.integers(0, 4)) for key, _c, _d, _a in entries})
    say(f"stats, {ns.node_count()} nodes, {len(entries)} entries, 4 maps, [150, 400, 800]")
    _compare(ns, entries, maps, [150, 400, 800], 1.33333)
    say("ten thresholds, spatial tolerance 40")
    dists = [100, 150, 200, 300, 400, 500, 600, 700, 800, 900]
    _compare(ns, entries, maps[:1], dists, 1.33333, spatial_tolerance=40)


def test_large_dedupe_group():
    lat = mock.perturbed_lattice(30, seed=404)
    ns = mock.network_structure_from_lattice(lat)
    rng = np.random.default_rng(405)
    n = ns.node_count()
    entries = [(f"p{e}", "x", "poly", [(int(rng.integers(0, n)), float(rng.uniform(0, 40)))]) for e in range(2000)]
    entries += _random_case(ns, 300, 406)
    vals = {key: float(rng.uniform(0, 100)) for key, _c, _d, _a in entries}
    say(f"stats with a dedupe group of 2,000 members, {n} nodes")
    _compare(ns, entries, [vals], [300, 700], 1.33333)


def test_groups_after_a_larger_graph():
    """Dedupe groups keep per-wave records just past the node records of the graph in hand, in
    words that a larger graph's searches used before. From freshly released workspaces, a
    betweenness run on a larger lattice leaves path counts (doubles) in the payload words of those
    records; statistics and accessibility on a smaller lattice, run next with the same waves, must
    still count every group through its nearest member."""
    from cityseer import _native

    from test_gpu_access import _compare as access_compare

    _native.check(_native.load().cs_release_workspaces())
    big = mock.network_structure_from_lattice(mock.perturbed_lattice(40, seed=407, diag_frac=0.1))
    small = mock.network_structure_from_lattice(mock.perturbed_lattice(22, seed=408, diag_frac=0.1))
    entries = _random_case(small, 700, 409)
    rng = np.random.default_rng(410)
    vals = {key: float(rng.normal(5.0, 3.0)) for key, _c, _d, _a in entries}
    for round_ in range(2):
        say(f"betweenness on {big.node_count()} nodes, then stats and accessibility on {small.node_count()}")
        big.centrality_shortest(distances=[4000], compute_closeness=False)
        _compare(small, entries, [vals], [150, 400, 800], 1.33333)
        big.centrality_shortest(distances=[4000], compute_closeness=False)
        access_compare(small, entries, ["shop", "cafe", "park"], [200, 500, 900], 1.33333)


def test_angular_lattice_dual_ten_thresholds():
    """Angular statistics (data.rs:538-548 through stats): simplest-path trees on a lattice dual,
    ten thresholds (two launches of the aggregation over one set of reached-node lists)."""
    ns = mock.network_structure_from_lattice(mock.lattice_dual(mock.perturbed_lattice(12, seed=411)))
    entries = _random_case(ns, 400, 412)
    rng = np.random.default_rng(413)
    vals = {key: float(rng.normal(5.0, 3.0)) for key, _c, _d, _a in entries}
    for key in rng.choice([e[0] for e in entries], 30, replace=False):
        vals[str(key)] = float("nan")
    ints = {key: float(rng.integers(0, 4)) for key, _c, _d, _a in entries}
    dists = [100, 150, 200, 300, 400, 500, 600, 700, 800, 900]
    say(f"angular stats, {ns.node_count()} dual nodes, 2 maps, ten thresholds")
    _compare(ns, entries, [vals, ints], dists, 1.33333, spatial_tolerance=30, angular=True)

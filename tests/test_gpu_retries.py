"""GPU: sources whose search outgrows the per-wave (per-lane, for simplest paths) workspace are
re-run with a larger one and give the same results. A tiny slot capacity forces the retry path of
the plain shortest call, the plain simplest call and a served slice that packs several simplest
requests (padded item lists)."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from cityseer import _native
from cityseer.serving import CentralityServer
from cityseer.tools import mock
import oracle.oracle as O
from oracle.oracle import OracleGraph
from test_gpu_shortest import assert_close, say, stack

pytestmark = pytest.mark.gpu

SHORTEST = ("node_density", "node_farness", "node_cycles", "node_harmonic", "node_beta", "node_betweenness",
            "node_betweenness_beta")
SIMPLEST = ("node_density", "node_farness", "node_harmonic", "node_betweenness")


@pytest.fixture(autouse=True)
def _keyed():
    O.set_heap_mode(False)


@pytest.fixture
def tiny_slots():
    lib = _native.load()
    old = _native.CsTuning()
    _native.check(lib.cs_get_tuning(ctypes.byref(old)))
    t = _native.CsTuning(old.block_threads, old.waves_per_simd, old.delta_scale, 48, old.xcd_split)
    _native.check(lib.cs_set_tuning(ctypes.byref(t)))
    yield
    _native.check(lib.cs_set_tuning(ctypes.byref(old)))


def test_shortest_retries_match_oracle(tiny_slots):
    ns = mock.network_structure_from_lattice(mock.perturbed_lattice(18, seed=51, diag_frac=0.1))
    og = OracleGraph.from_network_structure(ns)
    d = [300, 900]
    say("shortest with 48 slots per wave")
    res = ns.centrality_shortest(distances=d)
    assert _native.last_stats()["retries"] > 0
    ref, _ = og.centrality_shortest(d)
    got = stack(res, SHORTEST, d)
    assert np.array_equal(got[0], ref[0])
    assert_close(got, ref, "shortest after retries")


def test_simplest_retries_match_oracle_plain_and_served(tiny_slots):
    a = mock.network_structure_from_lattice(mock.lattice_dual(mock.perturbed_lattice(11, seed=52)))
    b = mock.network_structure_from_lattice(mock.lattice_dual(mock.perturbed_lattice(9, seed=53)))
    d = [400, 900]
    refs = []
    for ns in (a, b):
        ref, _ = OracleGraph.from_network_structure(ns).centrality_simplest(d)
        refs.append(ref)
    say("simplest with 48 states per lane")
    got = stack(a.centrality_simplest(distances=d), SIMPLEST, d)
    assert _native.last_stats()["retries"] > 0
    assert np.array_equal(got[0], refs[0][0])
    assert_close(got, refs[0], "simplest after retries")
    say("served simplest slice of two requests, 48 states per lane")
    with CentralityServer() as srv:
        ra = srv.submit_simplest(a, distances=d)
        rb = srv.submit_simplest(b, distances=d)
        for r, ref, what in ((ra, refs[0], "served a"), (rb, refs[1], "served b")):
            g = stack(r.result(), SIMPLEST, d)
            assert np.array_equal(g[0], ref[0]), what
            assert_close(g, ref, what)


def test_angular_access_retries_match_oracle(tiny_slots):
    """Angular aggregation with reached-node lists of 24 nodes per source at first: sources that
    reach more go round again with four times the room, until every list fits."""
    from cityseer.rustalgos import data
    from oracle.access import accessibility as oracle_access
    from test_gpu_access import _datamap, _random_case

    ns = mock.network_structure_from_lattice(mock.lattice_dual(mock.perturbed_lattice(13, seed=54)))
    entries = _random_case(ns, 300, 55)
    dm, lu = _datamap(entries)
    dists = [300, 1200]
    say("angular accessibility with 48 state slots per lane")
    res = dm.accessibility(ns, lu, ["shop", "cafe"], distances=dists, angular=True)
    assert _native.last_stats()["retries"] > 0
    og = OracleGraph.from_network_structure(ns)
    _d, bets, secs = O.thresholds(1.33333, distances=dists)
    ndm = {v: [(k[4:], d) for k, d in pairs] for v, pairs in dm.node_data_map.items()}
    dedupe_of = {key: (dd if dd is not None else key) for key, _c, dd, _a in entries}
    ref = oracle_access(og, [1] * og.n, ndm, dedupe_of, lu, ["shop", "cafe"], dists, bets,
                        data.clip_wts_curve(dists, bets, 0), max(secs), 1.33333, angular=True)
    for cls in ("shop", "cafe"):
        for i, d in enumerate(dists):
            assert np.array_equal(res.result[cls].unweighted[d], ref[cls][0][i]), (cls, d)
            np.testing.assert_allclose(res.result[cls].weighted[d], ref[cls][1][i], rtol=1e-5, atol=1e-12)
        assert np.array_equal(res.result[cls].distance[1200], ref[cls][2], equal_nan=True)

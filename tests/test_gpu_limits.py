"""GPU: calls beyond one launch's shape, against the oracle.

- Any number of thresholds: the kernels take eight per launch, so a call with more runs one
  launch per chunk of eight, each searching to the call's largest cutoff as the reference's single
  search does (centrality.rs:1624-1655 takes any list of distances). Ten and twenty thresholds,
  every entry point.
- Nodes with more than 64 out-edges (the reference keeps uncapped predecessor lists,
  centrality.rs:396): predecessor sets at such rows hold up to four positions anywhere in the row.
  A wheel of 100 spokes, and a hub whose three tied predecessors sit at positions 100-102 of its
  103-edge row; five tied predecessors at a wide row are refused with a ValueError.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from cityseer import _native, rustalgos
from cityseer.rustalgos.centrality import OdMatrix
from cityseer.tools import mock
import oracle.oracle as O
from oracle.oracle import OracleGraph
from parity import assert_close
from test_gpu_shortest import SEGMENT_KEYS, SHORTEST_KEYS, SPEED, say, stack

pytestmark = pytest.mark.gpu

SIMPLEST_KEYS = ("node_density", "node_farness", "node_harmonic", "node_betweenness")
TEN = [100, 150, 200, 300, 400, 500, 600, 700, 850, 1000]
TWENTY = list(range(50, 1001, 50))


@pytest.fixture(autouse=True)
def _keyed():
    O.set_heap_mode(False)


@pytest.fixture(scope="module")
def lattice():
    lat = mock.perturbed_lattice(24, seed=61, diag_frac=0.05)
    ns = mock.network_structure_from_lattice(lat)
    return ns, OracleGraph.from_network_structure(ns)


@pytest.mark.parametrize("dists", [TEN, TWENTY], ids=["10", "20"])
def test_many_thresholds_shortest_and_segment(lattice, dists):
    ns, og = lattice
    say(f"shortest, {len(dists)} thresholds")
    res = ns.centrality_shortest(distances=dists)
    ref, _ = og.centrality_shortest(dists)
    got = stack(res, SHORTEST_KEYS, dists)
    assert np.array_equal(got[0], ref[0])
    assert_close(got, ref, f"shortest, {len(dists)} thresholds")
    say("source subset: reachability totals per threshold")
    src = np.random.default_rng(62).choice(ns.node_count(), 50, replace=False).tolist()
    res = ns.centrality_shortest(distances=dists, source_indices=src)
    plan = ns._sampling_plan(None, None, None, src)
    ref, reach = og.centrality_shortest(dists, sources=plan["sources"], wt=plan["wt"], eligible=plan["eligible"],
                                        track=True, scale=plan["scale"])
    assert_close(stack(res, SHORTEST_KEYS, dists), ref, "shortest subset")
    assert res.reachability_totals == reach.tolist()
    say(f"segment, {len(dists)} thresholds")
    seg = ns.segment_centrality(distances=dists)
    assert_close(stack(seg, SEGMENT_KEYS, dists), og.segment_centrality(dists), "segment")


def test_many_thresholds_simplest():
    dual = mock.lattice_dual(mock.perturbed_lattice(12, seed=63))
    ns = mock.network_structure_from_lattice(dual)
    og = OracleGraph.from_network_structure(ns)
    say("simplest, 10 thresholds")
    res = ns.centrality_simplest(distances=TEN)
    ref, _ = og.centrality_simplest(TEN)
    got = stack(res, SIMPLEST_KEYS, TEN)
    assert np.array_equal(got[0], ref[0])
    assert_close(got, ref, "simplest, 10 thresholds")


def test_many_thresholds_od_and_device(lattice):
    ns, og = lattice
    n = ns.node_count()
    rng = np.random.default_rng(64)
    o, d = rng.integers(0, n, 400), rng.integers(0, n, 400)
    w = rng.uniform(0.5, 3.0, 400).astype(np.float32)
    say("OD betweenness, 10 thresholds")
    res = ns.betweenness_od_shortest(OdMatrix(o.tolist(), d.tolist(), w.tolist()), distances=TEN)
    ref = og.betweenness_od_shortest((o, d, w), TEN)
    got = np.stack([np.stack([getattr(res, k)[x] for x in TEN]) for k in ("node_betweenness",
                                                                          "node_betweenness_beta")])
    assert_close(got, ref, "OD, 10 thresholds")
    say("device entry point, 20 thresholds")
    from test_gpu_node_order import _Hip

    want = stack(ns.centrality_shortest(distances=TWENTY), SHORTEST_KEYS, TWENTY)
    dd, bb, ss = rustalgos.pair_distances_betas_time(SPEED, distances=TWENTY)
    hd, hb, hs = np.asarray(dd, np.uint32), np.asarray(bb, np.float32), np.asarray(ss, np.uint32)
    hip = _Hip()
    try:
        acc0 = np.zeros((7, len(TWENTY), n), np.float64)
        d_src, d_wt = hip.put(np.arange(n, dtype=np.uint32)), hip.put(np.ones(n, np.float32))
        d_el, d_acc = hip.put(np.ones(n, np.uint8)), hip.put(acc0)
        p = _native.CsParams()
        p.n_thresholds = len(TWENTY)
        p.distances = _native.ptr(hd, ctypes.c_uint32)
        p.betas = _native.ptr(hb, ctypes.c_float)
        p.seconds = _native.ptr(hs, ctypes.c_uint32)
        p.speed_m_s = SPEED
        p.tolerance = 1e-4
        p.compute_closeness = 1
        p.compute_betweenness = 1
        p.n_sources = n
        p.sources = ctypes.cast(d_src, ctypes.POINTER(ctypes.c_uint32))
        p.source_wt = ctypes.cast(d_wt, ctypes.POINTER(ctypes.c_float))
        p.eligible = ctypes.cast(d_el, ctypes.POINTER(ctypes.c_uint8))
        p.betw_scale = 1.0
        p.angular_scaling_unit = 180.0
        p.farness_scaling_offset = 1.0
        lib = _native.load()
        _native.check(lib.cs_centrality_shortest_device(ns._native_handle(), ctypes.byref(p), d_wt, d_acc, None,
                                                        None))
        got = hip.get(d_acc, acc0)
        assert np.array_equal(got[0], want[0])
        assert_close(got, want, "device entry point, 20 thresholds")
    finally:
        hip.free()


# ---------------------------------------------------------------- wide rows
def _hub(n_tied: int, n_spokes: int) -> mock.Lattice:
    """Source S, n_tied middle nodes at 100 m from S and from hub H, and n_spokes spokes on H.

    The middle-to-hub edges go in first, so they sit at the end of the hub's out-row (rows list
    the newest edge first): positions n_spokes .. n_spokes + n_tied - 1."""
    S, H = 0, 1
    mids = list(range(2, 2 + n_tied))
    spokes = list(range(2 + n_tied, 2 + n_tied + n_spokes))
    n = 2 + n_tied + n_spokes
    ang = np.linspace(0, 2 * np.pi, n_spokes, endpoint=False)
    x = np.concatenate([[0.0, 200.0], np.full(n_tied, 100.0), 200.0 + 60.0 * np.cos(ang)])
    y = np.concatenate([[0.0, 0.0], np.linspace(-50, 50, n_tied), 60.0 * np.sin(ang)])
    edges = [(m, H) for m in mids] + [(S, m) for m in mids] + [(H, k) for k in spokes]
    length = [100.0] * (2 * n_tied) + [50.0 + 0.37 * i for i in range(n_spokes)]
    # a ring through the spokes gives the hub's neighbours other routes than the hub
    edges += [(spokes[i], spokes[(i + 1) % n_spokes]) for i in range(n_spokes)]
    length += [37.0 + 0.11 * i for i in range(n_spokes)]
    return mock.Lattice(x=x, y=y, edges=np.asarray(edges, np.int64), length=np.asarray(length, np.float32))


def _wheel(n_spokes: int, seed: int) -> mock.Lattice:
    rng = np.random.default_rng(seed)
    ang = np.sort(rng.uniform(0, 2 * np.pi, n_spokes))
    r = rng.uniform(150, 250, n_spokes)
    x = np.concatenate([[0.0], r * np.cos(ang)])
    y = np.concatenate([[0.0], r * np.sin(ang)])
    edges = [(0, k + 1) for k in range(n_spokes)] + [(k + 1, (k + 1) % n_spokes + 1) for k in range(n_spokes)]
    e = np.asarray(edges, np.int64)
    length = (np.hypot(x[e[:, 1]] - x[e[:, 0]], y[e[:, 1]] - y[e[:, 0]]) * rng.uniform(1.0, 1.3, len(e)))
    return mock.Lattice(x=x, y=y, edges=e, length=length.astype(np.float32))


@pytest.mark.parametrize("case", ["wheel100", "hub3"])
def test_wide_rows_match_oracle(case):
    lat = _wheel(100, 65) if case == "wheel100" else _hub(3, 100)
    ns = mock.network_structure_from_lattice(lat)
    og = OracleGraph.from_network_structure(ns)
    dists = [150, 300, 600]
    say(f"{ns.node_count()} nodes, widest row {lat.n_edges} edges: shortest")
    res = ns.centrality_shortest(distances=dists)
    ref, _ = og.centrality_shortest(dists)
    got = stack(res, SHORTEST_KEYS, dists)
    assert np.array_equal(got[0], ref[0])
    assert_close(got, ref, "wide rows")
    say("Brandes state of every source")
    lib = _native.load()
    h = ns._native_handle()
    n = ns.node_count()
    max_s = int(round(600 / SPEED))
    p = _native.ptr
    for src in range(n):
        secs, route = np.zeros(n, np.float32), np.zeros(n, np.float32)
        sigma = np.zeros(n, np.float64)
        npred = np.zeros(n, np.uint32)
        preds = np.zeros(n * 16, np.uint32)
        _native.check(lib.cs_brandes_shortest(h, src, max_s, SPEED, 1e-4, p(secs, ctypes.c_float),
                                              p(route, ctypes.c_float), p(sigma, ctypes.c_double),
                                              p(npred, ctypes.c_uint32), p(preds, ctypes.c_uint32), 16))
        want = og.brandes_shortest(src, max_s)
        assert np.array_equal(sigma, want["sigma"]), src
        assert np.array_equal(npred, want["npred"]), src
        assert [sorted(preds[v * 16: v * 16 + min(npred[v], 16)].tolist()) for v in range(n)] == want["preds"], src
    if case == "hub3":  # from S, the hub's three tied predecessors
        hub_state = og.brandes_shortest(0, max_s)
        assert hub_state["preds"][1] == [2, 3, 4] and hub_state["sigma"][1] == 3.0


def test_five_tied_predecessors_at_a_wide_row_are_refused():
    ns = mock.network_structure_from_lattice(_hub(5, 100))
    say("hub with five tied predecessors: refused")
    with pytest.raises(ValueError, match="tied predecessors"):
        ns.centrality_shortest(distances=[600])
    # closeness alone keeps no predecessors
    og = OracleGraph.from_network_structure(ns)
    res = ns.centrality_shortest(distances=[600], compute_betweenness=False)
    ref, _ = og.centrality_shortest([600], betweenness=False)
    assert_close(stack(res, SHORTEST_KEYS, [600]), ref, "closeness at a wide row")

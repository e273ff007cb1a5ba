"""GPU: the device node order (nodes laid out along a Hilbert curve of their coordinates)
changes no result. Graphs here are inserted in random order, so the device order differs
from the original one almost everywhere; every entry point must still match the oracle
(which works in original order) and the same graph run in its original order."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from cityseer import _native, rustalgos
from cityseer.tools import mock
import oracle.oracle as O
from oracle.oracle import OracleGraph
from test_gpu_shortest import SEGMENT_KEYS, SHORTEST_KEYS, SPEED, assert_close, say, stack

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _keyed():
    O.set_heap_mode(False)


def shuffled(lat: mock.Lattice, seed: int) -> mock.Lattice:
    """The same network with its nodes inserted in a random order (edges keep theirs)."""
    rng = np.random.default_rng(seed)
    p = rng.permutation(lat.n_nodes)  # new node i is old node p[i]
    inv = np.empty_like(p)
    inv[p] = np.arange(len(p))
    return mock.Lattice(x=lat.x[p], y=lat.y[p], edges=inv[lat.edges], length=lat.length, meta=lat.meta)


def pair(lat):
    """The network twice: with the device node order and in insertion order."""
    a = mock.network_structure_from_lattice(lat)
    b = mock.network_structure_from_lattice(lat)
    b._spatial_order = False
    return a, b


def node_order(ns) -> np.ndarray:
    order = np.empty(ns.node_count(), np.uint32)
    _native.check(_native.load().cs_graph_node_order(ns._native_handle(), _native.ptr(order, ctypes.c_uint32)))
    return order


@pytest.fixture(scope="module")
def jittered():
    lat = shuffled(mock.perturbed_lattice(26, seed=81, diag_frac=0.05), seed=82)
    a, b = pair(lat)
    return a, b, OracleGraph.from_network_structure(a)


@pytest.fixture(scope="module")
def ties():
    # unit spacing, no jitter, equal lengths: exact distance ties everywhere
    lat = shuffled(mock.perturbed_lattice(18, jitter=0.0, len_range=(1.0, 1.0), seed=83), seed=84)
    a, b = pair(lat)
    return a, b, OracleGraph.from_network_structure(a)


def test_device_order_is_a_real_permutation(jittered):
    a, b, _og = jittered
    n = a.node_count()
    order = node_order(a)
    assert sorted(order.tolist()) == list(range(n))
    moved = int((order != np.arange(n)).sum())
    say(f"{moved} of {n} nodes move")
    assert moved > n // 2
    assert np.array_equal(node_order(b), np.arange(n))


@pytest.mark.parametrize("which", ["jittered", "ties"])
def test_shortest_matches_oracle_and_original_order(which, request):
    a, b, og = request.getfixturevalue(which)
    dists = [300, 700, 1500]
    say(f"{which}: centrality_shortest {dists}")
    ga = a.centrality_shortest(distances=dists)
    gb = b.centrality_shortest(distances=dists)
    ref, reach = og.centrality_shortest(dists)
    sa, sb = stack(ga, SHORTEST_KEYS, dists), stack(gb, SHORTEST_KEYS, dists)
    assert np.array_equal(sa[0], ref[0]) and np.array_equal(sb[0], ref[0])
    assert_close(sa, ref, f"{which} device order vs oracle")
    assert_close(sa, sb, f"{which} device order vs insertion order")
    assert ga.reachability_totals == gb.reachability_totals


@pytest.mark.parametrize("which", ["jittered", "ties"])
def test_subset_sources_and_weights(which, request):
    a, b, _og = request.getfixturevalue(which)
    n = a.node_count()
    rng = np.random.default_rng(85)
    sub = rng.choice(n, n // 3, replace=False).tolist()
    say(f"{which}: {len(sub)} sources")
    ga = a.centrality_shortest(distances=[900], source_indices=sub)
    gb = b.centrality_shortest(distances=[900], source_indices=sub)
    sa, sb = stack(ga, SHORTEST_KEYS, [900]), stack(gb, SHORTEST_KEYS, [900])
    assert np.array_equal(sa[0], sb[0])
    assert_close(sa, sb, f"{which} subset")


@pytest.mark.parametrize("which", ["jittered", "ties"])
def test_segment_matches_oracle(which, request):
    a, b, og = request.getfixturevalue(which)
    dists = [400, 900]
    say(f"{which}: segment {dists}")
    sa = stack(a.segment_centrality(distances=dists), SEGMENT_KEYS, dists)
    sb = stack(b.segment_centrality(distances=dists), SEGMENT_KEYS, dists)
    ref = og.segment_centrality(dists)
    assert_close(sa, ref, f"{which} segment vs oracle")
    assert_close(sa, sb, f"{which} segment device vs insertion order")


@pytest.mark.parametrize("which", ["jittered", "ties"])
def test_trees_and_brandes_state_bit_exact(which, request):
    a, _b, og = request.getfixturevalue(which)
    n = a.node_count()
    lib = _native.load()
    h = a._native_handle()
    p = _native.ptr
    max_seconds = int(round(700 / SPEED))
    rng = np.random.default_rng(86)
    for src in rng.choice(n, 6, replace=False).tolist():
        say(f"{which}: trees from {src}")
        order, visits = a.dijkstra_tree_shortest(src, max_seconds, SPEED)
        ref = og.tree_shortest(src, max_seconds)
        assert order == ref["order"].tolist()
        assert [(-1 if v.pred is None else v.pred) for v in visits] == ref["pred"].tolist()
        assert np.array_equal(np.float32([v.agg_seconds for v in visits]), ref["secs"])
        order, edges, visits, _ev = a.dijkstra_tree_segment(src, max_seconds, SPEED)
        ref = og.tree_segment(src, max_seconds)
        assert order == ref["order"].tolist() and edges == ref["edges"].tolist()
        assert [(-1 if v.origin_seg is None else v.origin_seg) for v in visits] == ref["origin_seg"].tolist()
        assert [(-1 if v.last_seg is None else v.last_seg) for v in visits] == ref["last_seg"].tolist()
        secs, route = np.zeros(n, np.float32), np.zeros(n, np.float32)
        sigma = np.zeros(n, np.float64)
        npred = np.zeros(n, np.uint32)
        preds = np.zeros(n * 16, np.uint32)
        _native.check(lib.cs_brandes_shortest(h, src, max_seconds, SPEED, 1e-4, p(secs, ctypes.c_float),
                                              p(route, ctypes.c_float), p(sigma, ctypes.c_double),
                                              p(npred, ctypes.c_uint32), p(preds, ctypes.c_uint32), 16))
        ref = og.brandes_shortest(src, max_seconds)
        assert np.array_equal(secs, ref["secs"])
        assert np.array_equal(sigma, ref["sigma"])
        assert np.array_equal(npred, ref["npred"])
        got = [sorted(preds[v * 16: v * 16 + min(npred[v], 16)].tolist()) for v in range(n)]
        assert got == ref["preds"]


def test_od_betweenness_matches_oracle(jittered):
    a, b, og = jittered
    n = a.node_count()
    rng = np.random.default_rng(87)
    o, d = rng.integers(0, n, 1500), rng.integers(0, n, 1500)
    w = rng.uniform(0.5, 3.0, 1500).astype(np.float32)
    dists = [500, 1200]
    say("OD betweenness, 1500 pairs")
    od = rustalgos.centrality.OdMatrix(o.tolist(), d.tolist(), w.tolist())
    ga, gb = a.betweenness_od_shortest(od, distances=dists), b.betweenness_od_shortest(od, distances=dists)
    got = np.stack([np.stack([ga.node_betweenness[x] for x in dists]),
                    np.stack([ga.node_betweenness_beta[x] for x in dists])])
    other = np.stack([np.stack([gb.node_betweenness[x] for x in dists]),
                      np.stack([gb.node_betweenness_beta[x] for x in dists])])
    ref = og.betweenness_od_shortest((o.tolist(), d.tolist(), w.tolist()), dists)
    assert ref.max() > 0
    assert_close(got, ref, "OD device order vs oracle")
    assert_close(got, other, "OD device vs insertion order")


class _Hip:
    """Device buffers through the HIP runtime the library itself links (torch bundles a
    runtime of its own, which cannot initialise once this one has)."""

    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.held = []

    def put(self, arr: np.ndarray) -> int:
        arr = np.ascontiguousarray(arr)
        ptr = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(ptr), max(arr.nbytes, 8)) == 0
        self.held.append(ptr)
        assert self.rt.hipMemcpy(ptr, arr.ctypes.data, arr.nbytes, 1) == 0  # host to device
        return ptr.value

    def get(self, ptr: int, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.rt.hipDeviceSynchronize() == 0
        assert self.rt.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for ptr in self.held:
            self.rt.hipFree(ptr)
        self.held = []


def test_device_entry_point_takes_original_order(jittered):
    """cs_centrality_shortest_device: buffers in original order, any source listing."""
    a, _b, _og = jittered
    n = a.node_count()
    dists = [400, 1000]
    d, bt, s = rustalgos.pair_distances_betas_time(SPEED, distances=dists)
    want = stack(a.centrality_shortest(distances=dists), SHORTEST_KEYS, dists)  # all sources: scale 1
    lib = _native.load()
    hip = _Hip()
    rng = np.random.default_rng(88)
    hd, hb, hs = np.asarray(d, np.uint32), np.asarray(bt, np.float32), np.asarray(s, np.uint32)
    try:
        for listing in ("original", "device", "random"):
            src = {"original": np.arange(n, dtype=np.uint32), "device": node_order(a),
                   "random": rng.permutation(n).astype(np.uint32)}[listing]
            say(f"device entry point, sources listed in {listing} order")
            acc0 = np.zeros((7, len(dists), n), np.float64)
            d_src, d_wt = hip.put(src), hip.put(np.ones(n, np.float32))
            d_el, d_acc = hip.put(np.ones(n, np.uint8)), hip.put(acc0)
            p = _native.CsParams()
            p.n_thresholds = len(dists)
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
            _native.check(lib.cs_centrality_shortest_device(a._native_handle(), ctypes.byref(p), d_wt, d_acc, None,
                                                            None))
            got = hip.get(d_acc, acc0)
            assert np.array_equal(got[0], want[0]), f"{listing}: density"
            assert_close(got, want, f"device entry point, {listing} listing")
    finally:
        hip.free()


def test_upload_stages_the_graph_and_freezes_its_order():
    lat = shuffled(mock.perturbed_lattice(14, seed=89), seed=90)
    a, b = pair(lat)
    lib = _native.load()
    say("upload, then compute; coordinates after the upload are refused")
    a.upload(0)
    with pytest.raises(ValueError, match="does not exist"):
        a.upload(99)
    x = np.ascontiguousarray(lat.x)
    y = np.ascontiguousarray(lat.y)
    rc = lib.cs_graph_set_xy(a._native_handle(), _native.ptr(x, ctypes.c_double), _native.ptr(y, ctypes.c_double))
    assert rc == _native.CS_EINVAL
    ga, gb = a.centrality_shortest(distances=[500, 900]), b.centrality_shortest(distances=[500, 900])
    sa, sb = stack(ga, SHORTEST_KEYS, [500, 900]), stack(gb, SHORTEST_KEYS, [500, 900])
    assert np.array_equal(sa[0], sb[0])
    assert_close(sa, sb, "uploaded graph vs insertion order")

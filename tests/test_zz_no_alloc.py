"""GPU: cs_centrality_shortest_device_noalloc, the device-resident shortest-path entry point
that allocates nothing (its working memory is a static device array), against
cs_centrality_shortest_device on the same inputs: two calls in a row, the second on a smaller
graph with fewer sources and thresholds, with the device node order (translated through the
static array) and without it (a dual graph keeps its order)."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from cityseer import _native, rustalgos
from cityseer.tools import mock
from parity import assert_close
from test_gpu_node_order import _Hip

pytestmark = pytest.mark.gpu

SPEED = 1.33333


def say(msg: str) -> None:
    print(msg, flush=True)


def _params(hd, hb, hs, n_src, d_src, d_wt, d_el) -> _native.CsParams:
    p = _native.CsParams()
    p.n_thresholds = len(hd)
    p.distances = _native.ptr(hd, ctypes.c_uint32)
    p.betas = _native.ptr(hb, ctypes.c_float)
    p.seconds = _native.ptr(hs, ctypes.c_uint32)
    p.speed_m_s = SPEED
    p.tolerance = 1e-4
    p.compute_closeness = 1
    p.compute_betweenness = 1
    p.n_sources = n_src
    p.sources = ctypes.cast(d_src, ctypes.POINTER(ctypes.c_uint32))
    p.source_wt = ctypes.cast(d_wt, ctypes.POINTER(ctypes.c_float))
    p.eligible = ctypes.cast(d_el, ctypes.POINTER(ctypes.c_uint8))
    p.betw_scale = 1.0
    p.angular_scaling_unit = 180.0
    p.farness_scaling_offset = 1.0
    return p


def _both(hip, ns, dists, sources):
    """(existing entry point, allocation-free entry point) results on the same device inputs."""
    lib = _native.load()
    n = ns.node_count()
    d, bt, s = rustalgos.pair_distances_betas_time(SPEED, distances=dists)
    hd, hb, hs = np.asarray(d, np.uint32), np.asarray(bt, np.float32), np.asarray(s, np.uint32)
    src = np.asarray(sources, np.uint32)
    wt = np.ones(len(src), np.float32)
    acc0 = np.zeros((7, len(dists), n), np.float64)
    reach0 = np.zeros(len(dists), np.uint32)
    d_src, d_wt, d_el = hip.put(src), hip.put(wt), hip.put(np.ones(n, np.uint8))
    out = []
    for name in ("cs_centrality_shortest_device", "cs_centrality_shortest_device_noalloc"):
        d_acc, d_reach = hip.put(acc0), hip.put(reach0)
        p = _params(hd, hb, hs, len(src), d_src, d_wt, d_el)
        say(f"{name}: {n} nodes, {len(src)} sources, {dists}")
        _native.check(getattr(lib, name)(ns._native_handle(), ctypes.byref(p), d_wt, d_acc, d_reach, None))
        out.append(hip.get(d_acc, acc0))
    return out


def test_noalloc_matches_device_entry_point():
    hip = _Hip()
    try:
        big = mock.network_structure_from_lattice(mock.perturbed_lattice(40, seed=31, diag_frac=0.05))
        small = mock.network_structure_from_lattice(mock.perturbed_lattice(12, seed=32))
        dual = mock.network_structure_from_lattice(mock.lattice_dual(mock.perturbed_lattice(9, seed=33)))
        lib = _native.load()
        say("not resident: refused")
        n = small.node_count()
        d_src, d_wt, d_el = hip.put(np.arange(n, dtype=np.uint32)), hip.put(np.ones(n, np.float32)), hip.put(
            np.ones(n, np.uint8))
        d_acc = hip.put(np.zeros((7, 1, n), np.float64))
        hd, hb, hs = (np.asarray([400], np.uint32), np.asarray([0.01], np.float32), np.asarray([300], np.uint32))
        p = _params(hd, hb, hs, n, d_src, d_wt, d_el)
        rc = lib.cs_centrality_shortest_device_noalloc(small._native_handle(), ctypes.byref(p), d_wt, d_acc, None,
                                                       None)
        assert rc == _native.CS_EINVAL and b"cs_graph_upload" in lib.cs_last_error()
        for ns in (big, small, dual):
            ns.upload(0, SPEED)
        cases = [
            (big, [300, 600, 1000, 2000], np.arange(big.node_count())),
            (small, [250, 700], np.random.default_rng(5).choice(small.node_count(), 40, replace=False)),
            (dual, [300, 800, 1500], np.arange(dual.node_count())),
            (small, [500], np.arange(small.node_count())[::-1]),
        ]
        for ns, dists, src in cases:
            want, got = _both(hip, ns, dists, src)
            assert np.array_equal(got[0], want[0]), "density"
            assert_close(got, want, f"allocation-free vs device entry point on {ns.node_count()} nodes")
    finally:
        hip.free()

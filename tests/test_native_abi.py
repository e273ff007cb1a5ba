"""The C ABI library: loads, exports every declared symbol, host-side conversions (CPU only)."""

from __future__ import annotations

import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from cityseer import _native, rustalgos
from cityseer.tools import graphs, io
import oracle.oracle as O

HEADER = os.path.join(ROOT, "include", "cityseer_amd.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cs_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _native.load()
    names = _declared()
    assert "cs_centrality_shortest" in names and "cs_graph_create" in names
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) <= set(_native._SIGNATURES), set(names) - set(_native._SIGNATURES)
    assert b"gfx950" in lib.cs_version()


def test_threshold_conversion_matches_oracle_bitwise():
    speed = 1.33333
    for dists in ([50, 150, 250], [400, 800], [500, 1000, 2000, 5000], list(range(100, 20001, 100))):
        got = rustalgos.pair_distances_betas_time(speed, distances=dists)
        want = O.thresholds(speed, distances=dists)
        assert got[0] == want[0]
        assert np.array_equal(np.float32(got[1]), np.float32(want[1]))
        assert got[2] == want[2]
    betas = [0.04, 0.02, 0.01, 0.0025, 0.0008]
    assert rustalgos.pair_distances_betas_time(speed, betas=betas)[0] == O.thresholds(speed, betas=betas)[0]
    minutes = [5, 7.5, 10, 20, 25, 30]
    assert rustalgos.pair_distances_betas_time(speed, minutes=minutes) == tuple(O.thresholds(speed, minutes=minutes))


def test_threshold_errors_match_reference(golden):
    for b in ([-0.04], [0], [0.0], []):
        with pytest.raises(ValueError):
            rustalgos.distances_from_betas(b)
    with pytest.raises(ValueError, match="strictly decreasing"):
        rustalgos.distances_from_betas([0.0025, 0.04])
    with pytest.raises(ValueError, match="strictly increasing"):
        rustalgos.betas_from_distances([100, 50])
    for d in ([0], []):
        with pytest.raises(ValueError):
            rustalgos.betas_from_distances(d)
    for bad in ("boo", None, [None]):
        with pytest.raises(TypeError):
            rustalgos.distances_from_betas(bad)
    with pytest.raises(ValueError, match="exactly one"):
        rustalgos.pair_distances_betas_time(1.33333, distances=[100], betas=[0.04])
    t = golden["thresholds"]
    assert rustalgos.distances_from_betas([0.04]) == [100]
    assert np.allclose(rustalgos.betas_from_distances([173], min_threshold_wt=0.001), [0.03992922231554985],
                       atol=0.01, rtol=1e-4)
    ds = t["distances_from_seconds"]
    secs = [round(m * 60) for m in ds["minutes"]]
    assert rustalgos.distances_from_seconds(secs, 1.33333) == ds["distances"]


def test_api_errors_before_any_gpu_work(primal_graph):
    _n, _e, ns = io.network_structure_from_nx(primal_graph)
    with pytest.raises(ValueError, match="dual graph"):
        ns.centrality_simplest(distances=[500], compute_closeness=True, compute_betweenness=False)
    with pytest.raises(ValueError, match="dual graph"):
        ns.dijkstra_tree_simplest(0, 1000, 1.33333)
    with pytest.raises(ValueError, match="both parameters are False"):
        ns.centrality_shortest(distances=[500], compute_closeness=False, compute_betweenness=False)
    with pytest.raises(ValueError, match="Tolerance"):
        ns.centrality_shortest(distances=[500], tolerance=101)
    with pytest.raises(ValueError, match="does not exist"):
        ns.centrality_shortest(distances=[500], source_indices=[999])
    with pytest.raises(ValueError, match="mutually exclusive"):
        ns.centrality_shortest(distances=[500], source_indices=[0], sampling_weights=[1.0] * ns.node_count())
    with pytest.raises(ValueError, match="sample_probability"):
        ns.centrality_shortest(distances=[500], sample_probability=0.0)


def test_graph_create_rejects_bad_input():
    lib = _native.load()
    h = ctypes.c_void_p()
    src = np.array([0, 5], np.uint32)
    dst = np.array([1, 0], np.uint32)
    length = np.array([1.0, 1.0], np.float32)
    p = _native.ptr
    rc = lib.cs_graph_create(2, None, None, None, 2, p(src, ctypes.c_uint32), p(dst, ctypes.c_uint32), None,
                             p(length, ctypes.c_float), None, None, None, None, 0, ctypes.byref(h))
    assert rc == _native.CS_EINVAL
    assert b"does not exist" in lib.cs_last_error()


def _grid_graph(side: int, perm: np.ndarray, is_dual: int = 0):
    """A side x side grid whose node i sits at old grid cell perm[i]; a chain of edges."""
    lib = _native.load()
    n = side * side
    src = np.arange(n - 1, dtype=np.uint32)
    dst = np.arange(1, n, dtype=np.uint32)
    length = np.ones(n - 1, np.float32)
    p = _native.ptr
    h = ctypes.c_void_p()
    pk = np.zeros(n - 1, np.int32)
    _native.check(lib.cs_graph_create(n, None, None, None, n - 1, p(src, ctypes.c_uint32), p(dst, ctypes.c_uint32),
                                      None, p(length, ctypes.c_float), None, None, None,
                                      p(pk, ctypes.c_int32) if is_dual else None, is_dual, ctypes.byref(h)))
    x = (perm % side).astype(np.float64) * 37.0 + 1000.0
    y = (perm // side).astype(np.float64) * 37.0 - 500.0
    return lib, h, x, y


def _order(lib, h, n):
    out = np.empty(n, np.uint32)
    _native.check(lib.cs_graph_node_order(h, _native.ptr(out, ctypes.c_uint32)))
    return out


def test_device_node_order_follows_a_hilbert_curve():
    """cs_graph_set_xy orders a shuffled 8 x 8 grid along a Hilbert curve: a permutation whose
    every step moves to an edge-adjacent grid cell, starting in a corner (the curve's defining
    properties; a wrong rotation or quadrant code breaks adjacency)."""
    side, n = 8, 64
    perm = np.random.default_rng(5).permutation(n)
    lib, h, x, y = _grid_graph(side, perm)
    try:
        assert np.array_equal(_order(lib, h, n), np.arange(n))  # identity until coordinates arrive
        _native.check(lib.cs_graph_set_xy(h, _native.ptr(x, ctypes.c_double), _native.ptr(y, ctypes.c_double)))
        order = _order(lib, h, n)
        assert sorted(order.tolist()) == list(range(n))
        cells = perm[order]
        cx, cy = cells % side, cells // side
        steps = np.abs(np.diff(cx)) + np.abs(np.diff(cy))
        assert np.all(steps == 1), steps
        assert (cx[0], cy[0]) == (0, 0) and cy[-1] == 0 and cx[-1] == side - 1
        assert lib.cs_graph_set_xy(h, None, _native.ptr(y, ctypes.c_double)) == _native.CS_EINVAL
    finally:
        lib.cs_graph_destroy(h)


def test_device_node_order_kept_for_dual_and_missing_coordinates():
    side, n = 4, 16
    perm = np.random.default_rng(6).permutation(n)
    lib, h, x, y = _grid_graph(side, perm, is_dual=1)
    try:
        _native.check(lib.cs_graph_set_xy(h, _native.ptr(x, ctypes.c_double), _native.ptr(y, ctypes.c_double)))
        assert np.array_equal(_order(lib, h, n), np.arange(n))
    finally:
        lib.cs_graph_destroy(h)
    lib, h, x, y = _grid_graph(side, perm)
    try:
        x[3] = np.nan
        _native.check(lib.cs_graph_set_xy(h, _native.ptr(x, ctypes.c_double), _native.ptr(y, ctypes.c_double)))
        assert np.array_equal(_order(lib, h, n), np.arange(n))
    finally:
        lib.cs_graph_destroy(h)


def test_compute_fails_loudly_without_gpu(diamond_graph):
    """No CPU fallback: without a HIP device the compute entry points raise."""
    try:
        import torch

        has_gpu = torch.cuda.is_available()
    except Exception:  # noqa: BLE001
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    _n, _e, ns = io.network_structure_from_nx(diamond_graph)
    with pytest.raises(RuntimeError, match="HIP device"):
        ns.centrality_shortest(distances=[150])
    with pytest.raises(RuntimeError, match="HIP device"):
        ns.segment_centrality(distances=[150])
    with pytest.raises(RuntimeError, match="HIP device"):
        ns.upload()


def test_dual_builder_keys_and_slots(diamond_graph):
    dual = graphs.nx_to_dual(diamond_graph)
    nodes, _edges, ns = io.network_structure_from_nx(dual)
    assert ns.is_dual
    a = ns.arrays()
    assert (a["pk"] >= 0).all()
    # every dual node touches at most two primal endpoints
    ends = {}
    for s, d, k in zip(a["src"], a["dst"], a["pk"], strict=True):
        ends.setdefault(int(s), set()).add(int(k))
        ends.setdefault(int(d), set()).add(int(k))
    assert max(len(v) for v in ends.values()) <= 2

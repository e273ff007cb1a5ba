"""Street-network ingest through the native library (host only, no GPU).

``edge_geometry`` derives the edge payload ``add_street_edge`` computes from a
polyline (rust/src/graph.rs:764-857) for many edges in one call;
``decompose`` splits polylines as ``tools.graphs.nx_decompose`` cuts edges
(pysrc/cityseer/tools/graphs.py:1911-1961); ``halve`` cuts them at half length as
``tools.graphs.nx_to_dual`` does (graphs.py:1965-2149).
"""

from __future__ import annotations

import ctypes
from collections.abc import Sequence

import numpy as np

from . import _native


def _flatten(polylines: Sequence[Sequence[Sequence[float]]], dims: int) -> tuple[np.ndarray, np.ndarray]:
    counts = np.fromiter((len(p) for p in polylines), dtype=np.int64, count=len(polylines))
    off = np.zeros(len(polylines) + 1, dtype=np.uint32)
    if len(polylines):
        off[1:] = np.cumsum(counts)
    pts = np.empty((int(off[-1]), dims), dtype=np.float64)
    i = 0
    for p in polylines:
        for c in p:
            pts[i, 0] = c[0]
            pts[i, 1] = c[1]
            if dims == 3:
                pts[i, 2] = c[2] if len(c) > 2 else np.nan
            i += 1
    return off, pts


def edge_geometry(polylines: Sequence[Sequence[Sequence[float]]]) -> tuple[np.ndarray, np.ndarray]:
    """(length f32[m], angle_sum f32[m]) of each polyline, as ``add_street_edge`` stores them."""
    m = len(polylines)
    off, pts = _flatten(polylines, 2)
    length = np.empty(m, dtype=np.float32)
    angle = np.empty(m, dtype=np.float32)
    p = _native.ptr
    _native.check(_native.load().cs_edge_geometry(m, p(off, ctypes.c_uint32), p(pts, ctypes.c_double),
                                                  p(length, ctypes.c_float), p(angle, ctypes.c_float), None, None))
    return length, angle


def edge_bearings(polylines: Sequence[Sequence[Sequence[float]]]) -> tuple[np.ndarray, np.ndarray]:
    """(in_bearing f32[m], out_bearing f32[m]) of each polyline, in degrees, as ``add_street_edge``
    stores them (graph.rs:764-857)."""
    m = len(polylines)
    off, pts = _flatten(polylines, 2)
    ib = np.empty(m, dtype=np.float32)
    ob = np.empty(m, dtype=np.float32)
    p = _native.ptr
    _native.check(_native.load().cs_edge_geometry(m, p(off, ctypes.c_uint32), p(pts, ctypes.c_double), None, None,
                                                  p(ib, ctypes.c_float), p(ob, ctypes.c_float)))
    return ib, ob


def decompose(polylines: Sequence[Sequence[Sequence[float]]], decompose_max: float) -> list[list[np.ndarray]]:
    """Each polyline cut into ceil(length / decompose_max) equal pieces: per polyline, the list of
    piece point arrays ([k][2], or [k][3] when every input point carries z)."""
    return _cut(polylines, float(decompose_max))


def halve(polylines: Sequence[Sequence[Sequence[float]]]) -> list[list[np.ndarray]]:
    """The two halves of each polyline, cut at half its length (same rules as ``decompose``)."""
    return _cut(polylines, None)


def _cut(polylines, decompose_max: float | None) -> list[list[np.ndarray]]:
    """Polylines whose every point has z are cut in 3-D (z interpolated), the rest in 2-D."""
    has_z = [len(pl) > 0 and all(len(c) > 2 for c in pl) for pl in polylines]
    if any(has_z) and not all(has_z):
        out: list[list[np.ndarray]] = [[] for _ in polylines]
        for flag in (True, False):
            idx = [i for i, z in enumerate(has_z) if z == flag]
            for i, pieces in zip(idx, _cut_dims([polylines[i] for i in idx], decompose_max, 3 if flag else 2),
                                 strict=True):
                out[i] = pieces
        return out
    return _cut_dims(polylines, decompose_max, 3 if polylines and all(has_z) else 2)


def _cut_dims(polylines, decompose_max: float | None, dims: int) -> list[list[np.ndarray]]:
    off, pts = _flatten(polylines, dims)
    lib = _native.load()
    h = ctypes.c_void_p()
    n_pieces, n_points = ctypes.c_uint32(), ctypes.c_uint32()
    p = _native.ptr
    if decompose_max is None:
        rc = lib.cs_halve(len(polylines), p(off, ctypes.c_uint32), p(pts, ctypes.c_double), dims, ctypes.byref(h),
                          ctypes.byref(n_pieces), ctypes.byref(n_points))
    else:
        rc = lib.cs_decompose(len(polylines), p(off, ctypes.c_uint32), p(pts, ctypes.c_double), dims,
                              decompose_max, ctypes.byref(h), ctypes.byref(n_pieces), ctypes.byref(n_points))
    _native.check(rc)
    try:
        cuts = np.empty(len(polylines), dtype=np.uint32)
        poff = np.empty(n_pieces.value + 1, dtype=np.uint32)
        out_pts = np.empty((n_points.value, dims), dtype=np.float64)
        _native.check(lib.cs_pieces_copy(h, p(cuts, ctypes.c_uint32), p(poff, ctypes.c_uint32),
                                         p(out_pts, ctypes.c_double)))
    finally:
        lib.cs_pieces_free(h)
    result, k = [], 0
    for c in cuts.tolist():
        result.append([out_pts[poff[k + i]: poff[k + i + 1]] for i in range(c)])
        k += c
    return result

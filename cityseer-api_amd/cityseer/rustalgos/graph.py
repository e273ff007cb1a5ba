"""``NetworkStructure``: the graph the centrality kernels run on.

Mirrors ``cityseer.rustalgos.graph.NetworkStructure`` (rust/src/graph.rs:381-889,
rust/src/centrality.rs:1497-2407) for the network-centrality path. Nodes and
edges are stored as columns on the host; the first compute call hands them to
the native library (``cs_graph_create``), which builds CSR adjacency in the
reference's edge order and keeps it device-resident until the graph changes.

Edges are directed, as in the reference: ``io.network_structure_from_nx``
inserts each undirected networkX edge once per direction.
"""

from __future__ import annotations

import ctypes
import math
from collections.abc import Sequence
from typing import Any

import numpy as np

from .. import _ingest, _native
from ..tools import util
from . import centrality as _res

TIE_EPSILON = 1e-4
TOLERANCE_WARN_PCT = 2.0
WALKING_SPEED = 1.33333


class NodeVisit:
    """Per-node visit state of a single-source tree (graph.rs NodeVisit)."""

    __slots__ = ("visited", "discovered", "pred", "short_dist", "simpl_dist", "origin_seg", "last_seg", "agg_seconds")

    def __init__(self):
        self.visited = False
        self.discovered = False
        self.pred: int | None = None
        self.short_dist = math.inf
        self.simpl_dist = math.inf
        self.origin_seg: int | None = None
        self.last_seg: int | None = None
        self.agg_seconds = math.inf

    @classmethod
    def new(cls) -> NodeVisit:
        return cls()

    def __repr__(self) -> str:  # pragma: no cover - debugging aid
        return (
            f"NodeVisit(visited={self.visited}, pred={self.pred}, short_dist={self.short_dist}, "
            f"simpl_dist={self.simpl_dist}, agg_seconds={self.agg_seconds})"
        )


class EdgeVisit:
    """Per-edge visit state of a segment tree (graph.rs EdgeVisit)."""

    __slots__ = ("visited", "start_nd_idx", "end_nd_idx", "edge_idx")

    def __init__(self):
        self.visited = False
        self.start_nd_idx: int | None = None
        self.end_nd_idx: int | None = None
        self.edge_idx: int | None = None

    @classmethod
    def new(cls) -> EdgeVisit:
        return cls()


def _wkt_num(v: float) -> str:
    """A coordinate as WKT writers print it: whole numbers without a fraction, others in the
    shortest form that reads back exactly."""
    r = repr(float(v))
    return r[:-2] if r.endswith(".0") else r


class EdgePayload:
    """Read-only edge payload view (graph.rs:88-200): what ``get_edge_payload_py`` returns."""

    def __init__(self, start_nd_key_py, end_nd_key_py, edge_idx, length, angle_sum, imp_factor, in_bearing,
                 out_bearing, seconds, geom_wkt, shared_primal_node_key):
        self.start_nd_key_py = start_nd_key_py
        self.end_nd_key_py = end_nd_key_py
        self.edge_idx = edge_idx
        self.length = length
        self.angle_sum = angle_sum
        self.imp_factor = imp_factor
        self.in_bearing = in_bearing
        self.out_bearing = out_bearing
        self.seconds = seconds
        self.geom_wkt = geom_wkt
        self.shared_primal_node_key = shared_primal_node_key
        self.is_transport = False


class NodePayload:
    """Read-only node payload view (graph.rs:24-86)."""

    def __init__(self, node_key, x, y, z, live, weight):
        self.node_key = node_key
        self._x = x
        self._y = y
        self.z = z
        self.live = live
        self.weight = weight
        self.is_transport = False

    @property
    def coord(self) -> tuple[float, float]:
        return (self._x, self._y)

    @property
    def coord_z(self) -> tuple[float, float, float | None]:
        return (self._x, self._y, self.z)

    def validate(self) -> None:
        if not math.isfinite(self.weight) or self.weight < 0.0:
            raise ValueError(
                f"Invalid street node payload: weight must be finite and non-negative (>= 0.0). "
                f"Found {self.weight}. Node key: {self.node_key!r}"
            )


class _Column:
    """Append-only column: scalar appends buffered in a list, bulk appends as arrays."""

    def __init__(self, dtype):
        self.dtype = dtype
        self.chunks: list[np.ndarray] = []
        self.pending: list = []
        self._cache: np.ndarray | None = None

    def append(self, v) -> None:
        self.pending.append(v)
        self._cache = None

    def extend(self, arr: np.ndarray) -> None:
        self._flush()
        self.chunks.append(np.asarray(arr, dtype=self.dtype))
        self._cache = None

    def _flush(self) -> None:
        if self.pending:
            self.chunks.append(np.asarray(self.pending, dtype=self.dtype))
            self.pending = []

    def array(self) -> np.ndarray:
        if self._cache is None:
            self._flush()
            self._cache = (
                np.concatenate(self.chunks) if len(self.chunks) != 1 else self.chunks[0]
            ) if self.chunks else np.zeros(0, dtype=self.dtype)
            self.chunks = [self._cache] if len(self._cache) else []
        return self._cache

    def __len__(self) -> int:
        return sum(len(c) for c in self.chunks) + len(self.pending)

    def set(self, i: int, v) -> None:
        arr = self.array()
        arr[i] = v


def validate_tolerance(tolerance: float | None) -> float:
    """User percent -> decimal, clamped to >= TIE_EPSILON (centrality.rs:34-48)."""
    pct = np.float32(0.0 if tolerance is None else tolerance)
    if pct < 0.0 or pct > 100.0:
        raise ValueError(f"Tolerance must be between 0 and 100 (percent), got {pct}")
    if pct > TOLERANCE_WARN_PCT:
        import logging

        logging.getLogger(__name__).warning(
            f"Tolerance {pct:.1f}% is high — values above {TOLERANCE_WARN_PCT}% increasingly "
            "diffuse route concentration, especially at larger distance thresholds."
        )
    return float(max(np.float32(pct / np.float32(100.0)), np.float32(TIE_EPSILON)))


class NetworkStructure:
    """Street network for centrality analysis (graph.rs:381-405)."""

    def __init__(self):
        self._is_dual = False
        self._progress = 0
        self._keys: list[Any] = []
        self._x = _Column(np.float64)
        self._y = _Column(np.float64)
        self._z = _Column(np.float64)  # NaN = no elevation
        self._live = _Column(np.bool_)
        self._weight = _Column(np.float32)
        self._esrc = _Column(np.uint32)
        self._edst = _Column(np.uint32)
        self._ekey = _Column(np.uint32)
        self._elen = _Column(np.float32)
        self._eang = _Column(np.float32)
        self._eimp = _Column(np.float32)
        self._esec = _Column(np.float32)
        self._epk = _Column(np.int32)
        # edge polylines (for the data assignment): edge e's coordinates are
        # _gxy[2 * _goff[e] : 2 * (_goff[e] + _gn[e])]; _gn 0 = none given (a
        # straight line between the end nodes stands in)
        self._goff = _Column(np.uint64)
        self._gn = _Column(np.uint32)
        self._gxy = _Column(np.float64)
        self._barriers: list | None = None
        self._pk_ids: dict[str, int] = {}
        self._pk_names: list[str] = []
        self._handle: ctypes.c_void_p | None = None
        # device memory lays nodes out along a Hilbert curve of their
        # coordinates (results are unaffected); False keeps insertion order
        self._spatial_order = True

    @classmethod
    def new(cls) -> NetworkStructure:
        return cls()

    def __del__(self):
        try:
            self._release()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    # -- progress (graph.rs:407-415) ---------------------------------------
    def progress_init(self) -> None:
        self._progress = 0

    def progress(self) -> int:
        if self._handle is not None:
            return int(_native.load().cs_graph_progress(self._handle))
        return self._progress

    def upload(self, device: int = 0, speed_m_s: float | None = None) -> None:
        """Stage the graph on HIP device ``device`` now (adjacency, node order and, for
        ``speed_m_s``, edge travel seconds; default the walking speed) instead of on
        its first computation there. Not in the reference, which has no device."""
        from ..config import SPEED_M_S

        speed = SPEED_M_S if speed_m_s is None else float(speed_m_s)
        _native.check(_native.load().cs_graph_upload(self._native_handle(), int(device), speed))

    @property
    def is_dual(self) -> bool:
        return self._is_dual

    def set_is_dual(self, is_dual: bool) -> None:
        self._is_dual = bool(is_dual)
        self._release()

    # -- nodes -------------------------------------------------------------
    def add_street_node(self, node_key: Any, x: float, y: float, live: bool, weight: float, z: float | None = None) -> int:
        """Add a node; returns its index (graph.rs:426-452)."""
        w = float(np.float32(weight))
        if not math.isfinite(w) or w < 0.0:
            raise ValueError(
                "Invalid street node payload: weight must be finite and non-negative (>= 0.0). "
                f"Found {weight}. Node key: {node_key!r}"
            )
        idx = len(self._keys)
        self._keys.append(node_key)
        self._x.append(float(x))
        self._y.append(float(y))
        self._z.append(math.nan if z is None else float(z))
        self._live.append(bool(live))
        self._weight.append(w)
        self._release()
        return idx

    def add_street_nodes_bulk(
        self,
        node_keys: Sequence[Any] | None,
        x: np.ndarray,
        y: np.ndarray,
        live: np.ndarray | None = None,
        weight: np.ndarray | None = None,
        z: np.ndarray | None = None,
    ) -> np.ndarray:
        """Array form of ``add_street_node`` (keys default to the index as a string)."""
        n = len(x)
        start = len(self._keys)
        if node_keys is None:
            self._keys.extend(_LazyKeys(start, n))
        else:
            if len(node_keys) != n:
                raise ValueError("node_keys and x must have equal length")
            self._keys.extend(node_keys)
        w = np.ones(n, dtype=np.float32) if weight is None else np.asarray(weight, dtype=np.float32)
        if not np.all(np.isfinite(w)) or np.any(w < 0):
            raise ValueError("Invalid street node payload: weight must be finite and non-negative (>= 0.0).")
        self._x.extend(np.asarray(x, dtype=np.float64))
        self._y.extend(np.asarray(y, dtype=np.float64))
        self._z.extend(np.full(n, np.nan) if z is None else np.asarray(z, dtype=np.float64))
        self._live.extend(np.ones(n, dtype=bool) if live is None else np.asarray(live, dtype=bool))
        self._weight.extend(w)
        self._release()
        return np.arange(start, start + n)

    def add_transport_node(self, *args, **kwargs):  # pragma: no cover - out of scope
        raise NotImplementedError("transport nodes belong to the GTFS/data layer, outside the centrality path")

    def _check_node(self, node_idx: int, what: str = "node_idx") -> None:
        if not (0 <= int(node_idx) < len(self._keys)):
            raise ValueError(f"Node index {node_idx} does not exist in the graph.")

    def get_node_payload_py(self, node_idx: int) -> NodePayload:
        self._check_node(node_idx)
        z = self._z.array()[node_idx]
        return NodePayload(
            self._keys[node_idx], float(self._x.array()[node_idx]), float(self._y.array()[node_idx]),
            None if math.isnan(z) else float(z), bool(self._live.array()[node_idx]),
            float(self._weight.array()[node_idx]),
        )

    def get_node_weight(self, node_idx: int) -> float:
        self._check_node(node_idx)
        return float(self._weight.array()[node_idx])

    def is_node_live(self, node_idx: int) -> bool:
        self._check_node(node_idx)
        return bool(self._live.array()[node_idx])

    def set_node_live(self, node_idx: int, live: bool) -> None:
        self._check_node(node_idx)
        self._live.set(node_idx, bool(live))

    def node_count(self) -> int:
        return len(self._keys)

    def node_bound(self) -> int:
        return len(self._keys)

    def edge_bound(self) -> int:
        return len(self._esrc)

    def street_node_count(self) -> int:
        return len(self._keys)

    def node_indices(self) -> list[int]:
        return list(range(len(self._keys)))

    def node_keys_py(self) -> list[Any]:
        return list(self._keys)

    def street_node_indices(self) -> list[int]:
        return list(range(len(self._keys)))

    # -- barriers and edge polylines for the data assignment (graph.rs:1154-1235) --
    def set_barriers(self, barriers_wkt: list[str]) -> None:
        """Barrier geometries (WKT) that assignment lines may not cross; replaces any set
        before. Geometries that fail to parse or have no coordinates are skipped with a
        warning, as the reference skips them."""
        from .._geom import parse_wkt

        geoms = []
        for i, w in enumerate(barriers_wkt):
            try:
                parts = parse_wkt(w)
            except ValueError as err:
                logging.getLogger(__name__).warning(f"Failed to parse WKT barrier (input index {i}): {err}")
                continue
            if not parts:
                logging.getLogger(__name__).warning(f"Skipping barrier geom (input index {i}) with no bounding box")
                continue
            geoms.append(parts)
        self._barriers = geoms or None

    def unset_barriers(self) -> None:
        self._barriers = None

    def _street_polylines(self):
        """(start, end, offsets, xy) of every edge, in edge order; an edge added without a
        polyline runs straight between its end nodes."""
        src, dst = self._esrc.array(), self._edst.array()
        goff, gn, gxy = self._goff.array(), self._gn.array(), self._gxy.array().reshape(-1, 2)
        xs, ys = self._x.array(), self._y.array()
        n = np.where(gn > 0, gn, 2).astype(np.uint64)
        off = np.zeros(len(src) + 1, dtype=np.uint64)
        np.cumsum(n, out=off[1:])
        xy = np.empty((int(off[-1]), 2), dtype=np.float64)
        has = gn > 0
        if has.any():
            cnt = gn[has].astype(np.int64)
            within = np.arange(int(cnt.sum())) - np.repeat(np.cumsum(cnt) - cnt, cnt)
            xy[np.repeat(off[:-1][has].astype(np.int64), cnt) + within] = gxy[
                np.repeat(goff[has].astype(np.int64), cnt) + within]
        if (~has).any():
            a = off[:-1][~has].astype(np.int64)
            xy[a] = np.stack([xs[src[~has]], ys[src[~has]]], axis=1)
            xy[a + 1] = np.stack([xs[dst[~has]], ys[dst[~has]]], axis=1)
        if off[-1] >= 2**32:
            raise ValueError("too many edge coordinates for the data assignment")
        return src, dst, off.astype(np.uint32), xy

    @property
    def node_xs(self) -> list[float]:
        return self._x.array().tolist()

    @property
    def node_ys(self) -> list[float]:
        return self._y.array().tolist()

    @property
    def node_xys(self) -> list[tuple[float, float]]:
        return list(zip(self.node_xs, self.node_ys, strict=True))

    @property
    def node_zs(self) -> list[float | None]:
        return [None if math.isnan(z) else z for z in self._z.array().tolist()]

    @property
    def node_xyzs(self) -> list[tuple[float, float, float | None]]:
        return list(zip(self.node_xs, self.node_ys, self.node_zs, strict=True))

    @property
    def node_lives(self) -> list[bool]:
        return self._live.array().tolist()

    @property
    def street_node_lives(self) -> list[bool]:
        return self.node_lives

    @property
    def edge_count(self) -> int:
        return len(self._esrc)

    # -- edges -------------------------------------------------------------
    def _primal_key_id(self, key: str | None) -> int:
        if key is None:
            return -1
        key = str(key)
        pid = self._pk_ids.get(key)
        if pid is None:
            pid = len(self._pk_names)
            self._pk_ids[key] = pid
            self._pk_names.append(key)
        return pid

    def add_street_edge(
        self,
        start_nd_idx: int,
        end_nd_idx: int,
        edge_idx: int,
        start_nd_key_py: Any,
        end_nd_key_py: Any,
        geom_wkt: str | Sequence[Sequence[float]],
        imp_factor: float | None = None,
        shared_primal_node_key: str | None = None,
    ) -> int:
        """Add a directed street edge; length and angle come from its geometry (graph.rs:725-889)."""
        self._check_node(start_nd_idx, "start_nd_idx")
        self._check_node(end_nd_idx, "end_nd_idx")
        coords = util.parse_linestring_wkt(geom_wkt) if isinstance(geom_wkt, str) else list(geom_wkt)
        if len(coords) < 2:
            raise ValueError(
                f"Street edge geometry (idx {edge_idx}) between nodes {start_nd_idx} and {end_nd_idx} "
                f"must have at least 2 coordinates. Found {len(coords)}."
            )
        length, angle = (v[0] for v in _ingest.edge_geometry([coords]))
        imp = np.float32(1.0 if imp_factor is None else imp_factor)
        if not math.isfinite(float(imp)) or imp <= 0.0:
            raise ValueError(
                f"Invalid impedance factor ({imp}) for edge (idx {edge_idx}) between nodes "
                f"{start_nd_idx} and {end_nd_idx}.\nImpedance must be finite and positive (> 0.0)."
            )
        eid = len(self._esrc)
        self._goff.append(len(self._gxy) // 2)
        self._gn.append(len(coords))
        self._gxy.extend(np.asarray([c[:2] for c in coords], dtype=np.float64).reshape(-1))
        self._esrc.append(int(start_nd_idx))
        self._edst.append(int(end_nd_idx))
        self._ekey.append(int(edge_idx))
        self._elen.append(length)
        self._eang.append(angle)
        self._eimp.append(imp)
        self._esec.append(np.float32(np.nan))
        self._epk.append(self._primal_key_id(shared_primal_node_key))
        self._release()
        return eid

    def add_street_edges_bulk(
        self,
        start_nd_idx: np.ndarray,
        end_nd_idx: np.ndarray,
        edge_idx: np.ndarray | None,
        length: np.ndarray,
        angle_sum: np.ndarray | None = None,
        imp_factor: np.ndarray | None = None,
        shared_primal_node_key: np.ndarray | None = None,
        seconds: np.ndarray | None = None,
    ) -> np.ndarray:
        """Array form of ``add_street_edge`` with lengths (f32) given directly.

        ``shared_primal_node_key`` takes integer ids (``-1`` = none).
        """
        m = len(start_nd_idx)
        n = len(self._keys)
        s = np.asarray(start_nd_idx, dtype=np.int64)
        d = np.asarray(end_nd_idx, dtype=np.int64)
        if m and (s.min() < 0 or d.min() < 0 or s.max() >= n or d.max() >= n):
            raise ValueError("edge references a node index that does not exist in the graph")
        imp = np.ones(m, dtype=np.float32) if imp_factor is None else np.asarray(imp_factor, dtype=np.float32)
        if m and (not np.all(np.isfinite(imp)) or np.any(imp <= 0)):
            raise ValueError("Invalid impedance factor: impedance must be finite and positive (> 0.0).")
        start = len(self._esrc)
        self._goff.extend(np.full(m, len(self._gxy) // 2, dtype=np.uint64))
        self._gn.extend(np.zeros(m, dtype=np.uint32))
        self._esrc.extend(s.astype(np.uint32))
        self._edst.extend(d.astype(np.uint32))
        self._ekey.extend(np.zeros(m, dtype=np.uint32) if edge_idx is None else np.asarray(edge_idx, dtype=np.uint32))
        self._elen.extend(np.asarray(length, dtype=np.float32))
        self._eang.extend(np.zeros(m, dtype=np.float32) if angle_sum is None else np.asarray(angle_sum, dtype=np.float32))
        self._eimp.extend(imp)
        self._esec.extend(np.full(m, np.nan, dtype=np.float32) if seconds is None else np.asarray(seconds, dtype=np.float32))
        if shared_primal_node_key is None:
            self._epk.extend(np.full(m, -1, dtype=np.int32))
        else:
            self._epk.extend(np.asarray(shared_primal_node_key, dtype=np.int32))
        self._release()
        return np.arange(start, start + m)

    def add_street_edges_from_lines(
        self,
        start_nd_idx: np.ndarray,
        end_nd_idx: np.ndarray,
        edge_idx: np.ndarray | None,
        lines: Sequence[Sequence[Sequence[float]]],
        imp_factor: np.ndarray | None = None,
        shared_primal_node_key: Sequence[str | None] | None = None,
    ) -> np.ndarray:
        """Array form of ``add_street_edge``: lengths and angle sums derived from the polylines
        natively in one call (rust/src/graph.rs:764-857). ``shared_primal_node_key`` takes the
        keys themselves (``None`` = none)."""
        if len(lines) != len(start_nd_idx) or len(lines) != len(end_nd_idx):
            raise ValueError("start_nd_idx, end_nd_idx and lines must have the same length")
        length, angle = _ingest.edge_geometry(lines)
        pk = None
        if shared_primal_node_key is not None:
            pk = np.fromiter((self._primal_key_id(k) for k in shared_primal_node_key), dtype=np.int32,
                             count=len(lines))
        eids = self.add_street_edges_bulk(start_nd_idx, end_nd_idx, edge_idx, length, angle, imp_factor, pk)
        # keep the polylines (x, y) for the data assignment
        counts = np.fromiter((len(c) for c in lines), dtype=np.uint32, count=len(lines))
        base = len(self._gxy) // 2
        offs = base + np.concatenate([[0], np.cumsum(counts[:-1], dtype=np.uint64)]) if len(lines) else counts
        self._goff.array()[eids] = offs
        self._gn.array()[eids] = counts
        if len(lines):
            self._gxy.extend(np.concatenate([np.asarray(c, dtype=np.float64)[:, :2].reshape(-1) for c in lines]))
        return eids

    def edge_references(self) -> list[tuple[int, int, int]]:
        return list(zip(self._esrc.array().tolist(), self._edst.array().tolist(), self._ekey.array().tolist(), strict=True))

    def _find_edge(self, start_nd_idx: int, end_nd_idx: int, edge_idx: int) -> int:
        # edges_connecting(start, end).find(edge_idx): newest edge first (graph.rs:1278-1292)
        s, d, k = self._esrc.array(), self._edst.array(), self._ekey.array()
        hits = np.nonzero((s == start_nd_idx) & (d == end_nd_idx) & (k == edge_idx))[0]
        if len(hits) == 0:
            raise ValueError(f"Edge not found: {start_nd_idx} -> {end_nd_idx} (edge_idx {edge_idx})")
        return int(hits[-1])

    def get_edge_payload_py(self, start_nd_idx: int, end_nd_idx: int, edge_idx: int) -> EdgePayload:
        """The edge's payload (graph.rs:1001-1011). Bearings come from its stored polyline (a
        straight line between its nodes where none was given); the WKT is rendered from it."""
        e = self._find_edge(start_nd_idx, end_nd_idx, edge_idx)
        n = int(self._gn.array()[e])
        if n:
            o = int(self._goff.array()[e])
            coords = self._gxy.array()[2 * o: 2 * (o + n)].reshape(-1, 2)
        else:
            a, b = int(self._esrc.array()[e]), int(self._edst.array()[e])
            coords = np.array([[self._x.array()[a], self._y.array()[a]], [self._x.array()[b], self._y.array()[b]]])
        ib, ob = _ingest.edge_bearings([coords.tolist()])
        pk = int(self._epk.array()[e])
        return EdgePayload(
            self._keys[int(self._esrc.array()[e])], self._keys[int(self._edst.array()[e])], int(self._ekey.array()[e]),
            float(self._elen.array()[e]), float(self._eang.array()[e]), float(self._eimp.array()[e]),
            float(ib[0]), float(ob[0]), float(self._esec.array()[e]),
            "LINESTRING (" + ", ".join(f"{_wkt_num(x)} {_wkt_num(y)}" for x, y in coords.tolist()) + ")",
            None if pk < 0 else self._pk_names[pk],
        )

    def get_edge_length(self, start_nd_idx: int, end_nd_idx: int, edge_idx: int) -> float:
        return float(self._elen.array()[self._find_edge(start_nd_idx, end_nd_idx, edge_idx)])

    def get_edge_impedance(self, start_nd_idx: int, end_nd_idx: int, edge_idx: int) -> float:
        return float(self._eimp.array()[self._find_edge(start_nd_idx, end_nd_idx, edge_idx)])

    def validate(self) -> None:
        """Payload checks (graph.rs validate): node weights finite and non-negative, edges in range."""
        w = self._weight.array()
        if len(w) and (not np.all(np.isfinite(w)) or np.any(w < 0)):
            raise ValueError("Invalid street node payload: weight must be finite and non-negative (>= 0.0).")

    def build_edge_rtree(self) -> None:
        """Kept for the reference's API (graph.rs:1064): the data assignment indexes the street
        edges itself, in a grid built per call, so there is nothing to prepare here."""
        return None

    # -- native handle -------------------------------------------------------
    def arrays(self) -> dict[str, np.ndarray]:
        """Contiguous node and edge columns (insertion order)."""
        return {
            "x": np.ascontiguousarray(self._x.array()),
            "y": np.ascontiguousarray(self._y.array()),
            "z": np.ascontiguousarray(self._z.array()),
            "live": np.ascontiguousarray(self._live.array().astype(np.uint8)),
            "weight": np.ascontiguousarray(self._weight.array()),
            "src": np.ascontiguousarray(self._esrc.array()),
            "dst": np.ascontiguousarray(self._edst.array()),
            "key": np.ascontiguousarray(self._ekey.array()),
            "length": np.ascontiguousarray(self._elen.array()),
            "angle": np.ascontiguousarray(self._eang.array()),
            "imp": np.ascontiguousarray(self._eimp.array()),
            "seconds": np.ascontiguousarray(self._esec.array()),
            "pk": np.ascontiguousarray(self._epk.array()),
        }

    def _release(self) -> None:
        if self._handle is not None:
            h, self._handle = self._handle, None
            _native.load().cs_graph_destroy(h)

    def _native_handle(self) -> ctypes.c_void_p:
        if self._handle is None:
            a = self.arrays()
            lib = _native.load()
            h = ctypes.c_void_p()
            p = _native.ptr
            _native.check(
                lib.cs_graph_create(
                    len(self._keys), p(a["z"], ctypes.c_double), p(a["live"], ctypes.c_uint8),
                    p(a["weight"], ctypes.c_float), len(a["src"]), p(a["src"], ctypes.c_uint32),
                    p(a["dst"], ctypes.c_uint32), p(a["key"], ctypes.c_uint32), p(a["length"], ctypes.c_float),
                    p(a["imp"], ctypes.c_float), p(a["angle"], ctypes.c_float), p(a["seconds"], ctypes.c_float),
                    p(a["pk"], ctypes.c_int32), int(self._is_dual), ctypes.byref(h),
                )
            )
            if self._spatial_order and len(self._keys):
                rc = lib.cs_graph_set_xy(h, p(a["x"], ctypes.c_double), p(a["y"], ctypes.c_double))
                if rc != 0:
                    lib.cs_graph_destroy(h)
                    _native.check(rc)
            self._handle = h
        return self._handle

    # -- validation helpers (centrality.rs:895-959,1009-1030) ----------------
    def validate_dual_for_angular(self, context: str) -> None:
        if not self._is_dual:
            raise ValueError(
                f"{context} requires a dual graph for angular analysis. Convert the graph with "
                "cityseer.tools.graphs.nx_to_dual(...) before ingesting it into NetworkStructure."
            )

    def _validate_dijkstra_inputs(self, src_idx: int, speed_m_s: float) -> None:
        if not (0 <= int(src_idx) < len(self._keys)):
            raise ValueError(f"src_idx {src_idx} out of range for network with node_bound {len(self._keys)}")
        if not math.isfinite(speed_m_s) or speed_m_s <= 0.0:
            raise ValueError(f"speed_m_s must be finite and positive, got {speed_m_s}")

    # -- single-source trees (centrality.rs:1497-1611) -----------------------
    def _tree_visits(self, vis, disc, pred, dist, secs, simplest: bool) -> list[NodeVisit]:
        out = []
        for i in range(len(vis)):
            nv = NodeVisit()
            nv.visited = bool(vis[i])
            nv.discovered = bool(disc[i])
            nv.pred = None if pred[i] < 0 else int(pred[i])
            if simplest:
                nv.simpl_dist = float(dist[i])
            else:
                nv.short_dist = float(dist[i])
            nv.agg_seconds = float(secs[i])
            out.append(nv)
        return out

    def dijkstra_tree_shortest(self, src_idx: int, max_seconds: int, speed_m_s: float) -> tuple[list[int], list[NodeVisit]]:
        self._validate_dijkstra_inputs(src_idx, speed_m_s)
        n = len(self._keys)
        vis, disc = np.zeros(n, np.uint8), np.zeros(n, np.uint8)
        pred = np.zeros(n, np.int32)
        dist, secs = np.zeros(n, np.float32), np.zeros(n, np.float32)
        order = np.zeros(max(n, 1), np.uint32)
        no = ctypes.c_uint32()
        p = _native.ptr
        _native.check(
            _native.load().cs_tree_shortest(
                self._native_handle(), int(src_idx), int(max_seconds), float(speed_m_s), p(vis, ctypes.c_uint8),
                p(disc, ctypes.c_uint8), p(pred, ctypes.c_int32), p(dist, ctypes.c_float), p(secs, ctypes.c_float),
                p(order, ctypes.c_uint32), ctypes.byref(no),
            )
        )
        return order[: no.value].tolist(), self._tree_visits(vis, disc, pred, dist, secs, False)

    def dijkstra_tree_simplest(self, src_idx: int, max_seconds: int, speed_m_s: float) -> tuple[list[int], list[NodeVisit]]:
        self.validate_dual_for_angular("dijkstra_tree_simplest")
        self._validate_dijkstra_inputs(src_idx, speed_m_s)
        n = len(self._keys)
        vis, disc = np.zeros(n, np.uint8), np.zeros(n, np.uint8)
        pred = np.zeros(n, np.int32)
        dist, secs = np.zeros(n, np.float32), np.zeros(n, np.float32)
        order = np.zeros(max(n, 1), np.uint32)
        no = ctypes.c_uint32()
        p = _native.ptr
        _native.check(
            _native.load().cs_tree_simplest(
                self._native_handle(), int(src_idx), int(max_seconds), float(speed_m_s), p(vis, ctypes.c_uint8),
                p(disc, ctypes.c_uint8), p(pred, ctypes.c_int32), p(dist, ctypes.c_float), p(secs, ctypes.c_float),
                p(order, ctypes.c_uint32), ctypes.byref(no),
            )
        )
        return order[: no.value].tolist(), self._tree_visits(vis, disc, pred, dist, secs, True)

    def dijkstra_tree_segment(self, src_idx: int, max_seconds: int, speed_m_s: float):
        self._validate_dijkstra_inputs(src_idx, speed_m_s)
        n, m = len(self._keys), len(self._esrc)
        vis, disc = np.zeros(n, np.uint8), np.zeros(n, np.uint8)
        pred, oseg, lseg = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n, np.int32)
        dist, secs = np.zeros(n, np.float32), np.zeros(n, np.float32)
        order = np.zeros(max(n, 1), np.uint32)
        eorder = np.zeros(max(m, 1), np.uint32)
        no, ne = ctypes.c_uint32(), ctypes.c_uint32()
        p = _native.ptr
        _native.check(
            _native.load().cs_tree_segment(
                self._native_handle(), int(src_idx), int(max_seconds), float(speed_m_s), p(vis, ctypes.c_uint8),
                p(disc, ctypes.c_uint8), p(pred, ctypes.c_int32), p(dist, ctypes.c_float), p(secs, ctypes.c_float),
                p(oseg, ctypes.c_int32), p(lseg, ctypes.c_int32), p(order, ctypes.c_uint32), ctypes.byref(no),
                p(eorder, ctypes.c_uint32), ctypes.byref(ne),
            )
        )
        visits = self._tree_visits(vis, disc, pred, dist, secs, False)
        for i, nv in enumerate(visits):
            nv.origin_seg = None if oseg[i] < 0 else int(oseg[i])
            nv.last_seg = None if lseg[i] < 0 else int(lseg[i])
        edges = [EdgeVisit() for _ in range(m)]
        src, dst, key = self._esrc.array(), self._edst.array(), self._ekey.array()
        for e in eorder[: ne.value].tolist():
            ev = edges[e]
            ev.visited = True
            ev.start_nd_idx = int(dst[e])
            ev.end_nd_idx = int(src[e])
            ev.edge_idx = int(key[e])
        return order[: no.value].tolist(), eorder[: ne.value].tolist(), visits, edges

    # -- sampling plan (centrality.rs:1032-1139) ---------------------------
    def _sampling_plan(
        self,
        sample_probability: float | None,
        sampling_weights: Sequence[float] | None,
        random_seed: int | None,
        source_indices: Sequence[int] | None,
    ) -> dict:
        nb = len(self._keys)
        weights = None
        if sampling_weights is not None:
            w = np.asarray(sampling_weights, dtype=np.float32)
            if len(w) not in (nb,):
                raise ValueError(
                    f"sampling_weights length ({len(w)}) must match node_count ({nb}) or node_bound ({nb})"
                )
            bad = np.nonzero((w < 0.0) | (w > 1.0))[0]
            if len(bad):
                raise ValueError(f"sampling_weights[{bad[0]}] = {w[bad[0]]} is out of range [0.0, 1.0]")
            weights = w
        if sample_probability is not None and (sample_probability <= 0.0 or sample_probability > 1.0):
            raise ValueError("sample_probability must be in (0.0, 1.0]")
        if source_indices is not None and weights is not None:
            raise ValueError("source_indices and sampling_weights are mutually exclusive")
        live = self._live.array()
        node_w = self._weight.array()
        if source_indices is not None:
            src = np.asarray(list(source_indices), dtype=np.int64)
            for s in src.tolist():
                if not (0 <= s < nb):
                    raise ValueError(f"node index {s} does not exist in the graph")
        n_live = int(live.sum())
        indexed = source_indices is not None
        sources = src.astype(np.uint32) if indexed else np.arange(nb, dtype=np.uint32)
        eligible = np.zeros(nb, dtype=np.uint8)
        if indexed:
            eligible[sources] = 1
        else:
            eligible[live] = 1
        wt = node_w[sources].astype(np.float32).copy()
        prob = None if sample_probability is None else np.float32(sample_probability)
        sampled = 0
        if indexed:
            if prob is not None:
                wt = (wt / prob).astype(np.float32)
            sampled = int(eligible[sources].sum())
        elif prob is not None:
            # the reference draws with rand's StdRng; this build uses numpy's PCG64 (seeded)
            rng = np.random.default_rng(random_seed)
            randoms = rng.random(nb).astype(np.float32)
            p = np.full(nb, prob, dtype=np.float32)
            if weights is not None:
                p = (p * weights).astype(np.float32)
            p_s = p[sources]
            take = (p_s > 0.0) & (randoms[sources] < p_s) & (eligible[sources] == 1)
            wt = np.where(take, (wt / np.where(p_s > 0, p_s, 1)).astype(np.float32), np.float32(np.nan))
            sampled = int(take.sum())
        track = prob is not None or indexed
        if indexed and prob is None:
            scale = n_live / len(sources) if len(sources) else 1.0
        else:
            scale = 1.0
        return {
            "sources": np.ascontiguousarray(sources, dtype=np.uint32),
            "wt": np.ascontiguousarray(wt, dtype=np.float32),
            "eligible": eligible,
            "track": track,
            "sampled": sampled,
            "scale": float(scale),
        }

    def _params(self, distances, betas, seconds, speed, tolerance, closeness, betweenness, plan, unit=180.0, offset=1.0):
        d = np.asarray(distances, dtype=np.uint32)
        b = np.asarray(betas, dtype=np.float32)
        s = np.asarray(seconds, dtype=np.uint32)
        keep = [d, b, s, plan]
        p = _native.CsParams()
        p.n_thresholds = len(d)
        p.distances = _native.ptr(d, ctypes.c_uint32)
        p.betas = _native.ptr(b, ctypes.c_float)
        p.seconds = _native.ptr(s, ctypes.c_uint32)
        p.speed_m_s = float(speed)
        p.tolerance = float(tolerance)
        p.compute_closeness = int(bool(closeness))
        p.compute_betweenness = int(bool(betweenness))
        p.n_sources = len(plan["sources"])
        p.sources = _native.ptr(plan["sources"], ctypes.c_uint32)
        p.source_wt = _native.ptr(plan["wt"], ctypes.c_float)
        p.eligible = _native.ptr(plan["eligible"], ctypes.c_uint8)
        p.track_reach = int(plan["track"])
        p.betw_scale = float(plan["scale"])
        p.angular_scaling_unit = float(unit)
        p.farness_scaling_offset = float(offset)
        return p, keep

    # -- centrality (centrality.rs:1624-2407) --------------------------------
    def centrality_shortest(
        self,
        distances: list[int] | None = None,
        betas: list[float] | None = None,
        minutes: list[float] | None = None,
        compute_closeness: bool | None = None,
        compute_betweenness: bool | None = None,
        min_threshold_wt: float | None = None,
        speed_m_s: float | None = None,
        tolerance: float | None = None,
        sample_probability: float | None = None,
        sampling_weights: list[float] | None = None,
        random_seed: int | None = None,
        source_indices: list[int] | None = None,
        pbar_disabled: bool | None = None,
    ) -> _res.CentralityShortestResult:
        """Closeness and betweenness from one Brandes search per source (centrality.rs:1624-1874)."""
        from . import pair_distances_betas_time

        closeness = True if compute_closeness is None else bool(compute_closeness)
        betweenness = True if compute_betweenness is None else bool(compute_betweenness)
        if not closeness and not betweenness:
            raise ValueError(
                "Either or both closeness and betweenness flags is required, but both parameters are False."
            )
        speed = WALKING_SPEED if speed_m_s is None else float(speed_m_s)
        tol = validate_tolerance(tolerance)
        dists, bets, secs = pair_distances_betas_time(speed, distances, betas, minutes, min_threshold_wt)
        plan = self._sampling_plan(sample_probability, sampling_weights, random_seed, source_indices)
        n, t = len(self._keys), len(dists)
        out = np.zeros((7, t, n), dtype=np.float64)
        reach = np.zeros(max(t, 1), dtype=np.uint32)
        params, _keep = self._params(dists, bets, secs, speed, tol, closeness, betweenness, plan)
        self.progress_init()
        _native.check(
            _native.load().cs_centrality_shortest(
                self._native_handle(), ctypes.byref(params), _native.ptr(out, ctypes.c_double),
                _native.ptr(reach, ctypes.c_uint32),
            )
        )
        return _res.CentralityShortestResult(
            dists, self.node_keys_py(), self.node_indices(), out,
            reach[:t].tolist() if plan["track"] else [], plan["sampled"] if plan["track"] else 0,
        )

    def betweenness_od_shortest(
        self,
        od_matrix: _res.OdMatrix,
        distances: list[int] | None = None,
        betas: list[float] | None = None,
        minutes: list[float] | None = None,
        min_threshold_wt: float | None = None,
        speed_m_s: float | None = None,
        tolerance: float | None = None,
        pbar_disabled: bool | None = None,
    ) -> _res.BetweennessShortestResult:
        """OD-weighted betweenness from a Brandes search per origin with outbound trips, each
        destination seeding its trip weight (centrality.rs:2413-2540)."""
        from . import pair_distances_betas_time

        speed = WALKING_SPEED if speed_m_s is None else float(speed_m_s)
        tol = validate_tolerance(tolerance)
        dists, bets, secs = pair_distances_betas_time(speed, distances, betas, minutes, min_threshold_wt)
        n, t = len(self._keys), len(dists)
        out = np.zeros((2, t, n), dtype=np.float64)
        plan = {"sources": np.zeros(0, np.uint32), "wt": np.zeros(0, np.float32), "eligible": np.zeros(n, np.uint8),
                "track": False, "scale": 1.0}
        params, _keep = self._params(dists, bets, secs, speed, tol, False, True, plan)
        o, d, w = od_matrix.origins, od_matrix.destinations, od_matrix.weights
        self.progress_init()
        _native.check(
            _native.load().cs_betweenness_od_shortest(
                self._native_handle(), ctypes.byref(params), len(o), _native.ptr(o, ctypes.c_uint32),
                _native.ptr(d, ctypes.c_uint32), _native.ptr(w, ctypes.c_float), _native.ptr(out, ctypes.c_double),
            )
        )
        return _res.BetweennessShortestResult(dists, self.node_keys_py(), self.node_indices(), out)

    def centrality_simplest(
        self,
        distances: list[int] | None = None,
        betas: list[float] | None = None,
        minutes: list[float] | None = None,
        compute_closeness: bool | None = None,
        compute_betweenness: bool | None = None,
        min_threshold_wt: float | None = None,
        speed_m_s: float | None = None,
        tolerance: float | None = None,
        angular_scaling_unit: float | None = None,
        farness_scaling_offset: float | None = None,
        sample_probability: float | None = None,
        sampling_weights: list[float] | None = None,
        random_seed: int | None = None,
        source_indices: list[int] | None = None,
        pbar_disabled: bool | None = None,
    ) -> _res.CentralitySimplestResult:
        """Angular closeness and betweenness on the dual graph (centrality.rs:1876-2132)."""
        from . import pair_distances_betas_time

        self.validate_dual_for_angular("centrality_simplest")
        closeness = True if compute_closeness is None else bool(compute_closeness)
        betweenness = True if compute_betweenness is None else bool(compute_betweenness)
        tol = validate_tolerance(tolerance)
        if not closeness and not betweenness:
            raise ValueError(
                "Either or both closeness and betweenness flags is required, but both parameters are False."
            )
        speed = WALKING_SPEED if speed_m_s is None else float(speed_m_s)
        unit = 180.0 if angular_scaling_unit is None else float(angular_scaling_unit)
        offset = 1.0 if farness_scaling_offset is None else float(farness_scaling_offset)
        dists, bets, secs = pair_distances_betas_time(speed, distances, betas, minutes, min_threshold_wt)
        plan = self._sampling_plan(sample_probability, sampling_weights, random_seed, source_indices)
        n, t = len(self._keys), len(dists)
        out = np.zeros((4, t, n), dtype=np.float64)
        reach = np.zeros(max(t, 1), dtype=np.uint32)
        params, _keep = self._params(dists, bets, secs, speed, tol, closeness, betweenness, plan, unit, offset)
        self.progress_init()
        _native.check(
            _native.load().cs_centrality_simplest(
                self._native_handle(), ctypes.byref(params), _native.ptr(out, ctypes.c_double),
                _native.ptr(reach, ctypes.c_uint32),
            )
        )
        return _res.CentralitySimplestResult(
            dists, self.node_keys_py(), self.node_indices(), out,
            reach[:t].tolist() if plan["track"] else [], plan["sampled"] if plan["track"] else 0,
        )

    def segment_centrality(
        self,
        distances: list[int] | None = None,
        betas: list[float] | None = None,
        minutes: list[float] | None = None,
        compute_closeness: bool | None = None,
        compute_betweenness: bool | None = None,
        min_threshold_wt: float | None = None,
        speed_m_s: float | None = None,
        pbar_disabled: bool | None = None,
    ) -> _res.CentralitySegmentResult:
        """Segment closeness and betweenness (centrality.rs:2134-2407)."""
        from . import pair_distances_betas_time

        speed = WALKING_SPEED if speed_m_s is None else float(speed_m_s)
        dists, bets, secs = pair_distances_betas_time(speed, distances, betas, minutes, min_threshold_wt)
        closeness = True if compute_closeness is None else bool(compute_closeness)
        betweenness = True if compute_betweenness is None else bool(compute_betweenness)
        if not closeness and not betweenness:
            raise ValueError(
                "Either or both closeness and betweenness flags is required, but both parameters are False."
            )
        n, t = len(self._keys), len(dists)
        plan = {
            "sources": np.arange(n, dtype=np.uint32),
            "wt": np.ones(n, dtype=np.float32),
            "eligible": self._live.array().astype(np.uint8),
            "track": False,
            "sampled": 0,
            "scale": 1.0,
        }
        out = np.zeros((4, t, n), dtype=np.float64)
        params, _keep = self._params(dists, bets, secs, speed, TIE_EPSILON, closeness, betweenness, plan)
        self.progress_init()
        _native.check(
            _native.load().cs_segment_centrality(
                self._native_handle(), ctypes.byref(params), _native.ptr(out, ctypes.c_double)
            )
        )
        return _res.CentralitySegmentResult(dists, self.node_keys_py(), self.node_indices(), out)


class _LazyKeys(Sequence):
    """Node keys ``str(i)`` for a bulk range, materialised on access."""

    def __init__(self, start: int, n: int):
        self.start = start
        self.n = n

    def __len__(self) -> int:
        return self.n

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [str(self.start + j) for j in range(*i.indices(self.n))]
        if i < 0:
            i += self.n
        if not 0 <= i < self.n:
            raise IndexError(i)
        return str(self.start + i)

"""Data layer: ``DataMap`` and land-use accessibility (rust/src/data.rs).

``DataMap.accessibility`` (data.rs:617-747) runs on the device through
``cs_accessibility``: one shortest-path search per live node, then, in the
same launch, the data entries assigned to the reached nodes, each dedupe group
counted once through its nearest member (data.rs:520-615).

The assignment of entries to nodes is ``node_data_map``: ``{node index:
[(data key, distance), ...]}``, the map ``assign_data_to_network`` leaves
(data.rs:455-512), computed natively on the host (``cs_assign_data``: nearest
street edges, barrier and street-crossing checks, graph.rs:1346-1535); a caller
may also set it directly. With ``angular`` the searches are simplest-path trees
on a dual graph (data.rs:538-548): the reached nodes of each source come from
a simplest-path reach kernel and the same aggregation runs over them.
``stats`` (data.rs:1013-1233)
runs the same launch with each source's wave summarising the numeric values of
the entries it reaches.
``mixed_uses`` (data.rs:749-944) shares the launch and adds a per-node
diversity kernel over the class counts.
"""

from __future__ import annotations

import ctypes
from typing import Any

import numpy as np

from .. import _native
from .graph import WALKING_SPEED


def clip_wts_curve(distances: list[int], betas: list[float], spatial_tolerance: int) -> list[float]:
    """exp(-beta * spatial_tolerance) per threshold, in f32 (common.rs:311-334)."""
    if len(distances) != len(betas):
        raise ValueError("Input 'distances' and 'betas' must have the same length.")
    out = []
    for d, b in zip(distances, betas, strict=True):
        if spatial_tolerance > d:
            raise ValueError(
                f"Clipping distance ({spatial_tolerance}) cannot be greater than the distance threshold ({d})."
            )
        out.append(float(np.exp(np.float32(-np.float32(b)) * np.float32(spatial_tolerance), dtype=np.float32)))
    return out


def clipped_beta_wt(beta: float, max_curve_wt: float, data_dist: float) -> float:
    """min(exp(-beta * d), max_curve_wt) / max_curve_wt in f32 (common.rs:339-348)."""
    if not 0.0 <= max_curve_wt <= 1.0:
        raise ValueError("Max curve weight must be in the range [0, 1].")
    raw = np.exp(np.float32(-np.float32(beta)) * np.float32(data_dist), dtype=np.float32)
    return float(np.float32(min(raw, np.float32(max_curve_wt)) / np.float32(max_curve_wt)))


class DataEntry:
    """One data point (data.rs:306-377): its key and dedupe key, both as given and as composite
    strings, and its WKT text, parsed here so that a bad geometry fails at insertion with
    ``ValueError`` as in the reference."""

    def __init__(self, data_key_py: Any, geom_wkt: str, dedupe_key_py: Any = None):
        from .._geom import parse_wkt

        self.data_key_py = data_key_py
        self.data_key = _composite_key(data_key_py)
        self.dedupe_key_py = dedupe_key_py if dedupe_key_py is not None else data_key_py
        self.dedupe_key = _composite_key(self.dedupe_key_py)
        self.geom_wkt = geom_wkt
        try:
            self.parts = parse_wkt(geom_wkt)
        except ValueError as e:
            raise ValueError(f"Failed to parse WKT for key '{self.data_key}': {e}") from None


class LanduseAccess:
    """Per-class accessibility rows (data.rs:24-44): getters return ``{distance: f64[N]}``."""

    def __init__(self, distances, node_indices, count, weighted, nearest, max_dist):
        self._distances = list(distances)
        self._idx = np.asarray(node_indices, dtype=np.int64)
        self._count, self._weighted, self._nearest, self._max = count, weighted, nearest, int(max_dist)

    @property
    def unweighted(self) -> dict[int, np.ndarray]:
        return {int(d): self._count[i][self._idx].copy() for i, d in enumerate(self._distances)}

    @property
    def weighted(self) -> dict[int, np.ndarray]:
        return {int(d): self._weighted[i][self._idx].copy() for i, d in enumerate(self._distances)}

    @property
    def distance(self) -> dict[int, np.ndarray]:
        return {self._max: self._nearest[self._idx].copy()}


class MixedUsesResult:
    """Result of ``DataMap.mixed_uses`` (data.rs:101-175): ``hill`` / ``hill_weighted`` as
    ``{q: {distance: f64[N]}}`` for q in 0, 1, 2; ``shannon`` / ``gini`` as ``{distance: f64[N]}``."""

    def __init__(self, distances, node_keys_py, node_indices, rows: np.ndarray):
        self.distances = list(distances)
        self.node_keys_py = list(node_keys_py)
        self.node_indices = list(node_indices)
        self._rows = rows  # [8][T][N]
        self._idx = np.asarray(self.node_indices, dtype=np.int64)

    def _load(self, m: int) -> dict[int, np.ndarray]:
        return {int(d): self._rows[m][i][self._idx].copy() for i, d in enumerate(self.distances)}

    @property
    def hill(self) -> dict[int, dict[int, np.ndarray]]:
        return {q: self._load(q) for q in (0, 1, 2)}

    @property
    def hill_weighted(self) -> dict[int, dict[int, np.ndarray]]:
        return {q: self._load(3 + q) for q in (0, 1, 2)}

    @property
    def shannon(self) -> dict[int, np.ndarray]:
        return self._load(6)

    @property
    def gini(self) -> dict[int, np.ndarray]:
        return self._load(7)


class AccessibilityResult:
    """Result of ``DataMap.accessibility`` (data.rs:46-99)."""

    def __init__(self, distances, node_keys_py, node_indices, lu_map: dict[str, LanduseAccess]):
        self.distances = list(distances)
        self.node_keys_py = list(node_keys_py)
        self.node_indices = list(node_indices)
        self._lu_map = lu_map

    @property
    def result(self) -> dict[str, LanduseAccess]:
        return dict(self._lu_map)


STAT_METRICS = 14
_STAT_NAMES = ("sum", "sum_wt", "mean", "mean_wt", "median", "median_wt", "count", "count_wt", "variance",
               "variance_wt", "mad", "mad_wt", "max", "min")


class Stats:
    """One numeric map's statistics (data.rs:170-248): each getter returns ``{distance: f64[N]}``."""

    def __init__(self, distances, node_indices, rows: np.ndarray):
        self._distances = list(distances)
        self._idx = np.asarray(node_indices, dtype=np.int64)
        self._rows = rows  # [14][T][N]

    def _load(self, m: int) -> dict[int, np.ndarray]:
        return {int(d): self._rows[m][i][self._idx].copy() for i, d in enumerate(self._distances)}


for _m, _name in enumerate(_STAT_NAMES):
    setattr(Stats, _name, property(lambda self, m=_m: self._load(m)))


class StatsResult:
    """Result of ``DataMap.stats`` (data.rs:250-304): ``result`` is one ``Stats`` per numeric map."""

    def __init__(self, distances, node_keys_py, node_indices, stats: list[Stats]):
        self.distances = list(distances)
        self.node_keys_py = list(node_keys_py)
        self.node_indices = list(node_indices)
        self._stats = stats

    @property
    def result(self) -> list[Stats]:
        return list(self._stats)


def _composite_key(key: Any) -> str:
    """Keys are stored as "type name:str(value)" (common.rs:91-96)."""
    return f"{type(key).__name__}:{key}"


class DataMap:
    """Map of data entries (data.rs:379-520)."""

    def __init__(self):
        self.entries: dict[str, DataEntry] = {}
        self.node_data_map: dict[int, list[tuple[str, float]]] = {}
        self._progress = 0
        self._running = None  # network of the call in progress
        self._node_bound: int | None = None  # of the network the current call runs on

    def insert(self, data_key_py: Any, geom_wkt: str, dedupe_key_py: Any = None) -> None:
        entry = DataEntry(data_key_py, geom_wkt, dedupe_key_py)
        self.entries[entry.data_key] = entry

    def entry_keys(self) -> list[str]:
        return list(self.entries.keys())

    def get_entry(self, data_key: str) -> DataEntry | None:
        return self.entries.get(data_key)

    def count(self) -> int:
        return len(self.entries)

    def is_empty(self) -> bool:
        return not self.entries

    def progress(self) -> int:
        """Sources finished by the running call (read live from its network's counter), or by
        the last one."""
        ns = self._running
        return int(ns.progress()) if ns is not None else self._progress

    def progress_init(self) -> None:
        self._progress = 0

    def _run(self, network_structure, call) -> None:
        """Runs the native call with progress() following the network's live counter."""
        self.progress_init()
        self._running = network_structure
        try:
            _native.check(call())
        finally:
            self._running = None
        self._progress = network_structure.node_bound()

    def aggregate_to_src_idx(self, netw_src_idx: int, network_structure, max_walk_seconds: int, speed_m_s: float,
                             angular: bool | None = None) -> dict[str, float]:
        """{data key: network distance + data distance} of the entries one source reaches, each
        dedupe group through its nearest member (data.rs:520-615): the source's shortest-path
        tree (simplest-path with ``angular``) comes from the device, the walk over its visited
        nodes runs here in f32 as the reference's does."""
        if angular:
            network_structure.validate_dual_for_angular("aggregate_to_src_idx")
            order, visits = network_structure.dijkstra_tree_simplest(netw_src_idx, max_walk_seconds, speed_m_s)
        else:
            order, visits = network_structure.dijkstra_tree_shortest(netw_src_idx, max_walk_seconds, speed_m_s)
        speed = np.float32(speed_m_s)
        max_walk_dist = np.float32(np.float32(max_walk_seconds) * speed)
        result: dict[str, float] = {}
        nearest: dict[str, tuple[str, np.float32]] = {}
        for node_idx in order:
            agg = np.float32(visits[node_idx].agg_seconds)
            if agg >= np.float32(max_walk_seconds):
                continue
            for data_key, data_dist in self.node_data_map.get(node_idx, []):
                entry = self.entries.get(data_key)
                if entry is None:
                    continue
                total = np.float32(np.float32(agg * speed) + np.float32(data_dist))
                if not total <= max_walk_dist:
                    continue
                held = nearest.get(entry.dedupe_key)
                if held is None:
                    nearest[entry.dedupe_key] = (data_key, total)
                    result[data_key] = float(total)
                elif total < held[1]:
                    result.pop(held[0], None)
                    nearest[entry.dedupe_key] = (data_key, total)
                    result[data_key] = float(total)
        return result

    def assign_data_to_network(self, network_structure, max_assignment_dist: float, n_nearest_candidates: int):
        """Assign every entry to the end nodes of its nearest street edges whose line of sight
        crosses no barrier and no street (data.rs:455-512, graph.rs:1346-1535), natively on
        the host (``cs_assign_data``). Replaces ``node_data_map``."""
        from .._geom import GeomArrays

        keys = list(self.entries.keys())
        data = GeomArrays([self.entries[k].parts for k in keys])
        bars = None if network_structure._barriers is None else GeomArrays(network_structure._barriers)
        src, dst, off, xy = network_structure._street_polylines()
        nxy = np.ascontiguousarray(np.stack([network_structure._x.array(), network_structure._y.array()], axis=1),
                                   dtype=np.float64)
        p = _native.ptr
        lib = _native.load()
        handle = ctypes.c_void_p()
        n_out = ctypes.c_uint64()
        d_struct = data.struct(_native.CsGeoms)
        b_struct = bars.struct(_native.CsGeoms) if bars is not None else None
        src32 = np.ascontiguousarray(src, dtype=np.uint32)
        dst32 = np.ascontiguousarray(dst, dtype=np.uint32)
        _native.check(lib.cs_assign_data(
            len(nxy), p(nxy, ctypes.c_double), len(src32), p(src32, ctypes.c_uint32), p(dst32, ctypes.c_uint32),
            p(off, ctypes.c_uint32), p(np.ascontiguousarray(xy), ctypes.c_double),
            ctypes.byref(b_struct) if b_struct is not None else None, ctypes.byref(d_struct),
            float(max_assignment_dist), int(n_nearest_candidates), ctypes.byref(handle), ctypes.byref(n_out)))
        try:
            n = int(n_out.value)
            di = np.empty(n, dtype=np.uint32)
            nodes = np.empty(n, dtype=np.uint32)
            dist = np.empty(n, dtype=np.float64)
            if n:
                _native.check(lib.cs_assignment_copy(handle, p(di, ctypes.c_uint32), p(nodes, ctypes.c_uint32),
                                                     p(dist, ctypes.c_double)))
        finally:
            lib.cs_assignment_free(handle)
        self.node_data_map = {}
        for d, v, dd in zip(di.tolist(), nodes.tolist(), dist.tolist(), strict=True):
            self.node_data_map.setdefault(v, []).append((keys[d], dd))

    def stats(
        self,
        network_structure,
        numerical_maps: list[dict],
        distances: list[int] | None = None,
        betas: list[float] | None = None,
        minutes: list[float] | None = None,
        angular: bool | None = None,
        spatial_tolerance: int | None = None,
        min_threshold_wt: float | None = None,
        speed_m_s: float | None = None,
        pbar_disabled: bool | None = None,
    ) -> StatsResult:
        """Numerical summaries of the data each node reaches (data.rs:1013-1233): per map and
        distance, the sum, mean, count, variance, median, median absolute deviation, max and min of
        the values of the entries within the distance (one per dedupe group, through its nearest
        member), and their distance-weighted forms."""
        from . import pair_distances_betas_time

        if angular:
            network_structure.validate_dual_for_angular("stats")
        speed = WALKING_SPEED if speed_m_s is None else float(speed_m_s)
        dists, bets, secs = pair_distances_betas_time(speed, distances, betas, minutes, min_threshold_wt)
        self._node_bound = network_structure.node_bound()
        keys, entry_group, a_node, a_ent, a_dist = self._assignment_arrays()
        index = {k: i for i, k in enumerate(keys)}
        values = np.full((len(numerical_maps), len(keys)), np.nan, dtype=np.float32)
        for mp, num_map in enumerate(numerical_maps):
            if len(num_map) != self.count():
                raise ValueError("The number of numeric data points must match the number of data points")
            for k, v in num_map.items():
                ck = _composite_key(k)
                if ck not in index:
                    raise KeyError(f"Data entries key missing: {ck}")
                values[mp, index[ck]] = np.float32(v)
        mcw = clip_wts_curve(dists, bets, int(spatial_tolerance or 0))
        n, T = network_structure.node_bound(), len(dists)
        out = np.empty((len(numerical_maps), STAT_METRICS, T, n), dtype=np.float64)
        p = _native.ptr
        self._run(network_structure, lambda: _native.load().cs_data_stats(
            network_structure._native_handle(), len(keys), p(entry_group, ctypes.c_uint32), len(a_node),
            p(a_node, ctypes.c_uint32), p(a_ent, ctypes.c_uint32), p(a_dist, ctypes.c_double),
            len(numerical_maps), p(values, ctypes.c_float), T,
            p(np.asarray(dists, dtype=np.uint32), ctypes.c_uint32),
            p(np.asarray(bets, dtype=np.float32), ctypes.c_float),
            p(np.asarray(mcw, dtype=np.float32), ctypes.c_float), int(max(secs)), speed, int(bool(angular)),
            p(out, ctypes.c_double),
        ))
        idx = network_structure.node_indices()
        return StatsResult(dists, network_structure.node_keys_py(), idx, [Stats(dists, idx, out[m]) for m in
                                                                          range(len(numerical_maps))])

    def _assignment_arrays(self):
        keys = list(self.entries.keys())
        index = {k: i for i, k in enumerate(keys)}
        groups: dict[str, int] = {}
        entry_group = np.empty(len(keys), dtype=np.uint32)
        for i, k in enumerate(keys):
            entry_group[i] = groups.setdefault(self.entries[k].dedupe_key, len(groups))
        nodes, ents, dists = [], [], []
        n_nodes = self._node_bound
        for node_idx, pairs in self.node_data_map.items():
            # the reference only looks up nodes a search reaches (data.rs:555-559): others are ignored
            if n_nodes is not None and not 0 <= int(node_idx) < n_nodes:
                continue
            for data_key, dist in pairs:
                e = index.get(data_key)
                if e is None:  # data.rs:563-566: unknown keys are skipped
                    continue
                nodes.append(int(node_idx))
                ents.append(e)
                dists.append(float(dist))
        return (keys, entry_group, np.asarray(nodes, dtype=np.uint32), np.asarray(ents, dtype=np.uint32),
                np.asarray(dists, dtype=np.float64))

    def _landuses(self, landuses_map: dict) -> dict[str, str]:
        if len(landuses_map) != self.count():
            raise ValueError("The number of landuse encodings must match the number of data points")
        lu: dict[str, str] = {}
        for k, v in landuses_map.items():
            ck = _composite_key(k)
            if not isinstance(v, str):  # the reference extracts a String first (data.rs:650-660)
                raise TypeError(f"land-use values must be strings, got {type(v).__name__} for {ck}")
            if ck not in self.entries:
                raise KeyError(f"Data entries key missing: {ck}")
            lu[ck] = v
        return lu

    def mixed_uses(
        self,
        network_structure,
        landuses_map: dict,
        distances: list[int] | None = None,
        betas: list[float] | None = None,
        minutes: list[float] | None = None,
        compute_hill: bool | None = None,
        compute_hill_weighted: bool | None = None,
        compute_shannon: bool | None = None,
        compute_gini: bool | None = None,
        angular: bool | None = None,
        spatial_tolerance: int | None = None,
        min_threshold_wt: float | None = None,
        speed_m_s: float | None = None,
        pbar_disabled: bool | None = None,
    ) -> MixedUsesResult:
        """Land-use diversity of what each node reaches (data.rs:749-944)."""
        from . import pair_distances_betas_time

        if angular:
            network_structure.validate_dual_for_angular("mixed_uses")
        speed = WALKING_SPEED if speed_m_s is None else float(speed_m_s)
        dists, bets, secs = pair_distances_betas_time(speed, distances, betas, minutes, min_threshold_wt)
        lu = self._landuses(landuses_map)
        flags = [True if compute_hill is None else bool(compute_hill),
                 True if compute_hill_weighted is None else bool(compute_hill_weighted),
                 bool(compute_shannon), bool(compute_gini)]
        if not any(flags):
            raise ValueError("One of the compute_<measure> flags must be True, but all are currently False.")
        mcw = clip_wts_curve(dists, bets, int(spatial_tolerance or 0))
        classes = sorted(set(lu.values()))
        cls_index = {k: i for i, k in enumerate(classes)}
        self._node_bound = network_structure.node_bound()
        keys, entry_group, a_node, a_ent, a_dist = self._assignment_arrays()
        entry_class = np.asarray([cls_index[lu[k]] if k in lu else -1 for k in keys], dtype=np.int32)
        n, T = network_structure.node_bound(), len(dists)
        rows = np.zeros((8, T, n), dtype=np.float64)
        p = _native.ptr
        self._run(network_structure, lambda: _native.load().cs_mixed_uses(
            network_structure._native_handle(), len(keys), p(entry_group, ctypes.c_uint32),
            p(entry_class, ctypes.c_int32), len(classes), len(a_node), p(a_node, ctypes.c_uint32),
            p(a_ent, ctypes.c_uint32), p(a_dist, ctypes.c_double), T,
            p(np.asarray(dists, dtype=np.uint32), ctypes.c_uint32),
            p(np.asarray(bets, dtype=np.float32), ctypes.c_float),
            p(np.asarray(mcw, dtype=np.float32), ctypes.c_float), int(max(secs)), speed, int(bool(angular)),
            p(rows, ctypes.c_double),
        ))
        # measures not asked for stay zero (data.rs:866-941 only adds the requested ones)
        for on, ms in zip(flags, ((0, 1, 2), (3, 4, 5), (6,), (7,)), strict=True):
            if not on:
                rows[list(ms)] = 0.0
        return MixedUsesResult(dists, network_structure.node_keys_py(), network_structure.node_indices(), rows)

    def accessibility(
        self,
        network_structure,
        landuses_map: dict,
        accessibility_keys: list[str],
        distances: list[int] | None = None,
        betas: list[float] | None = None,
        minutes: list[float] | None = None,
        angular: bool | None = None,
        spatial_tolerance: int | None = None,
        min_threshold_wt: float | None = None,
        speed_m_s: float | None = None,
        pbar_disabled: bool | None = None,
    ) -> AccessibilityResult:
        """Counts and weights of reachable land uses per node (data.rs:617-747)."""
        from . import pair_distances_betas_time

        if angular:
            network_structure.validate_dual_for_angular("accessibility")
        speed = WALKING_SPEED if speed_m_s is None else float(speed_m_s)
        dists, bets, secs = pair_distances_betas_time(speed, distances, betas, minutes, min_threshold_wt)
        lu = self._landuses(landuses_map)
        mcw = clip_wts_curve(dists, bets, int(spatial_tolerance or 0))
        acc_keys = list(accessibility_keys)
        cls_index = {k: i for i, k in enumerate(acc_keys)}
        self._node_bound = network_structure.node_bound()
        keys, entry_group, a_node, a_ent, a_dist = self._assignment_arrays()
        entry_class = np.asarray([cls_index.get(lu.get(k, ""), -1) for k in keys], dtype=np.int32)
        n = network_structure.node_bound()
        C, T = len(acc_keys), len(dists)
        count = np.zeros((C, T, n), dtype=np.float64)
        weighted = np.zeros((C, T, n), dtype=np.float64)
        nearest = np.full((C, n), np.nan, dtype=np.float64)
        p = _native.ptr
        self._run(network_structure, lambda: _native.load().cs_accessibility(
            network_structure._native_handle(), len(keys), p(entry_group, ctypes.c_uint32),
            p(entry_class, ctypes.c_int32), C, len(a_node), p(a_node, ctypes.c_uint32),
            p(a_ent, ctypes.c_uint32), p(a_dist, ctypes.c_double), T,
            p(np.asarray(dists, dtype=np.uint32), ctypes.c_uint32),
            p(np.asarray(bets, dtype=np.float32), ctypes.c_float),
            p(np.asarray(mcw, dtype=np.float32), ctypes.c_float), int(max(secs)), speed,
            int(bool(angular)), p(count, ctypes.c_double), p(weighted, ctypes.c_double),
            p(nearest, ctypes.c_double),
        ))
        idx = network_structure.node_indices()
        lu_map = {k: LanduseAccess(dists, idx, count[c], weighted[c], nearest[c], max(dists))
                  for k, c in cls_index.items()}
        return AccessibilityResult(dists, network_structure.node_keys_py(), idx, lu_map)

"""ctypes wrapper of the CPU oracle (liboracle.so). TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the parity checker and the CPU baseline; never by the product package.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_u32p = ctypes.POINTER(ctypes.c_uint32)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i32p = ctypes.POINTER(ctypes.c_int32)


class OgParams(ctypes.Structure):
    _fields_ = [
        ("n_thresholds", ctypes.c_uint32),
        ("distances", _u32p),
        ("betas", _f32p),
        ("seconds", _u32p),
        ("speed_m_s", ctypes.c_float),
        ("tolerance", ctypes.c_float),
        ("compute_closeness", ctypes.c_int),
        ("compute_betweenness", ctypes.c_int),
        ("n_sources", ctypes.c_uint32),
        ("sources", _u32p),
        ("source_wt", _f32p),
        ("eligible", _u8p),
        ("track_reach", ctypes.c_int),
        ("betw_scale", ctypes.c_double),
        ("angular_scaling_unit", ctypes.c_float),
        ("farness_scaling_offset", ctypes.c_float),
        ("n_threads", ctypes.c_int),
    ]


_lib = None


def od_rows(origins, destinations, weights):
    """OdMatrix::new (rust/src/centrality.rs:61-82): a map origin -> {destination: weight}, a later
    pair overwriting an earlier one; returned as unique origins and a CSR of their rows."""
    rows: dict[int, dict[int, float]] = {}
    for o, d_, w in zip(origins, destinations, weights):
        rows.setdefault(int(o), {})[int(d_)] = float(w)
    orig = np.array(sorted(rows), dtype=np.uint32)
    off = np.zeros(len(orig) + 1, dtype=np.uint32)
    dest, wts = [], []
    for i, o in enumerate(orig):
        for d_, w in rows[int(o)].items():
            dest.append(d_)
            wts.append(w)
        off[i + 1] = len(dest)
    return orig, off, np.asarray(dest, dtype=np.uint32), np.asarray(wts, dtype=np.float32)


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    lib.og_graph_new.restype = ctypes.c_void_p
    lib.og_graph_new.argtypes = [
        ctypes.c_uint32, _f64p, _u8p, _f32p, ctypes.c_uint32, _u32p, _u32p, _u32p, _f32p, _f32p, _f32p, _f32p,
        _i32p, ctypes.c_int,
    ]
    lib.og_graph_free.argtypes = [ctypes.c_void_p]
    lib.og_set_heap_mode.restype = None
    lib.og_set_heap_mode.argtypes = [ctypes.c_int]
    lib.og_set_mutation.restype = None
    lib.og_set_mutation.argtypes = [ctypes.c_int]
    lib.og_pair_thresholds.restype = ctypes.c_int
    lib.og_pair_thresholds.argtypes = [
        ctypes.c_float, ctypes.c_int, _u32p, _f32p, _f32p, ctypes.c_uint32, ctypes.c_float, _u32p, _f32p, _u32p,
    ]
    for name in ("og_centrality_shortest", "og_centrality_simplest"):
        fn = getattr(lib, name)
        fn.restype = ctypes.c_int
        fn.argtypes = [ctypes.c_void_p, ctypes.POINTER(OgParams), _f64p, _u32p]
    lib.og_betweenness_od_shortest.restype = ctypes.c_int
    lib.og_betweenness_od_shortest.argtypes = [ctypes.c_void_p, ctypes.POINTER(OgParams), _u32p, _u32p,
                                               ctypes.POINTER(ctypes.c_float), _f64p]
    lib.og_accessibility.restype = ctypes.c_int
    lib.og_accessibility.argtypes = [
        ctypes.c_void_p, ctypes.c_uint32, _u32p, _i32p, ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p, _f64p,
        ctypes.c_uint32, _u32p, _f32p, _f32p, ctypes.c_uint32, ctypes.c_float, ctypes.c_int, _u8p, ctypes.c_int,
        _f64p, _f64p, _f64p,
    ]
    lib.og_segment_centrality.restype = ctypes.c_int
    lib.og_segment_centrality.argtypes = [ctypes.c_void_p, ctypes.POINTER(OgParams), _f64p]
    lib.og_brandes_shortest.restype = ctypes.c_int
    lib.og_brandes_shortest.argtypes = [
        ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_float, _f32p, _f32p, _f64p,
        _u32p, _u32p, ctypes.c_uint32, _u32p, _u32p,
    ]
    lib.og_brandes_angular.restype = ctypes.c_int
    lib.og_brandes_angular.argtypes = [
        ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_float, _f32p, _f32p, _f64p,
        _u32p, _u32p, ctypes.c_uint32, _u32p, _u32p, _f32p, _f32p,
    ]
    for name in ("og_tree_shortest", "og_tree_simplest"):
        fn = getattr(lib, name)
        fn.restype = ctypes.c_int
        fn.argtypes = [
            ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, _u8p, _u8p, _i32p, _f32p, _f32p,
            _u32p, _u32p,
        ]
    lib.og_tree_segment.restype = ctypes.c_int
    lib.og_tree_segment.argtypes = [
        ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, _u8p, _u8p, _i32p, _f32p, _f32p, _i32p,
        _i32p, _u32p, _u32p, _u32p, _u32p,
    ]
    _lib = lib
    return lib


def set_heap_mode(rust_order: bool) -> None:
    """Pop equal route costs in Rust BinaryHeap order (True) or keyed (seconds, index) order (False)."""
    load().og_set_heap_mode(int(bool(rust_order)))


MUTATIONS = {
    "beta_source_route": 1,
    "beta_full_seed": 2,
    "ang_far_swap": 4,
    "ang_far_no_offset": 5,
    "ang_full_seed": 6,
    "ang_no_split": 7,
}


class mutation:
    """Context manager planting one named mistake in the oracle (pinning tests only)."""

    def __init__(self, name: str):
        self.code = MUTATIONS[name]

    def __enter__(self):
        load().og_set_mutation(self.code)
        return self

    def __exit__(self, *exc):
        load().og_set_mutation(0)
        return False


def _p(a, t):
    if a is None:
        return ctypes.cast(None, ctypes.POINTER(t))
    return a.ctypes.data_as(ctypes.POINTER(t))


def thresholds(speed: float, distances=None, betas=None, minutes=None, min_threshold_wt=0.01831563888873418):
    """Oracle restatement of pair_distances_betas_time; returns (d, b, s) or raises ValueError."""
    lib = load()
    if distances is not None:
        mode, src = 0, np.asarray(distances, dtype=np.uint32)
    elif betas is not None:
        mode, src = 1, np.asarray(betas, dtype=np.float32)
    else:
        mode, src = 2, np.asarray(minutes, dtype=np.float32)
    n = len(src)
    od = np.zeros(max(n, 1), np.uint32)
    ob = np.zeros(max(n, 1), np.float32)
    os_ = np.zeros(max(n, 1), np.uint32)
    rc = lib.og_pair_thresholds(
        float(speed), mode, _p(src if mode == 0 else None, ctypes.c_uint32), _p(src if mode == 1 else None, ctypes.c_float),
        _p(src if mode == 2 else None, ctypes.c_float), n, float(min_threshold_wt), _p(od, ctypes.c_uint32),
        _p(ob, ctypes.c_float), _p(os_, ctypes.c_uint32),
    )
    if rc != 0:
        raise ValueError("invalid thresholds")
    return od[:n].tolist(), ob[:n].tolist(), os_[:n].tolist()


class OracleGraph:
    """A graph handle of the oracle built from NetworkStructure-style arrays."""

    def __init__(self, n_nodes: int, arrays: dict, is_dual: bool):
        lib = load()
        self.n = int(n_nodes)
        self.m = len(arrays["src"])
        self._keep = {k: np.ascontiguousarray(v) for k, v in arrays.items()}
        a = self._keep
        self.weight = a["weight"]
        self.live = a["live"]
        self.h = lib.og_graph_new(
            self.n, _p(a["z"], ctypes.c_double), _p(a["live"], ctypes.c_uint8), _p(a["weight"], ctypes.c_float), self.m,
            _p(a["src"], ctypes.c_uint32), _p(a["dst"], ctypes.c_uint32), _p(a["key"], ctypes.c_uint32),
            _p(a["length"], ctypes.c_float), _p(a["imp"], ctypes.c_float), _p(a["angle"], ctypes.c_float),
            _p(a["seconds"], ctypes.c_float), _p(a["pk"], ctypes.c_int32), int(is_dual),
        )

    @classmethod
    def from_network_structure(cls, ns) -> OracleGraph:
        return cls(ns.node_count(), ns.arrays(), ns.is_dual)

    def __del__(self):
        try:
            load().og_graph_free(self.h)
        except Exception:  # noqa: BLE001
            pass

    def _params(self, distances, betas, seconds, speed, tolerance, closeness, betweenness, sources, wt, eligible,
                track, scale, unit, offset, n_threads):
        d = np.asarray(distances, dtype=np.uint32)
        b = np.asarray(betas, dtype=np.float32)
        s = np.asarray(seconds, dtype=np.uint32)
        src = np.asarray(np.arange(self.n) if sources is None else sources, dtype=np.uint32)
        w = np.asarray(self.weight[src] if wt is None else wt, dtype=np.float32)
        el = np.asarray(self.live if eligible is None else eligible, dtype=np.uint8)
        keep = [d, b, s, src, w, el]
        p = OgParams()
        p.n_thresholds = len(d)
        p.distances, p.betas, p.seconds = _p(d, ctypes.c_uint32), _p(b, ctypes.c_float), _p(s, ctypes.c_uint32)
        p.speed_m_s = float(speed)
        p.tolerance = float(tolerance)
        p.compute_closeness = int(closeness)
        p.compute_betweenness = int(betweenness)
        p.n_sources = len(src)
        p.sources, p.source_wt, p.eligible = _p(src, ctypes.c_uint32), _p(w, ctypes.c_float), _p(el, ctypes.c_uint8)
        p.track_reach = int(track)
        p.betw_scale = float(scale)
        p.angular_scaling_unit = float(unit)
        p.farness_scaling_offset = float(offset)
        p.n_threads = int(n_threads)
        return p, keep

    def centrality_shortest(self, distances, speed=1.33333, tolerance=1e-4, closeness=True, betweenness=True,
                            sources=None, wt=None, eligible=None, track=False, scale=1.0, n_threads=1,
                            min_threshold_wt=0.01831563888873418):
        d, b, s = thresholds(speed, distances=distances, min_threshold_wt=min_threshold_wt)
        p, _keep = self._params(d, b, s, speed, tolerance, closeness, betweenness, sources, wt, eligible, track,
                                scale, 180.0, 1.0, n_threads)
        out = np.zeros((7, len(d), self.n), dtype=np.float64)
        reach = np.zeros(32, dtype=np.uint32)
        rc = load().og_centrality_shortest(self.h, ctypes.byref(p), _p(out, ctypes.c_double), _p(reach, ctypes.c_uint32))
        if rc != 0:
            raise ValueError(f"oracle error {rc}")
        return out, reach[: len(d)]

    def centrality_simplest(self, distances, speed=1.33333, tolerance=1e-4, closeness=True, betweenness=True,
                            sources=None, wt=None, eligible=None, track=False, scale=1.0, unit=180.0, offset=1.0,
                            n_threads=1):
        d, b, s = thresholds(speed, distances=distances)
        p, _keep = self._params(d, b, s, speed, tolerance, closeness, betweenness, sources, wt, eligible, track,
                                scale, unit, offset, n_threads)
        out = np.zeros((4, len(d), self.n), dtype=np.float64)
        reach = np.zeros(32, dtype=np.uint32)
        rc = load().og_centrality_simplest(self.h, ctypes.byref(p), _p(out, ctypes.c_double), _p(reach, ctypes.c_uint32))
        if rc != 0:
            raise ValueError(f"oracle error {rc}")
        return out, reach[: len(d)]

    def betweenness_od_shortest(self, od, distances, speed=1.33333, tolerance=1e-4, n_threads=1,
                                min_threshold_wt=0.01831563888873418):
        """rust/src/centrality.rs:2413-2540; od = (origins, destinations, weights) as given to OdMatrix."""
        d, b, s = thresholds(speed, distances=distances, min_threshold_wt=min_threshold_wt)
        origins, off, dest, w = od_rows(*od)
        p, _keep = self._params(d, b, s, speed, tolerance, False, True, origins, None, None, False, 1.0, 180.0, 1.0,
                                n_threads)
        out = np.zeros((2, len(d), self.n), dtype=np.float64)
        rc = load().og_betweenness_od_shortest(self.h, ctypes.byref(p), _p(off, ctypes.c_uint32),
                                               _p(dest, ctypes.c_uint32), _p(w, ctypes.c_float),
                                               _p(out, ctypes.c_double))
        if rc != 0:
            raise ValueError(f"oracle error {rc}")
        return out

    def segment_centrality(self, distances, speed=1.33333, closeness=True, betweenness=True, sources=None,
                           n_threads=1):
        d, b, s = thresholds(speed, distances=distances)
        p, _keep = self._params(d, b, s, speed, 1e-4, closeness, betweenness, sources, None, None, False, 1.0, 180.0,
                                1.0, n_threads)
        out = np.zeros((4, len(d), self.n), dtype=np.float64)
        rc = load().og_segment_centrality(self.h, ctypes.byref(p), _p(out, ctypes.c_double))
        if rc != 0:
            raise ValueError(f"oracle error {rc}")
        return out

    def brandes_shortest(self, src, max_seconds, speed=1.33333, tolerance=1e-4, max_preds=16):
        n = self.n
        secs, route = np.zeros(n, np.float32), np.zeros(n, np.float32)
        sigma = np.zeros(n, np.float64)
        npred = np.zeros(n, np.uint32)
        preds = np.zeros(n * max_preds, np.uint32)
        order = np.zeros(n, np.uint32)
        no = np.zeros(1, np.uint32)
        load().og_brandes_shortest(
            self.h, int(src), int(max_seconds), float(speed), float(tolerance), _p(secs, ctypes.c_float),
            _p(route, ctypes.c_float), _p(sigma, ctypes.c_double), _p(npred, ctypes.c_uint32),
            _p(preds, ctypes.c_uint32), max_preds, _p(order, ctypes.c_uint32), _p(no, ctypes.c_uint32),
        )
        plist = [sorted(preds[v * max_preds: v * max_preds + min(npred[v], max_preds)].tolist()) for v in range(n)]
        return {"secs": secs, "route": route, "sigma": sigma, "npred": npred, "preds": plist, "order": order[: no[0]]}

    def brandes_angular(self, src, max_seconds, speed=1.33333, tolerance=1e-4, max_preds=16):
        n = self.n
        s2 = 2 * n
        secs, route = np.zeros(s2, np.float32), np.zeros(s2, np.float32)
        sigma = np.zeros(s2, np.float64)
        npred = np.zeros(s2, np.uint32)
        preds = np.zeros(s2 * max_preds, np.uint32)
        order = np.zeros(s2, np.uint32)
        no = np.zeros(1, np.uint32)
        best_r, best_s = np.zeros(n, np.float32), np.zeros(n, np.float32)
        load().og_brandes_angular(
            self.h, int(src), int(max_seconds), float(speed), float(tolerance), _p(secs, ctypes.c_float),
            _p(route, ctypes.c_float), _p(sigma, ctypes.c_double), _p(npred, ctypes.c_uint32),
            _p(preds, ctypes.c_uint32), max_preds, _p(order, ctypes.c_uint32), _p(no, ctypes.c_uint32),
            _p(best_r, ctypes.c_float), _p(best_s, ctypes.c_float),
        )
        plist = [sorted(preds[v * max_preds: v * max_preds + min(npred[v], max_preds)].tolist()) for v in range(s2)]
        return {"secs": secs, "route": route, "sigma": sigma, "npred": npred, "preds": plist,
                "order": order[: no[0]], "best_route": best_r, "best_secs": best_s}

    def _tree(self, fn, src, max_seconds, speed):
        n = self.n
        vis, disc = np.zeros(n, np.uint8), np.zeros(n, np.uint8)
        pred = np.zeros(n, np.int32)
        dist, secs = np.zeros(n, np.float32), np.zeros(n, np.float32)
        order = np.zeros(max(n, 1), np.uint32)
        no = np.zeros(1, np.uint32)
        rc = fn(self.h, int(src), int(max_seconds), float(speed), _p(vis, ctypes.c_uint8), _p(disc, ctypes.c_uint8),
                _p(pred, ctypes.c_int32), _p(dist, ctypes.c_float), _p(secs, ctypes.c_float), _p(order, ctypes.c_uint32),
                _p(no, ctypes.c_uint32))
        if rc != 0:
            raise ValueError(f"oracle error {rc}")
        return {"visited": vis, "discovered": disc, "pred": pred, "dist": dist, "secs": secs, "order": order[: no[0]]}

    def tree_shortest(self, src, max_seconds, speed=1.33333):
        return self._tree(load().og_tree_shortest, src, max_seconds, speed)

    def tree_simplest(self, src, max_seconds, speed=1.33333):
        return self._tree(load().og_tree_simplest, src, max_seconds, speed)

    def tree_segment(self, src, max_seconds, speed=1.33333):
        n, m = self.n, self.m
        vis, disc = np.zeros(n, np.uint8), np.zeros(n, np.uint8)
        pred, oseg, lseg = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n, np.int32)
        dist, secs = np.zeros(n, np.float32), np.zeros(n, np.float32)
        order = np.zeros(max(n, 1), np.uint32)
        eorder = np.zeros(max(m, 1), np.uint32)
        no, ne = np.zeros(1, np.uint32), np.zeros(1, np.uint32)
        rc = load().og_tree_segment(
            self.h, int(src), int(max_seconds), float(speed), _p(vis, ctypes.c_uint8), _p(disc, ctypes.c_uint8),
            _p(pred, ctypes.c_int32), _p(dist, ctypes.c_float), _p(secs, ctypes.c_float), _p(oseg, ctypes.c_int32),
            _p(lseg, ctypes.c_int32), _p(order, ctypes.c_uint32), _p(no, ctypes.c_uint32), _p(eorder, ctypes.c_uint32),
            _p(ne, ctypes.c_uint32),
        )
        if rc != 0:
            raise ValueError(f"oracle error {rc}")
        return {"visited": vis, "pred": pred, "dist": dist, "secs": secs, "origin_seg": oseg, "last_seg": lseg,
                "order": order[: no[0]], "edges": eorder[: ne[0]]}

"""ctypes binding to ``libcityseer_amd.so`` (the C ABI declared in include/cityseer_amd.h).

The library holds the HIP kernels and their host launchers. There is no CPU
fallback: compute calls raise if the library is missing or no GPU is present.
"""

from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

LIB_NAME = "libcityseer_amd.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

CS_OK = 0
CS_EINVAL = 1
CS_EDEVICE = 2
CS_ENOMEM = 3
CS_EINTERNAL = 4
CS_EBUSY = 5
CS_ETIMEDOUT = 6
CS_ECANCELED = 7
CS_MAX_THRESHOLDS = 8


class RequestCancelled(RuntimeError):
    """A served request was cancelled before it completed."""

_u32p = ctypes.POINTER(ctypes.c_uint32)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i32p = ctypes.POINTER(ctypes.c_int32)


class CsParams(ctypes.Structure):
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
    ]


class CsStats(ctypes.Structure):
    _fields_ = [
        ("kernel_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double),
        ("sources", ctypes.c_uint64),
        ("nodes_reached", ctypes.c_uint64),
        ("edges_scanned", ctypes.c_uint64),
        ("within_sum", ctypes.c_uint64),
        ("waves", ctypes.c_uint32),
        ("retries", ctypes.c_uint32),
    ]


class CsGeoms(ctypes.Structure):
    _fields_ = [
        ("n_geoms", ctypes.c_uint32),
        ("geom_part_off", ctypes.POINTER(ctypes.c_uint32)),
        ("part_kind", ctypes.POINTER(ctypes.c_uint8)),
        ("part_ring_off", ctypes.POINTER(ctypes.c_uint32)),
        ("ring_off", ctypes.POINTER(ctypes.c_uint32)),
        ("xy", ctypes.POINTER(ctypes.c_double)),
    ]


class CsTuning(ctypes.Structure):
    _fields_ = [
        ("block_threads", ctypes.c_uint32),
        ("waves_per_simd", ctypes.c_uint32),
        ("delta_scale", ctypes.c_float),
        ("slot_capacity", ctypes.c_uint32),
        ("xcd_split", ctypes.c_uint32),
        ("simplest_lanes", ctypes.c_uint32),
    ]


class CsServerConfig(ctypes.Structure):
    _fields_ = [
        ("hbm_budget_bytes", ctypes.c_uint64),
        ("max_active", ctypes.c_uint32),
        ("slice_ms", ctypes.c_double),
    ]


class CsRequestInfo(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32),
        ("queued_ms", ctypes.c_double),
        ("latency_ms", ctypes.c_double),
        ("sources", ctypes.c_uint64),
        ("launches", ctypes.c_uint32),
    ]


CS_REQ_SHORTEST = 0
CS_REQ_SIMPLEST = 1

# every symbol the header declares, with (restype, argtypes)
_SIGNATURES = {
    "cs_last_error": (ctypes.c_char_p, []),
    "cs_version": (ctypes.c_char_p, []),
    "cs_pair_distances_betas_time": (
        ctypes.c_int,
        [ctypes.c_float, ctypes.c_int, _u32p, _f32p, _f32p, ctypes.c_uint32, ctypes.c_float, _u32p, _f32p, _u32p],
    ),
    "cs_distances_from_betas": (ctypes.c_int, [_f32p, ctypes.c_uint32, ctypes.c_float, _u32p]),
    "cs_betas_from_distances": (ctypes.c_int, [_u32p, ctypes.c_uint32, ctypes.c_float, _f32p]),
    "cs_distances_from_seconds": (ctypes.c_int, [_u32p, ctypes.c_uint32, ctypes.c_float, _u32p]),
    "cs_seconds_from_distances": (ctypes.c_int, [_u32p, ctypes.c_uint32, ctypes.c_float, _u32p]),
    "cs_graph_create": (
        ctypes.c_int,
        [
            ctypes.c_uint32, _f64p, _u8p, _f32p, ctypes.c_uint32, _u32p, _u32p, _u32p, _f32p, _f32p, _f32p,
            _f32p, _i32p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p),
        ],
    ),
    "cs_graph_destroy": (None, [ctypes.c_void_p]),
    "cs_graph_set_xy": (ctypes.c_int, [ctypes.c_void_p, _f64p, _f64p]),
    "cs_graph_node_order": (ctypes.c_int, [ctypes.c_void_p, _u32p]),
    "cs_graph_upload": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_float]),
    "cs_edge_geometry": (ctypes.c_int, [ctypes.c_uint32, _u32p, _f64p, _f32p, _f32p, _f32p, _f32p]),
    "cs_decompose": (
        ctypes.c_int,
        [ctypes.c_uint32, _u32p, _f64p, ctypes.c_uint32, ctypes.c_double, ctypes.POINTER(ctypes.c_void_p),
         _u32p, _u32p],
    ),
    "cs_halve": (
        ctypes.c_int,
        [ctypes.c_uint32, _u32p, _f64p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p), _u32p, _u32p],
    ),
    "cs_pieces_copy": (ctypes.c_int, [ctypes.c_void_p, _u32p, _u32p, _f64p]),
    "cs_pieces_free": (None, [ctypes.c_void_p]),
    "cs_assign_data": (
        ctypes.c_int,
        [ctypes.c_uint32, ctypes.POINTER(ctypes.c_double), ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32),
         ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_double),
         ctypes.POINTER(CsGeoms), ctypes.POINTER(CsGeoms), ctypes.c_double, ctypes.c_uint32,
         ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_uint64)],
    ),
    "cs_assignment_copy": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
         ctypes.POINTER(ctypes.c_double)],
    ),
    "cs_assignment_free": (None, [ctypes.c_void_p]),
    "cs_graph_info": (ctypes.c_int, [ctypes.c_void_p, _u32p, _u32p, _u32p, _u32p]),
    "cs_graph_progress": (ctypes.c_uint64, [ctypes.c_void_p]),
    "cs_centrality_shortest": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(CsParams), _f64p, _u32p]),
    "cs_betweenness_od_shortest": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.POINTER(CsParams), ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32),
         ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)],
    ),
    "cs_centrality_shortest_device": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.POINTER(CsParams), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    ),
    "cs_centrality_shortest_device_noalloc": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.POINTER(CsParams), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    ),
    "cs_centrality_simplest": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(CsParams), _f64p, _u32p]),
    "cs_segment_centrality": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(CsParams), _f64p]),
    "cs_accessibility": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_uint32, _u32p, _i32p, ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p, _f64p,
         ctypes.c_uint32, _u32p, _f32p, _f32p, ctypes.c_uint32, ctypes.c_float, ctypes.c_int, _f64p, _f64p, _f64p],
    ),
    "cs_mixed_uses": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_uint32, _u32p, _i32p, ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p, _f64p,
         ctypes.c_uint32, _u32p, _f32p, _f32p, ctypes.c_uint32, ctypes.c_float, ctypes.c_int, _f64p],
    ),
    "cs_data_stats": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_uint32, _u32p, ctypes.c_uint32, _u32p, _u32p, _f64p, ctypes.c_uint32, _f32p,
         ctypes.c_uint32, _u32p, _f32p, _f32p, ctypes.c_uint32, ctypes.c_float, ctypes.c_int, _f64p],
    ),
    "cs_tree_shortest": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, _u8p, _u8p, _i32p, _f32p, _f32p,
         _u32p, _u32p],
    ),
    "cs_tree_simplest": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, _u8p, _u8p, _i32p, _f32p, _f32p,
         _u32p, _u32p],
    ),
    "cs_tree_segment": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, _u8p, _u8p, _i32p, _f32p, _f32p,
         _i32p, _i32p, _u32p, _u32p, _u32p, _u32p],
    ),
    "cs_brandes_shortest": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_float, _f32p, _f32p, _f64p,
         _u32p, _u32p, ctypes.c_uint32],
    ),
    "cs_last_stats": (ctypes.c_int, [ctypes.POINTER(CsStats)]),
    "cs_set_tuning": (ctypes.c_int, [ctypes.POINTER(CsTuning)]),
    "cs_get_tuning": (ctypes.c_int, [ctypes.POINTER(CsTuning)]),
    "cs_release_workspaces": (ctypes.c_int, []),
    "cs_batch_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p)]),
    "cs_batch_destroy": (None, [ctypes.c_void_p]),
    "cs_batch_add_shortest": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(CsParams), _f64p, _u32p]),
    "cs_batch_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(CsTuning)]),
    "cs_batch_set_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(CsTuning)]),
    "cs_batch_run": (ctypes.c_int, [ctypes.c_void_p]),
    "cs_batch_results": (ctypes.c_int, [ctypes.c_void_p]),
    "cs_server_create": (ctypes.c_int, [ctypes.POINTER(CsServerConfig), ctypes.POINTER(ctypes.c_void_p)]),
    "cs_server_destroy": (None, [ctypes.c_void_p]),
    "cs_server_submit": (
        ctypes.c_int64, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(CsParams), _f64p, _u32p]
    ),
    "cs_server_poll": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(CsRequestInfo)]),
    "cs_server_wait": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(CsRequestInfo)]),
    "cs_server_wait_for": (
        ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.POINTER(CsRequestInfo)],
    ),
    "cs_server_cancel": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    "cs_server_inject_fault": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "cs_server_release": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
}

_lib = None
_lock = threading.Lock()


def load() -> ctypes.CDLL:
    """Load the native library (raises ``ImportError`` with build instructions if missing)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_NAME} not found at {LIB_PATH}: build the HIP extension first "
                "(python -c 'import __graft_entry__; __graft_entry__.build()')."
            )
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def check(rc: int) -> None:
    """Raise the reference's exception type for a non-zero status."""
    if rc == CS_OK:
        return
    msg = load().cs_last_error().decode("utf-8", "replace")
    if rc == CS_EINVAL:
        raise ValueError(msg)
    if rc == CS_ENOMEM:
        raise MemoryError(msg)
    if rc == CS_ETIMEDOUT:
        raise TimeoutError(msg)
    if rc == CS_ECANCELED:
        raise RequestCancelled(msg)
    raise RuntimeError(msg)


def ptr(arr: np.ndarray | None, ctype):
    """Pointer to a contiguous numpy array (or NULL)."""
    if arr is None:
        return ctypes.cast(None, ctypes.POINTER(ctype))
    return arr.ctypes.data_as(ctypes.POINTER(ctype))


def last_stats() -> dict:
    st = CsStats()
    check(load().cs_last_stats(ctypes.byref(st)))
    return {name: getattr(st, name) for name, _ in CsStats._fields_}

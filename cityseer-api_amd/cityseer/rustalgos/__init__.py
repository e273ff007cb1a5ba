"""Drop-in for ``cityseer.rustalgos`` on the network-centrality path.

The reference's PyO3 module (rust/src/lib.rs:37-164) exposes the threshold
conversions and weight helpers at top level and the ``graph``,
``centrality``, ``data`` and ``diversity`` submodules. Here the conversions run
in the native library's host code (f32 arithmetic and Rust rounding,
rust/src/common.rs:98-270) and the graph algorithms run as HIP kernels behind
the same classes.
"""

from __future__ import annotations

from collections.abc import Sequence

import numpy as np

from .. import _native
from . import centrality, data, diversity, graph
from .data import clip_wts_curve, clipped_beta_wt

__all__ = [
    "centrality",
    "data",
    "diversity",
    "graph",
    "avg_distances_for_betas",
    "check_numerical_data",
    "clip_wts_curve",
    "clipped_beta_wt",
    "pair_distances_betas_time",
    "distances_from_betas",
    "betas_from_distances",
    "distances_from_seconds",
    "seconds_from_distances",
]

MIN_THRESH_WT = 0.01831563888873418


def _as_u32_list(values, name: str) -> np.ndarray:
    if values is None or isinstance(values, (str, bytes)) or not isinstance(values, Sequence | np.ndarray):
        raise TypeError(f"'{name}' must be a list of non-negative integers.")
    out = []
    for v in values:
        if v is None or isinstance(v, (str, bytes, bool)):
            raise TypeError(f"'{name}' must be a list of non-negative integers.")
        if isinstance(v, (float, np.floating)) and not float(v).is_integer():
            raise TypeError(f"'{name}' must be a list of non-negative integers.")
        iv = int(v)
        if iv < 0 or iv > 0xFFFFFFFF:
            raise OverflowError(f"'{name}' values must fit an unsigned 32-bit integer.")
        out.append(iv)
    return np.asarray(out, dtype=np.uint32)


def _as_f32_list(values, name: str) -> np.ndarray:
    if values is None or isinstance(values, (str, bytes)) or not isinstance(values, Sequence | np.ndarray):
        raise TypeError(f"'{name}' must be a list of floats.")
    out = []
    for v in values:
        if v is None or isinstance(v, (str, bytes)):
            raise TypeError(f"'{name}' must be a list of floats.")
        out.append(float(v))
    return np.asarray(out, dtype=np.float32)


def _mtw(min_threshold_wt) -> float:
    return float(MIN_THRESH_WT if min_threshold_wt is None else min_threshold_wt)


def distances_from_betas(betas: list[float], min_threshold_wt: float | None = None) -> list[int]:
    """Distances for betas: round(ln(min_threshold_wt) / -beta) (common.rs:98-131)."""
    b = _as_f32_list(betas, "betas")
    out = np.zeros(len(b), dtype=np.uint32)
    lib = _native.load()
    _native.check(
        lib.cs_distances_from_betas(_native.ptr(b, _native.ctypes.c_float), len(b), _mtw(min_threshold_wt),
                                    _native.ptr(out, _native.ctypes.c_uint32))
    )
    return [int(x) for x in out]


def betas_from_distances(distances: list[int], min_threshold_wt: float | None = None) -> list[float]:
    """Betas for distances: -ln(min_threshold_wt) / d rounded to 5 decimals (common.rs:133-165)."""
    d = _as_u32_list(distances, "distances")
    out = np.zeros(len(d), dtype=np.float32)
    lib = _native.load()
    _native.check(
        lib.cs_betas_from_distances(_native.ptr(d, _native.ctypes.c_uint32), len(d), _mtw(min_threshold_wt),
                                    _native.ptr(out, _native.ctypes.c_float))
    )
    return [float(x) for x in out]


def distances_from_seconds(seconds: list[int], speed_m_s: float) -> list[int]:
    """Distances for walking times (common.rs:167-201)."""
    s = _as_u32_list(seconds, "seconds")
    out = np.zeros(len(s), dtype=np.uint32)
    lib = _native.load()
    _native.check(
        lib.cs_distances_from_seconds(_native.ptr(s, _native.ctypes.c_uint32), len(s), float(speed_m_s),
                                      _native.ptr(out, _native.ctypes.c_uint32))
    )
    return [int(x) for x in out]


def seconds_from_distances(distances: list[int], speed_m_s: float) -> list[int]:
    """Walking times for distances (common.rs:203-237)."""
    d = _as_u32_list(distances, "distances")
    out = np.zeros(len(d), dtype=np.uint32)
    lib = _native.load()
    _native.check(
        lib.cs_seconds_from_distances(_native.ptr(d, _native.ctypes.c_uint32), len(d), float(speed_m_s),
                                      _native.ptr(out, _native.ctypes.c_uint32))
    )
    return [int(x) for x in out]


def pair_distances_betas_time(
    speed_m_s: float,
    distances: list[int] | None = None,
    betas: list[float] | None = None,
    minutes: list[float] | None = None,
    min_threshold_wt: float | None = None,
) -> tuple[list[int], list[float], list[int]]:
    """Resolve exactly one of distances / betas / minutes into all three (common.rs:239-270)."""
    given = [x is not None for x in (distances, betas, minutes)]
    if sum(given) != 1:
        raise ValueError(
            "Please provide exactly one of the following arguments: 'distances', 'betas', or 'minutes'."
        )
    if distances is not None:
        mode, n = 0, len(_as_u32_list(distances, "distances"))
    elif betas is not None:
        mode, n = 1, len(_as_f32_list(betas, "betas"))
    else:
        mode, n = 2, len(_as_f32_list(minutes, "minutes"))
    d_in = _as_u32_list(distances, "distances") if mode == 0 else None
    b_in = _as_f32_list(betas, "betas") if mode == 1 else None
    m_in = _as_f32_list(minutes, "minutes") if mode == 2 else None
    od = np.zeros(max(n, 1), dtype=np.uint32)
    ob = np.zeros(max(n, 1), dtype=np.float32)
    os_ = np.zeros(max(n, 1), dtype=np.uint32)
    lib = _native.load()
    c = _native.ctypes
    _native.check(
        lib.cs_pair_distances_betas_time(
            float(speed_m_s), mode, _native.ptr(d_in, c.c_uint32), _native.ptr(b_in, c.c_float),
            _native.ptr(m_in, c.c_float), n, _mtw(min_threshold_wt), _native.ptr(od, c.c_uint32),
            _native.ptr(ob, c.c_float), _native.ptr(os_, c.c_uint32),
        )
    )
    return [int(x) for x in od[:n]], [float(x) for x in ob[:n]], [int(x) for x in os_[:n]]


def check_numerical_data(data_arr) -> None:
    """Every value of a 2-D f32 array must be finite (common.rs:75-88)."""
    a = np.asarray(data_arr, dtype=np.float32)
    if a.ndim != 2:
        raise TypeError("data_arr must be a 2-D array")
    if not np.all(np.isfinite(a)):
        raise ValueError("The numeric data values must be finite.")


def avg_distances_for_betas(betas: list[float], min_threshold_wt: float | None = None) -> list[float]:
    """The average distance each beta's weight curve gives up to its own threshold distance:
    -ln(auc / d) / beta with auc = (exp(-beta d) - 1) / -beta, in f32 (common.rs:271-307)."""
    b = _as_f32_list(betas, "betas")
    if len(b) == 0:
        raise ValueError("Input 'betas' cannot be empty.")
    dists = distances_from_betas(b.tolist(), min_threshold_wt)
    f32 = np.float32
    out = []
    with np.errstate(all="ignore"):
        for beta, d in zip(b, dists, strict=True):
            beta = f32(beta)
            if abs(beta) < np.finfo(np.float32).eps:
                raise ValueError(f"Beta ({beta}) must not be zero during average distance calculation.")
            df = f32(d)
            auc = f32(f32(np.exp(f32(-beta * df), dtype=np.float32) - f32(1.0)) / f32(-beta))
            wt = f32(auc / df)
            if not wt > 0.0:
                raise ValueError(f"Invalid weight ({wt}) encountered during average distance calculation.")
            out.append(float(f32(f32(-np.log(wt, dtype=np.float32)) / beta)))
    return out

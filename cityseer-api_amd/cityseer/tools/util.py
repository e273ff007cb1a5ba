"""Geometry helpers for graph ingest.

Restates the bearing/angle helpers of ``pysrc/cityseer/tools/util.py:120-193``
and the linestring alignment of ``util.py:264-312``, plus the f64 geometry that
``rust/src/graph.rs:326-379,774`` applies when a street edge is added
(Euclidean length cast to f32, cumulative turn angle in degrees cast to f32).

``LineString`` is a minimal stand-in for the shapely geometry the reference
stores on networkX edges (shapely is not a dependency of this package): it
keeps the coordinate list and offers the few operations ingest needs.
"""

from __future__ import annotations

import math
import re
from collections.abc import Sequence

import numpy as np

Coord = tuple[float, ...]


class LineString:
    """A polyline with ``coords``, ``length``, ``wkt``, ``interpolate`` and ``has_z``."""

    __slots__ = ("coords",)

    def __init__(self, coords: Sequence[Sequence[float]]):
        self.coords: list[Coord] = [tuple(float(c) for c in xy) for xy in coords]
        if len(self.coords) < 2:
            raise ValueError("LineString requires at least 2 coordinates.")

    @property
    def has_z(self) -> bool:
        return len(self.coords[0]) == 3

    @property
    def length(self) -> float:
        """2D Euclidean length in f64 (geo ``Euclidean.length``)."""
        total = 0.0
        for a, b in zip(self.coords[:-1], self.coords[1:], strict=True):
            total += math.hypot(b[0] - a[0], b[1] - a[1])
        return total

    @property
    def is_empty(self) -> bool:
        return False

    @property
    def is_valid(self) -> bool:
        return True

    @property
    def geom_type(self) -> str:
        return "LineString"

    @property
    def wkt(self) -> str:
        if self.has_z:
            body = ", ".join(f"{c[0]!r} {c[1]!r} {c[2]!r}" for c in self.coords)
            return f"LINESTRING Z ({body})"
        body = ", ".join(f"{c[0]!r} {c[1]!r}" for c in self.coords)
        return f"LINESTRING ({body})"

    def interpolate(self, frac: float, normalized: bool = False) -> Point:
        """Point at ``frac`` along the line (fraction of length if ``normalized``)."""
        target = frac * self.length if normalized else frac
        walked = 0.0
        for a, b in zip(self.coords[:-1], self.coords[1:], strict=True):
            seg = math.hypot(b[0] - a[0], b[1] - a[1])
            if walked + seg >= target or b is self.coords[-1]:
                t = 0.0 if seg == 0 else (target - walked) / seg
                t = min(max(t, 0.0), 1.0)
                xyz = [a[i] + (b[i] - a[i]) * t for i in range(len(a))]
                return Point(*xyz)
            walked += seg
        return Point(*self.coords[-1])

    def __repr__(self) -> str:  # pragma: no cover - debugging aid
        return self.wkt


class Point:
    """Minimal point with ``x``, ``y``, optional ``z``."""

    __slots__ = ("x", "y", "z")

    def __init__(self, x: float, y: float, z: float | None = None):
        self.x = float(x)
        self.y = float(y)
        self.z = None if z is None else float(z)

    @property
    def has_z(self) -> bool:
        return self.z is not None


_WKT_RE = re.compile(r"^\s*LINESTRING\s*(Z|M|ZM)?\s*\((.*)\)\s*$", re.IGNORECASE | re.DOTALL)


def parse_linestring_wkt(wkt: str) -> list[Coord]:
    """Parse ``LINESTRING [Z] (x y [z], ...)`` into coordinate tuples (2D or 3D)."""
    m = _WKT_RE.match(wkt)
    if m is None:
        raise ValueError(f"Failed to parse WKT for street edge: {wkt[:200]}")
    coords: list[Coord] = []
    for part in m.group(2).split(","):
        vals = part.split()
        if len(vals) < 2:
            raise ValueError(f"Failed to parse WKT coordinate '{part}'.")
        coords.append(tuple(float(v) for v in vals[:3]))
    return coords


def measure_bearing(xy_1: Sequence[float], xy_2: Sequence[float]) -> float:
    """Bearing in degrees from ``xy_1`` to ``xy_2`` (util.py:120-124; graph.rs:326-332)."""
    if xy_1[0] == xy_2[0] and xy_1[1] == xy_2[1]:
        return 0.0
    return math.degrees(math.atan2(xy_2[1] - xy_1[1], xy_2[0] - xy_1[0]))


def _rem_euclid(x: float, m: float) -> float:
    r = math.fmod(x, m)
    return r + abs(m) if r < 0.0 else r


def measure_coords_angle(coords_1: Sequence[float], coords_2: Sequence[float], coords_3: Sequence[float]) -> float:
    """Turn angle (degrees) across ``coords_2`` (util.py:127-143; graph.rs:336-346)."""
    if (coords_1[0] == coords_2[0] and coords_1[1] == coords_2[1]) or (
        coords_2[0] == coords_3[0] and coords_2[1] == coords_3[1]
    ):
        return 0.0
    a_1 = measure_bearing(coords_2, coords_1)
    a_2 = measure_bearing(coords_3, coords_2)
    return abs(_rem_euclid(a_2 - a_1 + 180.0, 360.0) - 180.0)


def measure_cumulative_angle(linestring_coords: Sequence[Sequence[float]]) -> float:
    """Sum of turn angles along a linestring (util.py:177-183; graph.rs:368-379)."""
    if len(linestring_coords) < 3:
        return 0.0
    angle_sum = 0.0
    for c_idx in range(1, len(linestring_coords) - 1):
        angle_sum += measure_coords_angle(
            linestring_coords[c_idx - 1], linestring_coords[c_idx], linestring_coords[c_idx + 1]
        )
    return angle_sum


def align_linestring_coords(
    linestring_coords: Sequence[Sequence[float]],
    x_y: Sequence[float],
    reverse: bool = False,
    tolerance: float = 0.5,
) -> list[Coord]:
    """Orient coordinates to start (or end) at ``x_y`` (util.py:264-312)."""
    coords = [tuple(c) for c in linestring_coords]
    a_dist = np.hypot(coords[0][0] - x_y[0], coords[0][1] - x_y[1])
    b_dist = np.hypot(coords[-1][0] - x_y[0], coords[-1][1] - x_y[1])
    if not reverse:
        if a_dist > b_dist:
            coords = coords[::-1]
        tol_dist = np.hypot(coords[0][0] - x_y[0], coords[0][1] - x_y[1])
    else:
        if a_dist < b_dist:
            coords = coords[::-1]
        tol_dist = np.hypot(coords[-1][0] - x_y[0], coords[-1][1] - x_y[1])
    if tol_dist > tolerance:
        raise ValueError(
            f"Closest side of edge geom {coords} is {tol_dist} from node {x_y}, exceeding tolerance of {tolerance}."
        )
    return coords


def validate_cityseer_networkx_graph(nx_multigraph):
    """Light validation: a copy of the graph with string node keys checked (util.py validation subset)."""
    g = nx_multigraph.copy()
    for node_key in g.nodes():
        if not isinstance(node_key, str):
            raise TypeError(f"Node key must be of type string but encountered {type(node_key)}")
    return g

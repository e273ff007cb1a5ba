"""WKT geometries as the flat arrays of ``cs_geoms`` (include/cityseer_amd.h).

A geometry is a list of parts ``(kind, rings)``: kind 0 point, 1 line string, 2 polygon; each
ring an ``(n, 2)`` float64 array (points and lines have one ring, a polygon its exterior then
its holes). Multi-geometries and collections are several parts; Z and M values are dropped;
EMPTY geometries have no parts. The reference parses WKT with the ``wkt`` crate into ``geo``
geometries (data.rs:290-377); this covers the geometry types those hold.
"""

from __future__ import annotations

import ctypes
import re

import numpy as np

KIND = {"POINT": 0, "LINESTRING": 1, "POLYGON": 2}
_TOKEN = re.compile(r"\s*([A-Za-z]+|\(|\)|,|[-+0-9.eE]+)")


def _tokens(text: str) -> list[str]:
    out, pos = [], 0
    text = text.strip()
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if m is None:
            raise ValueError(f"Failed to parse WKT near: {text[pos:pos + 40]!r}")
        out.append(m.group(1))
        pos = m.end()
    return out


class _Reader:
    def __init__(self, text: str):
        self.t = _tokens(text)
        self.i = 0

    def peek(self) -> str | None:
        return self.t[self.i] if self.i < len(self.t) else None

    def take(self, want: str | None = None) -> str:
        if self.i >= len(self.t):
            raise ValueError("Unexpected end of WKT")
        tok = self.t[self.i]
        if want is not None and tok.upper() != want:
            raise ValueError(f"Expected {want!r} in WKT, got {tok!r}")
        self.i += 1
        return tok

    def coords(self) -> np.ndarray:
        """'(' x y [z [m]] {',' x y ...} ')' -> (n, 2)"""
        self.take("(")
        pts = []
        while True:
            vals = []
            while self.peek() not in (",", ")"):
                vals.append(float(self.take()))
            if len(vals) < 2:
                raise ValueError("WKT coordinate needs at least x and y")
            pts.append(vals[:2])
            if self.take() == ")":
                break
        return np.asarray(pts, dtype=np.float64)

    def empty(self) -> bool:
        if self.peek() is not None and self.peek().upper() == "EMPTY":
            self.take()
            return True
        return False

    def dims(self) -> None:
        while self.peek() is not None and self.peek().upper() in ("Z", "M", "ZM"):
            self.take()

    def geometry(self) -> list[tuple[int, list[np.ndarray]]]:
        name = self.take().upper()
        self.dims()
        if self.empty():
            return []
        if name in ("POINT", "LINESTRING"):
            return [(KIND[name], [self.coords()])]
        if name == "POLYGON":
            return [(2, self.rings())]
        if name == "MULTIPOINT":
            self.take("(")
            parts = []
            while True:
                if self.peek() == "(":
                    parts.append((0, [self.coords()]))
                else:  # MULTIPOINT (x y, x y) without inner parentheses
                    vals = []
                    while self.peek() not in (",", ")"):
                        vals.append(float(self.take()))
                    parts.append((0, [np.asarray([vals[:2]], dtype=np.float64)]))
                if self.take() == ")":
                    return parts
        if name == "MULTILINESTRING":
            return [(1, [r]) for r in self.rings()]
        if name == "MULTIPOLYGON":
            self.take("(")
            parts = []
            while True:
                parts.append((2, self.rings()))
                if self.take() == ")":
                    return parts
        if name == "GEOMETRYCOLLECTION":
            self.take("(")
            parts = []
            while True:
                parts.extend(self.geometry())
                if self.take() == ")":
                    return parts
        raise ValueError(f"Unsupported WKT geometry type: {name}")

    def rings(self) -> list[np.ndarray]:
        self.take("(")
        rings = []
        while True:
            rings.append(self.coords())
            if self.take() == ")":
                return rings


def parse_wkt(text: str) -> list[tuple[int, list[np.ndarray]]]:
    """Parse one WKT geometry into parts (kind, rings)."""
    r = _Reader(text)
    parts = r.geometry()
    if r.peek() is not None:
        raise ValueError(f"Trailing text in WKT: {text[:200]}")
    return parts


class GeomArrays:
    """Flat ``cs_geoms`` arrays of many geometries; keeps them alive for the native call."""

    def __init__(self, geoms: list[list[tuple[int, list[np.ndarray]]]]):
        part_off, kinds, ring_off_p, ring_off, xy = [0], [], [0], [0], []
        n_rings = n_coords = 0
        for parts in geoms:
            for kind, rings in parts:
                kinds.append(kind)
                for ring in rings:
                    xy.append(np.asarray(ring, dtype=np.float64).reshape(-1, 2))
                    n_coords += len(ring)
                    ring_off.append(n_coords)
                    n_rings += 1
                ring_off_p.append(n_rings)
            part_off.append(len(kinds))
        self.geom_part_off = np.asarray(part_off, dtype=np.uint32)
        self.part_kind = np.asarray(kinds, dtype=np.uint8)
        self.part_ring_off = np.asarray(ring_off_p, dtype=np.uint32)
        self.ring_off = np.asarray(ring_off, dtype=np.uint32)
        self.xy = np.ascontiguousarray(np.concatenate(xy) if xy else np.zeros((0, 2)), dtype=np.float64)
        self.n = len(geoms)

    def struct(self, cls):
        p = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t))  # noqa: E731
        return cls(self.n, p(self.geom_part_off, ctypes.c_uint32), p(self.part_kind, ctypes.c_uint8),
                   p(self.part_ring_off, ctypes.c_uint32), p(self.ring_off, ctypes.c_uint32),
                   p(self.xy, ctypes.c_double))

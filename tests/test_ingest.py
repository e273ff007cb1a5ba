"""Native ingest (host only): edge payloads from polylines and edge decomposition.

Pinned by closed forms (a 3-4-5 segment, right-angle and U-turn polylines, straight-line cuts
at exact multiples), by the reference's own assertions on ``nx_decompose``
(tests/tools/test_graphs.py:548-609: one edge per node pair, total length kept, the same nodes
whatever the direction of each geometry, errors on missing geometry and on geometry that does
not meet its nodes) and, for random polylines, by a Python statement of the same rules.
"""

from __future__ import annotations

import math

import networkx as nx
import numpy as np
import pytest

from cityseer import _ingest
from cityseer.tools import graphs, io, util


def _py_substring(coords: list[tuple[float, ...]], start: float, end: float) -> list[tuple[float, ...]]:
    """Part of a polyline between two distances along it, for start < end (shapely
    ``ops.substring`` over GEOS: segment lengths as the root of summed squares, a point
    at a distance taken on the first segment extending strictly beyond it)."""
    seglens = [math.sqrt((b[0] - a[0]) ** 2 + (b[1] - a[1]) ** 2) for a, b in zip(coords[:-1], coords[1:], strict=True)]
    total = sum(seglens)
    start = min(max(start, 0.0), total)
    end = min(max(end, 0.0), total)

    def point_at(dist: float) -> tuple[float, ...]:
        walked = 0.0
        for (a, b), sl in zip(zip(coords[:-1], coords[1:], strict=True), seglens, strict=True):
            if walked + sl > dist:
                t = (dist - walked) / sl
                return tuple(a[i] + t * (b[i] - a[i]) for i in range(len(a)))
            walked += sl
        return tuple(coords[-1])

    out = [point_at(start)]
    walked = 0.0
    for i, sl in enumerate(seglens):
        walked += sl
        if walked >= end:
            break
        if start < walked:
            out.append(tuple(coords[i + 1]))
    out.append(point_at(end))
    return out


def _py_payload(coords):
    """Length (Σ hypot, f64 → f32) and cumulative turn angle (f64 → f32) of a polyline."""
    length = 0.0
    for a, b in zip(coords[:-1], coords[1:], strict=True):
        length += math.hypot(b[0] - a[0], b[1] - a[1])
    return np.float32(length), np.float32(util.measure_cumulative_angle(coords))


def test_edge_geometry_closed_forms():
    lines = [
        [(0.0, 0.0), (3.0, 4.0)],  # 5 m, straight
        [(0.0, 0.0), (10.0, 0.0), (10.0, 10.0)],  # one right angle
        [(0.0, 0.0), (10.0, 0.0), (0.0, 0.0)],  # a U-turn
        [(0.0, 0.0), (0.0, 0.0), (5.0, 0.0), (5.0, 0.0)],  # repeated points: no turn, length 5
        [(1.0, 1.0), (1.0, 1.0)],  # degenerate: zero length
    ]
    length, angle = _ingest.edge_geometry(lines)
    assert length.tolist() == [5.0, 20.0, 20.0, 5.0, 0.0]
    assert angle.tolist() == [0.0, 90.0, 180.0, 0.0, 0.0]


def test_edge_geometry_matches_python_statement():
    rng = np.random.default_rng(7)
    lines = []
    for _ in range(400):
        k = int(rng.integers(2, 9))
        pts = rng.uniform(-500.0, 500.0, (k, 2)).round(int(rng.integers(0, 4)))
        if k > 2 and rng.random() < 0.2:
            pts[1] = pts[0]  # a repeated point
        lines.append([tuple(p) for p in pts.tolist()])
    length, angle = _ingest.edge_geometry(lines)
    for i, line in enumerate(lines):
        pl, pa = _py_payload(line)
        # hypot may differ in the last bit of the f64 sum between libraries; f32 rarely sees it
        assert abs(float(length[i]) - float(pl)) <= float(np.spacing(pl)), (i, length[i], pl)
        assert abs(float(angle[i]) - float(pa)) <= float(np.spacing(np.float32(max(pa, 1.0)))), (i, angle[i], pa)


def test_decompose_closed_forms():
    # 100 m straight, 20 m pieces: 5 pieces ending at exact multiples of 20 m
    pieces = _ingest.decompose([[(0.0, 0.0), (100.0, 0.0)]], 20.0)[0]
    assert len(pieces) == 5
    assert [p[-1, 0] for p in pieces] == [20.0, 40.0, 60.0, 80.0, 100.0]
    assert all(p.shape == (2, 2) for p in pieces)
    # an L of 60 + 60 m in 50 m pieces: 3 pieces of 40 m, the corner kept inside the second
    pieces = _ingest.decompose([[(0.0, 0.0), (60.0, 0.0), (60.0, 60.0)]], 50.0)[0]
    assert len(pieces) == 3
    assert pieces[1].tolist() == [[40.0, 0.0], [60.0, 0.0], [60.0, 20.0]]
    assert pieces[2].tolist() == [[60.0, 20.0], [60.0, 60.0]]
    # shorter than the limit: one piece, the line itself
    pieces = _ingest.decompose([[(0.0, 0.0), (3.0, 4.0), (3.0, 8.0)]], 50.0)[0]
    assert len(pieces) == 1 and pieces[0].tolist() == [[0.0, 0.0], [3.0, 4.0], [3.0, 8.0]]
    # z interpolated with x and y
    pieces = _ingest.decompose([[(0.0, 0.0, 10.0), (40.0, 0.0, 30.0)]], 20.0)[0]
    assert pieces[0].tolist() == [[0.0, 0.0, 10.0], [20.0, 0.0, 20.0]]
    with pytest.raises(ValueError, match="zero length"):
        _ingest.decompose([[(1.0, 1.0), (1.0, 1.0)]], 20.0)
    with pytest.raises(ValueError, match="decompose_max"):
        _ingest.decompose([[(0.0, 0.0), (1.0, 1.0)]], 0.0)


def test_decompose_random_polylines_keep_length_and_limit():
    rng = np.random.default_rng(8)
    lines = [[tuple(p) for p in rng.uniform(-300.0, 300.0, (int(rng.integers(2, 7)), 2)).tolist()] for _ in range(300)]
    for limit in (7.0, 20.0, 95.5):
        out = _ingest.decompose(lines, limit)
        for line, pieces in zip(lines, out, strict=True):
            total = sum(math.dist(a, b) for a, b in zip(line[:-1], line[1:], strict=True))
            assert len(pieces) == max(1, math.ceil(total / limit))
            lens = [sum(math.dist(a, b) for a, b in zip(p[:-1].tolist(), p[1:].tolist(), strict=True)) for p in pieces]
            assert math.isclose(sum(lens), total, rel_tol=1e-12, abs_tol=1e-9)
            assert max(lens) <= limit * (1 + 1e-12)
            # consecutive pieces meet; the ends are the line's own
            assert pieces[0][0].tolist() == list(line[0]) and pieces[-1][-1].tolist() == list(line[-1])
            for a, b in zip(pieces[:-1], pieces[1:], strict=True):
                assert a[-1].tolist() == b[0].tolist()
            # every interior vertex of the line appears in exactly one piece
            inner = {tuple(v) for p in pieces for v in p[1:-1].tolist()}
            assert all(tuple(v) in inner for v in line[1:-1] if tuple(v) not in {tuple(p[-1]) for p in pieces})


def test_decompose_matches_python_statement():
    rng = np.random.default_rng(9)
    lines = [[tuple(p) for p in rng.uniform(-300.0, 300.0, (int(rng.integers(2, 6)), 2)).tolist()] for _ in range(200)]
    out = _ingest.decompose(lines, 13.0)
    for line, pieces in zip(lines, out, strict=True):
        total = sum(math.sqrt((b[0] - a[0]) ** 2 + (b[1] - a[1]) ** 2) for a, b in zip(line[:-1], line[1:], strict=True))
        step = total / len(pieces)
        at = 0.0
        for k, p in enumerate(pieces):
            end = total if k + 1 == len(pieces) else at + step
            want = _py_substring(line, at, end)
            assert p.tolist() == [list(c) for c in want]
            at += step


def test_nx_decompose_reference_properties(primal_graph):
    g = graphs.nx_decompose(primal_graph, 20)
    for s, e in g.edges():
        assert g.number_of_edges(s, e) == 1
    total = sum(d["geom"].length for _s, _e, d in primal_graph.edges(data=True))
    total_d = sum(d["geom"].length for _s, _e, d in g.edges(data=True))
    assert np.isclose(total, total_d, rtol=1e-9)
    # flipping every third geometry changes no node
    back = primal_graph.copy()
    for i, (s, e, k, d) in enumerate(back.edges(data=True, keys=True)):
        if i % 3 == 0:
            back[s][e][k]["geom"] = util.LineString(list(d["geom"].coords)[::-1])
    gb = graphs.nx_decompose(back, 20)
    for n, d in g.nodes(data=True):
        assert d["x"] == gb.nodes[n]["x"] and d["y"] == gb.nodes[n]["y"]
    # a geometry that does not meet its node, and a missing geometry, are errors
    moved = primal_graph.copy()
    first = next(iter(moved.nodes()))
    moved.nodes[first]["x"] += 1
    with pytest.raises(ValueError):
        graphs.nx_decompose(moved, 20)
    missing = primal_graph.copy()
    del missing["0"]["1"][0]["geom"]
    with pytest.raises(KeyError):
        graphs.nx_decompose(missing, 20)


def test_network_structure_payloads_from_native_bulk(primal_graph):
    """The bulk ingest stores exactly what the per-edge add_street_edge stores."""
    _n, _e, ns = io.network_structure_from_nx(primal_graph)
    a = ns.arrays()
    g = nx.MultiGraph(primal_graph)
    assert len(a["length"]) == 2 * g.number_of_edges()
    from cityseer.rustalgos.graph import NetworkStructure

    one = NetworkStructure()
    for key, nd in primal_graph.nodes(data=True):
        one.add_street_node(key, nd["x"], nd["y"], True, 1.0)
    src, dst, key = a["src"], a["dst"], a["key"]
    keys = list(primal_graph.nodes())
    for s, d, k in zip(src.tolist(), dst.tolist(), key.tolist(), strict=True):
        geom = primal_graph[keys[s]][keys[d]][k]["geom"]
        coords = util.align_linestring_coords(list(geom.coords), (primal_graph.nodes[keys[s]]["x"],
                                                                  primal_graph.nodes[keys[s]]["y"]))
        one.add_street_edge(s, d, k, keys[s], keys[d], coords)
    b = one.arrays()
    assert np.array_equal(a["length"], b["length"]) and np.array_equal(a["angle"], b["angle"])


def test_halve_and_dual_geometry(diamond_graph):
    """Halves meet at half length; a dual node sits at its primal edge's middle and a dual edge
    is the two near halves welded at the shared node (graphs.py:1965-2149)."""
    (a, b), = _ingest.halve([[(0.0, 0.0), (10.0, 0.0), (10.0, 10.0)]])
    assert a.tolist() == [[0.0, 0.0], [10.0, 0.0]] and b.tolist() == [[10.0, 0.0], [10.0, 10.0]]
    (a, b), = _ingest.halve([[(0.0, 0.0, 2.0), (8.0, 6.0, 4.0)]])
    assert a[-1].tolist() == [4.0, 3.0, 3.0] and b[0].tolist() == [4.0, 3.0, 3.0]
    dual = graphs.nx_to_dual(diamond_graph)
    for _n, d in dual.nodes(data=True):
        coords = list(d["primal_edge"].coords)
        (h0, _h1), = _ingest.halve([coords])
        assert (d["x"], d["y"]) == (h0[-1][0], h0[-1][1])
    for s, e, d in dual.edges(data=True):
        half = lambda n: dual.nodes[n]["primal_edge"].length / 2  # noqa: E731
        assert math.isclose(d["geom"].length, half(s) + half(e), rel_tol=1e-12)
        shared = d["primal_node_id"]
        px, py = diamond_graph.nodes[shared]["x"], diamond_graph.nodes[shared]["y"]
        assert any(math.isclose(c[0], px) and math.isclose(c[1], py) for c in d["geom"].coords)


# ---------------------------------------------------------------- network_structure_from_gpd
def _same_structure(a, b):
    x, y = a.arrays(), b.arrays()
    for k in ("x", "y", "live", "weight", "src", "dst", "key", "length", "angle", "imp", "pk"):
        assert np.array_equal(x[k], y[k], equal_nan=True), k
    assert a.node_keys_py() == b.node_keys_py()


@pytest.mark.parametrize("dual", [False, True])
def test_network_structure_from_gpd_round_trip(primal_graph, dual, tmp_path):
    from cityseer.tools import graphs, io

    g = graphs.nx_to_dual(primal_graph) if dual else primal_graph
    nodes, edges, ns = io.network_structure_from_nx(g)
    again = io.network_structure_from_gpd(nodes, edges)
    _same_structure(ns, again)
    # through CSV files, geometries as WKT text (io.py:1213-1321 serves reloaded frames)
    edges_csv = edges.assign(geom=[geom.wkt for geom in edges["geom"]])
    nodes.to_csv(tmp_path / "n.csv")
    edges_csv.to_csv(tmp_path / "e.csv")
    import pandas as pd

    n2 = pd.read_csv(tmp_path / "n.csv", index_col=0)
    n2.index = n2.index.astype(str)
    e2 = pd.read_csv(tmp_path / "e.csv", index_col=0, dtype={"nx_start_node_key": str, "nx_end_node_key": str})
    _same_structure(ns, io.network_structure_from_gpd(n2, e2))


def test_network_structure_from_gpd_checks(primal_graph, caplog):
    from cityseer.tools import io

    nodes, edges, ns = io.network_structure_from_nx(primal_graph)
    with pytest.raises(ValueError, match="nodes GDF: weight"):
        io.network_structure_from_gpd(nodes.drop(columns=["weight"]), edges)
    with pytest.raises(ValueError, match="edges GDF: imp_factor"):
        io.network_structure_from_gpd(nodes, edges.drop(columns=["imp_factor"]))
    # edges naming a node the frame lacks are skipped
    gone = nodes.index[0]
    import logging

    with caplog.at_level(logging.INFO):
        part = io.network_structure_from_gpd(nodes, edges)
        fewer = io.network_structure_from_gpd(nodes.drop(index=[gone]), edges)
    assert part.edge_count == ns.edge_count
    dropped = int(((edges["nx_start_node_key"] == gone) | (edges["nx_end_node_key"] == gone)).sum())
    assert dropped > 0 and fewer.edge_count == ns.edge_count - dropped
    assert "node nx key not found" in caplog.text


def test_edge_payload_view_matches_the_polyline():
    """get_edge_payload_py (graph.rs:1001-1011) on the mock graph: keys, index and stored length
    of each edge, bearings of its first and last segments, and its polyline as WKT."""
    import math

    from cityseer._geom import parse_wkt
    from cityseer.tools import graphs, io, mock

    G = graphs.nx_simple_geoms(mock.mock_graph())
    _n, _e, ns = io.network_structure_from_nx(G)
    keys = ns.node_keys_py()
    for s, d, k in ns.edge_references():
        pl = ns.get_edge_payload_py(s, d, k)
        assert (pl.start_nd_key_py, pl.end_nd_key_py, pl.edge_idx) == (keys[s], keys[d], k)
        assert pl.length == ns.get_edge_length(s, d, k) and pl.imp_factor == 1.0 and math.isnan(pl.seconds)
        (kind, rings), = parse_wkt(pl.geom_wkt)
        xy = rings[0]
        assert kind == 1 and tuple(xy[0]) == ns.get_node_payload_py(s).coord
        want_in = math.degrees(math.atan2(xy[1][1] - xy[0][1], xy[1][0] - xy[0][0]))
        want_out = math.degrees(math.atan2(xy[-1][1] - xy[-2][1], xy[-1][0] - xy[-2][0]))
        assert abs(pl.in_bearing - want_in) < 1e-3 and abs(pl.out_bearing - want_out) < 1e-3
        assert abs(float(np.hypot(*np.diff(xy, axis=0).T).sum()) - pl.length) < 1e-2
    pl = ns.get_edge_payload_py(*ns.edge_references()[0])
    assert "LINESTRING (700700 5719700" in pl.geom_wkt

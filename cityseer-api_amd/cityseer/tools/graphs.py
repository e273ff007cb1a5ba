"""Graph preparation for the centrality path.

Restates ``nx_simple_geoms`` (pysrc/cityseer/tools/graphs.py:58-107),
``nx_decompose`` (graphs.py:1829-1962) and ``nx_to_dual`` (graphs.py:1965-2149)
on networkX ``MultiGraph`` inputs, with ``tools.util.LineString`` in place of
shapely geometries. Graph cleaning (graphs.py:110-1828) is out of scope.
"""

from __future__ import annotations

import logging

import networkx as nx

from .. import _ingest
from . import util
from .util import LineString

logger = logging.getLogger(__name__)


def _coords_of(geom) -> list[tuple[float, ...]]:
    return [tuple(c) for c in (geom.coords if hasattr(geom, "coords") else geom)]


def nx_simple_geoms(nx_multigraph: nx.MultiGraph) -> nx.MultiGraph:
    """Straight-line edge geometries between node coordinates (graphs.py:58-107)."""
    g = nx_multigraph.copy()
    remove = []
    for s, e, k in g.edges(keys=True):
        sd, ed = g.nodes[s], g.nodes[e]
        s_z = "z" in sd and sd["z"] is not None
        e_z = "z" in ed and ed["z"] is not None
        if s_z and e_z:
            seg = LineString([(sd["x"], sd["y"], sd["z"]), (ed["x"], ed["y"], ed["z"])])
        else:
            seg = LineString([(sd["x"], sd["y"]), (ed["x"], ed["y"])])
        if seg.length == 0:
            if s != e:
                logger.warning(f"Found zero length edge between {s} and {e}, removing from graph.")
            remove.append((s, e, k))
        else:
            g[s][e][k]["geom"] = seg
    for s, e, k in remove:
        g.remove_edge(s, e, key=k)
    return util.validate_cityseer_networkx_graph(g)


def nx_decompose(nx_multigraph: nx.MultiGraph, decompose_max: float) -> nx.MultiGraph:
    """Split edges so that none exceeds ``decompose_max`` (graphs.py:1829-1962).

    The geometry (cut count, piece polylines) comes from the native library in
    one call for all edges; this function names the new nodes and builds the graph.
    """
    g = util.validate_cityseer_networkx_graph(nx_multigraph)
    edges = list(nx_multigraph.edges(data=True))
    lines = []
    for s, _e, data in edges:
        sxy = (nx_multigraph.nodes[s]["x"], nx_multigraph.nodes[s]["y"])
        lines.append(util.align_linestring_coords(_coords_of(data["geom"]), sxy))
    all_pieces = _ingest.decompose(lines, decompose_max) if lines else []
    for (s, e, data), pieces in zip(edges, all_pieces, strict=True):
        cuts = len(pieces)
        g.remove_edge(s, e)
        prior = s
        for sub in range(cuts - 1):
            seg = [tuple(c) for c in pieces[sub].tolist()]
            end_coord = seg[-1]
            x, y = end_coord[0], end_coord[1]
            z = end_coord[2] if len(end_coord) == 3 else None
            names = []
            for name in [s, sub, e]:
                name = str(name)
                if len(name) > 10:
                    name = f"{name[:5]}|{name[-5:]}"
                names.append(name)
            new_name = "±".join(names)
            if new_name in g:
                raise ValueError(
                    f"Attempted to add a duplicate node at x: {x}, y:{y}. "
                    f"Check for existence of duplicate edges in the vicinity of {s}-{e}."
                )
            attrs = {"x": x, "y": y}
            if z is not None:
                attrs["z"] = z
            g.add_node(new_name, **attrs)
            if "live" in nx_multigraph.nodes[s] and "live" in nx_multigraph.nodes[e]:
                live = bool(nx_multigraph.nodes[s]["live"]) or bool(nx_multigraph.nodes[e]["live"])
                g.nodes[new_name]["live"] = live
            copy = {k: v for k, v in data.items() if k != "geom"}
            g.add_edge(prior, new_name, geom=LineString(seg), **copy)
            prior = new_name
        seg = [tuple(c) for c in pieces[-1].tolist()]
        copy = {k: v for k, v in data.items() if k != "geom"}
        g.add_edge(prior, e, geom=LineString(seg), **copy)
    return util.validate_cityseer_networkx_graph(g)


def nx_to_dual(nx_multigraph: nx.MultiGraph) -> nx.MultiGraph:
    """Primal -> dual: one node per primal edge, one edge per pair of edges meeting at a node (graphs.py:1965-2149).

    Dual edge geometry is the near half of each primal edge welded at the
    shared primal node, so its length is half + half and its angle sum is the
    turn angle at that node. Dual edges carry ``primal_node_id``.
    """
    util.validate_cityseer_networkx_graph(nx_multigraph)
    if "is_dual" in nx_multigraph.graph:
        raise ValueError("Graph is already a dual graph.")
    g_dual = nx.MultiGraph()
    g_dual.graph["is_dual"] = True
    if "crs" in nx_multigraph.graph:
        g_dual.graph["crs"] = nx_multigraph.graph["crs"]

    # every primal edge halved natively in one call: as stored (its dual node sits at the
    # middle), and aligned to each of its ends (the halves that dual edges weld)
    edges = list(nx_multigraph.edges(data=True, keys=True))
    lines, where = [], {}
    for i, (s, e, k, data) in enumerate(edges):
        raw = _coords_of(data["geom"])
        s_xy = (nx_multigraph.nodes[s]["x"], nx_multigraph.nodes[s]["y"])
        e_xy = (nx_multigraph.nodes[e]["x"], nx_multigraph.nodes[e]["y"])
        lines += [raw, util.align_linestring_coords(raw, s_xy), util.align_linestring_coords(raw, e_xy)]
        where[(s, e, k)] = (i, 1)
        where[(e, s, k)] = (i, 2)
    halves = _ingest.halve(lines) if lines else []

    def half_geoms(a, b, k):
        i, side = where[(a, b, k)]
        a_half = [tuple(c) for c in halves[3 * i + side][0].tolist()]
        b_half = [tuple(c) for c in halves[3 * i + side][1].tolist()]
        a_xy = (nx_multigraph.nodes[a]["x"], nx_multigraph.nodes[a]["y"])
        a_half[0] = tuple([a_xy[0], a_xy[1]] + list(a_half[0][2:]))
        mid = a_half[-1]
        b_half[0] = mid
        b_xy = (nx_multigraph.nodes[b]["x"], nx_multigraph.nodes[b]["y"])
        b_half[-1] = tuple([b_xy[0], b_xy[1]] + list(b_half[-1][2:]))
        return a_half, b_half

    def key_of(s, e, k) -> str:
        se = sorted([str(s), str(e)])
        return f"{se[0]}_{se[1]}_k{k}"

    for i, (s, e, k, data) in enumerate(edges):
        geom = data["geom"]
        mid = halves[3 * i][0][-1]  # the middle of the edge as stored
        dk = key_of(s, e, k)
        kwargs = {
            "x": float(mid[0]),
            "y": float(mid[1]),
            "primal_edge": geom,
            "primal_edge_node_a": s,
            "primal_edge_node_b": e,
            "primal_edge_idx": k,
        }
        if len(mid) > 2:
            kwargs["z"] = float(mid[2])
        g_dual.add_node(dk, **kwargs)
        if "live" in nx_multigraph.nodes[s] and "live" in nx_multigraph.nodes[e]:
            live = bool(nx_multigraph.nodes[s]["live"]) or bool(nx_multigraph.nodes[e]["live"])
            g_dual.nodes[dk]["live"] = live
    for s, e, k in nx_multigraph.edges(keys=True):
        hub = key_of(s, e, k)
        s_half, e_half = half_geoms(s, e, k)
        for n_side, m_side, half in ((s, e, s_half), (e, s, e_half)):
            for nb in nx.neighbors(nx_multigraph, n_side):
                if nb == m_side:
                    continue
                for k2 in nx_multigraph[n_side][nb]:
                    spoke = key_of(n_side, nb, k2)
                    if g_dual.has_edge(hub, spoke):
                        continue
                    spoke_half, _ = half_geoms(n_side, nb, k2)
                    # hub midpoint -> shared primal node -> spoke midpoint
                    hub_part = list(reversed(half)) if half[0][:2] == spoke_half[0][:2] else list(half)
                    merged = hub_part + spoke_half[1:]
                    g_dual.add_edge(hub, spoke, primal_node_id=n_side, geom=LineString(merged))
    return util.validate_cityseer_networkx_graph(g_dual)

"""networkX -> ``NetworkStructure`` ingest.

Restates ``network_structure_from_nx`` (pysrc/cityseer/tools/io.py:1018-1210) and
``network_structure_from_gpd`` (io.py:1213-1321), which rebuilds the structure from the
two frames the former returns.

``network_structure_from_nx``:
nodes in networkX order, then for every node, every neighbour and every
multi-edge key one directed street edge with coordinates aligned to the start
node -- so each undirected edge is inserted once per direction, in the same
order as the reference. Returns pandas DataFrames in place of GeoDataFrames
(geopandas is not a dependency of this package). OSM / GTFS / CRS loaders are
out of scope.
"""

from __future__ import annotations

import logging
import math
from typing import Any

import networkx as nx
import numpy as np
import pandas as pd

from ..rustalgos.graph import NetworkStructure
from . import util

logger = logging.getLogger(__name__)


def network_structure_from_nx(nx_multigraph: nx.MultiGraph, crs: None = None):
    """Build ``(nodes_df, edges_df, NetworkStructure)`` from a networkX MultiGraph (io.py:1018-1210)."""
    g = util.validate_cityseer_networkx_graph(nx_multigraph)
    ns = NetworkStructure()
    is_dual = bool(g.graph.get("is_dual", False))
    ns.set_is_dual(is_dual)
    node_rows: dict[str, tuple[Any, ...]] = {}
    node_dual: dict[str, tuple[Any, ...]] = {}
    for node_key, nd in g.nodes(data=True):
        if not isinstance(node_key, str):
            raise TypeError(f"Node key must be of type string but encountered {type(node_key)}")
        x, y = nd["x"], nd["y"]
        z = nd.get("z", None)
        live = bool(nd["live"]) if "live" in nd else True
        weight = nd["weight"] if "weight" in nd else 1
        idx = ns.add_street_node(node_key, x, y, live, weight, z=z)
        node_rows[node_key] = (idx, x, y, z, live, weight)
        if is_dual:
            node_dual[node_key] = (
                nd["primal_edge"],
                nd["primal_edge_node_a"],
                nd["primal_edge_node_b"],
                nd["primal_edge_idx"],
            )
    edge_rows: dict[str, tuple[Any, ...]] = {}
    # every directed edge in adjacency order, then one native call derives the payloads
    starts, ends, keys, lines, imps, pks, rows = [], [], [], [], [], [], []
    for start in g.nodes():
        s_idx, sx, sy = node_rows[start][0], node_rows[start][1], node_rows[start][2]
        for end in g.neighbors(start):
            e_idx = node_rows[end][0]
            for edge_idx, ed in g[start][end].items():
                geom = ed.get("geom")
                if geom is None:
                    raise ValueError(f"Edge has no geometry: start_node={start}, end_node={end}, edge_idx={edge_idx}")
                coords = util.align_linestring_coords(list(geom.coords), (sx, sy))
                imp = float(ed.get("imp_factor", 1.0))
                if not math.isfinite(imp) or imp <= 0.0:
                    raise ValueError(
                        f"Invalid impedance factor ({imp}) for edge (idx {edge_idx}) between nodes "
                        f"{s_idx} and {e_idx}.\nImpedance must be finite and positive (> 0.0)."
                    )
                starts.append(s_idx)
                ends.append(e_idx)
                keys.append(int(edge_idx))
                lines.append(coords)
                imps.append(imp)
                pks.append(ed.get("primal_node_id"))
                rows.append((f"{start}-{end}", s_idx, e_idx, edge_idx, start, end, ed.get("imp_factor", 1.0),
                             util.measure_bearing(np.array(coords[0]), np.array(coords[-1])), util.LineString(coords),
                             ed.get("primal_node_id")))
    eids = ns.add_street_edges_from_lines(
        np.asarray(starts, dtype=np.int64), np.asarray(ends, dtype=np.int64), np.asarray(keys, dtype=np.uint32),
        lines, np.asarray(imps, dtype=np.float32), pks,
    ) if lines else []
    for eid, (name, *rest) in zip(eids, rows, strict=True):
        edge_rows[name] = (int(eid), *rest)
    nodes_df = pd.DataFrame.from_dict(
        node_rows, orient="index", columns=["ns_node_idx", "x", "y", "z", "live", "weight"]
    )
    if len(nodes_df):
        nodes_df[["x", "y"]] = nodes_df[["x", "y"]].astype(float)
    edges_df = pd.DataFrame.from_dict(
        edge_rows,
        orient="index",
        columns=[
            "ns_edge_idx", "start_ns_node_idx", "end_ns_node_idx", "edge_idx", "nx_start_node_key",
            "nx_end_node_key", "imp_factor", "total_bearing", "geom", "primal_node_id",
        ],
    )
    if not is_dual:
        edges_df = edges_df.drop(columns=["primal_node_id"])
    if is_dual:
        dual_df = pd.DataFrame.from_dict(
            node_dual, orient="index",
            columns=["primal_edge", "primal_edge_node_a", "primal_edge_node_b", "primal_edge_idx"],
        )
        nodes_df = nodes_df.join(dual_df)
    ns.validate()
    return nodes_df, edges_df, ns


_NODE_COLUMNS = ("x", "y", "live", "weight")
_EDGE_COLUMNS = ("start_ns_node_idx", "end_ns_node_idx", "edge_idx", "nx_start_node_key", "nx_end_node_key",
                 "imp_factor")


def _geometry_column(edges_df: pd.DataFrame) -> str:
    """The active geometry column of a GeoDataFrame, else ``geom`` or ``geometry``."""
    name = getattr(getattr(edges_df, "geometry", None), "name", None)
    if isinstance(name, str) and name in edges_df.columns:
        return name
    for name in ("geom", "geometry"):
        if name in edges_df.columns:
            return name
    raise ValueError("Missing expected geometry column in edges GDF: geom")


def _line_coords(geom) -> list:
    """Coordinates of a line given as WKT text or as an object with ``coords``."""
    if isinstance(geom, str):
        return util.parse_linestring_wkt(geom)
    return [tuple(float(c) for c in xy) for xy in geom.coords]


def network_structure_from_gpd(nodes_gdf: pd.DataFrame, edges_gdf: pd.DataFrame) -> NetworkStructure:
    """Rebuild a ``NetworkStructure`` from the nodes and edges frames ``network_structure_from_nx``
    returns, for instance after a round trip through files (io.py:1213-1321).

    Nodes are added in frame order with their ``x``, ``y``, ``live``, ``weight`` (and ``z`` when
    present and not missing); then every edge whose two networkX node keys are in the nodes frame,
    in frame order, its geometry aligned to start at its start node, with its ``edge_idx``,
    ``imp_factor`` and, for dual graphs, ``primal_node_id``. Edges naming an unknown node are
    skipped with an info log, as the reference does. The payload of all edges is derived in one
    native call. Any pandas DataFrame with these columns serves (GeoDataFrames included); edge
    geometries may be WKT text or objects with ``coords``.
    """
    for col in _NODE_COLUMNS:
        if col not in nodes_gdf.columns:
            raise ValueError(f"Missing expected column in nodes GDF: {col}")
    for col in _EDGE_COLUMNS:
        if col not in edges_gdf.columns:
            raise ValueError(f"Missing expected column in edges GDF: {col}")
    ns = NetworkStructure()
    has_z = "z" in nodes_gdf.columns
    position: dict[Any, tuple[int, float, float]] = {}
    for nd_key, nd in nodes_gdf.iterrows():
        z = nd["z"] if has_z else None
        if z is not None and isinstance(z, float) and math.isnan(z):
            z = None
        x, y = float(nd["x"]), float(nd["y"])
        idx = ns.add_street_node(str(nd_key), x, y, bool(nd["live"]), float(nd["weight"]),
                                 z=None if z is None else float(z))
        position[nd_key] = (idx, x, y)
    geom_col = _geometry_column(edges_gdf)
    has_pk = "primal_node_id" in edges_gdf.columns
    starts, ends, keys, lines, imps, pks = [], [], [], [], [], []
    for _row_key, ed in edges_gdf.iterrows():
        a, b = ed["nx_start_node_key"], ed["nx_end_node_key"]
        missing = [k for k in (a, b) if k not in position]
        if missing:
            for k in missing:
                logger.info(f"Skipping edge as {'start' if k == a else 'end'} node nx key not found {k}")
            continue
        s_idx, sx, sy = position[a]
        starts.append(s_idx)
        ends.append(position[b][0])
        keys.append(int(ed["edge_idx"]))
        lines.append(util.align_linestring_coords(_line_coords(ed[geom_col]), (sx, sy)))
        imps.append(float(ed["imp_factor"]))
        pk = ed["primal_node_id"] if has_pk else None
        pks.append(str(pk) if pk is not None and pd.notna(pk) else None)
    for imp, k, s_idx, e_idx in zip(imps, keys, starts, ends, strict=True):
        if not math.isfinite(imp) or imp <= 0.0:
            raise ValueError(
                f"Invalid impedance factor ({imp}) for edge (idx {k}) between nodes {s_idx} and {e_idx}.\n"
                "Impedance must be finite and positive (> 0.0)."
            )
    if lines:
        ns.add_street_edges_from_lines(
            np.asarray(starts, dtype=np.int64), np.asarray(ends, dtype=np.int64), np.asarray(keys, dtype=np.uint32),
            lines, np.asarray(imps, dtype=np.float32), pks if has_pk else None,
        )
    ns.validate()
    ns.build_edge_rtree()
    return ns


def nx_from_cityseer_geopandas(nodes_df: pd.DataFrame, edges_df: pd.DataFrame) -> nx.MultiGraph:
    """Round-trip the DataFrames back to a networkX MultiGraph (io.py:1431-1500)."""
    g = nx.MultiGraph()
    for nd_key, nd in nodes_df.iterrows():
        kwargs: dict[str, Any] = {
            "x": nd["x"],
            "y": nd["y"],
            "live": nd["live"] if "live" in nd else True,
            "weight": nd["weight"] if "weight" in nd else 1,
        }
        z = nd.get("z", None)
        if z is not None and not (isinstance(z, float) and np.isnan(z)):
            kwargs["z"] = z
        g.add_node(str(nd_key), **kwargs)
    for _, row in edges_df.iterrows():
        s, e = str(row["nx_start_node_key"]), str(row["nx_end_node_key"])
        if not g.has_edge(s, e, row["edge_idx"]):
            g.add_edge(s, e, row["edge_idx"], imp_factor=row["imp_factor"], geom=row["geom"])
    return g

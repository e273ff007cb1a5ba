"""networkX -> ``NetworkStructure`` ingest.

Restates ``network_structure_from_nx`` (pysrc/cityseer/tools/io.py:1018-1210):
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
                             util.measure_bearing(np.array(coords[0]), np.array(coords[-1])), util.LineString(coords)))
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
            "nx_end_node_key", "imp_factor", "total_bearing", "geom",
        ],
    )
    if is_dual:
        dual_df = pd.DataFrame.from_dict(
            node_dual, orient="index",
            columns=["primal_edge", "primal_edge_node_a", "primal_edge_node_b", "primal_edge_idx"],
        )
        nodes_df = nodes_df.join(dual_df)
    ns.validate()
    return nodes_df, edges_df, ns


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

r"""Network centrality front-end.

Drop-in for ``cityseer.metrics.networks`` (pysrc/cityseer/metrics/networks.py):
``node_centrality_shortest``, ``node_centrality_simplest``,
``segment_centrality`` and the closeness-only / betweenness-only wrappers.
Each resolves its thresholds, splits exact and sampled distances, runs the
GPU implementation through ``NetworkStructure`` and writes the same ``cc_*``
columns into ``nodes_gdf`` (any pandas DataFrame indexed by node key,
GeoDataFrames included).
"""

from __future__ import annotations

import logging
from functools import partial

import numpy as np
import pandas as pd

from .. import config, rustalgos, sampling

logger = logging.getLogger(__name__)

MIN_THRESH_WT = config.MIN_THRESH_WT
SPEED_M_S = config.SPEED_M_S


def _require_dual_for_angular(network_structure, context: str) -> None:
    """networks.py:90-98"""
    if not network_structure.is_dual:
        raise ValueError(
            f"{context} requires a dual graph for angular analysis. "
            "Convert the graph with cityseer.tools.graphs.nx_to_dual(...) before ingesting it."
        )


def _split_distances(resolved: list[int], sample: bool, epsilon: float | None):
    """Exact vs sampled thresholds (networks.py:182-194)."""
    eps = epsilon if epsilon is not None else sampling.HOEFFDING_EPSILON
    full: list[int] = []
    sampled: list[tuple[int, float]] = []
    if not sample:
        full = sorted(resolved)
    else:
        logger.warning("Sampling is experimental: API and behaviour may change in future releases.")
        for d in sorted(resolved):
            p = sampling.compute_distance_p(d, epsilon=eps)
            if p >= 1.0:
                full.append(d)
            else:
                sampled.append((d, p))
    return full, sampled


def _assign(nodes_gdf: pd.DataFrame, node_keys_py, temp_data: dict) -> pd.DataFrame:
    gdf_idx = nodes_gdf.index.intersection(node_keys_py)
    temp_df = pd.DataFrame(temp_data, index=node_keys_py)
    for col in temp_df.columns:
        if col not in nodes_gdf.columns:
            nodes_gdf[col] = np.nan
    nodes_gdf.loc[gdf_idx, temp_df.columns] = temp_df.loc[gdf_idx, temp_df.columns]
    return nodes_gdf


def node_centrality_shortest(
    network_structure,
    nodes_gdf: pd.DataFrame,
    distances: list[int] | None = None,
    betas: list[float] | None = None,
    minutes: list[float] | None = None,
    compute_closeness: bool = True,
    compute_betweenness: bool = True,
    min_threshold_wt: float = MIN_THRESH_WT,
    speed_m_s: float = SPEED_M_S,
    tolerance: float | None = None,
    random_seed: int | None = None,
    sample: bool = False,
    epsilon: float | None = None,
) -> pd.DataFrame:
    """Shortest-path node centrality, one Brandes search per source (networks.py:101-275)."""
    logger.info("Computing node centrality (shortest).")
    resolved, _b, _s = rustalgos.pair_distances_betas_time(
        speed_m_s, distances, betas, minutes, min_threshold_wt=min_threshold_wt
    )
    node_count = network_structure.street_node_count()
    full, sampled = _split_distances(resolved, sample, epsilon)
    results = {}
    if full:
        label = ", ".join(f"{d}m" for d in full)
        func = partial(
            network_structure.centrality_shortest,
            distances=full,
            compute_closeness=compute_closeness,
            compute_betweenness=compute_betweenness,
            min_threshold_wt=min_threshold_wt,
            speed_m_s=speed_m_s,
            tolerance=tolerance,
            random_seed=random_seed,
        )
        result = config.wrap_progress(total=node_count, rust_struct=network_structure, partial_func=func,
                                      desc=f"centrality full: {label}")
        for d in full:
            results[d] = result
    for d, p in sampled:
        func = partial(
            network_structure.centrality_shortest,
            distances=[d],
            compute_closeness=compute_closeness,
            compute_betweenness=compute_betweenness,
            min_threshold_wt=min_threshold_wt,
            speed_m_s=speed_m_s,
            tolerance=tolerance,
            sample_probability=p,
            random_seed=random_seed,
        )
        results[d] = config.wrap_progress(total=node_count, rust_struct=network_structure, partial_func=func,
                                          desc=f"centrality p={p:.0%}: {d}m")
    if not results:
        return nodes_gdf
    node_keys_py = next(iter(results.values())).node_keys_py
    temp: dict = {}
    if compute_closeness:
        for measure_key, attr_key in [
            ("beta", "node_beta"),
            ("cycles", "node_cycles"),
            ("density", "node_density"),
            ("farness", "node_farness"),
            ("harmonic", "node_harmonic"),
        ]:
            for d, res in results.items():
                temp[config.prep_gdf_key(measure_key, d)] = getattr(res, attr_key)[d]
        for d, res in results.items():
            with np.errstate(divide="ignore", invalid="ignore"):
                temp[config.prep_gdf_key("hillier", d)] = res.node_density[d] ** 2 / res.node_farness[d]
    if compute_betweenness:
        for measure_key, attr_key in [("betweenness", "node_betweenness"), ("betweenness_beta", "node_betweenness_beta")]:
            for d, res in results.items():
                temp[config.prep_gdf_key(measure_key, d)] = getattr(res, attr_key)[d]
    return _assign(nodes_gdf, node_keys_py, temp)


def node_centrality_simplest(
    network_structure,
    nodes_gdf: pd.DataFrame,
    distances: list[int] | None = None,
    betas: list[float] | None = None,
    minutes: list[float] | None = None,
    compute_closeness: bool = True,
    compute_betweenness: bool = True,
    min_threshold_wt: float = MIN_THRESH_WT,
    speed_m_s: float = SPEED_M_S,
    tolerance: float | None = None,
    angular_scaling_unit: float = 90,
    farness_scaling_offset: float = 1,
    random_seed: int | None = None,
    sample: bool = False,
    epsilon: float | None = None,
) -> pd.DataFrame:
    """Simplest-path (angular) node centrality on a dual graph (networks.py:464-635)."""
    _require_dual_for_angular(network_structure, "node_centrality_simplest")
    logger.info("Computing node centrality (simplest).")
    resolved, _b, _s = rustalgos.pair_distances_betas_time(
        speed_m_s, distances, betas, minutes, min_threshold_wt=min_threshold_wt
    )
    node_count = network_structure.street_node_count()
    full, sampled = _split_distances(resolved, sample, epsilon)
    results = {}
    common = dict(
        compute_closeness=compute_closeness,
        compute_betweenness=compute_betweenness,
        min_threshold_wt=min_threshold_wt,
        speed_m_s=speed_m_s,
        tolerance=tolerance,
        angular_scaling_unit=angular_scaling_unit,
        farness_scaling_offset=farness_scaling_offset,
        random_seed=random_seed,
    )
    if full:
        label = ", ".join(f"{d}m" for d in full)
        func = partial(network_structure.centrality_simplest, distances=full, **common)
        result = config.wrap_progress(total=node_count, rust_struct=network_structure, partial_func=func,
                                      desc=f"centrality simplest full: {label}")
        for d in full:
            results[d] = result
    for d, p in sampled:
        func = partial(network_structure.centrality_simplest, distances=[d], sample_probability=p, **common)
        results[d] = config.wrap_progress(total=node_count, rust_struct=network_structure, partial_func=func,
                                          desc=f"centrality simplest p={p:.0%}: {d}m")
    if not results:
        return nodes_gdf
    node_keys_py = next(iter(results.values())).node_keys_py
    temp: dict = {}
    if compute_closeness:
        for d, res in results.items():
            temp[config.prep_gdf_key("density", d, angular=True)] = res.node_density[d]
            temp[config.prep_gdf_key("harmonic", d, angular=True)] = res.node_harmonic[d]
            temp[config.prep_gdf_key("farness", d, angular=True)] = res.node_farness[d]
            with np.errstate(divide="ignore", invalid="ignore"):
                temp[config.prep_gdf_key("hillier", d, angular=True)] = res.node_density[d] ** 2 / res.node_farness[d]
    if compute_betweenness:
        for d, res in results.items():
            temp[config.prep_gdf_key("betweenness", d, angular=True)] = res.node_betweenness[d]
    return _assign(nodes_gdf, node_keys_py, temp)


def segment_centrality(
    network_structure,
    nodes_gdf: pd.DataFrame,
    distances: list[int] | None = None,
    betas: list[float] | None = None,
    minutes: list[float] | None = None,
    compute_closeness: bool | None = True,
    compute_betweenness: bool | None = True,
    min_threshold_wt: float = MIN_THRESH_WT,
    speed_m_s: float = SPEED_M_S,
) -> pd.DataFrame:
    """Segment-based centrality with shortest paths (networks.py:638-755)."""
    logger.info("Computing shortest path segment centrality.")
    func = partial(
        network_structure.segment_centrality,
        distances=distances,
        betas=betas,
        minutes=minutes,
        compute_closeness=compute_closeness,
        compute_betweenness=compute_betweenness,
        min_threshold_wt=min_threshold_wt,
        speed_m_s=speed_m_s,
    )
    result = config.wrap_progress(total=network_structure.street_node_count(), rust_struct=network_structure,
                                  partial_func=func)
    distances = config.log_thresholds(
        distances=distances, betas=betas, minutes=minutes, min_threshold_wt=min_threshold_wt, speed_m_s=speed_m_s
    )
    temp: dict = {}
    if compute_closeness is True:
        for measure_key, attr_key in [
            ("seg_density", "segment_density"),
            ("seg_harmonic", "segment_harmonic"),
            ("seg_beta", "segment_beta"),
        ]:
            for d in distances:
                temp[config.prep_gdf_key(measure_key, d)] = getattr(result, attr_key)[d]
    if compute_betweenness is True:
        for d in distances:
            temp[config.prep_gdf_key("seg_betweenness", d)] = result.segment_betweenness[d]
    return _assign(nodes_gdf, result.node_keys_py, temp)


def closeness_shortest(network_structure, nodes_gdf, distances=None, betas=None, minutes=None,
                       min_threshold_wt=MIN_THRESH_WT, speed_m_s=SPEED_M_S, tolerance=None, random_seed=None,
                       sample=False, epsilon=None):
    """Closeness only (networks.py:763-791)."""
    return node_centrality_shortest(network_structure, nodes_gdf, distances, betas, minutes, True, False,
                                    min_threshold_wt, speed_m_s, tolerance, random_seed, sample, epsilon)


def closeness_simplest(network_structure, nodes_gdf, distances=None, betas=None, minutes=None,
                       min_threshold_wt=MIN_THRESH_WT, speed_m_s=SPEED_M_S, tolerance=None, angular_scaling_unit=90,
                       farness_scaling_offset=1, random_seed=None, sample=False, epsilon=None):
    """Angular closeness only (networks.py:794-826)."""
    return node_centrality_simplest(network_structure, nodes_gdf, distances, betas, minutes, True, False,
                                    min_threshold_wt, speed_m_s, tolerance, angular_scaling_unit,
                                    farness_scaling_offset, random_seed, sample, epsilon)


def betweenness_shortest(network_structure, nodes_gdf, distances=None, betas=None, minutes=None,
                         min_threshold_wt=MIN_THRESH_WT, speed_m_s=SPEED_M_S, tolerance=None, random_seed=None,
                         sample=False, epsilon=None):
    """Betweenness only (networks.py:829-857)."""
    return node_centrality_shortest(network_structure, nodes_gdf, distances, betas, minutes, False, True,
                                    min_threshold_wt, speed_m_s, tolerance, random_seed, sample, epsilon)


def betweenness_simplest(network_structure, nodes_gdf, distances=None, betas=None, minutes=None,
                         min_threshold_wt=MIN_THRESH_WT, speed_m_s=SPEED_M_S, tolerance=None, random_seed=None,
                         sample=False, epsilon=None):
    """Angular betweenness only (networks.py:860-888)."""
    return node_centrality_simplest(network_structure, nodes_gdf, distances, betas, minutes, False, True,
                                    min_threshold_wt, speed_m_s, tolerance, 90, 1, random_seed, sample, epsilon)


def _zone_points(zones) -> tuple[np.ndarray, np.ndarray]:
    """Zone identifiers (the index) and one (x, y) per zone: the centroid of a polygon geometry,
    a point geometry's coordinates, or columns ``x`` and ``y``."""
    ids = np.asarray(zones.index.values)
    if "geometry" in getattr(zones, "columns", []):
        pts = []
        for g in zones["geometry"]:
            c = g.centroid if getattr(g, "geom_type", "Point") in ("Polygon", "MultiPolygon") else g
            pts.append((float(c.x), float(c.y)))
        return ids, np.asarray(pts, dtype=np.float64).reshape(-1, 2)
    return ids, np.column_stack([np.asarray(zones["x"], np.float64), np.asarray(zones["y"], np.float64)])


def build_od_matrix(
    od_df: pd.DataFrame,
    zones_gdf,
    network_structure,
    origin_col: str,
    destination_col: str,
    weight_col: str,
    zone_id_col: str | None = None,
    max_snap_dist: float = 500.0,
) -> rustalgos.centrality.OdMatrix:
    """An ``OdMatrix`` from zone-to-zone flows (networks.py:278-373): each zone's centroid snaps
    to its nearest network node (KD-tree) unless farther than ``max_snap_dist``; flows with a
    missing, zero or negative weight, or an unsnapped zone, are dropped. Zones in projected
    coordinates matching the network (reprojection from EPSG:4326 is not supported here)."""
    from scipy.spatial import KDTree

    crs = getattr(zones_gdf, "crs", None)
    if crs is not None and getattr(crs, "to_epsg", lambda: None)() == 4326:
        raise ValueError("zones in EPSG:4326 must be projected to the network's CRS first")
    ids, xy = _zone_points(zones_gdf)
    if zone_id_col is not None:
        ids = np.asarray(zones_gdf[zone_id_col].values)
    tree = KDTree(np.asarray(network_structure.node_xys, dtype=np.float64))
    snap_d, snap_i = tree.query(xy)
    zone_to_node: dict = {}
    n_excluded = 0
    for i, zid in enumerate(ids):
        if snap_d[i] > max_snap_dist:
            n_excluded += 1
            continue
        zone_to_node[zid] = int(snap_i[i])
    if n_excluded:
        logger.warning(f"{n_excluded} zone centroids exceeded max_snap_dist={max_snap_dist}m and were excluded")
    origins: list[int] = []
    dests: list[int] = []
    weights: list[float] = []
    for o_zone, d_zone, w in zip(od_df[origin_col], od_df[destination_col], od_df[weight_col]):
        if pd.isna(w) or w <= 0:
            continue
        if o_zone not in zone_to_node or d_zone not in zone_to_node:
            continue
        origins.append(zone_to_node[o_zone])
        dests.append(zone_to_node[d_zone])
        weights.append(float(w))
    logger.info(f"Built OD matrix: {len(origins)} pairs, {sum(weights):.0f} total trips")
    return rustalgos.centrality.OdMatrix(origins, dests, weights)


def betweenness_od(
    network_structure,
    nodes_gdf: pd.DataFrame,
    od_matrix,
    distances: list[int] | None = None,
    betas: list[float] | None = None,
    minutes: list[float] | None = None,
    min_threshold_wt: float = MIN_THRESH_WT,
    speed_m_s: float = SPEED_M_S,
    tolerance: float | None = None,
) -> pd.DataFrame:
    """OD-weighted betweenness with shortest paths (networks.py:376-461): columns
    ``cc_betweenness_{d}`` and ``cc_betweenness_beta_{d}``; closeness is not computed."""
    logger.info("Computing OD-weighted betweenness centrality.")
    func = partial(
        network_structure.betweenness_od_shortest,
        od_matrix=od_matrix,
        distances=distances,
        betas=betas,
        minutes=minutes,
        min_threshold_wt=min_threshold_wt,
        speed_m_s=speed_m_s,
        tolerance=tolerance,
    )
    result = config.wrap_progress(total=network_structure.street_node_count(), rust_struct=network_structure,
                                  partial_func=func)
    distances = config.log_thresholds(
        distances=distances, betas=betas, minutes=minutes, min_threshold_wt=min_threshold_wt, speed_m_s=speed_m_s
    )
    temp: dict = {}
    for measure_key, attr_key in [("betweenness", "node_betweenness"), ("betweenness_beta", "node_betweenness_beta")]:
        for d in distances:
            temp[config.prep_gdf_key(measure_key, d)] = getattr(result, attr_key)[d]
    return _assign(nodes_gdf, result.node_keys_py, temp)

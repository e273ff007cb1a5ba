r"""Land-use and statistical layers front-end.

Drop-in for ``cityseer.metrics.layers`` (pysrc/cityseer/metrics/layers.py):
``build_data_map`` assigns a frame of geometries to the network (natively, on
the host) and ``compute_accessibilities``, ``compute_mixed_uses`` and
``compute_stats`` run the data layer on the GPU through ``DataMap`` and write
the same ``cc_*`` columns into ``nodes_gdf``.

Frames are pandas DataFrames (GeoDataFrames included) with unique indices. The
geometry column is the frame's ``active_geometry_name`` where it has one, else
``geometry``; its values are WKT strings or objects with a ``wkt`` attribute
(shapely geometries).
"""

from __future__ import annotations

import logging
from functools import partial

import pandas as pd

from .. import config, rustalgos
from .networks import _assign, _require_dual_for_angular

logger = logging.getLogger(__name__)

MIN_THRESH_WT = config.MIN_THRESH_WT
SPEED_M_S = config.SPEED_M_S


def _wkt_column(frame: pd.DataFrame) -> list[str]:
    name = getattr(frame, "active_geometry_name", None) or "geometry"
    if name not in frame.columns:
        raise ValueError(f"The frame has no geometry column ({name!r}).")
    return [g if isinstance(g, str) else g.wkt for g in frame[name]]


def build_data_map(
    data_gdf: pd.DataFrame,
    network_structure,
    max_netw_assign_dist: int = 100,
    data_id_col: str | None = None,
    barriers_gdf: pd.DataFrame | None = None,
    n_nearest_candidates: int = 50,
):
    """A ``DataMap`` of ``data_gdf``'s rows, keyed by index (deduplicated by ``data_id_col``
    when given), assigned to the nearest street edges within ``max_netw_assign_dist``, with
    ``barriers_gdf``'s geometries blocking lines of sight during the assignment
    (layers.py:30-101)."""
    if data_gdf.index.duplicated().any():
        raise ValueError("The data GeoDataFrame index must contain unique entries.")
    data_map = rustalgos.data.DataMap()
    logger.info("Assigning data to network.")
    wkts = _wkt_column(data_gdf)
    ids = [None] * len(wkts) if data_id_col is None else list(data_gdf[data_id_col])
    for key, wkt, data_id in zip(data_gdf.index, wkts, ids, strict=True):
        data_map.insert(key, wkt, data_id)
    if barriers_gdf is not None:
        network_structure.set_barriers(_wkt_column(barriers_gdf))
    try:
        data_map.assign_data_to_network(network_structure, max_netw_assign_dist, n_nearest_candidates)
    finally:
        network_structure.unset_barriers()
    return data_map


def _call(network_structure, data_map, func):
    return config.wrap_progress(total=network_structure.street_node_count(), rust_struct=data_map,
                                partial_func=func)


def compute_accessibilities(
    data_gdf: pd.DataFrame,
    landuse_column_label: str,
    accessibility_keys: list[str],
    nodes_gdf: pd.DataFrame,
    network_structure,
    max_netw_assign_dist: int = 100,
    distances: list[int] | None = None,
    betas: list[float] | None = None,
    minutes: list[float] | None = None,
    data_id_col: str | None = None,
    barriers_gdf: pd.DataFrame | None = None,
    angular: bool = False,
    n_nearest_candidates: int = 50,
    spatial_tolerance: int = 0,
    min_threshold_wt: float = MIN_THRESH_WT,
    speed_m_s: float = SPEED_M_S,
) -> tuple[pd.DataFrame, pd.DataFrame]:
    """Land-use accessibility per node and threshold (layers.py:104-292): for each key,
    ``cc_{key}_{d}_nw`` (count), ``cc_{key}_{d}_wt`` (distance-weighted count) and, at the largest
    threshold, ``cc_{key}_nearest_max_{d}`` (distance to the nearest)."""
    if angular:
        _require_dual_for_angular(network_structure, "compute_accessibilities")
    logger.info(f"Computing land-use accessibility for: {', '.join(accessibility_keys)}")
    data_map = build_data_map(data_gdf, network_structure, max_netw_assign_dist, data_id_col,
                              barriers_gdf=barriers_gdf, n_nearest_candidates=n_nearest_candidates)
    if landuse_column_label not in data_gdf.columns:
        raise ValueError("The specified landuse column name can't be found in the GeoDataFrame.")
    landuses_map = dict(data_gdf[landuse_column_label])
    result = _call(network_structure, data_map, partial(
        data_map.accessibility, network_structure=network_structure, landuses_map=landuses_map,
        accessibility_keys=accessibility_keys, distances=distances, betas=betas, minutes=minutes, angular=angular,
        spatial_tolerance=spatial_tolerance, min_threshold_wt=min_threshold_wt, speed_m_s=speed_m_s))
    distances = config.log_thresholds(distances=distances, betas=betas, minutes=minutes,
                                      min_threshold_wt=min_threshold_wt, speed_m_s=speed_m_s)
    temp_data = {}
    for key in accessibility_keys:
        lu = result.result[key]
        unweighted, weighted, nearest = lu.unweighted, lu.weighted, lu.distance
        for d in distances:
            temp_data[config.prep_gdf_key(key, d, angular, weighted=False)] = unweighted[d]
            temp_data[config.prep_gdf_key(key, d, angular, weighted=True)] = weighted[d]
            if d == max(distances):
                temp_data[config.prep_gdf_key(f"{key}_nearest_max", d, angular)] = nearest[d]
    return _assign(nodes_gdf, result.node_keys_py, temp_data), data_gdf


def compute_mixed_uses(
    data_gdf: pd.DataFrame,
    landuse_column_label: str,
    nodes_gdf: pd.DataFrame,
    network_structure,
    max_netw_assign_dist: int = 100,
    compute_hill: bool | None = True,
    compute_hill_weighted: bool | None = True,
    compute_shannon: bool | None = False,
    compute_gini: bool | None = False,
    distances: list[int] | None = None,
    betas: list[float] | None = None,
    minutes: list[float] | None = None,
    data_id_col: str | None = None,
    barriers_gdf: pd.DataFrame | None = None,
    angular: bool = False,
    n_nearest_candidates: int = 50,
    spatial_tolerance: int = 0,
    min_threshold_wt: float = MIN_THRESH_WT,
    speed_m_s: float = SPEED_M_S,
) -> tuple[pd.DataFrame, pd.DataFrame]:
    """Mixed-use measures per node and threshold (layers.py:295-538): Hill numbers for q = 0, 1,
    2 (``cc_hill_q{q}_{d}_nw`` / ``_wt``), Shannon (``cc_shannon_{d}``) and Gini-Simpson
    (``cc_gini_{d}``), each as its flag asks."""
    if angular:
        _require_dual_for_angular(network_structure, "compute_mixed_uses")
    logger.info("Computing mixed-use measures.")
    data_map = build_data_map(data_gdf, network_structure, max_netw_assign_dist, data_id_col,
                              barriers_gdf=barriers_gdf, n_nearest_candidates=n_nearest_candidates)
    if landuse_column_label not in data_gdf.columns:
        raise ValueError("The specified landuse column name can't be found in the GeoDataFrame.")
    landuses_map = dict(data_gdf[landuse_column_label])
    result = _call(network_structure, data_map, partial(
        data_map.mixed_uses, network_structure=network_structure, landuses_map=landuses_map, distances=distances,
        betas=betas, minutes=minutes, compute_hill=compute_hill, compute_hill_weighted=compute_hill_weighted,
        compute_shannon=compute_shannon, compute_gini=compute_gini, angular=angular,
        spatial_tolerance=spatial_tolerance, min_threshold_wt=min_threshold_wt, speed_m_s=speed_m_s))
    distances = config.log_thresholds(distances=distances, betas=betas, minutes=minutes,
                                      min_threshold_wt=min_threshold_wt, speed_m_s=speed_m_s)
    hill = result.hill if compute_hill else None
    hill_wt = result.hill_weighted if compute_hill_weighted else None
    shannon = result.shannon if compute_shannon else None
    gini = result.gini if compute_gini else None
    temp_data = {}
    for d in distances:
        for q in (0, 1, 2):
            if hill is not None:
                temp_data[config.prep_gdf_key(f"hill_q{q}", d, angular, weighted=False)] = hill[q][d]
            if hill_wt is not None:
                temp_data[config.prep_gdf_key(f"hill_q{q}", d, angular, weighted=True)] = hill_wt[q][d]
        if shannon is not None:
            temp_data[config.prep_gdf_key("shannon", d, angular)] = shannon[d]
        if gini is not None:
            temp_data[config.prep_gdf_key("gini", d, angular)] = gini[d]
    return _assign(nodes_gdf, result.node_keys_py, temp_data), data_gdf


# (column suffix, weighted flag, Stats getter) in the reference's column order (layers.py:722-751)
_STAT_COLUMNS = (
    ("sum", False, "sum"), ("sum", True, "sum_wt"), ("mean", False, "mean"), ("mean", True, "mean_wt"),
    ("count", False, "count"), ("count", True, "count_wt"), ("median", False, "median"),
    ("median", True, "median_wt"), ("var", False, "variance"), ("var", True, "variance_wt"),
    ("mad", False, "mad"), ("mad", True, "mad_wt"), ("max", None, "max"), ("min", None, "min"),
)


def compute_stats(
    data_gdf: pd.DataFrame,
    stats_column_labels: list[str],
    nodes_gdf: pd.DataFrame,
    network_structure,
    max_netw_assign_dist: int = 100,
    distances: list[int] | None = None,
    betas: list[float] | None = None,
    minutes: list[float] | None = None,
    data_id_col: str | None = None,
    barriers_gdf: pd.DataFrame | None = None,
    angular: bool = False,
    n_nearest_candidates: int = 50,
    spatial_tolerance: int = 0,
    min_threshold_wt: float = MIN_THRESH_WT,
    speed_m_s: float = SPEED_M_S,
) -> tuple[pd.DataFrame, pd.DataFrame]:
    """Numerical statistics per column, node and threshold (layers.py:541-756):
    ``cc_{col}_{stat}_{d}[_nw|_wt]`` for sum, mean, count, median, var and mad (plain and
    distance-weighted), max and min."""
    if angular:
        _require_dual_for_angular(network_structure, "compute_stats")
    logger.info("Computing statistics.")
    data_map = build_data_map(data_gdf, network_structure, max_netw_assign_dist, data_id_col,
                              barriers_gdf=barriers_gdf, n_nearest_candidates=n_nearest_candidates)
    stats_maps = []
    for label in stats_column_labels:
        if label not in data_gdf.columns:
            raise ValueError("The specified numerical stats column name can't be found in the GeoDataFrame.")
        stats_maps.append(dict(data_gdf[label]))
    result = _call(network_structure, data_map, partial(
        data_map.stats, network_structure=network_structure, numerical_maps=stats_maps, distances=distances,
        betas=betas, minutes=minutes, angular=angular, spatial_tolerance=spatial_tolerance,
        min_threshold_wt=min_threshold_wt, speed_m_s=speed_m_s))
    distances = config.log_thresholds(distances=distances, betas=betas, minutes=minutes,
                                      min_threshold_wt=min_threshold_wt, speed_m_s=speed_m_s)
    temp_data = {}
    for stats, label in zip(result.result, stats_column_labels, strict=True):
        rows = {getter: getattr(stats, getter) for _s, _w, getter in _STAT_COLUMNS}
        for d in distances:
            for suffix, weighted, getter in _STAT_COLUMNS:
                temp_data[config.prep_gdf_key(f"{label}_{suffix}", d, angular, weighted=weighted)] = rows[getter][d]
    return _assign(nodes_gdf, result.node_keys_py, temp_data), data_gdf

"""Accessibility restated on the CPU. TEST INFRASTRUCTURE ONLY.

Loaded by tests/ as the parity checker; never by the product package. Follows
DataMap.aggregate_to_src_idx (rust/src/data.rs:520-615) and the accumulation in
DataMap.accessibility (data.rs:680-741) loop for loop, over the oracle's own
dijkstra_tree_shortest (the visit order is the settle order). Values are f32
where the reference computes in f32 and f64 where it accumulates in f64.
"""

from __future__ import annotations

import numpy as np

f32 = np.float32


def aggregate_to_src_idx(tree, node_data_map, dedupe_of, max_walk_seconds, speed):
    """{data key: total distance} for one source (data.rs:520-615)."""
    entries_result: dict[str, np.float32] = {}
    nearest_ids: dict[str, tuple[str, np.float32]] = {}
    max_walk_dist = f32(f32(max_walk_seconds) * f32(speed))
    for node_idx in tree["order"].tolist():
        agg = tree["secs"][node_idx]
        if agg >= f32(max_walk_seconds):
            continue
        for data_key, data_dist in node_data_map.get(node_idx, []):
            if data_key not in dedupe_of:
                continue
            network_dist = f32(agg * f32(speed))
            total = f32(network_dist + f32(data_dist))
            if total <= max_walk_dist:
                dk = dedupe_of[data_key]
                if dk in nearest_ids:
                    cur_key, cur_dist = nearest_ids[dk]
                    if total < cur_dist:
                        entries_result.pop(cur_key, None)
                        nearest_ids[dk] = (data_key, total)
                        entries_result[data_key] = total
                else:
                    nearest_ids[dk] = (data_key, total)
                    entries_result[data_key] = total
    return entries_result


def accessibility(og, live, node_data_map, dedupe_of, lu_class, accessibility_keys, distances, betas, mcws,
                  max_walk_seconds, speed, sources=None):
    """{class: (unweighted f64[T][N], weighted f64[T][N], nearest f64[N])} (data.rs:680-741);
    `sources` restricts the rows computed (all nodes by default)."""
    n, T = og.n, len(distances)
    max_dist = max(distances)
    res = {k: (np.zeros((T, n)), np.zeros((T, n)), np.full(n, np.nan)) for k in accessibility_keys}
    for src in (range(n) if sources is None else sources):
        if not live[src]:
            continue
        tree = og.tree_shortest(src, max_walk_seconds, speed)
        reach = aggregate_to_src_idx(tree, node_data_map, dedupe_of, max_walk_seconds, speed)
        for data_key, data_dist in reach.items():
            cls = lu_class.get(data_key)
            if cls not in res:
                continue
            unw, wtd, near = res[cls]
            for i, (d, b, mcw) in enumerate(zip(distances, betas, mcws, strict=True)):
                if data_dist <= f32(d):
                    unw[i, src] += 1.0
                    raw = np.exp(f32(-f32(b)) * data_dist, dtype=np.float32)
                    wtd[i, src] += float(f32(min(raw, f32(mcw)) / f32(mcw)))
                    if d == max_dist and (np.isnan(near[src]) or float(data_dist) < near[src]):
                        near[src] = float(data_dist)
    return res

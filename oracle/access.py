"""Accessibility restated on the CPU. TEST INFRASTRUCTURE ONLY.

Loaded by tests/ as the parity checker; never by the product package. Follows
DataMap.aggregate_to_src_idx (rust/src/data.rs:520-615) and the accumulation in
DataMap.accessibility (data.rs:680-741) loop for loop, over the oracle's own
dijkstra_tree_shortest (the visit order is the settle order) or, with
``angular``, its dijkstra_tree_simplest (the visit order is the first-reach
order; data.rs:538-548). Values are f32
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


def _tree(og, src, max_walk_seconds, speed, angular):
    """The search aggregate_to_src_idx walks (data.rs:538-548)."""
    if angular:
        return og.tree_simplest(src, max_walk_seconds, speed)
    return og.tree_shortest(src, max_walk_seconds, speed)


def accessibility(og, live, node_data_map, dedupe_of, lu_class, accessibility_keys, distances, betas, mcws,
                  max_walk_seconds, speed, sources=None, angular=False):
    """{class: (unweighted f64[T][N], weighted f64[T][N], nearest f64[N])} (data.rs:680-741);
    `sources` restricts the rows computed (all nodes by default)."""
    n, T = og.n, len(distances)
    max_dist = max(distances)
    res = {k: (np.zeros((T, n)), np.zeros((T, n)), np.full(n, np.nan)) for k in accessibility_keys}
    for src in (range(n) if sources is None else sources):
        if not live[src]:
            continue
        tree = _tree(og, src, max_walk_seconds, speed, angular)
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


def accessibility_threads(og, live, node_data_map, dedupe_of, lu_class, accessibility_keys, distances, betas, mcws,
                          max_walk_seconds, speed, angular=False, n_threads=1):
    """The same as ``accessibility`` in C (og_accessibility), threads over sources: the
    restatement the CPU baseline times. Returns the same {class: (unweighted, weighted,
    nearest)} rows."""
    import ctypes

    from .oracle import _p, load

    keys = list(dedupe_of.keys())
    index = {k: i for i, k in enumerate(keys)}
    groups: dict = {}
    entry_group = np.asarray([groups.setdefault(dedupe_of[k], len(groups)) for k in keys], dtype=np.uint32)
    cls_index = {c: i for i, c in enumerate(accessibility_keys)}
    entry_class = np.asarray([cls_index.get(lu_class.get(k), -1) for k in keys], dtype=np.int32)
    a_node, a_ent, a_dist = [], [], []
    for v, pairs in node_data_map.items():
        for k, d in pairs:
            if k in index:
                a_node.append(int(v))
                a_ent.append(index[k])
                a_dist.append(float(d))
    a_node = np.asarray(a_node, dtype=np.uint32)
    a_ent = np.asarray(a_ent, dtype=np.uint32)
    a_dist = np.asarray(a_dist, dtype=np.float64)
    n, T, C = og.n, len(distances), len(accessibility_keys)
    cnt = np.zeros((C, T, n))
    wtd = np.zeros((C, T, n))
    near = np.full((C, n), np.nan)
    lv = np.asarray(live, dtype=np.uint8)
    rc = load().og_accessibility(
        og.h, len(keys), _p(entry_group, ctypes.c_uint32), _p(entry_class, ctypes.c_int32), C, len(a_node),
        _p(a_node, ctypes.c_uint32), _p(a_ent, ctypes.c_uint32), _p(a_dist, ctypes.c_double), T,
        _p(np.asarray(distances, dtype=np.uint32), ctypes.c_uint32), _p(np.asarray(betas, dtype=np.float32), ctypes.c_float),
        _p(np.asarray(mcws, dtype=np.float32), ctypes.c_float), int(max_walk_seconds), float(speed), int(bool(angular)),
        _p(lv, ctypes.c_uint8), int(n_threads), _p(cnt, ctypes.c_double), _p(wtd, ctypes.c_double),
        _p(near, ctypes.c_double))
    if rc != 0:
        raise ValueError(f"oracle error {rc}")
    return {c: (cnt[i], wtd[i], near[i]) for i, c in enumerate(accessibility_keys)}


# ---- diversity (rust/src/diversity.rs), f32 throughout, classes summed in order

def hill(counts, q):
    """hill_diversity_core (diversity.rs:6-72); errors read as 0 (data.rs:866-941 unwrap_or)."""
    num = int(sum(counts))
    if num == 0:
        return f32(0.0)
    if q == 0:
        return f32(sum(1 for c in counts if c > 0))
    numf = f32(num)
    if q == 1:
        ent = f32(0.0)
        for c in counts:
            if c > 0:
                p = f32(f32(c) / numf)
                ent = f32(ent + f32(p * np.log(p, dtype=np.float32)))
        r = np.exp(-ent, dtype=np.float32)
        return r if np.isfinite(r) else f32(0.0)
    s = f32(0.0)
    for c in counts:
        if c > 0:
            p = f32(f32(c) / numf)
            s = f32(s + np.power(p, f32(q), dtype=np.float32))
    r = np.power(s, f32(f32(1.0) / f32(1.0 - q)), dtype=np.float32)
    return r if np.isfinite(s) and np.isfinite(r) else f32(0.0)


def clipped_wt(beta, mcw, d):
    raw = np.exp(f32(-f32(beta)) * f32(d), dtype=np.float32)
    return f32(min(raw, f32(mcw)) / f32(mcw))


def hill_weighted(counts, nearest, q, beta, mcw):
    """hill_diversity_branch_distance_wt_core (diversity.rs:77-175)."""
    num = int(sum(counts))
    if num == 0:
        return f32(0.0)
    numf = f32(num)
    pw = [(f32(f32(c) / numf), clipped_wt(beta, mcw, d)) for c, d in zip(counts, nearest, strict=True) if c > 0]
    agg = f32(0.0)
    for p, w in pw:
        agg = f32(agg + f32(w * p))
    if abs(agg) < np.finfo(np.float32).eps:
        return f32(0.0)
    if q == 1:
        we = f32(0.0)
        for p, w in pw:
            e = f32(f32(w * p) / agg)
            if e > 0:
                we = f32(we + f32(e * np.log(e, dtype=np.float32)))
        r = np.exp(-we, dtype=np.float32)
        return r if np.isfinite(r) else f32(0.0)
    s = f32(0.0)
    for p, w in pw:
        s = f32(s + f32(w * np.power(f32(p / agg), f32(q), dtype=np.float32)))
    r = np.power(s, f32(f32(1.0) / f32(1.0 - q)), dtype=np.float32)
    return r if np.isfinite(s) and np.isfinite(r) else f32(0.0)


def shannon(counts):
    """shannon_diversity_core (diversity.rs:365-385)."""
    num = int(sum(counts))
    if num == 0:
        return f32(0.0)
    numf = f32(num)
    ent = f32(0.0)
    for c in counts:
        if c > 0:
            p = f32(f32(c) / numf)
            ent = f32(ent + f32(p * np.log(p, dtype=np.float32)))
    return f32(max(-ent, f32(0.0))) if np.isfinite(ent) else f32(0.0)


def gini(counts):
    """gini_simpson_diversity_core (diversity.rs:341-358)."""
    num = int(sum(counts))
    if num < 2:
        return f32(0.0)
    numf, nm1 = f32(num), f32(num - 1)
    lam = f32(0.0)
    for c in counts:
        if c > 0:
            x = f32(c)
            lam = f32(lam + f32(f32(x / numf) * f32(max(f32(x - f32(1.0)), f32(0.0)) / nm1)))
    return f32(max(f32(f32(1.0) - lam), f32(0.0)))


def mixed_uses(og, live, node_data_map, dedupe_of, lu_class, distances, betas, mcws, max_walk_seconds, speed,
               sources=None, angular=False):
    """f64 [8][T][N]: hill q0-2, weighted hill q0-2, shannon, gini (data.rs:818-941)."""
    n, T = og.n, len(distances)
    classes = sorted(set(lu_class.values()))
    ci = {c: i for i, c in enumerate(classes)}
    out = np.zeros((8, T, n))
    for src in (range(n) if sources is None else sources):
        if not live[src]:
            continue
        tree = _tree(og, src, max_walk_seconds, speed, angular)
        reach = aggregate_to_src_idx(tree, node_data_map, dedupe_of, max_walk_seconds, speed)
        counts = np.zeros((T, len(classes)), dtype=np.int64)
        nearest = np.full((T, len(classes)), np.inf, dtype=np.float32)
        for data_key, data_dist in reach.items():
            c = ci[lu_class[data_key]]
            for i, d in enumerate(distances):
                if data_dist <= f32(d):
                    counts[i, c] += 1
                    nearest[i, c] = min(nearest[i, c], data_dist)
        for i in range(T):
            cs, ns_ = counts[i].tolist(), nearest[i].tolist()
            vals = [hill(cs, 0), hill(cs, 1), hill(cs, 2)]
            vals += [hill_weighted(cs, ns_, q, betas[i], mcws[i]) for q in (0, 1, 2)]
            vals += [shannon(cs), gini(cs)]
            out[:, i, src] = [float(v) for v in vals]
    return out


# ---- statistics (rust/src/data.rs:946-1233), f32 throughout

def _f32_exp(x):
    """The reference's f32::exp (the host libm's expf, correctly rounded but for a sliver of
    inputs): the f64 exponential rounded to f32."""
    import math

    return f32(math.exp(float(x)))


def _total_key(x):
    """f32::total_cmp as an ascending integer key."""
    b = int(np.array([x], dtype=np.float32).view(np.uint32)[0])
    return (~b & 0xFFFFFFFF) if b & 0x80000000 else (b | 0x80000000)


def median(vals):
    """data.rs:946-959."""
    n = len(vals)
    if n == 0:
        return f32(np.nan)
    s = sorted(vals, key=_total_key)
    if n % 2 == 1:
        return s[n // 2]
    return f32(f32(s[n // 2 - 1] + s[n // 2]) / f32(2.0))


def weighted_median(pairs, total_wt):
    """data.rs:961-997 (Python's sort is stable, as Rust's sort_by)."""
    if not pairs:
        return f32(np.nan)
    s = sorted(pairs, key=lambda p: _total_key(p[0]))
    if total_wt == f32(0.0):
        return f32(np.nan)
    mid = f32(total_wt / f32(2.0))
    for v, w in s:
        if w > mid:
            return v
    agg = f32(0.0)
    for i, (v, w) in enumerate(s):
        agg = f32(agg + w)
        if agg == mid:
            return f32(f32(v + s[i + 1][0]) / f32(2.0)) if i + 1 < len(s) else v
        if agg > mid:
            return v
    return s[-1][0]


STAT_NAMES = ("sum", "sum_wt", "mean", "mean_wt", "median", "median_wt", "count", "count_wt", "variance",
              "variance_wt", "mad", "mad_wt", "max", "min")


def stats(og, live, node_data_map, dedupe_of, entry_order, num_maps, distances, betas, mcws, max_walk_seconds,
          speed, sources=None, angular=False):
    """f64 [maps][14][T][N] in STAT_NAMES order (data.rs:1073-1227). Each source's reachable entries
    are taken in ascending `entry_order` (the reference walks its hash map, whose order varies)."""
    n, T = og.n, len(distances)
    out = np.full((len(num_maps), 14, T, n), np.nan)
    out[:, [0, 1, 6, 7]] = 0.0
    for src in (range(n) if sources is None else sources):
        if not live[src]:
            continue
        tree = _tree(og, src, max_walk_seconds, speed, angular)
        reach = aggregate_to_src_idx(tree, node_data_map, dedupe_of, max_walk_seconds, speed)
        items = sorted(reach.items(), key=lambda kv: entry_order[kv[0]])
        for mp, num_map in enumerate(num_maps):
            for i, (d, b, mcw) in enumerate(zip(distances, betas, mcws, strict=True)):
                vals, pairs = [], []
                s = s_wt = cnt = cnt_wt = sq = sq_wt = f32(0.0)
                mn = mx = f32(np.nan)
                for key, dist in items:
                    if not dist <= f32(d):
                        continue
                    num = num_map.get(key)
                    if num is None or np.isnan(num):
                        continue
                    num = f32(num)
                    raw = _f32_exp(f32(-f32(b)) * dist)
                    wt = f32(min(raw, f32(mcw)) / f32(mcw))
                    s = f32(s + num)
                    s_wt = f32(s_wt + f32(num * wt))
                    cnt = f32(cnt + f32(1.0))
                    cnt_wt = f32(cnt_wt + wt)
                    sq = f32(sq + f32(num * num))
                    sq_wt = f32(sq_wt + f32(f32(wt * num) * num))
                    mx = num if np.isnan(mx) else max(mx, num)
                    mn = num if np.isnan(mn) else min(mn, num)
                    vals.append(num)
                    pairs.append((num, wt))
                mean = f32(s / cnt) if cnt > 0 else f32(np.nan)
                mean_wt = f32(s_wt / cnt_wt) if cnt_wt > 0 else f32(np.nan)
                var = max(f32(f32(sq / cnt) - f32(mean * mean)), f32(0.0)) if cnt > 0 else f32(np.nan)
                var_wt = (max(f32(f32(sq_wt / cnt_wt) - f32(mean_wt * mean_wt)), f32(0.0)) if cnt_wt > 0
                          else f32(np.nan))
                med = median(vals)
                med_wt = weighted_median(pairs, cnt_wt)
                mad = median([f32(abs(f32(v - med))) for v in vals]) if vals and not np.isnan(med) else f32(np.nan)
                mad_wt = (weighted_median([(f32(abs(f32(v - med_wt))), w) for v, w in pairs], cnt_wt)
                          if pairs and not np.isnan(med_wt) and cnt_wt > 0 else f32(np.nan))
                row = (s, s_wt, mean, mean_wt, med, med_wt, cnt, cnt_wt, var, var_wt, mad, mad_wt, mx, mn)
                for m, v in enumerate(row):
                    out[mp, m, i, src] = float(v)
    return out

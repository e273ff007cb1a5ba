"""Benchmark: source searches/sec (SSPS) of localised network centrality on one MI355X.

Default workload (``--mode plain``, the configuration the metric is quoted on): a
seeded 200x200 perturbed lattice (xy jitter +-20 m, 10% of edges dropped, 5%
diagonal shortcuts, largest component, f32 lengths = Euclidean x U(1.0, 1.3)),
node_centrality_shortest at [500, 1000, 2000, 5000] m, closeness and
betweenness, all ~40k sources. A step is one full centrality pass over every
source with the graph, sources and output accumulators resident in HBM
(cs_centrality_shortest_device); the timed region covers K steps between
device synchronisations.

``--mode simplest`` / ``--mode segment``: the dual (simplest-path, angular
impedance) and the primal (segment) of a 150x150 perturbed lattice at
[500, 1000, 2000] m through the NetworkStructure methods; ``value`` uses the
whole C call (graph resident; per-call tables in, rows out), with the kernel
time beside it.

``--mode access``: land-use accessibility (DataMap.accessibility) on the
200x200 lattice at [400, 800] m over about two synthetic data entries per node;
``value`` uses the whole C call. Nested legs: the same with one dedupe group of
2,000 members, and angular accessibility on the dual of a 150x150 lattice.

Prints one JSON line.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "cityseer-api_amd"), ROOT]

METRIC = "source searches/sec + p50/p99 request latency, batched/served vs plain GPU"
HBM_PEAK_GBS = 8000.0
DISTANCES = [500, 1000, 2000, 5000]


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def build_workload(side: int, diag: float, seed: int):
    from cityseer.tools import mock

    lat = mock.perturbed_lattice(side, seed=seed, diag_frac=diag)
    ns = mock.network_structure_from_lattice(lat)
    return lat, ns


def profiled_traffic(key: str = "traffic_bytes_per_launch", name: str = "k_search_brandes_cfg2_hbm.md"):
    """HBM bytes per launch of the dominant kernel on this workload, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (counters scaled by a plain-copy calibration); None if absent."""
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None


def cpu_threads() -> int:
    """Host threads this process may run on."""
    return len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)


def cpu_baseline(ns, target_s: float):
    """The oracle (a C restatement of the reference's CPU algorithm) on the host cores, on a
    bounded sample of sources of the same workload."""
    import numpy as np

    from oracle.oracle import OracleGraph

    og = OracleGraph.from_network_structure(ns)
    threads = cpu_threads()
    n = ns.node_count()
    rng = np.random.default_rng(123)

    def run(k):
        src = np.sort(rng.choice(n, k, replace=False)).astype(np.uint32)
        t0 = time.perf_counter()
        og.centrality_shortest(DISTANCES, sources=src, n_threads=threads)
        return time.perf_counter() - t0

    k0 = max(threads * 4, 64)
    t0 = run(k0)
    rate = k0 / max(t0, 1e-6)
    k = int(min(n, max(k0, rate * target_s)))
    t = run(k)
    return {
        "value": round(k / t, 2),
        "unit": "source searches/sec",
        "cores": threads,
        "kind": "port",
        "sample": f"{k} random sources of the same workload ({t:.1f} s wall, {threads} threads)",
    }


VARIANT_DISTANCES = [500, 1000, 2000]
VARIANT_SIDE = 150


def variant_cpu_baseline(ns, kind: str, target_s: float):
    """The oracle on the host cores over a bounded random sample of sources (simplest / segment)."""
    import numpy as np

    from oracle.oracle import OracleGraph

    og = OracleGraph.from_network_structure(ns)
    threads = cpu_threads()
    n = ns.node_count()
    rng = np.random.default_rng(321)

    def run(k):
        src = np.sort(rng.choice(n, k, replace=False)).astype(np.uint32)
        t0 = time.perf_counter()
        if kind == "simplest":
            og.centrality_simplest(VARIANT_DISTANCES, sources=src, n_threads=threads)
        else:
            og.segment_centrality(VARIANT_DISTANCES, sources=src, n_threads=threads)
        return time.perf_counter() - t0

    k0 = max(threads * 4, 64)
    rate = k0 / max(run(k0), 1e-6)
    k = int(min(n, max(k0, rate * target_s)))
    t = run(k)
    return {
        "value": round(k / t, 2),
        "unit": "source searches/sec",
        "cores": threads,
        "kind": "port",
        "sample": f"{k} random sources of the same workload ({t:.1f} s wall, {threads} threads)",
    }


def run_variant(args) -> int:
    """Simplest-path centrality on the dual, or segment centrality on the primal, of a 150x150
    perturbed lattice at [500, 1000, 2000] m, all sources."""
    import numpy as np

    from cityseer import _native
    from cityseer.tools import mock

    kind = args.mode
    log(f"building {kind} workload")
    lat = mock.perturbed_lattice(VARIANT_SIDE, seed=args.seed)
    if kind == "simplest":
        lat = mock.lattice_dual(lat)
    ns = mock.network_structure_from_lattice(lat)
    n, m = ns.node_count(), ns.edge_count

    def step():
        if kind == "simplest":
            ns.centrality_simplest(distances=VARIANT_DISTANCES)
        else:
            ns.segment_centrality(distances=VARIANT_DISTANCES)
        return _native.last_stats()

    for _ in range(args.warmup):
        step()
    kernel_ms, total_ms = [], []
    stats = None
    for i in range(args.steps):
        stats = step()
        kernel_ms.append(stats["kernel_ms"])
        total_ms.append(stats["total_ms"])
        log(f"step {i}: kernel {stats['kernel_ms']:.2f} ms, call {stats['total_ms']:.2f} ms")
    kms = float(np.mean(kernel_ms))
    # a step is the whole C call (host rows out; the graph stays uploaded)
    call_ms = float(np.mean(total_ms))
    metrics = 4
    algo_bytes = 12 * stats["edges_scanned"] + 16 * stats["nodes_reached"] + 16 * metrics * stats["within_sum"]
    achieved = algo_bytes / (kms / 1000.0) / 1e9
    cpu = None if args.no_cpu else variant_cpu_baseline(ns, kind, args.cpu_seconds)
    call = "node_centrality_simplest (dual, angular)" if kind == "simplest" else "segment_centrality (primal)"
    line = {
        "metric": METRIC,
        "value": round(n / (call_ms / 1000.0), 2),
        "unit": "source searches/sec",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(call_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 accumulators)",
        "data": "synthetic",
        "config": {
            "workload": f"{VARIANT_SIDE}x{VARIANT_SIDE} perturbed lattice, {call} {VARIANT_DISTANCES} m, "
                        "closeness+betweenness, all sources",
            "nodes": n,
            "directed_edges": m,
            "sources": n,
            "mode": kind,
            "call_ms": round(float(np.mean(total_ms)), 3),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": (profiled_traffic("simplest_traffic_bytes_per_launch", "k_angular_cfg3_dual_hbm.md")
                        if kind == "simplest" else None),
            "kernel_ms": round(kms, 3),
            "algorithmic_bytes": int(algo_bytes),
            "edges_scanned": int(stats["edges_scanned"]),
            "nodes_reached": int(stats["nodes_reached"]),
            "within_sum": int(stats["within_sum"]),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


BATCH_DISTANCES = [400, 800, 1600, 3200, 6400]
ACCESS_DISTANCES = [400, 800]
ACCESS_CLASSES = ["shop", "cafe", "park", "other"]


def access_workload(ns, lat, per_node: float, seed: int):
    """Synthetic data entries: about `per_node` per street node, each assigned to a random node and
    up to two of its neighbours (distances 0-60 m), 30 % sharing a dedupe key, four classes."""
    import numpy as np

    from cityseer.rustalgos import data

    rng = np.random.default_rng(seed)
    n = ns.node_count()
    nbrs: list[list[int]] = [[] for _ in range(n)]
    for a, b in lat.edges.tolist():
        nbrs[a].append(b)
        nbrs[b].append(a)
    n_entries = int(per_node * n)
    dm = data.DataMap()
    lu = {}
    for e in range(n_entries):
        key = f"e{e}"
        dd = f"g{int(rng.integers(0, n_entries // 4))}" if rng.random() < 0.3 else None
        dm.insert(key, "POINT (0 0)", dd)
        lu[key] = ACCESS_CLASSES[int(rng.integers(0, len(ACCESS_CLASSES)))]
        v = int(rng.integers(0, n))
        nodes = [v] + nbrs[v][: int(rng.integers(0, 3))]
        for u in nodes:
            dm.node_data_map.setdefault(u, []).append((f"str:{key}", float(rng.uniform(0.0, 60.0))))
    return dm, lu


def access_cpu_baseline(ns, dm, lu, target_s: float, angular: bool = False):
    """The oracle's C restatement of the aggregation (og_accessibility: the shortest-path tree,
    the simplest-path tree with `angular`, then aggregate_to_src_idx and the counts) on all host
    threads, over a bounded random sample of sources."""
    import numpy as np

    from oracle import oracle as O
    from oracle.access import accessibility_threads
    from oracle.oracle import OracleGraph

    O.set_heap_mode(False)
    og = OracleGraph.from_network_structure(ns)
    _d, bets, secs = O.thresholds(1.33333, distances=ACCESS_DISTANCES)
    ndm = {v: [(k[4:], d) for k, d in pairs] for v, pairs in dm.node_data_map.items()}
    dedupe_of = {k[4:]: e.dedupe_key[4:] for k, e in dm.entries.items()}
    rng = np.random.default_rng(321)
    n = ns.node_count()
    threads = cpu_threads()

    def run(k):
        live = np.zeros(n, dtype=np.uint8)
        live[rng.choice(n, k, replace=False)] = 1
        t0 = time.perf_counter()
        accessibility_threads(og, live, ndm, dedupe_of, lu, ACCESS_CLASSES[:3], ACCESS_DISTANCES, bets,
                              [1.0] * len(ACCESS_DISTANCES), max(secs), 1.33333, angular=angular, n_threads=threads)
        return time.perf_counter() - t0

    k0 = min(n, 64 * threads)
    rate = k0 / max(run(k0), 1e-6)
    k = int(min(n, max(k0, rate * target_s)))
    t = run(k)
    return {
        "value": round(k / t, 2),
        "unit": "source searches/sec",
        "cores": threads,
        "kind": "port",
        "sample": f"{k} random sources of the same workload ({t:.2f} s wall, {threads} threads, C restatement)",
    }


def run_access(args) -> int:
    """Land-use accessibility (DataMap.accessibility) on the 200x200 perturbed lattice at
    [400, 800] m: about two data entries per node, every live node a source."""
    import numpy as np

    from cityseer import _native
    from cityseer.tools import mock

    log("building access workload")
    lat = mock.perturbed_lattice(args.side, seed=args.seed, diag_frac=args.diag)
    ns = mock.network_structure_from_lattice(lat)
    n, m = ns.node_count(), ns.edge_count
    dm, lu = access_workload(ns, lat, 2.0, args.seed + 1)
    n_assign = sum(len(p) for p in dm.node_data_map.values())

    def timed(dmap, lus, label):
        for _ in range(args.warmup):
            dmap.accessibility(ns, lus, ACCESS_CLASSES[:3], distances=ACCESS_DISTANCES)
        kernel_ms, total_ms = [], []
        st = None
        for i in range(args.steps):
            dmap.accessibility(ns, lus, ACCESS_CLASSES[:3], distances=ACCESS_DISTANCES)
            st = _native.last_stats()
            kernel_ms.append(st["kernel_ms"])
            total_ms.append(st["total_ms"])
            log(f"{label} step {i}: kernel {st['kernel_ms']:.2f} ms, call {st['total_ms']:.2f} ms")
        return st, kernel_ms, total_ms

    stats, kernel_ms, total_ms = timed(dm, lu, "access")
    kms = float(np.mean(kernel_ms))
    # a step is the whole C call: host assignment tables in, device tables built,
    # the launch, the rows copied out (the Python layer's assembly of the tables
    # from its dicts is outside)
    call_ms = float(np.mean(total_ms))
    # search bytes only (the aggregation's reads of the assignment tables are not counted)
    algo_bytes = 12 * stats["edges_scanned"] + 16 * stats["nodes_reached"]
    achieved = algo_bytes / (kms / 1000.0) / 1e9
    cpu = None if args.no_cpu else access_cpu_baseline(ns, dm, lu, args.cpu_seconds)
    line = {
        "metric": METRIC,
        "value": round(n / (call_ms / 1000.0), 2),
        "unit": "source searches/sec",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(call_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 accumulators)",
        "data": "synthetic",
        "config": {
            "workload": f"{args.side}x{args.side} perturbed lattice, DataMap.accessibility {ACCESS_DISTANCES} m, "
                        f"{dm.count()} entries ({n_assign} node assignments), 3 of 4 classes, all sources",
            "nodes": n,
            "directed_edges": m,
            "sources": n,
            "mode": "access",
            "call_ms": round(float(np.mean(total_ms)), 3),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": profiled_traffic("traffic_bytes_per_launch", "k_search_access_cfg2_hbm.md"),
            "kernel_ms": round(kms, 3),
            "algorithmic_bytes": int(algo_bytes),
            "edges_scanned": int(stats["edges_scanned"]),
            "nodes_reached": int(stats["nodes_reached"]),
        },
        "cpu_baseline": cpu,
    }
    if args.access_group:
        # the same workload plus one dedupe group of many members spread over the
        # lattice (the vertices of one large feature): each group counts once per
        # source, through its nearest member
        rng = np.random.default_rng(args.seed + 2)
        for e in range(args.access_group):
            key = f"poly{e}"
            dm.insert(key, "POINT (0 0)", "poly")
            lu[key] = ACCESS_CLASSES[0]
            dm.node_data_map.setdefault(int(rng.integers(0, n)), []).append((f"str:{key}", float(rng.uniform(0.0, 40.0))))
        st, kg, tg = timed(dm, lu, f"access + group of {args.access_group}")
        line["large_group"] = {
            "workload": f"the same plus one dedupe group of {args.access_group} members at random nodes",
            "value": round(n / (float(np.mean(tg)) / 1000.0), 2),
            "unit": "source searches/sec",
            "ms_per_step": round(float(np.mean(tg)), 3),
            "kernel_ms": round(float(np.mean(kg)), 3),
            "kernel_over_plain": round(float(np.mean(kg)) / kms, 3),
        }
    if args.access_angular_side:
        # angular aggregation: simplest-path reach lists on the dual of a lattice,
        # then the same aggregation over them
        side = args.access_angular_side
        log(f"building angular access workload (dual of a {side}x{side} lattice)")
        dlat = mock.lattice_dual(mock.perturbed_lattice(side, seed=args.seed))
        dns = mock.network_structure_from_lattice(dlat)
        ddm, dlu = access_workload(dns, dlat, 2.0, args.seed + 3)
        dn = dns.node_count()

        def angular_call():
            ddm.accessibility(dns, dlu, ACCESS_CLASSES[:3], distances=ACCESS_DISTANCES, angular=True)
            return _native.last_stats()

        for _ in range(args.warmup):
            angular_call()
        ka, ta = [], []
        for i in range(args.steps):
            st = angular_call()
            ka.append(st["kernel_ms"])
            ta.append(st["total_ms"])
            log(f"angular access step {i}: kernels {st['kernel_ms']:.2f} ms, call {st['total_ms']:.2f} ms")
        line["angular"] = {
            "workload": f"dual of a {side}x{side} perturbed lattice, DataMap.accessibility(angular=True) "
                        f"{ACCESS_DISTANCES} m, {ddm.count()} entries, all {dn} sources",
            "value": round(dn / (float(np.mean(ta)) / 1000.0), 2),
            "unit": "source searches/sec",
            "ms_per_step": round(float(np.mean(ta)), 3),
            "kernel_ms": round(float(np.mean(ka)), 3),
            "cpu_baseline": None if args.no_cpu else access_cpu_baseline(dns, ddm, dlu, args.cpu_seconds / 2,
                                                                         angular=True),
        }
    print(json.dumps(line), flush=True)
    return 0


def batch_workload(n_graphs: int, seed: int = 0):
    """Graphs with node counts log-uniform in [1,000, 60,000] (side = round(sqrt(N))), seeds 0..n-1,
    each with 2-4 distances drawn from {400, 800, 1600, 3200, 6400}."""
    import numpy as np

    from cityseer.tools import mock

    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_graphs):
        n_target = float(np.exp(rng.uniform(np.log(1e3), np.log(6e4))))
        k = int(rng.integers(2, 5))
        dists = sorted(rng.choice(BATCH_DISTANCES, k, replace=False).tolist())
        side = max(3, int(round(np.sqrt(n_target))))
        lat = mock.perturbed_lattice(side, seed=i)
        out.append((mock.network_structure_from_lattice(lat), [int(d) for d in dists]))
    return out


def all_sources(ns, dev):
    """Every node as a source, listed in the graph's device node order
    (cs_graph_node_order): the same work as 0..N-1, with searches from
    neighbouring nodes running side by side. Unit weights and eligibility need
    no reordering."""
    import ctypes

    import numpy as np
    import torch

    from cityseer import _native

    order = np.empty(ns.node_count(), np.uint32)
    _native.check(_native.load().cs_graph_node_order(ns._native_handle(), _native.ptr(order, ctypes.c_uint32)))
    return torch.from_numpy(order.view(np.int32).copy()).to(dev)


def run_batch(args) -> int:
    print(json.dumps(batch_line(args)), flush=True)
    return 0


def batch_line(args) -> dict:
    """Many graphs of uneven size, each with its own thresholds: one autotuned persistent launch
    (hipGraph replay) against one plain device-resident launch per graph."""
    import numpy as np
    import torch

    from cityseer import _native, rustalgos
    from cityseer.batch import CentralityBatch

    log(f"building {args.graphs} graphs")
    work = batch_workload(args.graphs, args.seed)
    total_sources = sum(ns.node_count() for ns, _ in work)
    lib = _native.load()
    dev = torch.device("cuda:0")
    torch.cuda.init()
    stream = torch.cuda.current_stream(dev)

    # plain: one device-resident launch per graph
    plain = []
    for ns, dists in work:
        n = ns.node_count()
        d, b, s = rustalgos.pair_distances_betas_time(1.33333, distances=dists)
        t = len(d)
        bufs = dict(
            src=all_sources(ns, dev), wt=torch.ones(n, dtype=torch.float32, device=dev),
            el=torch.ones(n, dtype=torch.uint8, device=dev), acc=torch.zeros(n * 7 * t, dtype=torch.float64, device=dev),
            reach=torch.zeros(8, dtype=torch.int32, device=dev),
            d=np.asarray(d, np.uint32), b=np.asarray(b, np.float32), s=np.asarray(s, np.uint32))
        p = _native.CsParams()
        p.n_thresholds = t
        p.distances = _native.ptr(bufs["d"], ctypes.c_uint32)
        p.betas = _native.ptr(bufs["b"], ctypes.c_float)
        p.seconds = _native.ptr(bufs["s"], ctypes.c_uint32)
        p.speed_m_s = 1.33333
        p.tolerance = 1e-4
        p.compute_closeness = 1
        p.compute_betweenness = 1
        p.n_sources = n
        p.sources = ctypes.cast(bufs["src"].data_ptr(), ctypes.POINTER(ctypes.c_uint32))
        p.source_wt = ctypes.cast(bufs["wt"].data_ptr(), ctypes.POINTER(ctypes.c_float))
        p.eligible = ctypes.cast(bufs["el"].data_ptr(), ctypes.POINTER(ctypes.c_uint8))
        p.betw_scale = 1.0
        p.angular_scaling_unit = 180.0
        p.farness_scaling_offset = 1.0
        plain.append((ns._native_handle(), p, bufs))

    def plain_pass():
        kms = 0.0
        for h, p, bufs in plain:
            bufs["acc"].zero_()
            _native.check(lib.cs_centrality_shortest_device(h, ctypes.byref(p), bufs["wt"].data_ptr(),
                                                            bufs["acc"].data_ptr(), bufs["reach"].data_ptr(),
                                                            stream.cuda_stream))
            kms += _native.last_stats()["kernel_ms"]
        torch.cuda.synchronize()
        return kms

    log("plain warmup")
    plain_pass()
    t0 = time.perf_counter()
    plain_kms = 0.0
    for _ in range(args.steps):
        plain_kms += plain_pass()
    plain_s = (time.perf_counter() - t0) / args.steps
    log(f"plain: {plain_s * 1000:.1f} ms per pass over {len(work)} graphs")

    batch = CentralityBatch()
    for ns, dists in work:
        batch.add_shortest(ns, distances=dists)
    log("autotune")
    t0 = time.perf_counter()
    tune = batch.autotune(args.tune_fraction)
    tune_s = time.perf_counter() - t0
    log(f"chosen {tune} in {tune_s:.1f} s")
    for _ in range(max(args.warmup, 1)):
        batch.run()
    kernel_ms = []
    stats = None
    t0 = time.perf_counter()
    for i in range(args.steps):
        stats = batch.run()
        kernel_ms.append(stats["kernel_ms"])
        log(f"batched step {i}: {stats['kernel_ms']:.1f} ms kernel, {stats['total_ms']:.1f} ms call")
    batch_s = (time.perf_counter() - t0) / args.steps
    kms = float(np.mean(kernel_ms))
    algo_bytes = 12 * stats["edges_scanned"] + 16 * stats["nodes_reached"] + 16 * 7 * stats["within_sum"]
    achieved = algo_bytes / (kms / 1000.0) / 1e9
    value = total_sources / batch_s
    plain_value = total_sources / plain_s
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "source searches/sec",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(batch_s * 1000.0, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 accumulators)",
        "data": "synthetic",
        "config": {
            "workload": f"{len(work)} perturbed lattices, N log-uniform in [1000, 60000], 2-4 distances each "
                        f"from {BATCH_DISTANCES} m, node_centrality_shortest closeness+betweenness, all sources; "
                        "one autotuned persistent launch replayed from a hipGraph",
            "graphs": len(work),
            "nodes_total": total_sources,
            "sources": total_sources,
            "mode": "batch",
            "tuning": tune,
            "autotune_s": round(tune_s, 2),
            "plain_value": round(plain_value, 2),
            "plain_ms_per_pass": round(plain_s * 1000.0, 3),
            "plain_kernel_ms": round(plain_kms / args.steps, 3),
            "batched_over_plain": round(value / plain_value, 3),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel_ms": round(kms, 3),
            "algorithmic_bytes": int(algo_bytes),
            "edges_scanned": int(stats["edges_scanned"]),
            "nodes_reached": int(stats["nodes_reached"]),
            "within_sum": int(stats["within_sum"]),
        },
        "cpu_baseline": None,
    }
    return line


# thresholds of served requests (m) and the node visits one request may ask for at most
SERVE_DISTANCES = [400, 800, 1600, 3200, 6400, 10000]
SERVE_MAX_VISITS = {"shortest": 4e8, "simplest": 2.5e7}




def serve_workload(n_requests: int, seed: int = 0, max_nodes: float = 1e6):
    """Graphs decomposed into <= 20 m edges with node counts log-uniform in [2,000, max_nodes] (the
    three smallest also as duals for simplest-path requests); a stream of requests, each on a random
    graph, 80% shortest / 20% simplest, 1-3 distances from SERVE_DISTANCES and a random 1%-100%
    source subset, cut where needed so that sources x the nodes a source reaches at the request's
    largest distance (about 11 per 100 m x 100 m cell of the lattice) stays within
    SERVE_MAX_VISITS: long thresholds on large graphs take small subsets."""
    import numpy as np

    from cityseer.tools import mock

    rng = np.random.default_rng(seed)
    primal, dual = [], []
    for i in range(6):
        target = float(np.exp(np.log(2e3) + (np.log(max_nodes) - np.log(2e3)) * i / 5))
        side = max(4, int(round(np.sqrt(target / 11.0))))
        lat = mock.decompose_lattice(mock.perturbed_lattice(side, seed=100 + i), 20.0)
        primal.append(mock.network_structure_from_lattice(lat))
        if i < 3:
            dual.append(mock.network_structure_from_lattice(mock.lattice_dual(lat)))
    reqs = []
    for _ in range(n_requests):
        simplest = rng.random() < 0.2
        kind = "simplest" if simplest else "shortest"
        ns = dual[int(rng.integers(len(dual)))] if simplest else primal[int(rng.integers(len(primal)))]
        n = ns.node_count()
        nd = int(rng.integers(1, 4))
        dists = sorted(rng.choice(SERVE_DISTANCES, nd, replace=False).tolist())
        reach = min(float(n), 11.0 * np.pi * (max(dists) / 100.0) ** 2)
        frac = float(rng.uniform(0.01, 1.0))
        k = int(min(frac * n, SERVE_MAX_VISITS[kind] / reach))
        k = max(1, min(k, n))
        src = np.sort(rng.choice(n, k, replace=False)).tolist()
        reqs.append((kind, ns, dict(distances=[int(d) for d in dists], source_indices=src)))
    return primal + dual, reqs


def _fifo_latencies(arrivals, service):
    """Latency of each request in a single first-come first-served server."""
    free, lat = 0.0, []
    for a, s in zip(arrivals, service):
        start = max(a, free)
        free = start + s
        lat.append(free - a)
    return lat


def run_serve(args) -> int:
    print(json.dumps(serve_line(args)), flush=True)
    return 0


def serve_line(args) -> dict:
    """Continuous-batching server vs the plain one-call-per-request path on the same request stream:
    saturation throughput (all requests submitted at once) and p50/p99 latency under Poisson
    arrivals at a fraction of the served saturation rate."""
    import numpy as np

    from cityseer.serving import CentralityServer

    log("building serving workload")
    graphs, reqs = serve_workload(args.requests, args.seed, args.serve_max_nodes)
    total_sources = sum(len(kw["source_indices"]) for _, _, kw in reqs)
    for ns in graphs:  # device-resident graphs (and angular adjacency) before any timing
        if ns.is_dual:
            ns.centrality_simplest(distances=[400], source_indices=[0])
        else:
            ns.centrality_shortest(distances=[400], source_indices=[0])

    def plain_call(kind, ns, kw):
        return ns.centrality_shortest(**kw) if kind == "shortest" else ns.centrality_simplest(**kw)

    # the saturation pass is one burst whose launches interleave differently from run to run:
    # both sides are timed serve_passes times and the medians reported
    passes = []
    for rep in range(args.serve_passes):
        log(f"plain: one call per request, in order (pass {rep + 1}/{args.serve_passes})")
        service = []
        for kind, ns, kw in reqs:
            t0 = time.perf_counter()
            plain_call(kind, ns, kw)
            service.append(time.perf_counter() - t0)
        passes.append((float(np.sum(service)), service))
    passes.sort(key=lambda x: x[0])
    plain_s, service = passes[len(passes) // 2]

    with CentralityServer(slice_ms=args.slice_ms) as srv:
        sats = []
        for rep in range(args.serve_passes):
            log(f"served: all requests at once (saturation, pass {rep + 1}/{args.serve_passes})")
            t0 = time.perf_counter()
            handles = [(srv.submit_shortest if k == "shortest" else srv.submit_simplest)(ns, **kw) for k, ns, kw in reqs]
            for h in handles:
                h.wait()
            sats.append(time.perf_counter() - t0)
        sat_s = float(np.median(sats))
        sat_rate = len(reqs) / sat_s
        lam = args.load * sat_rate
        rng = np.random.default_rng(args.seed + 1)
        gaps = rng.exponential(1.0 / lam, len(reqs))
        arrivals = np.cumsum(gaps) - gaps[0]
        log(f"served: Poisson arrivals at {lam:.2f} requests/s ({args.load:.0%} of saturation)")
        handles = []
        t0 = time.perf_counter()
        for (k, ns, kw), at in zip(reqs, arrivals):
            delay = t0 + at - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            handles.append((srv.submit_shortest if k == "shortest" else srv.submit_simplest)(ns, **kw))
        infos = [h.wait() for h in handles]
        stream_s = time.perf_counter() - t0
    lat = np.array([i["latency_ms"] for i in infos])
    plain_lat = np.array(_fifo_latencies(arrivals, service)) * 1000.0
    launches = [i["launches"] for i in infos]
    value = total_sources / sat_s
    plain_value = total_sources / plain_s
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "source searches/sec",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": round(sat_s * 1000.0, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 accumulators)",
        "data": "synthetic",
        "config": {
            "workload": f"{len(reqs)} requests over {len(graphs)} graphs decomposed to <= 20 m edges "
                        f"({min(g.node_count() for g in graphs)}-{max(g.node_count() for g in graphs)} nodes), "
                        f"80% shortest / 20% simplest, 1-100% of sources (at most {SERVE_MAX_VISITS} node visits "
                        f"per request), 1-3 distances from {SERVE_DISTANCES} m; "
                        "continuous-batching server vs one plain call per request",
            "requests": len(reqs),
            "sources": total_sources,
            "mode": "serve",
            "plain_value": round(plain_value, 2),
            "served_over_plain": round(value / plain_value, 3),
            "served_s_passes": [round(x, 4) for x in sats],
            "plain_s_passes": sorted(round(p[0], 4) for p in passes),
            "arrival_rate_per_s": round(lam, 3),
            "p50_ms": round(float(np.percentile(lat, 50)), 2),
            "p99_ms": round(float(np.percentile(lat, 99)), 2),
            "plain_p50_ms": round(float(np.percentile(plain_lat, 50)), 2),
            "plain_p99_ms": round(float(np.percentile(plain_lat, 99)), 2),
            "stream_s": round(stream_s, 3),
            "mean_launches_per_request": round(float(np.mean(launches)), 2),
            "plain_s_by_kind": {k: round(float(sum(sv for (kk, _, _), sv in zip(reqs, service) if kk == k)), 3)
                                for k in ("shortest", "simplest")},
            "sources_by_kind": {k: int(sum(len(kw["source_indices"]) for kk, _, kw in reqs if kk == k))
                                for k in ("shortest", "simplest")},
        },
        "roofline": None,
        "cpu_baseline": None,
    }
    return line


def leg_summary(line: dict, ratio_key: str) -> dict:
    """The headline fields of a batch or serve line, nested in the plain line."""
    c = line["config"]
    out = {"workload": c["workload"], "value": line["value"], "unit": line["unit"],
           "ms_per_step": line["ms_per_step"], "sources": c["sources"], "plain_value": c["plain_value"],
           ratio_key: c[ratio_key]}
    for k in ("p50_ms", "p99_ms", "plain_p50_ms", "plain_p99_ms", "requests", "graphs", "tuning", "autotune_s"):
        if k in c:
            out[k] = c[k]
    if line.get("roofline"):
        out["roofline_frac"] = line["roofline"]["frac"]
        out["kernel_ms"] = line["roofline"]["kernel_ms"]
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--side", type=int, default=200)
    ap.add_argument("--diag", type=float, default=0.05)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mode", choices=["plain", "simplest", "segment", "batch", "serve", "access"], default="plain")
    ap.add_argument("--requests", type=int, default=40, help="serve mode: requests in the stream")
    ap.add_argument("--load", type=float, default=0.6,
                    help="serve mode: Poisson arrival rate as a fraction of the served saturation rate")
    ap.add_argument("--graphs", type=int, default=64, help="batch mode: number of graphs")
    ap.add_argument("--tune-fraction", type=float, default=0.03, help="batch mode: autotuner sample fraction")
    ap.add_argument("--serve-passes", type=int, default=3, help="serve mode: timed passes of each side")
    ap.add_argument("--serve-max-nodes", type=float, default=1e6, help="serve mode: largest graph")
    ap.add_argument("--slice-ms", type=float, default=0.0,
                    help="serve mode: target duration of one batched launch (0: the server's default)")
    ap.add_argument("--access-angular-side", type=int, default=150,
                    help="access mode: also time angular accessibility on the dual of a lattice this wide (0: no)")
    ap.add_argument("--access-group", type=int, default=2000,
                    help="access mode: also time the workload with one dedupe group of this many members (0: no)")
    ap.add_argument("--no-legs", action="store_true",
                    help="plain mode: skip the batched (64 graphs) and served legs nested in the line")
    ap.add_argument("--leg-steps", type=int, default=3, help="plain mode: timed passes of the nested legs")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device-to-device copy (3 times), to scale HBM counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (requested --gpus {args.gpus}); sharding sources over several "
              "GPUs is not a measured path.", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from cityseer import _native, rustalgos

    if args.copy_calibration:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": src.numel() * 8, "copies": 3}), flush=True)
        return 0

    if args.mode in ("simplest", "segment"):
        return run_variant(args)
    if args.mode == "batch":
        return run_batch(args)
    if args.mode == "serve":
        return run_serve(args)
    if args.mode == "access":
        return run_access(args)

    log("building workload")
    lat, ns = build_workload(args.side, args.diag, args.seed)
    n, m = ns.node_count(), ns.edge_count
    dists, betas, secs = rustalgos.pair_distances_betas_time(1.33333, distances=DISTANCES)
    t_count = len(dists)
    dev = torch.device("cuda:0")
    torch.cuda.init()
    d_src = all_sources(ns, dev)
    d_wt = torch.ones(n, dtype=torch.float32, device=dev)
    d_cyc = torch.ones(n, dtype=torch.float32, device=dev)
    d_elig = torch.ones(n, dtype=torch.uint8, device=dev)
    d_acc = torch.zeros(n * 7 * t_count, dtype=torch.float64, device=dev)
    d_reach = torch.zeros(8, dtype=torch.int32, device=dev)
    d = np.asarray(dists, np.uint32)
    b = np.asarray(betas, np.float32)
    s = np.asarray(secs, np.uint32)
    p = _native.CsParams()
    p.n_thresholds = t_count
    p.distances = _native.ptr(d, ctypes.c_uint32)
    p.betas = _native.ptr(b, ctypes.c_float)
    p.seconds = _native.ptr(s, ctypes.c_uint32)
    p.speed_m_s = 1.33333
    p.tolerance = 1e-4
    p.compute_closeness = 1
    p.compute_betweenness = 1
    p.n_sources = n
    p.sources = ctypes.cast(d_src.data_ptr(), ctypes.POINTER(ctypes.c_uint32))
    p.source_wt = ctypes.cast(d_wt.data_ptr(), ctypes.POINTER(ctypes.c_float))
    p.eligible = ctypes.cast(d_elig.data_ptr(), ctypes.POINTER(ctypes.c_uint8))
    p.track_reach = 0
    p.betw_scale = 1.0
    p.angular_scaling_unit = 180.0
    p.farness_scaling_offset = 1.0
    lib = _native.load()
    handle = ns._native_handle()
    stream = torch.cuda.current_stream(dev)

    def step():
        d_acc.zero_()
        _native.check(lib.cs_centrality_shortest_device(handle, ctypes.byref(p), d_cyc.data_ptr(),
                                                        d_acc.data_ptr(), d_reach.data_ptr(), stream.cuda_stream))
        return _native.last_stats()

    log(f"workload: N={n} directed E={m}; warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    kernel_ms = []
    stats = None
    log(f"timing {args.steps} steps")
    t0 = time.perf_counter()
    for i in range(args.steps):
        stats = step()
        kernel_ms.append(stats["kernel_ms"])
        log(f"step {i}: kernel {stats['kernel_ms']:.2f} ms")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1000.0 / args.steps
    value = n * args.steps / elapsed
    algo_bytes = 12 * stats["edges_scanned"] + 16 * stats["nodes_reached"] + 16 * 7 * stats["within_sum"]
    kms = float(np.mean(kernel_ms))
    achieved = algo_bytes / (kms / 1000.0) / 1e9
    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle, bounded sample)")
        cpu = cpu_baseline(ns, args.cpu_seconds)
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "source searches/sec",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 accumulators)",
        "data": "synthetic",
        "config": {
            "workload": f"plain GPU, one launch per request: {args.side}x{args.side} perturbed lattice "
                        f"+{int(args.diag * 100)}% diagonals, node_centrality_shortest {DISTANCES} m, "
                        "closeness+betweenness, all sources",
            "nodes": n,
            "directed_edges": m,
            "sources": n,
            "mode": "plain",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": profiled_traffic(),
            "kernel_ms": round(kms, 3),
            "algorithmic_bytes": int(algo_bytes),
            "edges_scanned": int(stats["edges_scanned"]),
            "nodes_reached": int(stats["nodes_reached"]),
            "within_sum": int(stats["within_sum"]),
            "teps": round(stats["edges_scanned"] / (kms / 1000.0), 1),
        },
        "cpu_baseline": cpu,
    }
    if not args.no_legs:
        # the metric's other half: batched and served throughput against their own plain paths
        legs = argparse.Namespace(**vars(args))
        legs.steps = args.leg_steps
        legs.warmup = 1
        legs.tune_fraction = 0.01
        legs.serve_passes = 3
        log("leg: batch (64 graphs)")
        line["batch"] = leg_summary(batch_line(legs), "batched_over_plain")
        line["batch"]["batched_over_plain_cfg2"] = round(line["batch"]["value"] / value, 3)
        log("leg: serve")
        line["serve"] = leg_summary(serve_line(legs), "served_over_plain")
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

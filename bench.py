"""Benchmark: source searches/sec of shortest-path centrality on one MI355X.

Workload (the configuration the metric is quoted on): a seeded 200x200
perturbed lattice (xy jitter +-20 m, 10% of edges dropped, 5% diagonal
shortcuts, largest component, f32 lengths = Euclidean x U(1.0, 1.3)),
node_centrality_shortest at [500, 1000, 2000, 5000] m, closeness and
betweenness, all ~40k sources.

A step is one full centrality pass over every source with the graph, sources
and output accumulators resident in HBM (cs_centrality_shortest_device); the
timed region covers K steps between device synchronisations. Prints one JSON
line.
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


def cpu_baseline(ns, target_s: float):
    """The oracle (a C restatement of the reference's CPU algorithm) on the host cores, on a
    bounded sample of sources of the same workload."""
    import numpy as np

    from oracle.oracle import OracleGraph

    og = OracleGraph.from_network_structure(ns)
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
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
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device-to-device copy (3 times), to scale HBM counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (requested --gpus {args.gpus}); the multi-GPU sharding is a "
              "design in DESIGN.md, not a measured path.", file=sys.stderr)
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

    log("building workload")
    lat, ns = build_workload(args.side, args.diag, args.seed)
    n, m = ns.node_count(), ns.edge_count
    dists, betas, secs = rustalgos.pair_distances_betas_time(1.33333, distances=DISTANCES)
    t_count = len(dists)
    dev = torch.device("cuda:0")
    torch.cuda.init()
    d_src = torch.arange(n, dtype=torch.int32, device=dev)
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
            "traffic": None,
            "kernel_ms": round(kms, 3),
            "algorithmic_bytes": int(algo_bytes),
            "edges_scanned": int(stats["edges_scanned"]),
            "nodes_reached": int(stats["nodes_reached"]),
            "within_sum": int(stats["within_sum"]),
            "teps": round(stats["edges_scanned"] / (kms / 1000.0), 1),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

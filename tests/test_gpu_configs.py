"""GPU: the benchmark configurations at their full sizes.

- 150x150 lattice: simplest-path centrality on its dual and segment centrality on the primal at
  [500, 1000, 2000] m; about 200 sampled sources against the oracle, then every source with
  size-independent checks (integer density, repeat runs agree, linearity in node weight).
- 64 graphs, node counts log-uniform in [1,000, 60,000], 2-4 distances each from
  {400, 800, 1600, 3200, 6400}: one batched launch equals one plain call per graph over all
  sources, and 32 sampled sources per graph equal the oracle.
- Serving on a lattice decomposed into <= 20 m edges (about a million nodes) and the dual of a
  smaller decomposed lattice: shortest and simplest requests on 0.1-100% source subsets and
  thresholds up to 10 km, each result equal to the same request alone, and sampled sources
  against the oracle.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from cityseer import _native
from cityseer.batch import CentralityBatch
from cityseer.serving import CentralityServer
from cityseer.tools import mock
import oracle.oracle as O
from oracle.oracle import OracleGraph
from parity import assert_close
from test_gpu_shortest import SEGMENT_KEYS, SHORTEST_KEYS, say, stack

pytestmark = pytest.mark.gpu

SIMPLEST_KEYS = ("node_density", "node_farness", "node_harmonic", "node_betweenness")
CFG3 = [500, 1000, 2000]


@pytest.fixture(autouse=True)
def _keyed():
    O.set_heap_mode(False)


def _sampled_vs_oracle(ns, og, kind, dists, sources, threads=8):
    call = ns.centrality_shortest if kind == "shortest" else ns.centrality_simplest
    res = call(distances=dists, source_indices=sources)
    plan = ns._sampling_plan(None, None, None, sources)
    ocall = og.centrality_shortest if kind == "shortest" else og.centrality_simplest
    ref, reach = ocall(dists, sources=plan["sources"], wt=plan["wt"], eligible=plan["eligible"], track=True,
                       scale=plan["scale"], n_threads=threads)
    got = stack(res, SHORTEST_KEYS if kind == "shortest" else SIMPLEST_KEYS, dists)
    assert np.array_equal(got[0], ref[0]), f"{kind}: density"
    assert_close(got, ref, f"{kind} sampled vs oracle")
    assert res.reachability_totals == reach.tolist()


def _segment_subset(ns, dists, sources):
    """cs_segment_centrality over a subset of sources (the Python method runs every node)."""
    from cityseer import rustalgos

    d, b, s = rustalgos.pair_distances_betas_time(1.33333, distances=dists)
    src = np.asarray(sources, np.uint32)
    plan = {"sources": src, "wt": np.ones(len(src), np.float32),
            "eligible": np.ones(ns.node_count(), np.uint8), "track": False, "sampled": 0, "scale": 1.0}
    params, _keep = ns._params(d, b, s, 1.33333, 1e-4, True, True, plan)
    out = np.zeros((4, len(d), ns.node_count()), np.float64)
    _native.check(_native.load().cs_segment_centrality(ns._native_handle(), ctypes.byref(params),
                                                        _native.ptr(out, ctypes.c_double)))
    return out


def test_cfg3_simplest_on_the_dual():
    dual = mock.lattice_dual(mock.perturbed_lattice(150, seed=0))
    ns = mock.network_structure_from_lattice(dual)
    og = OracleGraph.from_network_structure(ns)
    n = ns.node_count()
    src = np.sort(np.random.default_rng(301).choice(n, 200, replace=False)).tolist()
    say(f"cfg3 dual: {n} nodes, 200 sources vs the oracle")
    _sampled_vs_oracle(ns, og, "simplest", CFG3, src)
    say("cfg3 dual: all sources, twice, and with node weights x2")
    a = stack(ns.centrality_simplest(distances=CFG3), SIMPLEST_KEYS, CFG3)
    b = stack(ns.centrality_simplest(distances=CFG3), SIMPLEST_KEYS, CFG3)
    assert np.array_equal(a[0], np.round(a[0])) and np.array_equal(a[0], b[0])
    assert_close(b, a, "cfg3 simplest repeat", rtol=1e-12)
    ns2 = mock.network_structure_from_lattice(dual)
    ns2._weight.array()[:] = 2.0
    c = stack(ns2.centrality_simplest(distances=CFG3), SIMPLEST_KEYS, CFG3)
    assert_close(c, 2.0 * a, "cfg3 simplest, weights x2", rtol=1e-12)


def test_cfg3_segment_on_the_primal():
    lat = mock.perturbed_lattice(150, seed=0)
    ns = mock.network_structure_from_lattice(lat)
    og = OracleGraph.from_network_structure(ns)
    n = ns.node_count()
    src = np.sort(np.random.default_rng(302).choice(n, 200, replace=False))
    say(f"cfg3 primal: {n} nodes, segment, 200 sources vs the oracle")
    got = _segment_subset(ns, CFG3, src)
    ref = og.segment_centrality(CFG3, sources=src, n_threads=8)
    assert_close(got, ref, "cfg3 segment sampled vs oracle")
    say("cfg3 primal: segment over all sources, twice")
    a = stack(ns.segment_centrality(distances=CFG3), SEGMENT_KEYS, CFG3)
    b = stack(ns.segment_centrality(distances=CFG3), SEGMENT_KEYS, CFG3)
    assert_close(b, a, "cfg3 segment repeat", rtol=1e-12)
    # closeness at a source sums over its own search: the sampled rows equal the full run's
    assert_close(got[:3][:, :, src], a[:3][:, :, src], "segment closeness rows of the sampled sources")


def test_cfg4_batch_of_64_graphs():
    import bench

    work = bench.batch_workload(64, 0)
    total = sum(ns.node_count() for ns, _ in work)
    say(f"64 graphs, {total} sources: batched vs plain")
    batch = CentralityBatch()
    for ns, dists in work:
        batch.add_shortest(ns, distances=dists)
    batch.run()
    res = batch.results()
    rng = np.random.default_rng(303)
    for i, ((ns, dists), r) in enumerate(zip(work, res)):
        want = stack(ns.centrality_shortest(distances=dists), SHORTEST_KEYS, dists)
        got = stack(r, SHORTEST_KEYS, dists)
        assert np.array_equal(got[0], want[0]), f"graph {i}: density"
        assert_close(got, want, f"graph {i}: batch vs plain")
    say("32 sampled sources per graph vs the oracle")
    for i, (ns, dists) in enumerate(work):
        src = np.sort(rng.choice(ns.node_count(), 32, replace=False)).tolist()
        _sampled_vs_oracle(ns, OracleGraph.from_network_structure(ns), "shortest", dists, src)


def test_cfg5_serving_at_scale():
    big = mock.network_structure_from_lattice(
        mock.decompose_lattice(mock.perturbed_lattice(300, seed=304), 20.0))
    dual = mock.network_structure_from_lattice(
        mock.lattice_dual(mock.decompose_lattice(mock.perturbed_lattice(30, seed=305), 20.0)))
    n, nd = big.node_count(), dual.node_count()
    assert n > 900_000
    rng = np.random.default_rng(306)

    def sub(count, size):
        return np.sort(rng.choice(size, count, replace=False)).tolist()

    reqs = [
        ("shortest", big, dict(distances=[400], source_indices=list(range(n)))),  # 100%
        ("shortest", big, dict(distances=[400, 800], source_indices=sub(n // 100, n))),  # 1%
        ("shortest", big, dict(distances=[2000, 10000], source_indices=sub(n // 1000, n))),  # 0.1%
        ("simplest", dual, dict(distances=[400, 1000], source_indices=sub(nd // 10, nd))),  # 10%
        ("simplest", dual, dict(distances=[2000, 10000], source_indices=sub(nd // 100, nd))),  # 1%
    ]
    say(f"serving: {n}-node decomposed lattice and a {nd}-node dual, {len(reqs)} requests")
    alone = [(ns.centrality_shortest if k == "shortest" else ns.centrality_simplest)(**kw) for k, ns, kw in reqs]
    with CentralityServer(slice_ms=20.0) as srv:
        handles = [(srv.submit_shortest if k == "shortest" else srv.submit_simplest)(ns, **kw)
                   for k, ns, kw in reqs]
        for i, ((k, ns, kw), h, a) in enumerate(zip(reqs, handles, alone)):
            h.wait(timeout=600)
            keys = SHORTEST_KEYS if k == "shortest" else SIMPLEST_KEYS
            g, w = stack(h.result(), keys, kw["distances"]), stack(a, keys, kw["distances"])
            assert np.array_equal(g[0], w[0]), f"request {i}: density"
            assert_close(g, w, f"request {i} served vs alone")
            assert h.result().reachability_totals == a.reachability_totals
    say("sampled sources vs the oracle at full scale")
    _sampled_vs_oracle(big, OracleGraph.from_network_structure(big), "shortest", [2000, 10000], sub(8, n))
    _sampled_vs_oracle(dual, OracleGraph.from_network_structure(dual), "simplest", [2000, 10000], sub(16, nd))

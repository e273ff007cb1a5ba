"""GPU: batched shortest-path centrality (one persistent launch captured in a hipGraph) equals the
plain one-call-per-request path and the oracle, on every replay and after autotuning."""

from __future__ import annotations

import numpy as np
import pytest

from cityseer.batch import CentralityBatch
from cityseer.tools import io, mock
from oracle.oracle import OracleGraph
import oracle.oracle as O
from test_gpu_shortest import assert_close, say, stack

pytestmark = pytest.mark.gpu

KEYS = ("node_density", "node_farness", "node_cycles", "node_harmonic", "node_beta", "node_betweenness",
        "node_betweenness_beta")


@pytest.fixture(autouse=True)
def _keyed():
    O.set_heap_mode(False)


def _requests(primal_graph):
    _n, _e, mock_ns = io.network_structure_from_nx(primal_graph)
    small = mock.network_structure_from_lattice(mock.perturbed_lattice(14, seed=3))
    mid = mock.network_structure_from_lattice(mock.perturbed_lattice(24, seed=4, diag_frac=0.05))
    sub = np.random.default_rng(9).choice(mid.node_count(), 60, replace=False).tolist()
    return [
        (small, dict(distances=[400, 800])),
        (mid, dict(distances=[300, 600, 1200, 2400])),
        (mock_ns, dict(distances=[500, 2000, 5000], compute_betweenness=False)),
        (mid, dict(distances=[800], source_indices=sub)),
        (small, dict(distances=[250, 500, 1000], compute_closeness=False)),
    ]


def _check_equal(got, want, dists, what):
    g = stack(got, KEYS, dists)
    w = stack(want, KEYS, dists)
    assert np.array_equal(g[0], w[0]), f"{what}: density (reachable counts) must match bit-exactly"
    assert_close(g, w, what)


def test_batch_matches_plain_and_replays(primal_graph):
    reqs = _requests(primal_graph)
    say(f"batch of {len(reqs)} jobs over 3 graphs")
    b = CentralityBatch()
    for ns, kw in reqs:
        b.add_shortest(ns, **kw)
    plain = [ns.centrality_shortest(**kw) for ns, kw in reqs]
    for rep in range(3):
        say(f"replay {rep}")
        st = b.run()
        assert st["sources"] > 0
        res = b.results()
        for i, ((ns, kw), r, p) in enumerate(zip(reqs, res, plain)):
            _check_equal(r, p, kw["distances"], f"job {i} replay {rep}")
            assert r.reachability_totals == p.reachability_totals


def test_batch_matches_oracle():
    say("batch job vs oracle")
    lat = mock.perturbed_lattice(18, seed=11, diag_frac=0.05)
    ns = mock.network_structure_from_lattice(lat)
    og = OracleGraph.from_network_structure(ns)
    b = CentralityBatch()
    b.add_shortest(ns, distances=[400, 900])
    b.add_shortest(ns, distances=[1500])
    b.run()
    r0, r1 = b.results()
    for r, d in ((r0, [400, 900]), (r1, [1500])):
        ref, _reach = og.centrality_shortest(d)
        got = stack(r, KEYS, d)
        assert np.array_equal(got[0], ref[0])
        assert_close(got, ref, f"batch vs oracle {d}")


def test_autotune_keeps_results(primal_graph):
    reqs = _requests(primal_graph)[:3]
    b = CentralityBatch()
    for ns, kw in reqs:
        b.add_shortest(ns, **kw)
    b.run()
    before = b.results()
    say("autotune on a 20% sample")
    tune = b.autotune(0.2)
    assert tune["block_threads"] in (64, 128, 256)
    assert 0.0 < tune["delta_scale"] <= 1.0
    b.run()
    after = b.results()
    for (ns, kw), x, y in zip(reqs, before, after):
        _check_equal(y, x, kw["distances"], f"after autotune {tune}")


def test_explicit_tunings_agree():
    ns = mock.network_structure_from_lattice(mock.perturbed_lattice(16, seed=5))
    ref = ns.centrality_shortest(distances=[500, 1000])
    for blk, wps, ds in ((64, 2, 1.0), (128, 0, 0.75), (256, 3, 0.5)):
        say(f"tuning block={blk} waves/simd={wps} delta_scale={ds}")
        b = CentralityBatch()
        b.add_shortest(ns, distances=[500, 1000])
        b.set_tuning(blk, wps, ds)
        b.run()
        _check_equal(b.results()[0], ref, [500, 1000], f"tuning {blk}/{wps}/{ds}")


def test_batch_request_with_ten_thresholds():
    """A request of ten thresholds runs as two jobs (8 + 2) in the same launch, each searching to
    the request's largest cutoff, and fills one [7][10][N] result: equal to the plain call (which
    chunks the same way) and, for the chunk boundary's neighbours, to the oracle."""
    lat = mock.perturbed_lattice(16, seed=21, diag_frac=0.05)
    ns = mock.network_structure_from_lattice(lat)
    ten = [150, 300, 450, 600, 750, 900, 1050, 1200, 1350, 1500]
    b = CentralityBatch()
    assert b.add_shortest(ns, distances=[500]) == 0
    assert b.add_shortest(ns, distances=ten) == 1
    assert b.add_shortest(ns, distances=[250, 700], compute_closeness=False) == 2
    plain = [ns.centrality_shortest(distances=[500]), ns.centrality_shortest(distances=ten),
             ns.centrality_shortest(distances=[250, 700], compute_closeness=False)]
    for rep in range(2):
        say(f"ten-threshold batch, replay {rep}")
        b.run()
        res = b.results()
        assert len(res) == 3
        for i, (r, p, d) in enumerate(zip(res, plain, ([500], ten, [250, 700]))):
            _check_equal(r, p, d, f"job {i} replay {rep}")
            assert r.reachability_totals == p.reachability_totals
    og = OracleGraph.from_network_structure(ns)
    ref, _reach = og.centrality_shortest(ten)
    got = stack(res[1], KEYS, ten)
    assert np.array_equal(got[0], ref[0])
    assert_close(got, ref, "ten thresholds in a batch vs oracle")

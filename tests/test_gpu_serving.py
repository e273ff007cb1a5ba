"""GPU: the continuous-batching request server returns, for every request, the result the same
request gets alone through the plain call (density bit-exact, other metrics within 1e-5 relative),
whatever else shares its launches."""

from __future__ import annotations

import numpy as np
import pytest

from cityseer.serving import CentralityServer
from cityseer.tools import io, mock
import oracle.oracle as O
from test_gpu_shortest import assert_close, say, stack

pytestmark = pytest.mark.gpu

SHORTEST = ("node_density", "node_farness", "node_cycles", "node_harmonic", "node_beta", "node_betweenness",
            "node_betweenness_beta")
SIMPLEST = ("node_density", "node_farness", "node_harmonic", "node_betweenness")


@pytest.fixture(autouse=True)
def _keyed():
    O.set_heap_mode(False)


def _graphs(primal_graph):
    _n, _e, mock_ns = io.network_structure_from_nx(primal_graph)
    a = mock.network_structure_from_lattice(mock.perturbed_lattice(22, seed=21))
    b = mock.network_structure_from_lattice(mock.perturbed_lattice(30, seed=22, diag_frac=0.05))
    dual = mock.network_structure_from_lattice(mock.lattice_dual(mock.perturbed_lattice(12, seed=23)))
    return mock_ns, a, b, dual


def _mix(primal_graph):
    mock_ns, a, b, dual = _graphs(primal_graph)
    rng = np.random.default_rng(31)
    sub_b = rng.choice(b.node_count(), 90, replace=False).tolist()
    sub_d = rng.choice(dual.node_count(), 50, replace=False).tolist()
    return [
        ("shortest", a, dict(distances=[400, 800, 1600])),
        ("simplest", dual, dict(distances=[500, 1000])),
        ("shortest", b, dict(distances=[600], source_indices=sub_b)),
        ("shortest", mock_ns, dict(distances=[500, 5000], compute_betweenness=False)),
        ("shortest", b, dict(distances=[300, 900, 2700, 4000])),
        ("simplest", dual, dict(distances=[800], source_indices=sub_d, compute_closeness=False)),
        ("shortest", a, dict(distances=[1200], compute_closeness=False)),
    ]


def _alone(kind, ns, kw):
    return ns.centrality_shortest(**kw) if kind == "shortest" else ns.centrality_simplest(**kw)


def _same(got, want, kind, dists, what):
    keys = SHORTEST if kind == "shortest" else SIMPLEST
    g, w = stack(got, keys, dists), stack(want, keys, dists)
    assert np.array_equal(g[0], w[0]), f"{what}: density (reachable counts) must match bit-exactly"
    assert_close(g, w, what)
    assert got.reachability_totals == want.reachability_totals


def test_mix_equals_alone(primal_graph):
    mix = _mix(primal_graph)
    alone = [_alone(k, ns, kw) for k, ns, kw in mix]
    say(f"submitting {len(mix)} requests at once (slice 2 ms)")
    with CentralityServer(slice_ms=2.0) as srv:
        reqs = [(srv.submit_shortest if k == "shortest" else srv.submit_simplest)(ns, **kw) for k, ns, kw in mix]
        for i, ((k, ns, kw), r, a) in enumerate(zip(mix, reqs, alone)):
            info = r.wait()
            assert info["status"] == 2, info
            assert info["launches"] >= 1
            _same(r.result(), a, k, kw["distances"], f"request {i} ({k})")


def test_requests_span_several_launches_and_budget_queues():
    # more sources than the smallest slice (4096), so a request needs several launches
    big = mock.network_structure_from_lattice(mock.perturbed_lattice(100, seed=24))
    n = big.node_count()
    assert n > 2 * 4096
    want = big.centrality_shortest(distances=[500, 1000])
    one = 7 * 2 * n * 8 + n * 13 + 64
    say(f"{n} nodes; budget for one request at a time, slices of ~1 ms")
    with CentralityServer(hbm_budget_bytes=int(one * 1.5), max_active=4, slice_ms=1.0) as srv:
        reqs = [srv.submit_shortest(big, distances=[500, 1000]) for _ in range(3)]
        infos = [r.wait() for r in reqs]
        for r in reqs:
            _same(r.result(), want, "shortest", [500, 1000], f"queued request {r.id}")
        assert min(i["launches"] for i in infos) > 1, infos
        # one at a time: each later request waits for the one before it
        assert infos[1]["queued_ms"] > 0.0 and infos[2]["queued_ms"] >= infos[1]["queued_ms"]
        with pytest.raises(MemoryError):
            srv.submit_shortest(big, distances=[500, 1000, 2000, 4000, 8000])


def test_arrivals_during_running_launches(primal_graph):
    _m, a, b, dual = _graphs(primal_graph)
    with CentralityServer(slice_ms=3.0) as srv:
        first = srv.submit_shortest(b, distances=[3000])
        late = []
        for i in range(6):
            late.append((i, srv.submit_shortest(a, distances=[400 + 100 * i])))
        late.append((99, srv.submit_simplest(dual, distances=[700])))
        _same(first.result(), b.centrality_shortest(distances=[3000]), "shortest", [3000], "first")
        for i, r in late:
            if i == 99:
                _same(r.result(), dual.centrality_simplest(distances=[700]), "simplest", [700], "late simplest")
            else:
                d = [400 + 100 * i]
                _same(r.result(), a.centrality_shortest(distances=d), "shortest", d, f"late {i}")


def test_requests_with_ten_thresholds(primal_graph):
    """Served requests of more than 8 thresholds run one job per threshold chunk in each of their
    slices (each chunk searching to the request's largest cutoff) and equal the plain call, which
    chunks the same way, beside requests of one and three thresholds."""
    _mock_ns, a, _b, dual = _graphs(primal_graph)
    ten = [150, 300, 450, 600, 750, 900, 1050, 1200, 1350, 1500]
    mix = [
        ("shortest", a, dict(distances=ten)),
        ("simplest", dual, dict(distances=ten)),
        ("shortest", a, dict(distances=[400, 800, 1600])),
        ("simplest", dual, dict(distances=[700])),
    ]
    alone = [_alone(k, ns, kw) for k, ns, kw in mix]
    say("serving two ten-threshold requests in a mix of four (slice 2 ms)")
    with CentralityServer(slice_ms=2.0) as srv:
        reqs = [(srv.submit_shortest if k == "shortest" else srv.submit_simplest)(ns, **kw) for k, ns, kw in mix]
        for i, ((k, ns, kw), r, w) in enumerate(zip(mix, reqs, alone)):
            info = r.wait()
            assert info["status"] == 2, info
            _same(r.result(), w, k, kw["distances"], f"request {i} ({k}, {len(kw['distances'])} thresholds)")

"""GPU: the request server's failure handling, forced from the host (no kernel is made to fail):

- a slice whose host code throws a non-library exception (std::runtime_error, std::bad_alloc)
  fails the requests in that slice with a clear error, and the server keeps serving;
- a worker that ends on an exception outside the per-request handling fails its active and queued
  requests, and later submissions of its kind fail at once; the other kind keeps working;
- a request can be cancelled while queued or running, and waited for with a time limit.
No wait may block forever: every wait here carries a time limit.
"""

from __future__ import annotations

import numpy as np
import pytest

from cityseer import _native
from cityseer.serving import CentralityServer
from cityseer.tools import mock
from test_gpu_shortest import say, stack

pytestmark = pytest.mark.gpu

KEYS = ("node_density", "node_farness", "node_cycles", "node_harmonic", "node_beta", "node_betweenness",
        "node_betweenness_beta")
LIMIT = 60.0  # seconds: far above any request here


@pytest.fixture(scope="module")
def graphs():
    a = mock.network_structure_from_lattice(mock.perturbed_lattice(20, seed=41))
    big = mock.network_structure_from_lattice(mock.perturbed_lattice(120, seed=42))
    dual = mock.network_structure_from_lattice(mock.lattice_dual(mock.perturbed_lattice(10, seed=43)))
    return a, big, dual


def test_slice_exceptions_fail_their_requests_and_serving_goes_on(graphs):
    a, _big, _dual = graphs
    lib = _native.load()
    want = stack(a.centrality_shortest(distances=[500]), KEYS, [500])
    with CentralityServer() as srv:
        for what, exc in ((1, RuntimeError), (2, MemoryError)):
            say(f"slice failure {what}")
            _native.check(lib.cs_server_inject_fault(srv._h, _native.CS_REQ_SHORTEST, what))
            r = srv.submit_shortest(a, distances=[500])
            with pytest.raises(exc):
                r.wait(timeout=LIMIT)
            ok = srv.submit_shortest(a, distances=[500])
            ok.wait(timeout=LIMIT)
            got = stack(ok.result(), KEYS, [500])
            assert np.array_equal(got[0], want[0])


def test_worker_death_fails_its_requests_and_later_submissions(graphs):
    a, _big, dual = graphs
    lib = _native.load()
    with CentralityServer() as srv:
        say("shortest worker ends on an exception")
        _native.check(lib.cs_server_inject_fault(srv._h, _native.CS_REQ_SHORTEST, 3))
        r = srv.submit_shortest(a, distances=[500])
        with pytest.raises(RuntimeError, match="worker failed"):
            r.wait(timeout=LIMIT)
        with pytest.raises(RuntimeError, match="worker failed"):
            srv.submit_shortest(a, distances=[500])
        say("the simplest worker still serves")
        s = srv.submit_simplest(dual, distances=[600])
        s.wait(timeout=LIMIT)
        want = dual.centrality_simplest(distances=[600])
        assert np.array_equal(s.result().node_density[600], want.node_density[600])


def test_cancel_and_timed_wait(graphs):
    a, big, _dual = graphs
    n = big.node_count()
    # one request at a time, so the second queues behind the first
    one = 7 * 4 * n * 8 + n * 13 + 64
    with CentralityServer(hbm_budget_bytes=int(one * 1.5), slice_ms=1.0) as srv:
        say("a long request and one queued behind it")
        first = srv.submit_shortest(big, distances=[1000, 2000, 4000, 8000])
        queued = srv.submit_shortest(big, distances=[1000, 2000, 4000, 8000])
        with pytest.raises(TimeoutError):
            queued.wait(timeout=0.0)
        queued.cancel()
        with pytest.raises(_native.RequestCancelled):
            queued.wait(timeout=LIMIT)
        say("cancel the running one")
        first.cancel()
        try:
            first.wait(timeout=LIMIT)
            cancelled = False  # it had already finished
        except _native.RequestCancelled:
            cancelled = True
        info = first.info()
        assert info["status"] in (2, -_native.CS_ECANCELED) and cancelled == (info["status"] < 0)
        say("the server serves on after the cancellations")
        r = srv.submit_shortest(a, distances=[500])
        r.wait(timeout=LIMIT)
        want = a.centrality_shortest(distances=[500])
        assert np.array_equal(r.result().node_density[500], want.node_density[500])

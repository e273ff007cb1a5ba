"""GPU: NetworkStructure.progress() advances while a call runs (graph.rs:407-415; the
reference's rayon loop bumps it per source, centrality.rs:1703-1706, and config.wrap_progress polls
it from another thread). The kernels keep per-wave (per-lane in the angular search) counts of
finished sources in pinned host memory; progress() sums them."""

from __future__ import annotations

import threading
import time

import numpy as np
import pytest

from cityseer.tools import mock
from test_gpu_shortest import say

pytestmark = pytest.mark.gpu


def _poll_during(ns, call):
    seen = []
    out = {}

    def run():
        out["res"] = call()

    t = threading.Thread(target=run)
    t.start()
    while t.is_alive():
        seen.append(ns.progress())
        time.sleep(0.0002)
    t.join()
    seen.append(ns.progress())
    return seen, out["res"]


@pytest.mark.parametrize("kind", ["shortest", "simplest"])
def test_progress_advances_during_a_call(kind):
    if kind == "shortest":
        ns = mock.network_structure_from_lattice(mock.perturbed_lattice(200, seed=0, diag_frac=0.05))
        dists = list(range(1000, 10001, 1000))  # two launches of five-kilometre-plus searches
        call = lambda: ns.centrality_shortest(distances=dists)  # noqa: E731
    else:
        ns = mock.network_structure_from_lattice(mock.lattice_dual(mock.perturbed_lattice(120, seed=3)))
        call = lambda: ns.centrality_simplest(distances=[500, 1000, 2000])  # noqa: E731
    n = ns.node_count()
    call()  # first call stages the graph on the device
    say(f"{kind}: {n} sources, polling progress()")
    seen, _res = _poll_during(ns, call)
    seen = np.asarray(seen)
    # polls before the call starts still read the previous call's final count
    first = int(np.argmax(seen < n)) if np.any(seen < n) else len(seen)
    seen = seen[first:]
    mid = seen[(seen > 0) & (seen < n)]
    say(f"{len(seen)} polls, {len(np.unique(mid))} distinct values strictly between 0 and {n}")
    assert seen[-1] == n
    assert len(np.unique(mid)) >= 2, "progress did not advance while the call ran"
    assert np.all(np.diff(seen) >= 0), "progress went backwards"

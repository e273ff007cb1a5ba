"""GPU: the largest graphs the serving configuration names (about a million nodes), on a sample
of sources, against the oracle: index arithmetic, workspace sizing and the device node order at
full scale."""

from __future__ import annotations

import numpy as np
import pytest

from cityseer.tools import mock
import oracle.oracle as O
from oracle.oracle import OracleGraph
from test_gpu_shortest import SHORTEST_KEYS, assert_close, say, stack
from test_gpu_simplest import KEYS as SIMPLEST_KEYS

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _keyed():
    O.set_heap_mode(False)


def _sampled(ns, og, kind, dists, sources):
    res = (ns.centrality_shortest if kind == "shortest" else ns.centrality_simplest)(distances=dists,
                                                                                    source_indices=sources)
    plan = ns._sampling_plan(None, None, None, sources)
    call = og.centrality_shortest if kind == "shortest" else og.centrality_simplest
    ref, reach = call(dists, sources=plan["sources"], wt=plan["wt"], eligible=plan["eligible"], track=True,
                      scale=plan["scale"])
    got = stack(res, SHORTEST_KEYS if kind == "shortest" else SIMPLEST_KEYS, dists)
    assert np.array_equal(got[0], ref[0]), f"{kind}: density"
    assert_close(got, ref, f"large {kind}")
    assert res.reachability_totals == reach.tolist()


def test_million_node_primal():
    lat = mock.perturbed_lattice(1000, seed=111, diag_frac=0.02)
    ns = mock.network_structure_from_lattice(lat)
    n = ns.node_count()
    assert n > 900_000
    og = OracleGraph.from_network_structure(ns)
    src = np.sort(np.random.default_rng(112).choice(n, 192, replace=False)).tolist()
    say(f"{n} nodes, 192 sources, shortest [800, 3000]")
    _sampled(ns, og, "shortest", [800, 3000], src)


def test_large_dual():
    dual = mock.lattice_dual(mock.perturbed_lattice(260, seed=113))
    ns = mock.network_structure_from_lattice(dual)
    n = ns.node_count()
    assert n > 100_000
    og = OracleGraph.from_network_structure(ns)
    src = np.sort(np.random.default_rng(114).choice(n, 96, replace=False)).tolist()
    say(f"dual with {n} nodes, 96 sources, simplest [400, 1000]")
    _sampled(ns, og, "simplest", [400, 1000], src)

"""GPU: sources whose search outgrows the per-wave (per-lane, for simplest paths) workspace are
re-run with a larger one and give the same results. A tiny slot capacity forces the retry path of
the plain shortest call, the plain simplest call and a served slice that packs several simplest
requests (padded item lists)."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from cityseer import _native
from cityseer.serving import CentralityServer
from cityseer.tools import mock
import oracle.oracle as O
from oracle.oracle import OracleGraph
from test_gpu_shortest import assert_close, say, stack

pytestmark = pytest.mark.gpu

SHORTEST = ("node_density", "node_farness", "node_cycles", "node_harmonic", "node_beta", "node_betweenness",
            "node_betweenness_beta")
SIMPLEST = ("node_density", "node_farness", "node_harmonic", "node_betweenness")


@pytest.fixture(autouse=True)
def _keyed():
    O.set_heap_mode(False)


@pytest.fixture
def tiny_slots():
    lib = _native.load()
    old = _native.CsTuning()
    _native.check(lib.cs_get_tuning(ctypes.byref(old)))
    t = _native.CsTuning(old.block_threads, old.waves_per_simd, old.delta_scale, 48, old.xcd_split)
    _native.check(lib.cs_set_tuning(ctypes.byref(t)))
    yield
    _native.check(lib.cs_set_tuning(ctypes.byref(old)))


def test_shortest_retries_match_oracle(tiny_slots):
    ns = mock.network_structure_from_lattice(mock.perturbed_lattice(18, seed=51, diag_frac=0.1))
    og = OracleGraph.from_network_structure(ns)
    d = [300, 900]
    say("shortest with 48 slots per wave")
    res = ns.centrality_shortest(distances=d)
    assert _native.last_stats()["retries"] > 0
    ref, _ = og.centrality_shortest(d)
    got = stack(res, SHORTEST, d)
    assert np.array_equal(got[0], ref[0])
    assert_close(got, ref, "shortest after retries")


def test_simplest_retries_match_oracle_plain_and_served(tiny_slots):
    a = mock.network_structure_from_lattice(mock.lattice_dual(mock.perturbed_lattice(11, seed=52)))
    b = mock.network_structure_from_lattice(mock.lattice_dual(mock.perturbed_lattice(9, seed=53)))
    d = [400, 900]
    refs = []
    for ns in (a, b):
        ref, _ = OracleGraph.from_network_structure(ns).centrality_simplest(d)
        refs.append(ref)
    say("simplest with 48 states per lane")
    got = stack(a.centrality_simplest(distances=d), SIMPLEST, d)
    assert _native.last_stats()["retries"] > 0
    assert np.array_equal(got[0], refs[0][0])
    assert_close(got, refs[0], "simplest after retries")
    say("served simplest slice of two requests, 48 states per lane")
    with CentralityServer() as srv:
        ra = srv.submit_simplest(a, distances=d)
        rb = srv.submit_simplest(b, distances=d)
        for r, ref, what in ((ra, refs[0], "served a"), (rb, refs[1], "served b")):
            g = stack(r.result(), SIMPLEST, d)
            assert np.array_equal(g[0], ref[0]), what
            assert_close(g, ref, what)

"""GPU: OD-weighted betweenness (rust/src/centrality.rs:2413-2540) against the oracle, whose
restatement test_od.py pins; and against the GPU's own all-pairs betweenness through the
all-pairs-at-0.5 identity."""

from __future__ import annotations

import numpy as np
import pandas as pd
import pytest

from cityseer import rustalgos
from cityseer.metrics import networks
from cityseer.tools import mock
import oracle.oracle as O
from oracle.oracle import OracleGraph
from test_gpu_shortest import assert_close, say

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _keyed():
    O.set_heap_mode(False)


def _stack(res, dists):
    return np.stack([np.stack([res.node_betweenness[d] for d in dists]),
                     np.stack([res.node_betweenness_beta[d] for d in dists])])


@pytest.mark.parametrize("pct", [None, 5.0])
def test_od_matches_oracle(pct):
    lat = mock.perturbed_lattice(24, seed=62, diag_frac=0.08)
    ns = mock.network_structure_from_lattice(lat)
    ns._live.array()[7] = False  # a non-live origin is skipped
    og = OracleGraph.from_network_structure(ns)
    n = ns.node_count()
    rng = np.random.default_rng(63)
    o = rng.integers(0, n, 2000)
    o[:40] = 7
    d = rng.integers(0, n, 2000)
    w = rng.uniform(0.1, 5.0, 2000).astype(np.float32)
    o = np.concatenate([o, o[:25]])  # repeated pairs: the later weight wins
    d = np.concatenate([d, d[:25]])
    w = np.concatenate([w, np.full(25, 9.0, np.float32)])
    dist = [300, 800, 1600]
    say(f"OD betweenness, {len(o)} pairs, tolerance {pct}")
    od = rustalgos.centrality.OdMatrix(o.tolist(), d.tolist(), w.tolist())
    got = _stack(ns.betweenness_od_shortest(od, distances=dist, tolerance=pct), dist)
    ref = og.betweenness_od_shortest((o.tolist(), d.tolist(), w.tolist()), dist,
                                     tolerance=1e-4 if pct is None else pct / 100.0)
    assert ref.max() > 0
    assert_close(got, ref, "OD betweenness")


def test_od_all_pairs_half_equals_plain_betweenness():
    ns = mock.network_structure_from_lattice(mock.perturbed_lattice(16, seed=64, diag_frac=0.1))
    n = ns.node_count()
    o, d = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    keep = o != d
    od = rustalgos.centrality.OdMatrix(o[keep].tolist(), d[keep].tolist(), [0.5] * int(keep.sum()))
    dist = [400, 1000]
    say("OD all pairs at 0.5 vs plain betweenness")
    got = _stack(ns.betweenness_od_shortest(od, distances=dist), dist)
    plain = ns.centrality_shortest(distances=dist, compute_closeness=False)
    want = np.stack([np.stack([plain.node_betweenness[x] for x in dist]),
                     np.stack([plain.node_betweenness_beta[x] for x in dist])])
    assert_close(got, want, "all pairs at 0.5")


def test_reference_od_cases(primal_graph):
    """tests/test_sampling.py:678-722 of the reference: OD among the first five street nodes runs and
    gives non-negative, non-trivial betweenness; zero weights give zeros."""
    from cityseer.tools import io
    _n, _e, ns = io.network_structure_from_nx(primal_graph)
    idx = ns.street_node_indices()
    o, d, w = [], [], []
    for i in range(min(5, len(idx))):
        for j in range(min(5, len(idx))):
            if i != j:
                o.append(idx[i])
                d.append(idx[j])
                w.append(1.0)
    res = ns.betweenness_od_shortest(od_matrix=rustalgos.centrality.OdMatrix(o, d, w), distances=[500])
    betw = np.array(res.node_betweenness[500])
    assert len(betw) == len(idx) and np.all(betw >= 0) and np.any(betw > 0)
    zero = ns.betweenness_od_shortest(
        od_matrix=rustalgos.centrality.OdMatrix([idx[0], idx[1]], [idx[1], idx[2]], [0.0, 0.0]), distances=[500])
    assert np.allclose(np.array(zero.node_betweenness[500]), 0.0)


def test_networks_betweenness_od_columns():
    ns = mock.network_structure_from_lattice(mock.perturbed_lattice(10, seed=65))
    nodes = pd.DataFrame(index=ns.node_keys_py())
    od = rustalgos.centrality.OdMatrix([0, 5, 9], [50, 60, 99], [2.0, 1.0, 4.0])
    out = networks.betweenness_od(ns, nodes, od, distances=[600, 1200])
    for col in ("cc_betweenness_600", "cc_betweenness_1200", "cc_betweenness_beta_600", "cc_betweenness_beta_1200"):
        assert col in out.columns
    assert out["cc_betweenness_1200"].max() > 0

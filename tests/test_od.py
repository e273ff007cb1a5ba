"""OD-weighted betweenness (rust/src/centrality.rs:50-91, 2413-2540; networks.py:278-461), CPU side:
the OdMatrix surface (tests/test_sampling.py:666-722 of the reference), the oracle's restatement
pinned by identities that do not retype it, and the zone-snapping front-end."""

from __future__ import annotations

import numpy as np
import pandas as pd
import pytest

from cityseer import rustalgos
from cityseer.metrics import networks
from cityseer.tools import mock
import oracle.oracle as O
from oracle.oracle import OracleGraph


@pytest.fixture(autouse=True)
def _keyed():
    O.set_heap_mode(False)


def test_od_matrix_construction():
    od = rustalgos.centrality.OdMatrix([0, 0, 1], [1, 2, 2], [1.0, 2.0, 3.0])
    assert od.len() == 3
    assert od.n_origins() == 2
    with pytest.raises(ValueError, match="must have equal length"):
        rustalgos.centrality.OdMatrix([0, 1], [1], [1.0, 2.0])


def _lattice():
    ns = mock.network_structure_from_lattice(mock.perturbed_lattice(11, seed=61, diag_frac=0.1))
    return ns, OracleGraph.from_network_structure(ns)


def test_oracle_all_pairs_half_equals_brandes_betweenness():
    """Every ordered pair (s, t != s) at weight 0.5 is exactly the all-sources betweenness, whose
    seed is 0.5 for every eligible target (centrality.rs:1788-1811) and which is itself pinned to
    networkx in test_oracle.py."""
    ns, og = _lattice()
    n = ns.node_count()
    o, d = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    keep = o != d
    od = (o[keep].tolist(), d[keep].tolist(), [0.5] * int(keep.sum()))
    dist = [250, 600, 1100]
    got = og.betweenness_od_shortest(od, dist)
    ref, _ = og.centrality_shortest(dist, closeness=False)
    assert np.array_equal(got[0], ref[5])
    assert np.array_equal(got[1], ref[6])
    assert ref[5].max() > 0


def test_oracle_linear_in_weights_and_zero_weights():
    ns, og = _lattice()
    n = ns.node_count()
    rng = np.random.default_rng(7)
    o = rng.integers(0, n, 300).tolist()
    d = rng.integers(0, n, 300).tolist()
    w = rng.uniform(0.5, 4.0, 300).astype(np.float32)
    dist = [400, 900]
    a = og.betweenness_od_shortest((o, d, w.tolist()), dist)
    b = og.betweenness_od_shortest((o, d, (w * 4.0).tolist()), dist)
    assert np.allclose(b, 4.0 * a, rtol=1e-12, atol=0)
    assert a.max() > 0
    z = og.betweenness_od_shortest((o, d, [0.0] * 300), dist)
    assert not z.any()


def test_oracle_later_pair_overwrites():
    ns, og = _lattice()
    dist = [800]
    once = og.betweenness_od_shortest(([3, 3], [40, 77], [2.0, 1.5]), dist)
    twice = og.betweenness_od_shortest(([3, 3, 3], [40, 77, 40], [9.0, 1.5, 2.0]), dist)
    assert np.array_equal(once, twice)


def test_oracle_single_pair_is_the_path():
    """One OD pair on a tie-free lattice: the credit is the trip weight on every interior node of
    the unique shortest path (from the oracle's own tree), and nowhere else."""
    ns, og = _lattice()
    n = ns.node_count()
    src, dst = 0, n - 1
    dist = [5000]
    got = og.betweenness_od_shortest(([src], [dst], [3.0]), dist)[0][0]
    tree = og.tree_shortest(src, int(round(5000 / 1.33333)))
    path, v = [], dst
    while tree["pred"][v] >= 0 and v != src:
        v = int(tree["pred"][v])
        if v != src:
            path.append(v)
    want = np.zeros(n)
    want[path] = 3.0
    assert np.allclose(got, want, atol=1e-12)


def test_build_od_matrix_snaps_and_filters():
    ns, _og = _lattice()
    xy = np.asarray(ns.node_xys)
    zones = pd.DataFrame({"x": [xy[0, 0] + 5.0, xy[20, 0], xy[0, 0] + 1e6], "y": [xy[0, 1], xy[20, 1] - 3.0, 0.0]},
                         index=["a", "b", "far"])
    flows = pd.DataFrame({"o": ["a", "a", "b", "far", "b"], "d": ["b", "far", "a", "a", "b"],
                          "w": [5.0, 2.0, 0.0, 1.0, np.nan]})
    od = networks.build_od_matrix(flows, zones, ns, "o", "d", "w", max_snap_dist=50.0)
    assert od.len() == 1 and od.n_origins() == 1
    assert od.origins.tolist() == [0] and od.destinations.tolist() == [20] and od.weights.tolist() == [5.0]

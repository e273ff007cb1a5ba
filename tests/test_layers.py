"""CPU: the layers front-end's host side (pysrc/cityseer/metrics/layers.py:30-101 and the
argument checks of :104-756), the mock data generators it is tested with (tools/mock.py:314-403),
and DataEntry's keys and WKT validation (data.rs:306-377)."""

from __future__ import annotations

import math
import string

import numpy as np
import pandas as pd
import pytest

from cityseer.metrics import layers
from cityseer.rustalgos import data
from cityseer.tools import graphs, io, mock


@pytest.fixture(scope="module")
def primal():
    G = graphs.nx_simple_geoms(mock.mock_graph())
    nodes, _edges, ns = io.network_structure_from_nx(G)
    return G, nodes, ns


def _buffer(wkt, r=10.0, n=64):
    x, y = (float(v) for v in wkt[wkt.index("(") + 1:wkt.index(")")].split())
    ring = [(x + r * math.cos(2 * math.pi * k / n), y + r * math.sin(2 * math.pi * k / n)) for k in range(n)]
    ring.append(ring[0])
    return "POLYGON ((" + ", ".join(f"{a!r} {b!r}" for a, b in ring) + "))"


def test_build_data_map_keys_and_dedupe(primal):
    G, _nodes, ns = primal
    for typ in (int, float, str):
        for data_id_col in (None, "data_id"):
            df = mock.mock_data_gdf(G)
            df.index = df.index.astype(typ)
            for to_poly in (False, True):
                if to_poly:
                    df["geometry"] = [_buffer(w) for w in df["geometry"]]
                dm = layers.build_data_map(df, ns, max_netw_assign_dist=400, data_id_col=data_id_col)
                assert dm.count() == len(df)
                for key, entry in dm.entries.items():
                    row = df.loc[entry.data_key_py]
                    assert key == f"{typ.__name__}:{entry.data_key_py}"
                    assert entry.geom_wkt == row["geometry"]
                    if data_id_col is not None:
                        assert entry.dedupe_key_py == row[data_id_col]
                    else:
                        assert entry.dedupe_key_py == entry.data_key_py
                assigned = {k for pairs in dm.node_data_map.values() for k, _d in pairs}
                assert assigned and assigned <= set(dm.entries)


def test_build_data_map_geometry_objects_and_barriers(primal):
    G, _nodes, ns = primal

    class Geom:  # stands for a shapely geometry: only its ``wkt`` is read
        def __init__(self, wkt):
            self.wkt = wkt

    df = mock.mock_data_gdf(G)
    plain = layers.build_data_map(df, ns, 400)
    objs = df.copy()
    objs["geometry"] = [Geom(w) for w in df["geometry"]]
    assert layers.build_data_map(objs, ns, 400).node_data_map == plain.node_data_map
    bdf, _wkts = mock.mock_barriers()
    blocked = layers.build_data_map(df, ns, 400, barriers_gdf=bdf)
    assert blocked.node_data_map != plain.node_data_map
    # barriers hold for that call only
    assert layers.build_data_map(df, ns, 400).node_data_map == plain.node_data_map


def test_build_data_map_rejects_duplicate_index(primal):
    _G, _nodes, ns = primal
    df = pd.DataFrame({"data_idx": [1, 2, 2], "geometry": ["POINT (0 0)", "POINT (1 1)", "POINT (2 2)"]})
    with pytest.raises(ValueError, match="unique"):
        layers.build_data_map(df.set_index("data_idx"), ns, 400)


def test_bad_wkt_fails_at_insertion():
    dm = data.DataMap()
    with pytest.raises(ValueError, match="Failed to parse WKT for key 'int:3'"):
        dm.insert(3, "POINT (1 2")
    with pytest.raises(ValueError, match="Failed to parse WKT"):
        data.DataEntry("a", "TRIANGLE ((0 0, 1 0, 0 1, 0 0))")


def test_column_and_dual_checks(primal):
    G, nodes, ns = primal
    lu = mock.mock_landuse_categorical_data(G)
    num = mock.mock_numerical_data(G)
    with pytest.raises(ValueError, match="landuse column"):
        layers.compute_accessibilities(lu, "categorical_landuses-TYPO", ["c"], nodes.copy(), ns, 400, distances=[400])
    with pytest.raises(ValueError, match="landuse column"):
        layers.compute_mixed_uses(lu, "categorical_landuses-TYPO", nodes.copy(), ns, 400, distances=[400])
    with pytest.raises(ValueError, match="numerical stats column"):
        layers.compute_stats(num, ["mock_numerical_1", "typo"], nodes.copy(), ns, 400, distances=[400])
    for call in (lambda: layers.compute_accessibilities(lu, "categorical_landuses", ["a"], nodes, ns, angular=True),
                 lambda: layers.compute_mixed_uses(lu, "categorical_landuses", nodes, ns, angular=True),
                 lambda: layers.compute_stats(num, ["mock_numerical_1"], nodes, ns, angular=True)):
        with pytest.raises(ValueError, match="requires a dual graph"):
            call()


def test_mock_landuse_and_numerical_data(primal):
    G, _nodes, _ns = primal
    lu = mock.mock_landuse_categorical_data(G, length=60, num_classes=6, random_seed=3)
    assert list(lu["categorical_landuses"][-5:]) == ["z"] * 5
    assert set(lu["categorical_landuses"][:-5]) <= set(string.ascii_lowercase[:5])
    assert lu["geometry"].tolist() == mock.mock_data_gdf(G, length=60, random_seed=3)["geometry"].tolist()
    # the classes follow the coordinates in the seeded stream
    np.random.seed(3)
    np.random.uniform(size=120)
    want = [string.ascii_lowercase[int(np.random.randint(0, 5))] for _ in range(55)]
    assert list(lu["categorical_landuses"][:-5]) == want
    with pytest.raises(ValueError, match="exceeds"):
        mock.mock_landuse_categorical_data(G, num_classes=27)
    num = mock.mock_numerical_data(G, length=40, num_arrs=2, floating_pt=2, random_seed=4)
    for col in ("mock_numerical_1", "mock_numerical_2"):
        v = num[col].to_numpy()
        assert v.dtype == np.float32 and np.all(v[-5:] == v.max())
        assert np.all(v >= 0) and np.all(v < 1000.0)
        assert np.allclose(v * 100, np.round(v * 100), atol=1e-3)

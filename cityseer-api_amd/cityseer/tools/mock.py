"""Mock graphs and seeded synthetic street networks.

``mock_graph`` returns the reference's 57-node test graph (same node
coordinates and edge list as ``pysrc/cityseer/tools/mock.py:34-231``).

``perturbed_lattice``, ``lattice_dual`` and ``decompose_lattice`` generate the
seeded synthetic networks the benchmarks and parity tests run on: a jittered
square lattice with dropped edges and optional diagonal shortcuts (largest
connected component kept), its dual with turn-angle impedances, and its
decomposition into short segments. They return plain arrays (``Lattice``) so
that million-node graphs build in seconds; ``network_structure_from_lattice``
turns them into a ``NetworkStructure``.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import networkx as nx
import numpy as np

_MOCK_NODES = [
    (700700, 5719700), (700610, 5719780), (700460, 5719700), (700520, 5719820), (700620, 5719905),
    (700260, 5719700), (700320, 5719850), (700420, 5719880), (700460, 5719980), (700580, 5720030),
    (700100, 5719810), (700280, 5719980), (700400, 5720030), (700460, 5720130), (700190, 5720050),
    (700350, 5720200), (700800, 5719750), (700800, 5719920), (700900, 5719820), (700910, 5719690),
    (700905, 5720080), (701000, 5719870), (701040, 5719660), (701050, 5719760), (701000, 5719980),
    (701130, 5719950), (701130, 5719805), (701170, 5719700), (701100, 5720200), (701240, 5719990),
    (701300, 5719760), (700690, 5719590), (700570, 5719530), (700820, 5719500), (700700, 5719480),
    (700490, 5719440), (700580, 5719360), (700690, 5719370), (700920, 5719330), (700780, 5719300),
    (700680, 5719200), (700560, 5719280), (700450, 5719300), (700440, 5719150), (700650, 5719080),
    (700930, 5719110), (701015, 5719535), (701100, 5719480), (700917, 5719517), (700400, 5719550),
    (700700, 5720100), (700700, 5719900), (700400, 5719650), (700550, 5719550), (700410, 5719450),
    (700300, 5719550), (701300, 5719110),
]

_MOCK_EDGES = [
    (0, 1), (0, 16), (0, 31), (1, 2), (1, 4), (2, 3), (2, 5), (3, 4), (3, 7), (4, 9), (5, 6), (5, 10),
    (6, 7), (6, 11), (7, 8), (8, 9), (8, 12), (9, 13), (10, 14), (10, 43), (11, 12), (11, 14), (12, 13),
    (13, 15), (14, 15), (15, 28), (16, 17), (16, 19), (17, 18), (17, 20), (18, 19), (18, 21), (19, 22),
    (20, 24), (20, 28), (21, 23), (21, 24), (22, 23), (22, 27), (23, 26), (24, 25), (25, 26), (25, 29),
    (26, 27), (27, 30), (28, 29), (29, 30), (30, 45), (31, 32), (31, 33), (32, 34), (32, 35), (33, 34),
    (33, 38), (34, 37), (35, 36), (35, 42), (36, 37), (36, 41), (37, 39), (38, 39), (38, 45), (39, 40),
    (40, 41), (40, 44), (41, 42), (42, 43), (43, 44), (44, 45), (22, 46), (46, 47), (46, 48), (50, 51),
    (52, 53), (53, 54), (54, 55), (55, 52), (45, 56), (30, 56),
]


def mock_graph() -> nx.MultiGraph:
    """The reference's 57-node mock graph (mock.py:34-231), projected coordinates (EPSG:32630)."""
    g = nx.MultiGraph()
    for i, (x, y) in enumerate(_MOCK_NODES):
        g.add_node(str(i), x=x, y=y)
    for a, b in _MOCK_EDGES:
        g.add_edge(str(a), str(b))
    g.graph["crs"] = 32630
    return g


@dataclass
class Lattice:
    """Array form of an undirected street network.

    ``edges`` holds undirected node pairs; ``length`` one f32 length per pair,
    shared by both directions. Dual lattices also carry ``angle`` (f32 turn
    angle in degrees) and ``primal_node`` (the primal node each dual edge
    turns at).
    """

    x: np.ndarray
    y: np.ndarray
    edges: np.ndarray  # (E, 2) int64
    length: np.ndarray  # (E,) float32
    angle: np.ndarray | None = None  # (E,) float32 (dual only)
    primal_node: np.ndarray | None = None  # (E,) int64 (dual only)
    is_dual: bool = False
    meta: dict = field(default_factory=dict)

    @property
    def n_nodes(self) -> int:
        return int(self.x.shape[0])

    @property
    def n_edges(self) -> int:
        return int(self.edges.shape[0])

    def directed(self):
        """Directed edge arrays in insertion order: each pair as (a->b, b->a)."""
        e = self.edges
        src = np.empty(2 * len(e), dtype=np.uint32)
        dst = np.empty(2 * len(e), dtype=np.uint32)
        src[0::2], dst[0::2] = e[:, 0], e[:, 1]
        src[1::2], dst[1::2] = e[:, 1], e[:, 0]
        length = np.repeat(self.length.astype(np.float32), 2)
        angle = None if self.angle is None else np.repeat(self.angle.astype(np.float32), 2)
        pk = None if self.primal_node is None else np.repeat(self.primal_node.astype(np.int32), 2)
        return src, dst, length, angle, pk


def get_graph_extents(nx_multigraph: nx.MultiGraph) -> tuple[float, float, float, float]:
    """(min x, min y, max x, max y) of the graph's nodes."""
    xs = [d["x"] for _, d in nx_multigraph.nodes(data=True)]
    ys = [d["y"] for _, d in nx_multigraph.nodes(data=True)]
    return min(xs), min(ys), max(xs), max(ys)


def mock_data_gdf(nx_multigraph: nx.MultiGraph, length: int = 50, random_seed: int = 0):
    """Mock data points within the graph's extents (mock.py:274-311 of the reference): a
    ``pandas.DataFrame`` indexed by ``uid`` with WKT ``geometry`` and ``data_id`` columns
    (GeoPandas is not required). The points come from numpy's legacy global generator seeded
    with ``random_seed``, x then y, as the reference draws them; the last five share one
    ``data_id`` and sit 10 m apart on a diagonal, for deduplication checks."""
    import pandas as pd

    np.random.seed(seed=random_seed)
    min_x, min_y, max_x, max_y = get_graph_extents(nx_multigraph)
    xs = np.random.uniform(min_x, max_x, length)
    ys = np.random.uniform(min_y, max_y, length)
    pts = [(float(x), float(y)) for x, y in zip(xs, ys, strict=True)]
    data_id = list(range(length))
    for idx, loc in enumerate(range(length - 5, length)):
        data_id[loc] = length - 5
        pts[loc] = (700100.0 + idx * 10, 5719100.0 + idx * 10)
    df = pd.DataFrame({"uid": list(range(length)), "geometry": [f"POINT ({x!r} {y!r})" for x, y in pts],
                       "data_id": data_id})
    return df.set_index("uid")


def mock_landuse_categorical_data(nx_multigraph: nx.MultiGraph, length: int = 50, num_classes: int = 10,
                                  random_seed: int = 0):
    """``mock_data_gdf``'s points with a ``categorical_landuses`` column (mock.py:314-359 of the
    reference): letters drawn from the first ``num_classes - 1`` of the alphabet by the same
    seeded generator, after the coordinates; the last five rows (one dedupe group) are "z"."""
    import string

    letters = list(string.ascii_lowercase)
    if num_classes > len(letters):
        raise ValueError(
            f"The requested {num_classes} classes exceeds max available categorical classes of {len(letters)}"
        )
    np.random.seed(seed=random_seed)
    df = mock_data_gdf(nx_multigraph, length=length, random_seed=random_seed)
    letters = letters[: num_classes - 1]
    codes = ["z" if i >= length - 5 else letters[int(np.random.randint(0, len(letters)))] for i in range(len(df))]
    df["categorical_landuses"] = codes
    return df


def mock_numerical_data(nx_multigraph: nx.MultiGraph, length: int = 50, val_min: int = 0, val_max: int = 100000,
                        num_arrs: int = 1, floating_pt: int = 3, random_seed: int = 0):
    """``mock_data_gdf``'s points with ``mock_numerical_1`` .. ``mock_numerical_<num_arrs>``
    columns (mock.py:362-403 of the reference): f32 integers in [val_min, val_max) from the same
    seeded generator, divided by 10**floating_pt; the last five rows take the column's maximum."""
    np.random.seed(seed=random_seed)
    df = mock_data_gdf(nx_multigraph, length=length, random_seed=random_seed)
    for i in range(1, num_arrs + 1):
        arr = np.array(np.random.randint(val_min, high=val_max, size=length), dtype=np.float32)
        arr /= 10**floating_pt
        arr[-5:] = np.nanmax(arr)
        df[f"mock_numerical_{i}"] = arr
    return df


def mock_species_data(random_seed: int = 0):
    """Counts and probabilities of species classes for the diversity measures (mock.py:413-473
    of the reference): for n = 1, 6, ..., 46, n integers drawn in [1, 10) by numpy's legacy seeded
    generator, grouped by value."""
    np.random.seed(seed=random_seed)
    for n in range(1, 50, 5):
        draws = np.random.randint(1, 10, n)
        _vals, counts = np.unique(draws, return_counts=True)
        yield counts.astype(np.int_), counts / len(draws)


def mock_data_map(data_gdf, dedupe_key_col: str | None = None):
    """A ``DataMap`` of a frame like ``mock_data_gdf``'s (mock.py:540-560 of the reference);
    geometries may be WKT strings or objects with a ``wkt`` attribute."""
    from ..rustalgos import data

    dm = data.DataMap()
    for uid, row in data_gdf.iterrows():
        geom = row["geometry"]
        wkt = geom if isinstance(geom, str) else geom.wkt
        dm.insert(uid, wkt, row[dedupe_key_col] if dedupe_key_col is not None else None)
    return dm


def mock_barriers():
    """Mock barriers (mock.py:563-611 of the reference): a point, a line, a multi-line and a
    polygon, as a ``pandas.DataFrame`` of WKT and the list of their WKT."""
    import pandas as pd

    wkts = [
        "POINT (700900 5719350)",
        "LINESTRING (701020 5719150, 701250 5719600)",
        "MULTILINESTRING ((700400 5719150, 700150 5719600))",
        "POLYGON ((700900 5719140, 700900 5719160, 700750 5719160, 700750 5719140, 700900 5719140))",
    ]
    return pd.DataFrame({"geometry": wkts}), wkts


def _largest_component(n: int, edges: np.ndarray) -> np.ndarray:
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components

    if len(edges) == 0:
        return np.zeros(n, dtype=bool)
    m = coo_matrix((np.ones(len(edges)), (edges[:, 0], edges[:, 1])), shape=(n, n))
    _, labels = connected_components(m, directed=False)
    counts = np.bincount(labels)
    return labels == int(np.argmax(counts))


def perturbed_lattice(
    side: int,
    spacing: float = 100.0,
    jitter: float = 20.0,
    drop_frac: float = 0.10,
    diag_frac: float = 0.0,
    len_range: tuple[float, float] = (1.0, 1.3),
    seed: int = 0,
) -> Lattice:
    """Seeded side x side lattice with jittered nodes, dropped edges and diagonal shortcuts.

    Node xy = grid position +- U(-jitter, jitter); ``drop_frac`` of lattice
    edges removed; ``diag_frac`` of cells gain one diagonal; the largest
    connected component is kept; edge length = f32(Euclidean x U(len_range)).
    """
    rng = np.random.default_rng(seed)
    n = side * side
    ii, jj = np.divmod(np.arange(n), side)
    x = jj * spacing + rng.uniform(-jitter, jitter, n)
    y = ii * spacing + rng.uniform(-jitter, jitter, n)
    idx = np.arange(n).reshape(side, side)
    horiz = np.stack([idx[:, :-1].ravel(), idx[:, 1:].ravel()], axis=1)
    vert = np.stack([idx[:-1, :].ravel(), idx[1:, :].ravel()], axis=1)
    lat_edges = np.concatenate([horiz, vert])
    n_drop = int(round(drop_frac * len(lat_edges)))
    keep = np.ones(len(lat_edges), dtype=bool)
    if n_drop:
        keep[rng.choice(len(lat_edges), n_drop, replace=False)] = False
    lat_edges = lat_edges[keep]
    cells = idx[:-1, :-1].ravel()
    n_diag = int(round(diag_frac * len(cells)))
    if n_diag:
        pick = np.sort(rng.choice(len(cells), n_diag, replace=False))
        c = cells[pick]
        flip = rng.integers(0, 2, n_diag).astype(bool)
        a = np.where(flip, c + 1, c)
        b = np.where(flip, c + side, c + side + 1)
        lat_edges = np.concatenate([lat_edges, np.stack([a, b], axis=1)])
    mult = rng.uniform(len_range[0], len_range[1], len(lat_edges))
    keep_nodes = _largest_component(n, lat_edges)
    remap = -np.ones(n, dtype=np.int64)
    remap[keep_nodes] = np.arange(int(keep_nodes.sum()))
    emask = keep_nodes[lat_edges[:, 0]] & keep_nodes[lat_edges[:, 1]]
    e = lat_edges[emask]
    mult = mult[emask]
    euclid = np.hypot(x[e[:, 1]] - x[e[:, 0]], y[e[:, 1]] - y[e[:, 0]])
    length = (euclid * mult).astype(np.float32)
    return Lattice(
        x=x[keep_nodes],
        y=y[keep_nodes],
        edges=remap[e],
        length=length,
        meta={"side": side, "seed": seed, "spacing": spacing, "diag_frac": diag_frac},
    )


def lattice_dual(lat: Lattice) -> Lattice:
    """Dual of a primal lattice: one node per primal edge, one edge per pair of edges sharing a node.

    Dual edge length = f32(0.5 * len_a + 0.5 * len_b) and angle = f32 turn angle
    in degrees at the shared primal node between the two edge midpoints, as the
    welded half-geometries of ``graphs.nx_to_dual`` give (graphs.py:1965-2149).
    """
    from .util import measure_coords_angle

    e = lat.edges
    mx = 0.5 * (lat.x[e[:, 0]] + lat.x[e[:, 1]])
    my = 0.5 * (lat.y[e[:, 0]] + lat.y[e[:, 1]])
    # incident edge lists per primal node, in edge order
    order = np.argsort(np.concatenate([e[:, 0], e[:, 1]]), kind="stable")
    ends = np.concatenate([e[:, 0], e[:, 1]])[order]
    eids = np.concatenate([np.arange(len(e)), np.arange(len(e))])[order]
    bounds = np.searchsorted(ends, np.arange(lat.n_nodes + 1))
    d_edges, d_len, d_ang, d_pk = [], [], [], []
    len64 = lat.length.astype(np.float64)
    for p in range(lat.n_nodes):
        inc = eids[bounds[p] : bounds[p + 1]]
        px, py = lat.x[p], lat.y[p]
        for i in range(len(inc)):
            for j in range(i + 1, len(inc)):
                a, b = int(inc[i]), int(inc[j])
                d_edges.append((a, b))
                d_len.append(0.5 * len64[a] + 0.5 * len64[b])
                d_ang.append(measure_coords_angle((mx[a], my[a]), (px, py), (mx[b], my[b])))
                d_pk.append(p)
    return Lattice(
        x=mx,
        y=my,
        edges=np.asarray(d_edges, dtype=np.int64).reshape(-1, 2),
        length=np.asarray(d_len, dtype=np.float64).astype(np.float32),
        angle=np.asarray(d_ang, dtype=np.float64).astype(np.float32),
        primal_node=np.asarray(d_pk, dtype=np.int64),
        is_dual=True,
        meta=dict(lat.meta, dual=True),
    )


def decompose_lattice(lat: Lattice, decompose_max: float = 20.0) -> Lattice:
    """Split every edge into ceil(len / decompose_max) equal pieces (graphs.py:1829-1962).

    New nodes sit at evenly spaced points on the straight segment; each piece
    gets length f32(len / cuts).
    """
    e = lat.edges
    length = lat.length.astype(np.float64)
    cuts = np.maximum(np.ceil(length / decompose_max).astype(np.int64), 1)
    n0 = lat.n_nodes
    n_new = int((cuts - 1).sum())
    xs = [lat.x, np.empty(n_new)]
    ys = [lat.y, np.empty(n_new)]
    new_x, new_y = xs[1], ys[1]
    # vectorised construction
    starts = np.concatenate([[0], np.cumsum(cuts - 1)])[:-1]
    rep = np.repeat(np.arange(len(e)), cuts - 1)
    k = np.arange(n_new) - np.repeat(starts, cuts - 1) + 1
    t = k / np.repeat(cuts, cuts - 1)
    a, b = e[rep, 0], e[rep, 1]
    new_x[:] = lat.x[a] + (lat.x[b] - lat.x[a]) * t
    new_y[:] = lat.y[a] + (lat.y[b] - lat.y[a]) * t
    # chain per edge: a -> new_1 -> ... -> new_{c-1} -> b
    piece = (length / cuts).astype(np.float32)
    src_list, dst_list, len_list = [], [], []
    node_of_new = n0 + np.arange(n_new)
    new_ptr = n0 + starts
    c1 = cuts == 1
    src_list.append(e[c1, 0])
    dst_list.append(e[c1, 1])
    len_list.append(piece[c1])
    multi = np.nonzero(~c1)[0]
    if len(multi):
        cm = cuts[multi]
        # first piece a -> first new
        src_list.append(e[multi, 0])
        dst_list.append(new_ptr[multi])
        len_list.append(piece[multi])
        # interior pieces new_k -> new_{k+1}
        inner_counts = cm - 2
        if inner_counts.sum():
            base = np.repeat(new_ptr[multi], inner_counts)
            off = np.arange(int(inner_counts.sum())) - np.repeat(
                np.concatenate([[0], np.cumsum(inner_counts)])[:-1], inner_counts
            )
            src_list.append(base + off)
            dst_list.append(base + off + 1)
            len_list.append(np.repeat(piece[multi], inner_counts))
        # last piece last new -> b
        src_list.append(new_ptr[multi] + cm - 2)
        dst_list.append(e[multi, 1])
        len_list.append(piece[multi])
    del node_of_new
    edges = np.stack([np.concatenate(src_list), np.concatenate(dst_list)], axis=1).astype(np.int64)
    return Lattice(
        x=np.concatenate(xs),
        y=np.concatenate(ys),
        edges=edges,
        length=np.concatenate(len_list).astype(np.float32),
        meta=dict(lat.meta, decompose_max=decompose_max),
    )


def network_structure_from_lattice(lat: Lattice):
    """Build a ``NetworkStructure`` from a ``Lattice`` (two directed edges per pair, key 0)."""
    from ..rustalgos.graph import NetworkStructure

    ns = NetworkStructure()
    ns.set_is_dual(lat.is_dual)
    n = lat.n_nodes
    ns.add_street_nodes_bulk(
        node_keys=[str(i) for i in range(n)] if n <= 200_000 else None,
        x=lat.x,
        y=lat.y,
        live=np.ones(n, dtype=bool),
        weight=np.ones(n, dtype=np.float32),
    )
    src, dst, length, angle, pk = lat.directed()
    ns.add_street_edges_bulk(
        start_nd_idx=src,
        end_nd_idx=dst,
        edge_idx=np.zeros(len(src), dtype=np.uint32),
        length=length,
        angle_sum=angle,
        shared_primal_node_key=pk,
    )
    return ns


def lattice_side_for_nodes(n_nodes: float) -> int:
    """Side length of a square lattice with about ``n_nodes`` nodes."""
    return max(2, int(round(math.sqrt(n_nodes))))

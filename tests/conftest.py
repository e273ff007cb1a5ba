"""Test configuration: import paths, the ``gpu`` marker and shared graph fixtures."""

from __future__ import annotations

import json
import os
import subprocess
import sys

import networkx as nx
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cityseer-api_amd")
for _p in (PKG, ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

LIB = os.path.join(PKG, "cityseer", "libcityseer_amd.so")
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")
if not os.path.exists(ORACLE_LIB):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
if not os.path.exists(LIB):
    subprocess.run(["make", "-s", "-j", "8", "-C", os.path.join(PKG, "csrc")], check=True)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP kernels)")


def load_golden(name: str):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def _graph_from(nodes: dict, edges: list) -> nx.MultiGraph:
    from cityseer.tools import graphs

    g = nx.MultiGraph()
    for k, (x, y) in nodes.items():
        g.add_node(k, x=x, y=y)
    for a, b in edges:
        g.add_edge(a, b)
    g.graph["crs"] = 32630
    return graphs.nx_simple_geoms(g)


@pytest.fixture
def primal_graph() -> nx.MultiGraph:
    """The reference's mock graph with straight geometries (tests/conftest.py:11-24)."""
    from cityseer.tools import graphs, mock

    return graphs.nx_simple_geoms(mock.mock_graph())


@pytest.fixture
def dual_graph(primal_graph) -> nx.MultiGraph:
    from cityseer.tools import graphs

    return graphs.nx_to_dual(primal_graph)


@pytest.fixture
def diamond_graph() -> nx.MultiGraph:
    gold = load_golden("reference_expected.json")["diamond"]
    return _graph_from(gold["nodes"], gold["edges"])


@pytest.fixture
def golden():
    return load_golden("reference_expected.json")


def graph_from(nodes, edges):
    return _graph_from(nodes, edges)

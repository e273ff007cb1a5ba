"""Metrics front-end (drop-in for ``cityseer.metrics``): network centrality and the land-use and
statistical layers."""

from . import layers, networks

__all__ = ["layers", "networks"]

"""Centrality metrics front-end (drop-in for ``cityseer.metrics``)."""

from . import networks

__all__ = ["networks"]

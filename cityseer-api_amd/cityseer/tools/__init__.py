"""Graph ingest helpers (drop-in subset of ``cityseer.tools``), loaded lazily."""

from __future__ import annotations

import importlib

__all__ = ["graphs", "io", "mock", "util"]


def __getattr__(name):
    if name in __all__:
        return importlib.import_module(f".{name}", __name__)
    raise AttributeError(name)

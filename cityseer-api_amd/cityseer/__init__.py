"""MI355X implementation of cityseer's localised network-centrality path.

A drop-in for the ``cityseer`` import surface used by network centrality:
``cityseer.rustalgos`` (``NetworkStructure`` and the threshold conversions),
``cityseer.metrics.networks`` and the graph ingest helpers in
``cityseer.tools``. The graph algorithms run as HIP kernels on the GPU through
the C ABI in ``libcityseer_amd.so``.
"""

from __future__ import annotations

import importlib

__version__ = "0.1.0"

_SUBMODULES = ("batch", "config", "metrics", "rustalgos", "tools", "serving")


def __getattr__(name):
    if name in _SUBMODULES:
        return importlib.import_module(f".{name}", __name__)
    raise AttributeError(name)

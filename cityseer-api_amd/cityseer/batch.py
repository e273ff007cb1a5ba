"""Batched centrality: many shortest-path requests in one persistent GPU launch.

A ``CentralityBatch`` collects ``node_centrality_shortest`` requests over any
number of ``NetworkStructure`` graphs, each with its own thresholds, flags and
sources, and runs all their sources in one persistent launch captured in a
hipGraph (the first ``run`` captures it, later runs replay it). ``autotune``
times candidate launch shapes (workgroup size, resident waves per SIMD, bucket
width) on a sample of the batch's sources and keeps the fastest.

Each request is prepared exactly as ``NetworkStructure.centrality_shortest``
prepares it (rust/src/centrality.rs:1624-1874 semantics), so its results equal
that call's within floating-point summation order.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from .config import SPEED_M_S as WALKING_SPEED
from .rustalgos import centrality as _res
from .rustalgos.graph import validate_tolerance


class CentralityBatch:
    def __init__(self) -> None:
        lib = _native.load()
        h = ctypes.c_void_p()
        _native.check(lib.cs_batch_create(ctypes.byref(h)))
        self._h = h
        self._jobs: list[dict] = []

    def close(self) -> None:
        """Free the batch's device memory (also done when the object is collected)."""
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _native.load().cs_batch_destroy(h)
            self._h = None

    def __del__(self) -> None:
        self.close()

    def __enter__(self) -> "CentralityBatch":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def add_shortest(
        self,
        ns,
        distances: list[int] | None = None,
        betas: list[float] | None = None,
        minutes: list[float] | None = None,
        compute_closeness: bool | None = None,
        compute_betweenness: bool | None = None,
        min_threshold_wt: float | None = None,
        speed_m_s: float | None = None,
        tolerance: float | None = None,
        sample_probability: float | None = None,
        sampling_weights: list[float] | None = None,
        random_seed: int | None = None,
        source_indices: list[int] | None = None,
    ) -> int:
        """Queue ``ns.centrality_shortest(...)`` (same arguments); returns the job index."""
        from .rustalgos import pair_distances_betas_time

        closeness = True if compute_closeness is None else bool(compute_closeness)
        betweenness = True if compute_betweenness is None else bool(compute_betweenness)
        if not closeness and not betweenness:
            raise ValueError(
                "Either or both closeness and betweenness flags is required, but both parameters are False."
            )
        speed = WALKING_SPEED if speed_m_s is None else float(speed_m_s)
        tol = validate_tolerance(tolerance)
        dists, bets, secs = pair_distances_betas_time(speed, distances, betas, minutes, min_threshold_wt)
        plan = ns._sampling_plan(sample_probability, sampling_weights, random_seed, source_indices)
        n, t = ns.node_count(), len(dists)
        out = np.zeros((7, t, n), dtype=np.float64)
        reach = np.zeros(max(t, 1), dtype=np.uint32)
        params, _keep = ns._params(dists, bets, secs, speed, tol, closeness, betweenness, plan)
        rc = _native.load().cs_batch_add_shortest(
            self._h, ns._native_handle(), ctypes.byref(params), _native.ptr(out, ctypes.c_double),
            _native.ptr(reach, ctypes.c_uint32),
        )
        if rc < 0:
            _native.check(-rc)
        self._jobs.append(dict(ns=ns, out=out, reach=reach, dists=dists, plan=plan, t=t))
        return rc

    def set_tuning(self, block_threads: int = 256, waves_per_simd: int = 0, delta_scale: float = 1.0,
                   slot_capacity: int = 0, xcd_split: bool = True) -> None:
        t = _native.CsTuning(block_threads, waves_per_simd, delta_scale, slot_capacity, int(bool(xcd_split)))
        _native.check(_native.load().cs_batch_set_tuning(self._h, ctypes.byref(t)))

    def autotune(self, sample_fraction: float = 0.03) -> dict:
        """Time candidate launch shapes on a sample of the sources; keep and return the fastest."""
        t = _native.CsTuning()
        _native.check(_native.load().cs_batch_autotune(self._h, float(sample_fraction), ctypes.byref(t)))
        return {name: getattr(t, name) for name, _ in t._fields_}

    def run(self) -> dict:
        """Run every job (results stay on the device); returns the call's statistics."""
        _native.check(_native.load().cs_batch_run(self._h))
        return _native.last_stats()

    def results(self) -> list[_res.CentralityShortestResult]:
        """Copy the last run's results out, one result object per job in insertion order."""
        _native.check(_native.load().cs_batch_results(self._h))
        res = []
        for j in self._jobs:
            ns, plan, t = j["ns"], j["plan"], j["t"]
            res.append(
                _res.CentralityShortestResult(
                    j["dists"], ns.node_keys_py(), ns.node_indices(), j["out"].copy(),
                    j["reach"][:t].tolist() if plan["track"] else [], plan["sampled"] if plan["track"] else 0,
                )
            )
        return res

    def __len__(self) -> int:
        return len(self._jobs)

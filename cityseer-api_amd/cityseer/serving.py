"""Serving: a centrality request server with continuous batching on one GPU.

``CentralityServer`` keeps graphs device-resident and admits requests
(``node_centrality_shortest`` / ``node_centrality_simplest`` over any
``NetworkStructure``, each with its own thresholds and source subset) first-come
first-served under an HBM budget. Admitted shortest-path requests share
back-to-back batched GPU launches: every launch takes part of each active
request's remaining sources, so a request that arrives while a launch runs
joins the next one and a finished request leaves without waiting for the
others. Simplest-path requests run on a second stream, so the two kernel
families overlap on the device, and every simplest-path launch carries all the
simplest-path requests active at that moment.

Each request is prepared exactly as the corresponding ``NetworkStructure``
method prepares it, so its result equals the plain call's within
floating-point summation order.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from .config import SPEED_M_S as WALKING_SPEED
from .rustalgos import centrality as _res
from .rustalgos.graph import validate_tolerance


class Request:
    """A submitted request: ``result()`` blocks until it completes, ``info()`` reports its timing."""

    def __init__(self, server: "CentralityServer", rid: int, kind: int, ns, out, reach, dists, plan, t):
        self._server = server
        self.id = rid
        self.kind = kind
        self._ns, self._out, self._reach, self._dists, self._plan, self._t = ns, out, reach, dists, plan, t
        self._result = None

    def info(self) -> dict:
        inf = _native.CsRequestInfo()
        _native.check(_native.load().cs_server_poll(self._server._h, self.id, ctypes.byref(inf)))
        return {name: getattr(inf, name) for name, _ in inf._fields_}

    def done(self) -> bool:
        st = self.info()["status"]
        return st == 2 or st < 0

    def wait(self, timeout: float | None = None) -> dict:
        """Block until the request completes (raising its error), or ``TimeoutError`` once ``timeout``
        seconds have passed with the request still queued or running."""
        inf = _native.CsRequestInfo()
        ms = -1.0 if timeout is None else 1000.0 * float(timeout)
        _native.check(_native.load().cs_server_wait_for(self._server._h, self.id, ms, ctypes.byref(inf)))
        return {name: getattr(inf, name) for name, _ in inf._fields_}

    def cancel(self) -> None:
        """Cancel the request: ``wait``/``result`` then raise ``RequestCancelled`` (unless it had
        already completed). Its sources already on the device finish first."""
        _native.check(_native.load().cs_server_cancel(self._server._h, self.id))

    def result(self):
        if self._result is None:
            self.wait()
            cls = _res.CentralityShortestResult if self.kind == _native.CS_REQ_SHORTEST else _res.CentralitySimplestResult
            plan, t = self._plan, self._t
            self._result = cls(
                self._dists, self._ns.node_keys_py(), self._ns.node_indices(), self._out,
                self._reach[:t].tolist() if plan["track"] else [], plan["sampled"] if plan["track"] else 0,
            )
        return self._result


class CentralityServer:
    def __init__(self, hbm_budget_bytes: int = 0, max_active: int = 0, slice_ms: float = 0.0) -> None:
        cfg = _native.CsServerConfig(int(hbm_budget_bytes), int(max_active), float(slice_ms))
        h = ctypes.c_void_p()
        _native.check(_native.load().cs_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _native.load().cs_server_destroy(h)
            self._h = None

    def __del__(self) -> None:
        self.close()

    def __enter__(self) -> "CentralityServer":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def _submit(self, kind, ns, dists, bets, secs, speed, tol, closeness, betweenness, plan, unit=180.0,
                offset=1.0) -> Request:
        n, t = ns.node_count(), len(dists)
        m = 7 if kind == _native.CS_REQ_SHORTEST else 4
        out = np.zeros((m, t, n), dtype=np.float64)
        reach = np.zeros(max(t, 1), dtype=np.uint32)
        params, _keep = ns._params(dists, bets, secs, speed, tol, closeness, betweenness, plan, unit, offset)
        rid = _native.load().cs_server_submit(
            self._h, ns._native_handle(), kind, ctypes.byref(params), _native.ptr(out, ctypes.c_double),
            _native.ptr(reach, ctypes.c_uint32),
        )
        if rid < 0:
            _native.check(-int(rid))
        return Request(self, int(rid), kind, ns, out, reach, dists, plan, t)

    def submit_shortest(self, ns, distances=None, betas=None, minutes=None, compute_closeness=None,
                        compute_betweenness=None, min_threshold_wt=None, speed_m_s=None, tolerance=None,
                        sample_probability=None, sampling_weights=None, random_seed=None,
                        source_indices=None) -> Request:
        """Submit ``ns.centrality_shortest(...)`` (same arguments)."""
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
        return self._submit(_native.CS_REQ_SHORTEST, ns, dists, bets, secs, speed, tol, closeness, betweenness,
                            plan)

    def submit_simplest(self, ns, distances=None, betas=None, minutes=None, compute_closeness=None,
                        compute_betweenness=None, min_threshold_wt=None, speed_m_s=None, tolerance=None,
                        angular_scaling_unit=None, farness_scaling_offset=None, sample_probability=None,
                        sampling_weights=None, random_seed=None, source_indices=None) -> Request:
        """Submit ``ns.centrality_simplest(...)`` (same arguments)."""
        from .rustalgos import pair_distances_betas_time

        ns.validate_dual_for_angular("centrality_simplest")
        closeness = True if compute_closeness is None else bool(compute_closeness)
        betweenness = True if compute_betweenness is None else bool(compute_betweenness)
        tol = validate_tolerance(tolerance)
        if not closeness and not betweenness:
            raise ValueError(
                "Either or both closeness and betweenness flags is required, but both parameters are False."
            )
        speed = WALKING_SPEED if speed_m_s is None else float(speed_m_s)
        unit = 180.0 if angular_scaling_unit is None else float(angular_scaling_unit)
        offset = 1.0 if farness_scaling_offset is None else float(farness_scaling_offset)
        dists, bets, secs = pair_distances_betas_time(speed, distances, betas, minutes, min_threshold_wt)
        plan = ns._sampling_plan(sample_probability, sampling_weights, random_seed, source_indices)
        return self._submit(_native.CS_REQ_SIMPLEST, ns, dists, bets, secs, speed, tol, closeness, betweenness,
                            plan, unit, offset)

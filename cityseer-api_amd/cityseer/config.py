"""Configuration constants and helpers.

Mirrors ``pysrc/cityseer/config.py`` of the reference: column naming
(``prep_gdf_key``, config.py:21-32), the quiet/debug flags (config.py:35-58),
threshold constants (config.py:60-64) and the progress-polling wrapper used
around long-running native calls (config.py:96-130).
"""

from __future__ import annotations

import logging
import os
from collections.abc import Callable
from concurrent.futures import ThreadPoolExecutor, wait
from typing import Any

import numpy as np

logging.basicConfig(level=logging.INFO)
logger = logging.getLogger(__name__)

np.seterr(invalid="ignore")


def prep_gdf_key(key: str, dist: int, angular: bool = False, weighted: bool | None = None) -> str:
    """Format a column label (config.py:21-32)."""
    key = key.replace(".0", "")
    key = key.replace(".0_", "_")
    key = f"cc_{key}_{dist}"
    if angular is True:
        key += "_ang"
    if weighted is True:
        key += "_wt"
    elif weighted is False:
        key += "_nw"
    return key


_TRUTHY = frozenset({"true", "1"})


def _env_flag(name: str) -> bool:
    """True when environment variable ``name`` holds "true" or "1" (any case)."""
    return os.environ.get(name, "").strip().lower() in _TRUTHY


def check_quiet() -> bool:
    """Quiet mode (config.py:35-42): set by CITYSEER_QUIET_MODE, and always on in Google Cloud
    functions (where GCP_PROJECT is defined), which have no terminal for a progress bar."""
    return "GCP_PROJECT" in os.environ or _env_flag("CITYSEER_QUIET_MODE")


QUIET_MODE = check_quiet()


def check_debug() -> bool:
    """Debug mode (config.py:48-53), set by CITYSEER_DEBUG_MODE."""
    return _env_flag("CITYSEER_DEBUG_MODE")


DEBUG_MODE: bool = check_debug()
SKIP_VALIDATION: bool = False
MIN_THRESH_WT: float = 0.01831563888873418
SPEED_M_S = 1.33333
ATOL: float = 0.01
RTOL: float = 0.0001


def log_thresholds(
    distances: list[int] | None = None,
    betas: list[float] | None = None,
    minutes: list[float] | None = None,
    min_threshold_wt: float = MIN_THRESH_WT,
    speed_m_s: float = SPEED_M_S,
):
    """Resolve the threshold triplets from whichever of distances, betas or minutes is given, log
    one line per threshold and return the distances (config.py:67-82)."""
    from . import rustalgos

    resolved = rustalgos.pair_distances_betas_time(
        speed_m_s, distances, betas, minutes, min_threshold_wt=min_threshold_wt
    )
    lines = [
        f"Distance: {d}m, Beta: {round(b, 5)}, Walking Time: {s / 60} minutes."
        for d, b, s in zip(*resolved, strict=True)
    ]
    logger.info("\n".join(["Metrics computed for:", *lines]))
    return resolved[0]


_POLL_SECONDS = 0.1


def _progress_bar(total: int, desc: str | None):
    """A tqdm bar, or None in quiet mode or without tqdm."""
    if QUIET_MODE:
        return None
    try:
        from tqdm import tqdm
    except ImportError:  # pragma: no cover - tqdm is optional
        return None
    return tqdm(total=total, desc=desc)


def wrap_progress(
    total: int,
    rust_struct: Any,
    partial_func: Callable,
    desc: str | None = None,
) -> Any:
    """Run ``partial_func`` on a helper thread while this thread shows ``rust_struct.progress()``
    (sources finished so far, read while the native call runs) against ``total``
    (config.py:96-130). Returns the call's result or re-raises its exception here."""
    bar = _progress_bar(total, desc)
    with ThreadPoolExecutor(max_workers=1, thread_name_prefix="cityseer-call") as pool:
        call = pool.submit(partial_func)
        while True:
            done, _ = wait([call], timeout=_POLL_SECONDS)
            if bar is not None:
                bar.n = min(total, int(rust_struct.progress()))
                bar.refresh()
            if done:
                break
    if bar is not None:
        bar.n = total
        bar.close()
    return call.result()

"""Configuration constants and helpers.

Mirrors ``pysrc/cityseer/config.py`` of the reference: column naming
(``prep_gdf_key``, config.py:21-32), the quiet/debug flags (config.py:35-58),
threshold constants (config.py:60-64) and the progress-polling wrapper used
around long-running native calls (config.py:96-130).
"""

from __future__ import annotations

import logging
import os
import threading
import time
from collections.abc import Callable
from queue import Queue
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


def check_quiet() -> bool:
    """Quiet mode flag (config.py:35-42)."""
    if "GCP_PROJECT" in os.environ:
        return True
    return "CITYSEER_QUIET_MODE" in os.environ and os.environ["CITYSEER_QUIET_MODE"].lower() in ["true", "1"]


QUIET_MODE = check_quiet()


def check_debug() -> bool:
    """Debug mode flag (config.py:48-53)."""
    return "CITYSEER_DEBUG_MODE" in os.environ and os.environ["CITYSEER_DEBUG_MODE"].lower() in ["true", "1"]


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
    """Resolve and log the threshold triplets (config.py:67-82)."""
    from . import rustalgos

    distances, betas, seconds = rustalgos.pair_distances_betas_time(
        speed_m_s, distances, betas, minutes, min_threshold_wt=min_threshold_wt
    )
    logger.info("Metrics computed for:")
    for distance, beta, walking_time in zip(distances, betas, seconds, strict=True):
        logger.info(f"Distance: {distance}m, Beta: {round(beta, 5)}, Walking Time: {walking_time / 60} minutes.")
    return distances


def wrap_progress(
    total: int,
    rust_struct: Any,
    partial_func: Callable,
    desc: str | None = None,
) -> Any:
    """Run a long native call on a worker thread and poll its progress counter (config.py:96-130)."""

    def wrapper(queue: Queue):
        try:
            queue.put(partial_func())
        except Exception as e:  # noqa: BLE001 - re-raised on the calling thread
            queue.put(e)

    result_queue: Queue = Queue()
    thread = threading.Thread(target=wrapper, args=(result_queue,))
    pbar = None
    if not QUIET_MODE:
        try:
            from tqdm import tqdm

            pbar = tqdm(total=total, desc=desc)
        except ImportError:  # pragma: no cover - tqdm is optional
            pbar = None
    thread.start()
    while thread.is_alive():
        time.sleep(0.1)
        if pbar is not None:
            pbar.update(rust_struct.progress() - pbar.n)
    if pbar is not None:
        pbar.update(total - pbar.n)
        pbar.close()
    result = result_queue.get()
    thread.join()
    if isinstance(result, Exception):
        raise result
    return result

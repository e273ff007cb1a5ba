"""CPU: the oracle's threaded entry points (shortest closeness and betweenness in Phase 1 and
Phase 2, segment, OD-weighted betweenness; eight threads over f64 CAS accumulators; land-use
accessibility, eight threads each writing its own sources' rows) under ThreadSanitizer and under
AddressSanitizer with UndefinedBehaviorSanitizer, through oracle/sanitize_driver.c on a generated
lattice. Both builds must run clean and agree on the integer density and count rows."""

from __future__ import annotations

import os
import subprocess
from pathlib import Path

import pytest

ORACLE = Path(__file__).resolve().parent.parent / "oracle"


@pytest.fixture(scope="module")
def binaries():
    r = subprocess.run(["make", "-s", "-C", str(ORACLE), "sanitize"], capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        pytest.skip(f"sanitizer build unavailable here: {r.stderr.strip()[-300:]}")
    return {k: ORACLE / "_build" / f"oracle_{k}" for k in ("tsan", "asan")}


def _run(path, env_extra):
    env = dict(os.environ)
    env.update(env_extra)
    r = subprocess.run([str(path), "12"], capture_output=True, text=True, timeout=300, env=env)
    return r.returncode, r.stdout, r.stderr


def test_thread_sanitizer_clean(binaries):
    rc, out, err = _run(binaries["tsan"], {"TSAN_OPTIONS": "halt_on_error=1 exitcode=66"})
    assert "ThreadSanitizer" not in err, err[-2000:]
    assert rc == 0, err[-2000:]
    assert out.startswith("density checksum")


def test_address_and_undefined_sanitizers_clean(binaries):
    rc, out, err = _run(binaries["asan"], {"ASAN_OPTIONS": "detect_leaks=1:halt_on_error=1",
                                           "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"})
    assert "AddressSanitizer" not in err and "runtime error" not in err, err[-2000:]
    assert rc == 0, err[-2000:]
    assert out.startswith("density checksum")


def test_threaded_runs_agree(binaries):
    outs = {_run(binaries[k], {})[1] for k in ("tsan", "asan")}
    assert len(outs) == 1, outs

"""Parity assertions shared by the GPU tests.

The bar: reachable counts bit-exact; every other metric within 1e-5 relative of
the oracle. The tolerance is applied per (metric, threshold) row, so that a
row of small values (harmonic closeness at a short threshold, say) is held to
1e-5 of its own magnitude rather than to an absolute floor set by the largest
value of another row. The absolute floor is 1e-12 of the row's largest value:
it only admits rounding noise around values that are zero in the oracle.
"""

from __future__ import annotations

import numpy as np

RTOL = 1e-5
ATOL_REL = 1e-12


def assert_close(got, ref, what, rtol: float = RTOL):
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    assert got.shape == ref.shape, f"{what}: shape {got.shape} != {ref.shape}"
    if ref.size == 0:
        return
    g2 = got.reshape(-1, ref.shape[-1]) if ref.ndim > 1 else got.reshape(1, -1)
    r2 = ref.reshape(-1, ref.shape[-1]) if ref.ndim > 1 else ref.reshape(1, -1)
    lead = ref.shape[:-1] if ref.ndim > 1 else ()
    bad = []
    for k in range(r2.shape[0]):
        row_ref, row_got = r2[k], g2[k]
        atol = ATOL_REL * float(np.max(np.abs(row_ref))) if row_ref.size else 0.0
        ok = np.isclose(row_got, row_ref, rtol=rtol, atol=atol)
        ok &= np.isfinite(row_got) == np.isfinite(row_ref)
        if not ok.all():
            row = np.unravel_index(k, lead) if lead else ()
            idx = np.flatnonzero(~ok)[:4]
            bad.append((tuple(int(x) for x in row), int((~ok).sum()),
                        [(int(i), float(row_got[i]), float(row_ref[i])) for i in idx]))
    if bad:
        raise AssertionError(f"{what}: {len(bad)} rows differ beyond rtol {rtol}; first {bad[:3]}")

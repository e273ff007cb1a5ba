"""Diversity measures of one vector of class counts (rust/src/diversity.rs).

Scalar helpers with the reference's names, arguments and errors: Hill numbers
(plain, branch- and pairwise-distance weighted), Gini-Simpson, Shannon and
Rao's quadratic diversity. They compute in f32, summing classes in index
order as the reference does. The per-node diversity rows of
``DataMap.mixed_uses`` come from the device instead (``cs_mixed_uses``).
"""

from __future__ import annotations

from collections.abc import Sequence

import numpy as np

f32 = np.float32
_EPS = np.finfo(np.float32).eps


def _counts(class_counts: Sequence[int]) -> list[int]:
    out = [int(c) for c in class_counts]
    if any(c < 0 for c in out):
        raise OverflowError("class counts must be non-negative integers")
    return out


def _check_q(q: float) -> f32:
    q = f32(q)
    if q < 0.0:
        raise ValueError("Invalid value for q: must be non-negative.")
    return q


def _is_q1(q: f32) -> bool:
    return abs(f32(q - f32(1.0))) < _EPS


def _clipped_wt(beta: f32, mcw: f32, d: f32) -> f32:
    from .data import clipped_beta_wt

    return f32(clipped_beta_wt(float(beta), float(mcw), float(d)))


def _pow_root(total: f32, q: f32, what: str) -> f32:
    """total ** (1 / (1 - q)) with the reference's checks."""
    if not np.isfinite(total):
        raise ValueError(f"Intermediate {what} sum is invalid.")
    e = f32(f32(1.0) / f32(f32(1.0) - q))
    if total < 0.0 and float(e) % 1.0 != 0.0:
        raise ValueError(f"Cannot raise negative base to non-integer exponent in {what}.")
    with np.errstate(all="ignore"):
        r = f32(np.power(total, e, dtype=np.float32))
    return r


def hill_diversity(class_counts: Sequence[int], q: float) -> float:
    """Hill number of order q (diversity.rs:6-75)."""
    q = _check_q(q)
    counts = _counts(class_counts)
    num = sum(counts)
    if num == 0:
        return 0.0
    if q == 0.0:
        return float(sum(1 for c in counts if c > 0))
    n = f32(num)
    if _is_q1(q):
        ent = f32(0.0)
        for c in counts:
            if c > 0:
                p = f32(f32(c) / n)
                ent = f32(ent + f32(p * np.log(p, dtype=np.float32)))
        r = f32(np.exp(-ent, dtype=np.float32))
        if not np.isfinite(r):
            raise ValueError("Hill diversity calculation resulted in invalid value (q=1 case).")
        return float(r)
    s = f32(0.0)
    for c in counts:
        if c > 0:
            s = f32(s + f32(np.power(f32(f32(c) / n), q, dtype=np.float32)))
    r = _pow_root(s, q, "Hill diversity")
    if not np.isfinite(r):
        raise ValueError("Hill diversity calculation resulted in invalid value.")
    return float(r)


def _check_weighted(class_counts, class_distances, q, beta):
    counts = _counts(class_counts)
    dists = [f32(d) for d in class_distances]
    if len(counts) != len(dists):
        raise ValueError(
            f"Mismatch between class counts length ({len(counts)}) and distances length ({len(dists)})."
        )
    if f32(beta) < 0.0:
        raise ValueError("Beta must be non-negative.")
    return counts, dists, _check_q(q)


def hill_diversity_branch_distance_wt(class_counts: Sequence[int], class_distances: Sequence[float], q: float,
                                      beta: float, max_curve_wt: float) -> float:
    """Hill number weighted by each class's nearest distance (diversity.rs:77-187)."""
    counts, dists, q = _check_weighted(class_counts, class_distances, q, beta)
    num = sum(counts)
    if num == 0:
        return 0.0
    n, b, m = f32(num), f32(beta), f32(max_curve_wt)
    pw = [(f32(f32(c) / n), _clipped_wt(b, m, d)) for c, d in zip(counts, dists, strict=True) if c > 0]
    agg = f32(0.0)
    for p, w in pw:
        agg = f32(agg + f32(w * p))
    if abs(agg) < _EPS:
        return 0.0
    if _is_q1(q):
        ent = f32(0.0)
        for p, w in pw:
            e = f32(f32(w * p) / agg)
            if e > 0.0:
                ent = f32(ent + f32(e * np.log(e, dtype=np.float32)))
        r = f32(np.exp(-ent, dtype=np.float32))
        if not np.isfinite(r):
            raise ValueError("Weighted Hill diversity calculation resulted in invalid value (q=1 case).")
        return float(r)
    s = f32(0.0)
    for p, w in pw:
        s = f32(s + f32(w * f32(np.power(f32(p / agg), q, dtype=np.float32))))
    r = _pow_root(s, q, "weighted Hill diversity")
    if not np.isfinite(r):
        raise ValueError("Weighted Hill diversity calculation resulted in invalid value.")
    return float(r)


def hill_diversity_pairwise_distance_wt(class_counts: Sequence[int], class_distances: Sequence[float], q: float,
                                        beta: float, max_curve_wt: float) -> float:
    """Hill number weighted by the distance between each pair of classes, the square root of
    the pairwise form (diversity.rs:189-339)."""
    counts, dists, q = _check_weighted(class_counts, class_distances, q, beta)
    num = sum(counts)
    if num == 0:
        return 0.0
    n, b, m = f32(num), f32(beta), f32(max_curve_wt)
    probs = [f32(f32(c) / n) for c in counts]
    pairs = []  # (a_i, a_j, wt) over i, j <= i, both present, in the reference's order
    for i, ci in enumerate(counts):
        if ci == 0:
            continue
        for j in range(i + 1):
            if counts[j] == 0:
                continue
            pairs.append((probs[i], probs[j], _clipped_wt(b, m, f32(dists[i] + dists[j]))))
    agg = f32(0.0)
    for ai, aj, w in pairs:
        agg = f32(agg + f32(f32(w * ai) * aj))
    if abs(agg) < _EPS:
        return 0.0
    if _is_q1(q):
        ent = f32(0.0)
        for ai, aj, w in pairs:
            e = f32(f32(f32(w * ai) * aj) / agg)
            if e > 0.0:
                ent = f32(ent + f32(e * np.log(e, dtype=np.float32)))
        ex = f32(np.exp(-ent, dtype=np.float32))
        if not np.isfinite(ex) or ex < 0.0:
            raise ValueError(
                "Pairwise Hill diversity intermediate calculation resulted in invalid value (q=1 case).")
        return float(f32(np.sqrt(ex, dtype=np.float32)))
    s = f32(0.0)
    for ai, aj, w in pairs:
        s = f32(s + f32(w * f32(np.power(f32(f32(ai * aj) / agg), q, dtype=np.float32))))
    pv = _pow_root(s, q, "pairwise Hill diversity")
    if not np.isfinite(pv) or pv < 0.0:
        raise ValueError("Pairwise Hill diversity intermediate calculation resulted in invalid value (q!=1 case).")
    return float(f32(np.sqrt(pv, dtype=np.float32)))


def gini_simpson_diversity(class_counts: Sequence[int]) -> float:
    """1 - sum x (x - 1) / (n (n - 1)) over the classes (diversity.rs:341-363)."""
    counts = _counts(class_counts)
    num = sum(counts)
    if num < 2:
        return 0.0
    n, n1 = f32(num), f32(num - 1)
    lam = f32(0.0)
    for c in counts:
        if c > 0:
            x = f32(c)
            lam = f32(lam + f32(f32(x / n) * f32(max(f32(x - f32(1.0)), f32(0.0)) / n1)))
    return float(max(f32(f32(1.0) - lam), f32(0.0)))


def shannon_diversity(class_counts: Sequence[int]) -> float:
    """Shannon entropy of the class proportions (diversity.rs:365-391)."""
    counts = _counts(class_counts)
    num = sum(counts)
    if num == 0:
        return 0.0
    n = f32(num)
    ent = f32(0.0)
    for c in counts:
        if c > 0:
            p = f32(f32(c) / n)
            ent = f32(ent + f32(p * np.log(p, dtype=np.float32)))
    if not np.isfinite(ent):
        raise ValueError("Shannon entropy calculation resulted in invalid value.")
    return float(max(f32(-ent), f32(0.0)))


def raos_quadratic_diversity(class_counts: Sequence[int], wt_matrix: Sequence[Sequence[float]], alpha: float,
                             beta: float) -> float:
    """Rao's quadratic entropy over class pairs j <= i with disparity weights, bias-adjusted
    (diversity.rs:393-515)."""
    counts = _counts(class_counts)
    k = len(counts)
    if k != len(wt_matrix):
        raise ValueError(
            f"Mismatch between class counts length ({k}) and weights matrix rows ({len(wt_matrix)}).")
    if k > 0 and (len(wt_matrix) == 0 or len(wt_matrix[0]) != k):
        cols = 0 if len(wt_matrix) == 0 else len(wt_matrix[0])
        raise ValueError(f"Weights matrix must be square ({k}x{k}). Got {k}x{cols}.")
    a, b = f32(alpha), f32(beta)
    if a < 0.0:
        raise ValueError("Alpha must be non-negative.")
    if b < 0.0:
        raise ValueError("Beta must be non-negative.")
    num = sum(counts)
    if num < 2:
        return 0.0
    n, n1 = f32(num), f32(num - 1)
    raos = f32(0.0)
    with np.errstate(all="ignore"):
        for i, ci in enumerate(counts):
            if ci == 0:
                continue
            pi = f32(f32(ci) / n)
            for j in range(i + 1):
                if counts[j] == 0:
                    continue
                pj = f32(f32(counts[j]) / n1)
                wt = f32(wt_matrix[i][j])
                if not np.isfinite(wt):
                    raise ValueError(f"Invalid weight encountered in matrix at [{i}][{j}].")
                prob = f32(pi * pj)
                pb = f32(1.0) if (prob == 0.0 and b == 0.0) else f32(np.power(prob, b, dtype=np.float32))
                if wt < 0.0 and float(a) % 1.0 != 0.0:
                    raise ValueError(f"Cannot raise negative weight {wt} to non-integer exponent alpha {a}.")
                wa = f32(1.0) if (wt == 0.0 and a == 0.0) else f32(np.power(wt, a, dtype=np.float32))
                if not np.isfinite(pb) or not np.isfinite(wa):
                    raise ValueError(f"Invalid intermediate power calculation for pair ({i}, {j}).")
                raos = f32(raos + f32(wa * pb))
    if not np.isfinite(raos):
        raise ValueError("Rao's Q calculation resulted in invalid value.")
    return float(raos)

"""CPU: rustalgos.diversity (rust/src/diversity.rs) and the top-level helpers check_numerical_data and
avg_distances_for_betas (common.rs:75-88, 271-307), as the reference's tests/rustalgos/test_diversity.py
and test_common.py check them: against scipy's entropy, closed forms, plain restatements of
Gini-Simpson and Rao's Q, the reference's expected average distances, and the error cases."""

from __future__ import annotations

import math

import numpy as np
import pytest
from scipy.stats import entropy

from cityseer import config, rustalgos
from cityseer.rustalgos import diversity as D
from cityseer.tools import graphs, mock


def test_hill_diversity_against_entropy_and_closed_forms():
    for counts, probs in mock.mock_species_data():
        c = counts.tolist()
        assert np.isclose(D.hill_diversity(c, 1), np.exp(entropy(probs)), atol=config.ATOL, rtol=config.RTOL)
        # orders just either side of 1 converge on exp(entropy)
        for q in (0.99999999, 1.00000001):
            assert np.isclose(D.hill_diversity(c, q), np.exp(entropy(probs)), atol=config.ATOL, rtol=config.RTOL)
        assert D.hill_diversity(c, 0) == len(c)
        assert np.isclose(D.hill_diversity(c, 2), 1.0 / np.sum(probs**2), rtol=1e-5)
        with pytest.raises(ValueError, match="non-negative"):
            D.hill_diversity(c, -1)
    assert D.hill_diversity([], 1) == 0.0 and D.hill_diversity([0, 0], 2) == 0.0


def test_branch_and_pairwise_weighted_forms():
    for counts, _probs in mock.mock_species_data():
        c = counts.tolist()
        ones = [1.0] * len(c)
        for q in (0, 1, 2):
            # with no distance decay the branch form is the plain Hill number
            assert np.isclose(D.hill_diversity(c, q), D.hill_diversity_branch_distance_wt(c, ones, q, 0, 1),
                              atol=config.ATOL, rtol=config.RTOL)
            # the pairwise form over all ordered pairs j <= i, restated plainly
            p = counts / counts.sum()
            pairs = np.array([p[i] * p[j] for i in range(len(p)) for j in range(i + 1)])
            w = pairs / pairs.sum()
            want = math.sqrt(math.exp(-np.sum(w * np.log(w)))) if q == 1 else math.sqrt(np.sum(w**q) ** (1 / (1 - q)))
            assert np.isclose(D.hill_diversity_pairwise_distance_wt(c, ones, q, 0, 1), want, rtol=1e-4)
    # distance decay: one class near, one far; the far class counts for less
    near = D.hill_diversity_branch_distance_wt([5, 5], [0.0, 0.0], 0, 0.02, 1.0)
    far = D.hill_diversity_branch_distance_wt([5, 5], [0.0, 100.0], 0, 0.02, 1.0)
    assert near == 2.0 and np.isclose(far, 1.0 + math.exp(-2.0), rtol=1e-6)
    with pytest.raises(ValueError, match="Mismatch"):
        D.hill_diversity_branch_distance_wt([1, 2], [1.0], 1, 0.01, 1.0)
    with pytest.raises(ValueError, match="Beta must be non-negative"):
        D.hill_diversity_pairwise_distance_wt([1, 2], [1.0, 2.0], 1, -0.01, 1.0)


def test_gini_simpson_shannon_and_raos():
    for counts, probs in mock.mock_species_data():
        c = counts.tolist()
        n = counts.sum()
        gini = 0.0 if n < 2 else 1 - sum(x / n * (x - 1) / (n - 1) for x in c)
        assert np.isclose(D.gini_simpson_diversity(c), gini, atol=config.ATOL, rtol=config.RTOL)
        assert abs(entropy(probs) - D.shannon_diversity(c)) < config.ATOL
        ones = np.ones((len(c), len(c)))
        raos = 0.0 if n < 2 else sum(c[i] / n * c[j] / (n - 1) for i in range(len(c)) for j in range(i + 1))
        assert np.isclose(D.raos_quadratic_diversity(c, ones.tolist(), 1.0, 1.0), raos, atol=config.ATOL,
                          rtol=config.RTOL)
    with pytest.raises(ValueError, match="square"):
        D.raos_quadratic_diversity([1, 2], [[1.0], [1.0]], 1.0, 1.0)
    with pytest.raises(ValueError, match="Alpha"):
        D.raos_quadratic_diversity([1, 2], [[1.0, 1.0], [1.0, 1.0]], -1.0, 1.0)


def test_check_numerical_data_and_average_distances():
    G = graphs.nx_simple_geoms(mock.mock_graph())
    num = mock.mock_numerical_data(G, num_arrs=2)
    with pytest.raises(TypeError):
        rustalgos.check_numerical_data(num["mock_numerical_1"].values)
    ok = num[["mock_numerical_1", "mock_numerical_2"]].values
    rustalgos.check_numerical_data(ok)
    bad = ok.copy()
    bad[0, 0] = np.inf
    with pytest.raises(ValueError, match="finite"):
        rustalgos.check_numerical_data(bad)
    betas = rustalgos.betas_from_distances([100, 200, 400, 800, 1600])
    assert np.allclose(rustalgos.avg_distances_for_betas(betas), [35.11949, 70.23898, 140.47797, 280.95593, 561.91187],
                       atol=config.ATOL, rtol=config.RTOL)
    with pytest.raises(ValueError, match="cannot be empty"):
        rustalgos.avg_distances_for_betas([])

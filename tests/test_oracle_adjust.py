"""Pins of the oracle's adjust_moments (populate.cc:30-61, adjust_moments_host
populate_host.cc:8-44), the step that makes a populated bunch's sample mean
and second moments exactly the requested ones."""
import numpy as np
import pytest

import oracle as O


def random_cov(rng):
    A = rng.standard_normal((6, 6)) * np.array([1e-3, 2e-4, 3e-3, 4e-4, 1e-2, 1e-4])[:, None]
    return A @ A.T


def test_adjust_matrix_is_cholesky_ratio():
    rng = np.random.default_rng(5)
    C, X = random_cov(rng), random_cov(rng)
    rc, A = O.adjust_matrix(C, X)
    assert rc == 0
    want = np.linalg.cholesky(C) @ np.linalg.inv(np.linalg.cholesky(X))
    assert np.allclose(A, want, rtol=1e-12, atol=1e-14 * np.abs(want).max())
    # A maps X onto C: A X A^T = C
    assert np.allclose(A @ X @ A.T, C, rtol=1e-12, atol=1e-14 * np.abs(C).max())


@pytest.mark.parametrize("masked", [False, True])
def test_adjusted_moments_are_the_requested_ones(masked):
    rng = np.random.default_rng(6)
    n = 20000
    parts = np.zeros((7, n + 5))
    parts[:6, :n] = rng.standard_normal((6, n)) * 1.3 + 0.2
    parts[6, :n] = np.arange(n)
    masks = np.ones(n, np.uint8)
    if masked:
        masks[::9] = 0
    C = random_cov(rng)
    means = np.array([1e-4, -2e-5, 3e-4, 1e-6, -1e-3, 2e-6])
    ids = parts[6].copy()
    assert O.adjust_moments(parts, masks, means, C, total_num=int(masks.sum())) == 0
    mean, mom2 = O.moments(parts, masks, total_num=int(masks.sum()))
    assert np.allclose(mean, means, rtol=1e-12, atol=1e-16)
    assert np.allclose(mom2, C, rtol=1e-11, atol=1e-13 * np.abs(C).max())
    # numpy's own statistics of the active rows agree
    act = parts[:6, :n][:, masks == 1]
    assert np.allclose(act.mean(axis=1), means, rtol=1e-12, atol=1e-16)
    assert np.allclose(np.cov(act, bias=True), C, rtol=1e-10, atol=1e-12 * np.abs(C).max())
    assert np.array_equal(parts[6], ids)  # the id column is untouched
    assert np.all(parts[:, n:] == 0.0)    # slots past n are untouched


def test_adjust_rejects_bad_covariances():
    rng = np.random.default_rng(7)
    n = 1000
    parts = np.zeros((7, n))
    parts[:6] = rng.standard_normal((6, n))
    masks = np.ones(n, np.uint8)
    C = random_cov(rng)
    asym = C.copy()
    asym[0, 1] *= 1 + 1e-10
    assert O.adjust_moments(parts.copy(), masks, np.zeros(6), asym) == -1
    # a tiny asymmetry within floating_point_equal's 1e-14 is accepted
    near = C.copy()
    near[0, 1] *= 1 + 1e-16
    assert O.adjust_moments(parts.copy(), masks, np.zeros(6), near) == 0
    sing = C.copy()
    sing[5, :] = sing[:, 5] = 0.0
    assert O.adjust_moments(parts.copy(), masks, np.zeros(6), sing) == -2

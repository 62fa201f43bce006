"""Pins the EM-GMM oracle (oracle/gmm.c).

The reference stores no expected EM outputs, so the recurrence is pinned by:
  * scikit-learn's GaussianMixture (a library routine) run through the same recurrence: the
    same initial weights 1/K, means and variances 1e-10 (gaumixmod.h:14-27), diagonal
    covariances, no regularisation, a fixed number of E/M iterations;
  * K = 1: after the first M-step the mean and variance are the sample mean and (population)
    variance;
  * EM's monotone log-likelihood;
  * the reference's own data (distributions/dist_3.csv, R rnorm draws) recovers the generating
    means and weights with the initial means graphemgmm.cpp:299-314 uses.
"""
import os
import warnings

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load_dist(i):
    return np.loadtxt(os.path.join(GOLD, f"dist_{i}.txt")).reshape(-1, 1)


def sklearn_em(X, means_init, iters):
    from sklearn.mixture import GaussianMixture
    K, D = np.asarray(means_init).reshape(-1, X.shape[1]).shape
    gm = GaussianMixture(n_components=K, covariance_type="diag", reg_covar=0.0, tol=0.0,
                         max_iter=iters + 1, n_init=1, weights_init=np.full(K, 1.0 / K),
                         means_init=np.asarray(means_init, float).reshape(K, D),
                         precisions_init=np.full((K, D), 1e10))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        gm.fit(X)
    return gm


@pytest.mark.parametrize("N,D,K,iters", [(3000, 4, 3, 8), (1500, 2, 5, 12), (800, 1, 2, 20)])
def test_oracle_equals_sklearn_recurrence(N, D, K, iters):
    rng = np.random.default_rng(N + D)
    centers = rng.uniform(-10, 10, (K, D))
    comp = rng.integers(0, K, N)
    X = centers[comp] + rng.uniform(0.5, 2, (K, D))[comp] * rng.standard_normal((N, D))
    init = X[:: N // K][:K]
    mu, var, frac, ll = oracle.gmm_em(X, init, iters)
    gm = sklearn_em(X, init, iters)
    assert np.allclose(mu, gm.means_, rtol=1e-9, atol=1e-9)
    assert np.allclose(var, gm.covariances_, rtol=1e-8, atol=1e-12)
    assert np.allclose(frac, gm.weights_, rtol=1e-9, atol=1e-12)
    # sklearn's bound is the mean log-likelihood including -D/2 log(2 pi) per point
    ll_sk = N * (gm.lower_bound_ + 0.5 * D * np.log(2 * np.pi))
    assert abs(ll[-1] - ll_sk) <= 1e-9 * abs(ll_sk)


def test_single_component_gives_sample_moments():
    X = load_dist(3)
    mu, var, frac, ll = oracle.gmm_em(X, [[7.0]], 0)
    assert abs(mu[0, 0] - X.mean()) < 1e-12
    assert abs(var[0, 0] - X.var()) < 1e-12
    assert frac[0] == 1.0


def test_loglikelihood_is_monotone():
    X = load_dist(3)
    for init in ([[7.0], [12.0]], [[5.0], [9.0], [20.0]]):
        mu, var, frac, ll = oracle.gmm_em(X, init, 20)
        assert np.all(np.diff(ll[1:]) >= -1e-9 * np.abs(ll[1:-1]))


def test_reference_fixture_recovers_generating_mixture():
    """dist_3: 600 draws N(5, .5), 300 N(9, .5), 100 N(15, 1) (distributions/notes.txt)."""
    X = load_dist(3)
    assert X.shape == (1000, 1)
    mu, var, frac, ll = oracle.gmm_em(X, [[5.0], [9.0], [20.0]], 20)
    assert np.allclose(mu[:, 0], [5, 9, 15], atol=[0.1, 0.1, 0.35])
    assert np.allclose(np.sqrt(var[:, 0]), [0.5, 0.5, 1.0], atol=[0.08, 0.08, 0.25])
    assert np.allclose(frac, [0.6, 0.3, 0.1], atol=0.02)
    gm = sklearn_em(X, [[5.0], [9.0], [20.0]], 20)
    assert np.allclose(mu, gm.means_, rtol=1e-9)


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_threaded_form_is_bitwise_the_serial_one(threads):
    """oracle/parallel.c splits work only across points and components and keeps every sum in
    point order, so it equals the serial EM bit for bit for any thread count."""
    rng = np.random.default_rng(threads)
    N, D, K = 3001, 7, 5
    X = np.concatenate([rng.normal(c, 0.7, (N // K + 1, D)) for c in range(K)])[:N]
    init = X[np.arange(K) * (N // K)]
    a = oracle.gmm_em(X, init, 4)
    b = oracle.gmm_em(X, init, 4, threads=threads)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)

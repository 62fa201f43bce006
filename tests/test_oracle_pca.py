"""Pins the Gram-Schmidt PCA oracle (oracle/pca.c).

The reference holds no PCA fixtures (its inputs come from MSVC rand()), so the recurrence is
pinned by what the mathematics fixes:
  * at convergence on a matrix with a known, well-separated spectrum, L_k are numpy's singular
    values, P[:,k] the right and T[:,k]/L_k the left singular vectors (same orientation);
  * the deflation is exact algebra: X_centred = T P^T + R after any number of iterations;
  * the loads are orthonormal (classical Gram-Schmidt, then normalisation);
  * row centering removes each row's mean (pca_cpu.cpp:315-335 centres rows, not columns).
"""
import numpy as np
import pytest

import oracle


def known_spectrum(M, N, sigma, seed):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((M, len(sigma))))
    V, _ = np.linalg.qr(rng.standard_normal((N, len(sigma))))
    return (U * sigma) @ V.T, U, V


@pytest.mark.parametrize("M,N", [(300, 120), (64, 200)])
def test_converges_to_numpy_svd(M, N):
    sigma = np.array([50.0, 20.0, 8.0, 3.0, 1.0])
    X, U, V = known_spectrum(M, N, sigma, seed=M + N)
    T, P, R, L = oracle.pca_gs(X, K=3, J=300, center=False)
    assert np.allclose(L, sigma[:3], rtol=1e-10)
    for k in range(3):
        s = np.sign(P[:, k] @ V[:, k])
        assert np.allclose(P[:, k], s * V[:, k], atol=1e-9)
        assert np.allclose(T[:, k] / L[k], s * U[:, k], atol=1e-9)
    assert np.allclose(R, (U[:, 3:] * sigma[3:]) @ V[:, 3:].T, atol=1e-8)


def test_centred_spectrum_matches_svd_of_row_centred_matrix():
    rng = np.random.default_rng(5)
    X = rng.standard_normal((150, 60)) @ np.diag(np.linspace(6, 0.5, 60)) + 3.0
    T, P, R, L = oracle.pca_gs(X, K=2, J=2000, center=True)
    Xc = X - X.mean(axis=1, keepdims=True)
    s = np.linalg.svd(Xc, compute_uv=False)
    assert np.allclose(L, s[:2], rtol=1e-6)


@pytest.mark.parametrize("J", [1, 3, 10])
@pytest.mark.parametrize("center", [True, False])
def test_deflation_identity_and_orthonormal_loads(J, center):
    rng = np.random.default_rng(J)
    X = rng.uniform(0, 1, (90, 40))
    K = 5
    T, P, R, L = oracle.pca_gs(X, K=K, J=J, center=center)
    Xc = X - X.mean(axis=1, keepdims=True) if center else X
    assert np.allclose(Xc, T @ P.T + R, atol=1e-12)
    assert np.allclose(P.T @ P, np.eye(K), atol=1e-12)
    assert np.allclose(np.linalg.norm(T, axis=0), L, rtol=1e-12)
    if center:
        assert np.allclose(Xc.sum(axis=1), 0, atol=1e-12)


def test_first_component_one_iteration_by_hand():
    """J=1, K=1 on a 3 x 2 matrix small enough to follow by hand: t = R[:,0], p = R^T t / |R^T t|,
    t = R p, L = |t|."""
    X = np.array([[1.0, 2.0], [3.0, 5.0], [-1.0, 4.0]])
    T, P, R, L = oracle.pca_gs(X, K=1, J=1, center=False)
    t0 = X[:, 0]                              # (1, 3, -1)
    p = X.T @ t0                              # (1+9+1, 2+15-4) = (11, 13)
    assert np.allclose(p, [11.0, 13.0])
    p = p / np.hypot(11.0, 13.0)
    t = X @ p
    assert np.allclose(P[:, 0], p) and np.isclose(L[0], np.linalg.norm(t))
    assert np.allclose(T[:, 0], t)            # T = (t / L) * L
    assert np.allclose(R, X - np.outer(t, p))


@pytest.mark.parametrize("threads", [1, 3, 8])
@pytest.mark.parametrize("center", [True, False])
def test_threaded_form_is_bitwise_the_serial_one(threads, center):
    """oracle/parallel.c splits work only across matrix rows or columns and keeps every sum in
    index order, so it equals the serial PCA bit for bit (row blocks of 512 are crossed here)."""
    rng = np.random.default_rng(threads)
    X = rng.standard_normal((1500, 37))
    a = oracle.pca_gs(X, K=4, J=3, center=center)
    b = oracle.pca_gs(X, K=4, J=3, center=center, threads=threads)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)

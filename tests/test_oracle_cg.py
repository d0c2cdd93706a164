"""Pins of the oracle's CG (Eigen ConjugateGradient semantics as called by
IterativeSolver.h:19-40): dense solves, exact-arithmetic iteration bounds,
Eigen's iteration-count and stopping conventions."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle_py as O


def csr(A):
    A = sp.csr_matrix(A)
    A.sort_indices()
    return A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data.astype(np.float64)


def poisson1d(n):
    return sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1])


def test_matches_dense_solve():
    rng = np.random.default_rng(0)
    n = 60
    B = rng.normal(size=(n, n))
    A = B @ B.T + n * np.eye(n)
    b = rng.normal(size=n)
    x, it, err, rc = O.cg_csr(*csr(A), b, tol=1e-10)
    assert rc == 0
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-8, atol=1e-12)
    assert err < 1e-10
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-9


def test_poisson_converges_within_n():
    n = 50
    A = poisson1d(n)
    b = np.ones(n)
    x, it, err, rc = O.cg_csr(*csr(A), b, tol=1e-10)
    assert rc == 0 and it <= n
    # 1D Poisson solution of A x = 1 is x_i = i (n + 1 - i) / 2
    i = np.arange(1, n + 1)
    np.testing.assert_allclose(x, i * (n + 1 - i) / 2, rtol=1e-8)


def test_eigenvector_rhs_converges_in_first_iteration():
    # the residual vanishes after the first update: Eigen breaks before i++
    n = 30
    A = poisson1d(n).toarray()
    k = 3
    v = np.sin(np.pi * k * np.arange(1, n + 1) / (n + 1))
    x, it, err, rc = O.cg_csr(*csr(A), v)
    assert rc == 0 and it == 0
    lam = 2 - 2 * np.cos(np.pi * k / (n + 1))
    np.testing.assert_allclose(x, v / lam, rtol=1e-12)


def test_zero_rhs_and_exact_guess():
    A = poisson1d(20)
    x, it, err, rc = O.cg_csr(*csr(A), np.zeros(20), x0=np.ones(20))
    assert rc == 0 and it == 0 and err == 0 and np.all(x == 0)
    xs = np.linspace(1, 2, 20)
    b = A @ xs
    x, it, err, rc = O.cg_csr(*csr(A), b, x0=xs)
    assert rc == 0 and it == 0


def test_stopping_rule_and_max_iterations():
    n = 40
    A = poisson1d(n)
    rng = np.random.default_rng(1)
    b = rng.normal(size=n)
    _, it, err, rc = O.cg_csr(*csr(A), b, tol=1e-6)
    assert rc == 0 and err < 1e-6 and it > 2
    # the converging iteration has index `it`; with max_iter = it it is not run
    _, it2, err2, rc2 = O.cg_csr(*csr(A), b, tol=1e-6, max_iter=it)
    assert rc2 == 2 and it2 == it and err2 >= 1e-6
    _, it3, err3, rc3 = O.cg_csr(*csr(A), b, tol=1e-6, max_iter=it + 1)
    assert rc3 == 0 and it3 == it


def test_tolerance_floor_eps_over_norm_b():
    # tol = max(1e-6, eps / ||b||): a tiny right-hand side loosens it
    n = 20
    A = poisson1d(n)
    b = np.full(n, 1e-12)
    _, it, err, rc = O.cg_csr(*csr(A), b, tol=1e-6)
    assert rc == 0
    assert err <= np.finfo(float).eps / np.linalg.norm(b)

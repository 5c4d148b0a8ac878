"""Pins the linear-algebra oracle (oracle/la_oracle.py) to the expectations of
the reference's math/backend_test.clj and to LAPACK (numpy.linalg)."""
import numpy as np
import pytest

import la_oracle as lo

I2 = np.eye(2, dtype=complex)
CM = np.array([[1, -1j], [1j, 1]], dtype=complex)          # complex-matrix / hermitian-matrix of backend_test.clj
PZ = np.array([[1, 0], [0, -1]], dtype=complex)


def rand_c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_backend_test_expectations():
    assert lo.matrix_add(I2, I2)[0, 0] == 2
    assert lo.matrix_subtract(I2, I2)[0, 0] == 0
    assert lo.matrix_scale(I2, 2.0)[0, 0] == 2
    assert lo.matrix_negate(I2)[0, 0] == -1
    assert np.array_equal(lo.matrix_multiply(I2, CM), CM)
    assert lo.kronecker_product(I2, I2).shape == (4, 4)
    assert np.array_equal(lo.conjugate_transpose(CM), CM)
    assert lo.norm2([3, 4]) == 5.0
    assert lo.hermitian(CM, 1e-12) and lo.unitary(I2, 1e-12)
    e = lo.eigen_hermitian(PZ)
    assert e["eigenvalues"] == [-1, 1]
    for lam, v in zip(e["eigenvalues"], e["eigenvectors"]):
        assert np.linalg.norm(PZ @ v - lam * v) < 1e-10


def test_products_against_numpy():
    rng = np.random.default_rng(1)
    A, B, x, y = rand_c(rng, 5, 7), rand_c(rng, 7, 3), rand_c(rng, 7), rand_c(rng, 5)
    assert np.allclose(lo.matrix_multiply(A, B), A @ B, atol=1e-13)
    assert np.allclose(lo.matrix_vector_product(A, x), A @ x, atol=1e-13)
    assert np.allclose(lo.kronecker_product(A, B), np.kron(A, B), atol=0)
    assert np.allclose(lo.outer_product(y, y), np.outer(y, np.conj(y)))
    assert np.isclose(lo.inner_product(x, x), np.vdot(x, x))
    M = rand_c(rng, 4, 4)
    assert np.isclose(lo.trace(M), np.trace(M), atol=1e-14)
    assert np.array_equal(lo.transpose(A), A.T) and np.array_equal(lo.conjugate_transpose(A), A.conj().T)
    assert np.array_equal(lo.hadamard_product(A, A), A * A)


@pytest.mark.parametrize("n", [1, 2, 5, 17])
def test_solve_inverse_against_lapack(n):
    rng = np.random.default_rng(n)
    A = rand_c(rng, n, n) + n * np.eye(n)
    b = rand_c(rng, n)
    assert np.allclose(lo.solve_linear(A, b), np.linalg.solve(A, b), atol=1e-12)
    assert np.allclose(lo.inverse(A), np.linalg.inv(A), atol=1e-12)


def test_singular_and_pivot_rule():
    with pytest.raises(lo.OracleLAError):
        lo.inverse(np.array([[1, 2], [2, 4]], complex))
    with pytest.raises(lo.OracleLAError):
        lo.solve_linear(np.zeros((3, 3), complex), np.ones(3))
    # equal pivot candidates: the first row wins (no swap), so the result is exact here
    A = np.array([[0, 1, 0], [1, 0, 0], [-1, 0, 1]], complex)
    assert np.allclose(lo.inverse(A) @ A, np.eye(3))


@pytest.mark.parametrize("n", [1, 2, 3, 8])
def test_eigen_hermitian_against_lapack(n):
    rng = np.random.default_rng(10 + n)
    R = rng.standard_normal((n, n))
    S = R + R.T
    e = lo.eigen_hermitian(S)
    assert np.allclose([v.real for v in e["eigenvalues"]], np.linalg.eigvalsh(S), atol=1e-12)
    H = rand_c(rng, n, n)
    H = H + H.conj().T
    e = lo.eigen_hermitian(H)
    vals = np.array([v.real for v in e["eigenvalues"]])
    ref = np.linalg.eigvalsh(H)
    # the embedding doubles each eigenvalue; `distinct` keeps both copies unless bit-identical
    assert all(np.min(np.abs(ref - v)) < 1e-10 for v in vals)
    assert all(np.min(np.abs(vals - r)) < 1e-10 for r in ref)
    for lam, v in zip(e["eigenvalues"], e["eigenvectors"]):
        v = v / np.linalg.norm(v)
        assert np.linalg.norm(H @ v - lam * v) < 1e-9


def test_predicates_and_psd():
    assert not lo.hermitian(np.array([[1, 2], [3, 4]], complex), 1e-12)
    assert lo.diagonal(np.diag([1, 2, 3]).astype(complex), 1e-12)
    assert not lo.diagonal(np.array([[1, 1e-9], [0, 1]], complex), 1e-12)
    H = np.array([[1, 1j], [1j, 1]], complex)
    assert lo.unitary(H / np.sqrt(2), 1e-12)
    assert lo.positive_semidefinite(np.array([[2, 1], [1, 2]], complex), 1e-12)
    assert not lo.positive_semidefinite(np.array([[1, 2], [2, 1]], complex), 1e-12)
    with pytest.raises(lo.OracleLAError):
        lo.positive_semidefinite(np.array([[1, 2], [3, 1]], complex), 1e-12)

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


# ---- MatrixAnalysis, svd, matrix functions: the expectations of the reference's
# complex_linear_algebra_test.clj:156-326, then scipy / numpy on other inputs

def test_reference_function_expectations():
    assert np.allclose(lo.matrix_exp(np.zeros((2, 2))), np.eye(2), atol=1e-10)
    E = lo.matrix_exp(np.diag([1.0, 2.0]))
    assert abs(E[0, 0] - np.exp(1)) < 1e-10 and abs(E[1, 1] - np.exp(2)) < 1e-10 and abs(E[0, 1]) < 1e-10
    assert abs(lo.matrix_exp([[3.0]])[0, 0] - np.exp(3)) < 1e-10
    U = lo.matrix_exp(np.diag([-1j * np.pi / 4, 1j * np.pi / 4]))
    assert lo.unitary(U, 1e-12)
    c = np.sqrt(2) / 2
    assert abs(U[0, 0] - complex(c, -c)) < 1e-10 and abs(U[1, 1] - complex(c, c)) < 1e-10
    assert lo.unitary(lo.matrix_exp(np.diag([-1j, 1j])), 1e-12)
    assert np.allclose(lo.matrix_log(np.eye(2)), 0, atol=1e-12)
    L = lo.matrix_log(np.diag([2.0, 3.0]))
    assert abs(L[0, 0] - np.log(2)) < 1e-10 and abs(L[1, 1] - np.log(3)) < 1e-10 and abs(L[0, 1]) < 1e-12
    assert np.max(np.abs(lo.matrix_log(np.array([[1.0, 2.0], [2.0, -1.0]])).imag)) > 1e-10
    assert np.allclose(lo.matrix_sqrt(np.diag([4.0, 9.0])), np.diag([2.0, 3.0]), atol=1e-10)
    assert abs(lo.matrix_sqrt([[16.0]])[0, 0] - 4) < 1e-10
    assert np.allclose(lo.matrix_sqrt(np.eye(2)), np.eye(2), atol=1e-10)
    assert np.allclose(lo.matrix_sqrt(np.diag([1.0, -1.0])), np.diag([1.0, 1j]), atol=1e-10)
    assert abs(lo.spectral_norm(np.eye(2)) - 1) < 1e-12
    assert abs(lo.spectral_norm([[3.0, 4.0], [0.0, 5.0]]) - 6.708203932499369) < 1e-10
    assert abs(lo.spectral_norm(np.zeros((2, 2)))) < 1e-12
    assert abs(lo.condition_number(np.eye(2)) - 1) < 1e-12
    assert 1 < lo.condition_number([[4.0, 1.0], [1.0, 3.0]]) < 5
    assert lo.condition_number([[1.0, 1.0], [1.0, 1.001]]) > 100
    assert all(abs(s - 1) < 1e-10 for s in lo.svd(np.eye(2))["S"])
    s = lo.svd(np.diag([3.0, 4.0]))["S"]
    assert len(s) == 2 and s[0].real >= s[1].real and abs(s[0] - 4) < 1e-10 and abs(s[1] - 3) < 1e-10
    s = lo.svd(np.diag([2.0, 3.0, 1.0]))["S"]
    assert len(s) == 3 and s[0].real >= s[1].real >= s[2].real
    A = np.diag([4.0, 2.0])
    sv = [x.real for x in lo.svd(A)["S"]]
    assert abs(max(sv) - lo.spectral_norm(A)) < 1e-10
    assert abs(max(sv) / min(sv) - lo.condition_number(A)) < 1e-8


@pytest.mark.parametrize("m,n", [(1, 1), (3, 3), (5, 3), (3, 5), (8, 8)])
def test_svd_against_lapack(m, n):
    from scipy.linalg import svdvals
    rng = np.random.default_rng(100 + 10 * m + n)
    A = rand_c(rng, m, n)
    r = lo.svd(A)
    s = np.array([x.real for x in r["S"]])
    ref = svdvals(A)
    assert np.allclose(s[:min(m, n)], ref, atol=1e-10)
    k = min(m, n)
    # A = U[:, :k] diag(s[:k]) V†[:k, :]
    assert np.allclose(r["U"][:, :k] @ np.diag(s[:k]) @ r["V†"][:k, :], A, atol=1e-10)
    assert np.allclose(r["V†"] @ r["V†"].conj().T, np.eye(n), atol=1e-10)
    assert abs(lo.spectral_norm(A) - ref[0]) < 1e-10


@pytest.mark.parametrize("n", [2, 4, 9])
def test_matrix_functions_against_scipy(n):
    from scipy.linalg import expm, logm, sqrtm
    rng = np.random.default_rng(200 + n)
    H = rand_c(rng, n, n)
    H = (H + H.conj().T) / 2
    assert np.allclose(lo.matrix_exp(H), expm(H), atol=1e-10)
    assert np.allclose(lo.matrix_exp(1j * H), expm(1j * H), atol=1e-10)
    G = rand_c(rng, n, n) / n
    assert np.allclose(lo.matrix_exp(G), expm(G), atol=1e-10)
    P = H @ H.conj().T + np.eye(n)                     # Hermitian positive definite
    assert np.allclose(lo.matrix_log(P), logm(P), atol=1e-9)
    assert np.allclose(lo.matrix_sqrt(P), sqrtm(P), atol=1e-9)
    assert np.allclose(lo.matrix_exp(lo.matrix_log(P)), P, atol=1e-9)
    S = lo.matrix_sqrt(P)
    assert np.allclose(S @ S, P, atol=1e-9)
    Gp = G + 3 * np.eye(n)                               # eigenvalues off the negative real axis
    assert np.allclose(lo.matrix_log(Gp), logm(Gp), atol=1e-10)
    assert np.allclose(lo.matrix_sqrt(Gp) @ lo.matrix_sqrt(Gp), Gp, atol=1e-10)
    with pytest.raises(lo.OracleLAError):
        lo.matrix_log(np.array([[1.0, 2.0], [0.5, 1.0]]))   # singular


@pytest.mark.parametrize("n", [1, 3, 8])
def test_cholesky_lu_qr_against_lapack(n):
    from scipy.linalg import lu as sp_lu
    rng = np.random.default_rng(500 + n)
    B = rand_c(rng, n, n)
    P = B @ B.conj().T + n * np.eye(n)
    assert np.allclose(lo.cholesky(P)["L"], np.linalg.cholesky(P), atol=1e-12)
    # LU: real input has the reference's factors, the same pivots as LAPACK's getrf
    R = rng.standard_normal((n, n))
    f = lo.lu(R)
    p, l, u = sp_lu(R)
    assert np.allclose(f["P"], p.T) and np.allclose(f["L"], l, atol=1e-12) and np.allclose(f["U"], u, atol=1e-12)
    f = lo.lu(B)
    assert np.allclose(f["P"] @ B, f["L"] @ f["U"], atol=1e-12) and np.all(np.abs(np.tril(f["L"], -1)) <= 1 + 1e-15)
    # QR: A = QR, orthonormal Q, R upper with real positive diagonal (unique)
    A = rand_c(rng, n + 2, n)
    q = lo.qr(A)
    assert np.allclose(q["Q"] @ q["R"], A, atol=1e-12)
    assert np.allclose(q["Q"].conj().T @ q["Q"], np.eye(n), atol=1e-12)
    assert np.allclose(np.tril(q["R"], -1), 0) and np.all(np.diag(q["R"]).real > 0)
    Qn, Rn = np.linalg.qr(A)
    ph = np.diag(Rn) / np.abs(np.diag(Rn))
    assert np.allclose(q["R"], np.diag(ph.conj()) @ Rn, atol=1e-10)


def test_decomposition_errors_and_edge_cases():
    with pytest.raises(lo.OracleLAError, match="Hermitian"):
        lo.cholesky(np.array([[1, 2], [0, 1]], complex))
    with pytest.raises(lo.OracleLAError, match="positive definite"):
        lo.cholesky(np.array([[1, 2], [2, 1]], complex))
    with pytest.raises(lo.OracleLAError, match="singular"):
        lo.lu(np.array([[1, 2], [2, 4]], complex))
    # a dependent column: the reference puts e_j into Q
    q = lo.qr(np.array([[1, 2, 0], [1, 2, 1], [0, 0, 1]], complex))
    assert np.allclose(q["Q"][:, 1], [0, 1, 0]) and abs(q["R"][1, 1]) < 1e-12
    # equal pivot candidates: the first row stays
    assert np.array_equal(lo.lu(np.array([[1, 1], [1, 2]], complex))["P"], np.eye(2))

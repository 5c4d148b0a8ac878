"""The :hip complex linear-algebra backend: QClojure's math protocols over
libqsv's device matrices (src/org/soulspace/qclojure/domain/math/protocols.clj,
complex_linear_algebra.clj).

``HipBackend`` implements BackendAdapter (protocols.clj:12-76), MatrixAlgebra
(protocols.clj:81-353) and eigen-hermitian (protocols.clj:363-373) with one
libqsv call per operation, and on top of those device calls svd, the matrix
functions (exp, log, sqrt) and
MatrixAnalysis, and LU, QR and Cholesky as single-workgroup kernels
(protocols.clj:357-520) and eigen-general (n <= 64); its matrices are opaque device handles
(``DeviceMatrix``), which the protocol allows ("implementations may return
opaque handles"), so chained operations stay in HBM and only ``to_numpy()``
reads them back. ``kronecker-product`` results are kept as their factors
(``KronMatrix``): the reference's gate application builds I⊗…⊗G⊗…⊗I with
``(reduce kronecker-product …)`` and multiplies it into the state
(gate.clj:323-395); here that matrix-vector product applies each 2x2 factor to
the vector in O(2^n) (``qsv_la_kron_apply``) instead of materialising a
2^n x 2^n matrix. The facade functions at the bottom mirror
complex_linear_algebra.clj (``with_backend``, ``add``, ``matrix_multiply`` …).

Semantics follow the reference's fastmath backend
(math/fastmath/complex_linear_algebra.clj): inner products conjugate the
first vector, outer products the second; predicates compare the largest
element-wise deviation with the tolerance (default 1e-12); solve and inverse
fail on a pivot below 1e-14 ("Matrix is singular"); eigen-hermitian returns
eigenvalues in ascending order (as complex numbers) with unit eigenvectors.
"""
from __future__ import annotations

import contextlib
import ctypes
import threading

import numpy as np

from . import _abi
from ._abi import check

DEFAULT_TOLERANCE = 1.0e-12


class LinearAlgebraError(ValueError):
    """The reference's ex-info from a math backend."""

    def __init__(self, message: str, data: dict | None = None):
        super().__init__(message)
        self.data = data or {}


def _lib():
    from . import engine
    engine.init()
    return engine.lib()


def _call(fn, *args):
    """Run one libqsv call; QSV_E_INVALID becomes LinearAlgebraError."""
    code = fn(*args)
    if code == _abi.QSV_OK:
        return
    msg = _abi.load_library().qsv_last_error().decode()
    if code == -1:
        raise LinearAlgebraError(msg)
    check(code)


# ---- device matrices -------------------------------------------------------------------
class DeviceMatrix:
    """A rows x cols complex128 matrix in HBM (a ``qsv_mat``)."""

    def __init__(self, rows: int, cols: int, _handle=None, _owner=None):
        self.rows, self.cols = int(rows), int(cols)
        self._owner = _owner          # keeps a viewed DeviceState alive
        if _handle is None:
            h = ctypes.c_void_p()
            _call(_lib().qsv_mat_create, self.rows, self.cols, ctypes.byref(h))
            _handle = h
        self._h = _handle

    @property
    def handle(self) -> ctypes.c_void_p:
        if self._h is None:
            raise LinearAlgebraError("device matrix already released")
        return self._h

    @property
    def shape(self) -> tuple[int, int]:
        return (self.rows, self.cols)

    @classmethod
    def from_numpy(cls, a) -> "DeviceMatrix":
        arr = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
        if arr.ndim == 1:
            arr = arr.reshape(-1, 1)
        if arr.ndim != 2 or arr.size == 0:
            raise LinearAlgebraError("Cannot convert to backend matrix", {"shape": arr.shape})
        m = cls(arr.shape[0], arr.shape[1])
        _call(_lib().qsv_mat_upload, m.handle, arr.ctypes.data_as(ctypes.c_void_p))
        return m

    @classmethod
    def view_state(cls, dev) -> "DeviceMatrix":
        """The amplitudes of a complex128 ``DeviceState`` as a 2^n x 1 matrix (no copy)."""
        h = ctypes.c_void_p()
        _call(_lib().qsv_mat_view_state, dev.handle, ctypes.byref(h))
        return cls(1 << dev.num_qubits, 1, _handle=h, _owner=dev)

    def to_numpy(self) -> np.ndarray:
        out = np.empty((self.rows, self.cols), dtype=np.complex128)
        _call(_lib().qsv_mat_download, self.handle, out.ctypes.data_as(ctypes.c_void_p))
        return out

    def copy(self) -> "DeviceMatrix":
        m = DeviceMatrix(self.rows, self.cols)
        _call(_lib().qsv_mat_copy, self.handle, m.handle)
        return m

    def __array__(self, dtype=None, copy=None):
        a = self.to_numpy()
        return a.astype(dtype) if dtype is not None else a

    def tolist(self):
        return self.to_numpy().tolist()

    def __len__(self):
        return self.rows

    def __getitem__(self, i):
        a = self.to_numpy()
        return a[i] if self.cols > 1 else a[:, 0][i]

    def close(self) -> None:
        if self._h is not None:
            _abi.load_library().qsv_mat_destroy(self._h)
            self._h = None
            self._owner = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __repr__(self):
        return f"DeviceMatrix({self.rows}x{self.cols})"


def _is_identity2(f: np.ndarray) -> bool:
    return f.shape == (2, 2) and np.array_equal(f, np.eye(2, dtype=np.complex128))


class KronMatrix:
    """A Kronecker product kept as its factors (left factor = most significant
    index bits, i.e. reference qubit 0 for gate expansions)."""

    def __init__(self, factors):
        self.factors = [np.ascontiguousarray(np.asarray(f, dtype=np.complex128)) for f in factors]
        r, c = 1, 1
        for f in self.factors:
            if f.ndim != 2:
                raise LinearAlgebraError("Kronecker factor is not a matrix")
            r, c = r * f.shape[0], c * f.shape[1]
        self.rows, self.cols = r, c
        self._dense = None

    @property
    def shape(self) -> tuple[int, int]:
        return (self.rows, self.cols)

    def all_2x2(self) -> bool:
        return all(f.shape == (2, 2) for f in self.factors)

    def materialize(self) -> DeviceMatrix:
        if self._dense is None:
            acc = DeviceMatrix.from_numpy(self.factors[0])
            for f in self.factors[1:]:
                b = DeviceMatrix.from_numpy(f)
                out = DeviceMatrix(acc.rows * b.rows, acc.cols * b.cols)
                _call(_lib().qsv_la_kron, acc.handle, b.handle, out.handle)
                acc = out
            self._dense = acc
        return self._dense

    def to_numpy(self) -> np.ndarray:
        return self.materialize().to_numpy()

    def __array__(self, dtype=None, copy=None):
        a = self.to_numpy()
        return a.astype(dtype) if dtype is not None else a

    def __len__(self):
        return self.rows

    def __repr__(self):
        return f"KronMatrix({'⊗'.join(f'{f.shape[0]}x{f.shape[1]}' for f in self.factors)})"


# ---- conversions (BackendAdapter) ------------------------------------------------------------
def _scalar(x) -> complex:
    if isinstance(x, dict) and "real" in x and "imag" in x:
        return complex(float(x["real"]), float(x["imag"]))
    if isinstance(x, (int, float, complex, np.number)) and not isinstance(x, bool):
        return complex(x)
    raise LinearAlgebraError("Cannot convert to Vec2", {"value": x})


def _host_matrix(m) -> np.ndarray:
    if isinstance(m, dict) and "real" in m and "imag" in m:
        return np.asarray(m["real"], dtype=float) + 1j * np.asarray(m["imag"], dtype=float)
    if isinstance(m, np.ndarray):
        return m.astype(np.complex128)
    if isinstance(m, (list, tuple)) and m and all(isinstance(r, (list, tuple, np.ndarray)) for r in m):
        return np.array([[_scalar(x) for x in r] for r in m], dtype=np.complex128)
    raise LinearAlgebraError("Cannot convert to backend matrix", {"value": m})


def _host_vector(v) -> np.ndarray:
    if isinstance(v, np.ndarray):
        return v.astype(np.complex128).reshape(-1)
    if isinstance(v, (list, tuple)):
        return np.array([_scalar(x) for x in v], dtype=np.complex128)
    return np.array([_scalar(v)], dtype=np.complex128)


class HipBackend:
    """The :hip backend instance (tolerance and config as the reference's
    backend records, fastmath/backend.clj:85)."""

    def __init__(self, tolerance: float | None = None, config: dict | None = None):
        self.tolerance = float(tolerance if tolerance is not None else DEFAULT_TOLERANCE)
        self.config = dict(config or {})

    # -- BackendAdapter (protocols.clj:12-76)
    def vector_to_backend(self, v) -> DeviceMatrix:
        if isinstance(v, DeviceMatrix):
            return v
        if isinstance(v, KronMatrix):
            return v.materialize()
        return DeviceMatrix.from_numpy(_host_vector(v).reshape(-1, 1))

    def backend_to_vector(self, v):
        return v

    def matrix_to_backend(self, m):
        if isinstance(m, (DeviceMatrix, KronMatrix)):
            return m
        return DeviceMatrix.from_numpy(_host_matrix(m))

    def backend_to_matrix(self, m):
        return m

    def scalar_to_backend(self, s) -> complex:
        return _scalar(s)

    def backend_to_scalar(self, s):
        return s

    # -- helpers
    def _dense(self, A) -> DeviceMatrix:
        A = self.matrix_to_backend(A)
        return A.materialize() if isinstance(A, KronMatrix) else A

    def _same_shape(self, A: DeviceMatrix, B: DeviceMatrix, what: str):
        if A.shape != B.shape:
            raise LinearAlgebraError(f"{what}: shapes differ", {"A": A.shape, "B": B.shape})

    # -- MatrixAlgebra (protocols.clj:81-353)
    def shape(self, A) -> list[int]:
        A = self.matrix_to_backend(A)
        return [A.rows, A.cols]

    def _axpby(self, a: complex, A, b: complex, B):
        A = self._dense(A)
        Bd = self._dense(B) if B is not None else None
        if Bd is not None:
            self._same_shape(A, Bd, "matrix-add")
        C = DeviceMatrix(A.rows, A.cols)
        _call(_lib().qsv_la_axpby, a.real, a.imag, A.handle, b.real, b.imag,
              Bd.handle if Bd is not None else None, C.handle)
        return C

    def add(self, A, B):
        return self._axpby(1 + 0j, A, 1 + 0j, B)

    def subtract(self, A, B):
        return self._axpby(1 + 0j, A, -1 + 0j, B)

    def scale(self, A, alpha):
        return self._axpby(self.scalar_to_backend(alpha), A, 0j, None)

    def negate(self, A):
        return self._axpby(-1 + 0j, A, 0j, None)

    def matrix_multiply(self, A, B):
        A, B = self.matrix_to_backend(A), self.matrix_to_backend(B)
        if isinstance(A, KronMatrix) and isinstance(B, KronMatrix) and len(A.factors) == len(B.factors) and \
                all(fa.shape[1] == fb.shape[0] for fa, fb in zip(A.factors, B.factors)):
            return KronMatrix([fa @ fb for fa, fb in zip(A.factors, B.factors)])   # mixed-product rule
        A, B = self._dense(A), self._dense(B)
        if A.cols != B.rows:
            raise LinearAlgebraError("Matrix dimensions incompatible for multiplication",
                                     {"A": A.shape, "B": B.shape})
        C = DeviceMatrix(A.rows, B.cols)
        _call(_lib().qsv_la_gemm, 0, A.handle, B.handle, C.handle)
        return C

    def matrix_vector_product(self, A, x):
        A = self.matrix_to_backend(A)
        xv = self.vector_to_backend(x)
        n = xv.rows * xv.cols
        if isinstance(A, KronMatrix) and A.all_2x2() and A.cols == n:
            y = xv.copy()
            fs = [None if _is_identity2(f) else f for f in A.factors]
            keep = [np.ascontiguousarray(f).view(np.float64) if f is not None else None for f in fs]
            arr = (ctypes.POINTER(ctypes.c_double) * len(keep))(
                *[k.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if k is not None else None for k in keep])
            _call(_lib().qsv_la_kron_apply, arr, len(keep), y.handle)
            return y
        A = self._dense(A)
        if A.cols != n:
            raise LinearAlgebraError("Matrix and vector dimensions incompatible", {"A": A.shape, "x": n})
        y = DeviceMatrix(A.rows, 1)
        _call(_lib().qsv_la_gemv, A.handle, xv.handle, y.handle)
        return y

    def inner_product(self, x, y) -> complex:
        xv, yv = self.vector_to_backend(x), self.vector_to_backend(y)
        if xv.rows * xv.cols != yv.rows * yv.cols:
            raise LinearAlgebraError("Vectors must have same length")
        out = (ctypes.c_double * 2)()
        _call(_lib().qsv_la_dot, 1, xv.handle, yv.handle, out)
        return complex(out[0], out[1])

    def outer_product(self, x, y):
        xv, yv = self.vector_to_backend(x), self.vector_to_backend(y)
        C = DeviceMatrix(xv.rows * xv.cols, yv.rows * yv.cols)
        _call(_lib().qsv_la_outer, xv.handle, yv.handle, C.handle)
        return C

    def hadamard_product(self, A, B):
        A, B = self._dense(A), self._dense(B)
        self._same_shape(A, B, "hadamard-product")
        C = DeviceMatrix(A.rows, A.cols)
        _call(_lib().qsv_la_hadamard, A.handle, B.handle, C.handle)
        return C

    def kronecker_product(self, A, B):
        def factors(M):
            if isinstance(M, KronMatrix):
                return list(M.factors)
            if isinstance(M, DeviceMatrix):
                return [M.to_numpy()] if M.rows * M.cols <= 64 else None
            return [_host_matrix(M)]
        fa, fb = factors(A), factors(B)
        if fa is not None and fb is not None:
            return KronMatrix(fa + fb)
        A, B = self._dense(A), self._dense(B)
        C = DeviceMatrix(A.rows * B.rows, A.cols * B.cols)
        _call(_lib().qsv_la_kron, A.handle, B.handle, C.handle)
        return C

    def _transpose(self, A, conj: int):
        A = self._dense(A)
        C = DeviceMatrix(A.cols, A.rows)
        _call(_lib().qsv_la_transpose, conj, A.handle, C.handle)
        return C

    def transpose(self, A):
        return self._transpose(A, 0)

    def conjugate_transpose(self, A):
        return self._transpose(A, 1)

    def trace(self, A) -> complex:
        A = self._dense(A)
        out = (ctypes.c_double * 2)()
        _call(_lib().qsv_la_trace, A.handle, out)
        return complex(out[0], out[1])

    def norm2(self, x) -> float:
        xv = self.vector_to_backend(x)
        out = ctypes.c_double()
        _call(_lib().qsv_la_norm2, xv.handle, ctypes.byref(out))
        return out.value

    def solve_linear_system(self, A, b):
        A = self._dense(A)
        bv = self.vector_to_backend(b)
        if A.rows != A.cols:
            raise LinearAlgebraError("Matrix must be square for solving", {"shape": A.shape})
        if bv.rows * bv.cols != A.rows:
            raise LinearAlgebraError("Incompatible dimensions", {"matrix-rows": A.rows,
                                                                 "vector-length": bv.rows * bv.cols})
        x = DeviceMatrix(A.rows, 1)
        _call(_lib().qsv_la_solve, A.handle, bv.handle, x.handle)
        return x

    def inverse(self, A):
        A = self._dense(A)
        if A.rows != A.cols:
            raise LinearAlgebraError("Matrix must be square for inverse", {"shape": A.shape})
        X = DeviceMatrix(A.rows, A.cols)
        _call(_lib().qsv_la_solve, A.handle, None, X.handle)
        return X

    def _deviation(self, kind: int, A) -> float:
        out = ctypes.c_double()
        _call(_lib().qsv_la_deviation, kind, A.handle, ctypes.byref(out))
        return out.value

    def hermitian(self, A, eps: float | None = None) -> bool:
        eps = self.tolerance if eps is None else eps
        return self._deviation(_abi.QSV_LA_DEV_HERMITIAN, self._dense(A)) < eps

    def diagonal(self, A, eps: float | None = None) -> bool:
        eps = self.tolerance if eps is None else eps
        return self._deviation(_abi.QSV_LA_DEV_OFFDIAG, self._dense(A)) < eps

    def unitary(self, U, eps: float | None = None) -> bool:
        eps = self.tolerance if eps is None else eps
        U = self._dense(U)
        P = DeviceMatrix(U.cols, U.cols)
        _call(_lib().qsv_la_gemm, 1, U.handle, U.handle, P.handle)
        return self._deviation(_abi.QSV_LA_DEV_IDENTITY, P) < eps

    def positive_semidefinite(self, A, eps: float | None = None) -> bool:
        eps = self.tolerance if eps is None else eps
        try:
            A = self._dense(A)
            if A.rows != A.cols:
                raise LinearAlgebraError("Matrix must be square for positive-semidefinite check")
            if not self.hermitian(A, eps):
                raise LinearAlgebraError("Matrix must be Hermitian for positive-semidefinite check")
            return all(ev.real >= -eps for ev in self.eigen_hermitian(A)["eigenvalues"])
        except LinearAlgebraError as e:
            raise LinearAlgebraError("Error checking positive semidefinite property",
                                     {"original-error": str(e)}) from None

    # -- MatrixDecompositions: eigen-hermitian (protocols.clj:363-373)
    def eigen_hermitian(self, A) -> dict:
        A = self._dense(A)
        if A.rows != A.cols:
            raise LinearAlgebraError("Matrix must be square for eigendecomposition", {"shape": A.shape})
        n = A.rows
        w = np.empty(n, dtype=np.float64)
        V = DeviceMatrix(n, n)
        sweeps = ctypes.c_int(0)
        _call(_lib().qsv_la_eigh, A.handle, w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), V.handle,
              ctypes.byref(sweeps))
        order = np.argsort(w, kind="stable")
        vecs = V.to_numpy()
        return {"eigenvalues": [complex(w[i], 0.0) for i in order],
                "eigenvectors": [vecs[:, i].copy() for i in order],
                "sweeps": sweeps.value}


    # -- device helpers for the decompositions, functions and analysis below
    def _gemm(self, A: DeviceMatrix, B: DeviceMatrix, conja: int = 0) -> DeviceMatrix:
        C = DeviceMatrix(A.cols if conja else A.rows, B.cols)
        _call(_lib().qsv_la_gemm, conja, A.handle, B.handle, C.handle)
        return C

    def _eigh_dev(self, A: DeviceMatrix):
        """Eigenvalues (host, unsorted) and unit eigenvector columns (device)."""
        n = A.rows
        w = np.empty(n, dtype=np.float64)
        V = DeviceMatrix(n, n)
        sweeps = ctypes.c_int(0)
        _call(_lib().qsv_la_eigh, A.handle, w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), V.handle,
              ctypes.byref(sweeps))
        return w, V

    def _mul_cols(self, V: DeviceMatrix, d) -> DeviceMatrix:
        """V · diag(d): the elementwise product with the rank-one matrix 1 dᵀ
        (an outer product conjugates its second vector, so it gets conj(d))."""
        ones = DeviceMatrix.from_numpy(np.ones((V.rows, 1), dtype=np.complex128))
        dc = DeviceMatrix.from_numpy(np.conj(np.asarray(d, dtype=np.complex128)).reshape(-1, 1))
        D = DeviceMatrix(V.rows, V.cols)
        _call(_lib().qsv_la_outer, ones.handle, dc.handle, D.handle)
        return self.hadamard_product(V, D)

    def _spectral_fn(self, w, V: DeviceMatrix, f) -> DeviceMatrix:
        """V f(Λ) Vᴴ for unit eigenvector columns V of a normal matrix."""
        return self._gemm(self._mul_cols(V, [f(x) for x in w]), self.conjugate_transpose(V))

    def _ata_eigen(self, A):
        A = self._dense(A)
        return self._eigh_dev(self._gemm(A, A, conja=1))

    def _square(self, A, what: str) -> DeviceMatrix:
        A = self._dense(A)
        if A.rows != A.cols:
            raise LinearAlgebraError(f"{what} requires square matrix", {"shape": A.shape})
        return A

    # -- MatrixAnalysis (protocols.clj:494-520)
    def spectral_norm(self, A) -> float:
        """spectral-norm (fastmath/complex_linear_algebra.clj:1409-1436): the
        square root of the largest eigenvalue of AᴴA (clamped at 0, so that a
        zero matrix gives 0)."""
        w, _ = self._ata_eigen(A)
        return float(np.sqrt(max(0.0, float(np.max(w)))))

    def condition_number(self, A) -> float:
        """condition-number (:1438-1470): σ_max / σ_min from the eigenvalues of
        AᴴA, +inf when σ_min < 1e-14."""
        w, _ = self._ata_eigen(A)
        sv = np.sqrt(np.maximum(0.0, w))
        smax, smin = float(np.max(sv)), float(np.min(sv))
        return float("inf") if smin < 1e-14 else smax / smin

    # -- MatrixDecompositions (protocols.clj:357-444)
    def svd(self, A) -> dict:
        """svd (:883-963) as the reference computes it: V and σ from the
        eigendecomposition of AᴴA (σ = sqrt(max(0, λ)), descending, ties in
        eigenvalue order), U's column j = A v_j / σ_j (zero when σ_j <= 1e-12)
        for j < min(m, n), then for m > n the unit vectors e_n … e_{m-1}.
        Returns {"U": m x m, "S": σ (n values, complex), "V†": Vᴴ}."""
        A = self._dense(A)
        m, n = A.rows, A.cols
        w, V = self._ata_eigen(A)
        sv = np.sqrt(np.maximum(0.0, w))
        order = sorted(range(n), key=lambda j: -sv[j])            # stable: ties keep eigenvalue order
        P = np.zeros((n, n), dtype=np.complex128)                  # V P = V with sorted columns
        P[order, np.arange(n)] = 1.0
        Vs = self._gemm(V, DeviceMatrix.from_numpy(P))
        k = min(m, n)
        PS = np.zeros((n, m), dtype=np.complex128)                 # column j < k of U scaled by 1/σ_j
        for j in range(k):
            s = sv[order[j]]
            if s > 1e-12:
                PS[j, j] = 1.0 / s
        U = self._gemm(A, self._gemm(Vs, DeviceMatrix.from_numpy(PS)))
        if m > n:
            E = np.zeros((m, m), dtype=np.complex128)
            E[np.arange(n, m), np.arange(n, m)] = 1.0
            U = self.add(U, DeviceMatrix.from_numpy(E))
        return {"U": U, "S": [complex(sv[j], 0.0) for j in order], "V†": self.conjugate_transpose(Vs)}

    def eigen_general(self, A) -> dict:
        """eigen-general (protocols.clj:375-387; fastmath/complex_linear_algebra.clj:792-870):
        a Hermitian A (within 1e-10) goes to eigen-hermitian, as the
        reference's does; any other square A up to 64 x 64 is reduced to
        Hessenberg form and iterated to its Schur form on the device
        (qsv_la_eig). Returns the n eigenvalues in ascending (re, im) order
        with unit eigenvector columns. (The reference instead takes the n
        smallest of the 2n eigenvalues of its real embedding, which repeats or
        conjugates eigenvalues of a non-Hermitian A; that output is not
        reproduced.)"""
        A = self._dense(A)
        if A.rows != A.cols:
            raise LinearAlgebraError("Matrix must be square for eigendecomposition", {"shape": A.shape})
        if self.hermitian(A, 1e-10):
            return self.eigen_hermitian(A)
        n = A.rows
        w = np.empty(2 * n, dtype=np.float64)
        V = DeviceMatrix(n, n)
        _call(_lib().qsv_la_eig, A.handle, w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), V.handle)
        lam = w[0::2] + 1j * w[1::2]
        order = sorted(range(n), key=lambda i: (lam[i].real, lam[i].imag))
        vecs = V.to_numpy()
        return {"eigenvalues": [complex(lam[i]) for i in order],
                "eigenvectors": [vecs[:, i].copy() for i in order]}

    def lu_decomposition(self, A) -> dict:
        """lu-decomposition (:965-1046): {"L", "U", "P"} with P·A = L·U,
        partial pivoting on the first row of largest |a_ik| (the pivot rule of
        Apache Commons Math's LUDecomposition, which the reference applies to
        the real embedding of A: the same factors for real A), L unit lower."""
        A = self._square(A, "LU decomposition")
        n = A.rows
        L, U = DeviceMatrix(n, n), DeviceMatrix(n, n)
        perm = np.zeros(max(n, 1), dtype=np.int64)
        _call(_lib().qsv_la_lu, A.handle, L.handle, U.handle, perm.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        P = np.zeros((n, n), dtype=np.complex128)
        P[np.arange(n), perm[:n]] = 1.0
        return {"L": L, "U": U, "P": DeviceMatrix.from_numpy(P)}

    def qr_decomposition(self, A) -> dict:
        """qr-decomposition (:1048-1102): modified Gram-Schmidt, {"Q": m x n,
        "R": n x n}."""
        A = self._dense(A)
        Q, R = DeviceMatrix(A.rows, A.cols), DeviceMatrix(A.cols, A.cols)
        _call(_lib().qsv_la_qr, A.handle, Q.handle, R.handle)
        return {"Q": Q, "R": R}

    def cholesky_decomposition(self, A, eps: float | None = None) -> dict:
        """cholesky-decomposition (:1104-1166): {"L"} with A = L Lᴴ; the
        reference's checks and messages (square, Hermitian within eps, real
        positive diagonal residuals)."""
        eps = self.tolerance if eps is None else float(eps)
        A = self._square(A, "Cholesky decomposition")
        if not self.hermitian(A, eps):
            raise LinearAlgebraError("Cholesky decomposition requires Hermitian matrix")
        L = DeviceMatrix(A.rows, A.rows)
        _call(_lib().qsv_la_cholesky, A.handle, eps, L.handle)
        return {"L": L}

    # -- MatrixFunctions (protocols.clj:446-492)
    def matrix_exp(self, A):
        """matrix-exp (:1168-1258). Hermitian A (tolerance 1e-10): V e^Λ Vᴴ;
        skew-Hermitian A = iH: V e^{iΛ} Vᴴ from H's eigensystem; any other A:
        scaling and squaring of the Taylor series, all on the device."""
        A = self._square(A, "Matrix exponential")
        n = A.rows
        if self.hermitian(A, 1e-10):
            w, V = self._eigh_dev(A)
            return self._spectral_fn(w, V, lambda x: complex(np.exp(x)))
        H = self.scale(A, -1j)
        if self.hermitian(H, 1e-10):
            w, V = self._eigh_dev(H)
            return self._spectral_fn(w, V, lambda x: complex(np.cos(x), np.sin(x)))
        nrm = self.norm2(A)
        s = max(0, int(np.ceil(np.log2(nrm))) + 1) if nrm > 0.5 else 0
        B = self.scale(A, 2.0 ** -s)
        I = DeviceMatrix.from_numpy(np.eye(n, dtype=np.complex128))
        T = self.add(I, B)
        term = B
        for k in range(2, 40):
            term = self.scale(self._gemm(term, B), 1.0 / k)
            T = self.add(T, term)
            if self.norm2(term) <= 1e-17 * self.norm2(T):
                break
        for _ in range(s):
            T = self._gemm(T, T)
        return T

    def _diag_fn(self, A: DeviceMatrix, f, what: str):
        d = np.diag(A.to_numpy())
        if np.any(np.abs(d) < 1e-14):
            raise LinearAlgebraError(f"Cannot compute {what} of zero diagonal element")
        return DeviceMatrix.from_numpy(np.diag([f(z) for z in d]).astype(np.complex128))

    def _hermitian_fn(self, A: DeviceMatrix, f, what: str, failed: str):
        w, V = self._eigh_dev(A)
        if np.any(np.abs(w) < 1e-14):
            raise LinearAlgebraError(failed, {"original-error": f"Cannot compute {what} of zero eigenvalue",
                                              "matrix-size": A.rows})
        return self._spectral_fn(w, V, f)

    @staticmethod
    def _log(z) -> complex:
        z = complex(z)
        return complex(np.log(abs(z)), np.angle(z))

    @staticmethod
    def _sqrt(z) -> complex:
        z = complex(z)
        r, th = np.sqrt(abs(z)), np.angle(z) / 2.0
        return complex(r * np.cos(th), r * np.sin(th))

    def _db_sqrt(self, A: DeviceMatrix, failed: str) -> DeviceMatrix:
        """Principal square root of a general matrix by the Denman-Beavers
        iteration Y <- (Y + Z^-1)/2, Z <- (Z + Y^-1)/2 from Y = A, Z = I
        (device inverses, sums and norms). Fails for a singular A or one with
        eigenvalues on the closed negative real axis, where it does not settle."""
        n = A.rows
        Y, Z = A, DeviceMatrix.from_numpy(np.eye(n, dtype=np.complex128))
        for _ in range(60):
            try:
                Yi, Zi = self.inverse(Y), self.inverse(Z)
            except LinearAlgebraError as e:
                raise LinearAlgebraError(failed, {"original-error": str(e), "matrix-size": n}) from None
            Yn = self._axpby(0.5 + 0j, Y, 0.5 + 0j, Zi)
            Z = self._axpby(0.5 + 0j, Z, 0.5 + 0j, Yi)
            step = self.norm2(self.subtract(Yn, Y))
            Y = Yn
            if step <= 1e-14 * self.norm2(Y):
                return Y
        raise LinearAlgebraError(failed, {"original-error": "the square-root iteration did not converge "
                                          "(an eigenvalue on the closed negative real axis)", "matrix-size": n})

    def _iss_log(self, A: DeviceMatrix, failed: str) -> DeviceMatrix:
        """Principal logarithm of a general matrix by inverse scaling and
        squaring: k square roots until R = A^(1/2^k) is within 1/4 of I, then
        2^k log(I + X), X = R - I, from its Taylor series (all on the device)."""
        n = A.rows
        I = DeviceMatrix.from_numpy(np.eye(n, dtype=np.complex128))
        R, k = A, 0
        while self.norm2(self.subtract(R, I)) > 0.25:
            if k == 40:
                raise LinearAlgebraError(failed, {"original-error": "no convergence", "matrix-size": n})
            R = self._db_sqrt(R, failed)
            k += 1
        X = self.subtract(R, I)
        L, term = X, X
        for j in range(2, 60):
            term = self._gemm(term, X)
            c = (-1.0) ** (j + 1) / j
            L = self._axpby(1 + 0j, L, complex(c), term)
            if abs(c) * self.norm2(term) <= 1e-17 * self.norm2(L):
                break
        return self.scale(L, float(2 ** k))

    def matrix_log(self, A):
        """matrix-log (:1260-1329): principal branch, log|λ| + i arg λ; a
        diagonal A (1e-10) entry by entry, a Hermitian A through its
        eigensystem, any other A by inverse scaling and squaring (where the
        reference goes through eigen-general, mostly failing)."""
        A = self._square(A, "Matrix logarithm")
        if self.diagonal(A, 1e-10):
            return self._diag_fn(A, self._log, "log")
        if self.hermitian(A, 1e-10):
            return self._hermitian_fn(A, self._log, "log", "Complex matrix logarithm failed")
        return self._iss_log(A, "Complex matrix logarithm failed")

    def matrix_sqrt(self, A):
        """matrix-sqrt (:1331-1407): principal branch, sqrt|λ| e^{i arg(λ)/2};
        a diagonal A (1e-10) entry by entry, a Hermitian A through its
        eigensystem, any other A by the Denman-Beavers iteration."""
        A = self._square(A, "Matrix square root")
        if self.diagonal(A, 1e-10):
            return self._diag_fn(A, self._sqrt, "sqrt")
        if self.hermitian(A, 1e-10):
            return self._hermitian_fn(A, self._sqrt, "sqrt", "Complex matrix square root failed")
        return self._db_sqrt(A, "Complex matrix square root failed")


# ---- facade (complex_linear_algebra.clj:18-740) ----------------------------------------------
_BACKENDS = {"hip": HipBackend}
_state = threading.local()


def available_backends() -> set:
    """available-backends: the backend keys this host offers."""
    return set(_BACKENDS)


def create_backend(key: str = "hip", opts: dict | None = None):
    """create-backend (complex_linear_algebra.clj:95-113)."""
    ctor = _BACKENDS.get(key)
    if ctor is None:
        raise LinearAlgebraError(f"Unknown backend: {key}", {"available": available_backends()})
    opts = opts or {}
    return ctor(opts.get("tolerance"), opts)


def _current():
    b = getattr(_state, "backend", None)
    if b is None:
        b = _state.backend = (create_backend("hip"), "hip")
    return b


def get_backend() -> str:
    return _current()[1]


def get_backend_instance():
    return _current()[0]


def set_backend(key: str) -> str:
    """set-backend! (complex_linear_algebra.clj:115-133)."""
    _state.backend = (create_backend(key), key)
    return key


@contextlib.contextmanager
def with_backend(backend):
    """with-backend (complex_linear_algebra.clj:135-155): a key or an instance."""
    prev = _current()
    _state.backend = (create_backend(backend), backend) if isinstance(backend, str) else (backend, prev[1])
    try:
        yield _state.backend[0]
    finally:
        _state.backend = prev


def complex_vector(real_part, imag_part) -> np.ndarray:
    """complex-vector: a vector from separate real and imaginary parts."""
    if len(real_part) != len(imag_part):
        raise LinearAlgebraError("real and imaginary parts differ in length")
    return np.asarray(real_part, float) + 1j * np.asarray(imag_part, float)


def complex_matrix(real_m, imag_m) -> np.ndarray:
    """complex-matrix: a matrix from separate real and imaginary parts."""
    r, i = np.asarray(real_m, float), np.asarray(imag_m, float)
    if r.shape != i.shape:
        raise LinearAlgebraError("real and imaginary parts differ in shape")
    return r + 1j * i


def _facade(name):
    def f(*args):
        return getattr(get_backend_instance(), name)(*args)
    f.__name__ = name
    f.__doc__ = f"{name.replace('_', '-')} on the current backend."
    return f


for _n in ("shape", "add", "subtract", "scale", "negate", "matrix_multiply", "matrix_vector_product",
           "inner_product", "outer_product", "hadamard_product", "kronecker_product", "transpose",
           "conjugate_transpose", "trace", "norm2", "solve_linear_system", "inverse", "hermitian", "diagonal",
           "unitary", "positive_semidefinite", "eigen_hermitian", "eigen_general", "lu_decomposition",
           "qr_decomposition", "cholesky_decomposition", "matrix_exp", "matrix_log", "matrix_sqrt",
           "spectral_norm", "condition_number"):
    globals()[_n] = _facade(_n)
del _n


def svd(A) -> dict:
    """svd through the facade (complex_linear_algebra.clj:568-588): the
    backend's {:U :S :V†} renamed to {"U", "singular-values", "Vt"}."""
    r = get_backend_instance().svd(A)
    return {"U": r["U"], "singular-values": r["S"], "Vt": r["V†"]}

"""Result extraction on device-resident states
(src/org/soulspace/qclojure/domain/result.clj:179-639).

Every quantity that reads amplitudes (probabilities, shots, expectation
values, overlaps) is one libqsv call on the final state in HBM; this module
only shapes the returned maps. Large vectors (the 2^n probability vector and
the state vector) are materialised lazily on first access, so a 30-qubit job
does not copy 8-16 GB to the host unless asked.
"""
from __future__ import annotations

from collections import Counter
from fractions import Fraction

import numpy as np

from . import hamiltonian as ham
from .engine import DeviceState

PAULI_I = np.eye(2, dtype=np.complex128)


class LazyVector:
    """A vector computed on first use (len/indexing/np.asarray work)."""

    def __init__(self, n: int, fn):
        self._n = n
        self._fn = fn
        self._v = None

    def materialize(self) -> np.ndarray:
        if self._v is None:
            self._v = self._fn()
        return self._v

    def __len__(self):
        return self._n

    def __getitem__(self, i):
        return self.materialize()[i]

    def __iter__(self):
        return iter(self.materialize())

    def __array__(self, dtype=None, copy=None):
        v = self.materialize()
        return v.astype(dtype) if dtype is not None else v


class FinalState:
    """The :final-state of a job: a state map whose amplitudes stay in HBM
    until "state-vector" is read."""

    def __init__(self, dev: DeviceState):
        self.device_state = dev
        self.num_qubits = dev.num_qubits
        self._sv = LazyVector(dev.size, dev.download)

    def __getitem__(self, key):
        if key == "state-vector":
            return self._sv
        if key == "num-qubits":
            return self.num_qubits
        raise KeyError(key)

    def get(self, key, default=None):
        try:
            return self[key]
        except KeyError:
            return default

    def keys(self):
        return ["state-vector", "num-qubits"]

    def __contains__(self, key):
        return key in ("state-vector", "num-qubits")

    @property
    def on_device(self) -> bool:
        return self.device_state is not None

    def detach(self) -> None:
        """Copy the amplitudes to host memory and free the device state."""
        if self.device_state is not None:
            self._sv.materialize()
            self.device_state.destroy()
            self.device_state = None

    def release(self) -> None:
        """Free the device state without keeping the amplitudes (a job table
        dropping the result)."""
        if self.device_state is not None:
            self.device_state.destroy()
            self.device_state = None


def detach_results(results: dict) -> None:
    """Materialise every lazily computed vector of a result map into host
    memory, then free the final state's device memory, so that the map no
    longer holds HBM."""
    def walk(v):
        if isinstance(v, LazyVector):
            v.materialize()
        elif isinstance(v, dict):
            for x in v.values():
                walk(x)
    walk(results)
    fs = results.get("final-state")
    if isinstance(fs, FinalState):
        fs.detach()


def release_results(results: dict) -> None:
    """Free the device memory a result map holds (its final state)."""
    fs = results.get("final-state") if isinstance(results, dict) else None
    if isinstance(fs, FinalState):
        fs.release()


def basis_label(i: int, n: int) -> str:
    return "|" + format(i, f"0{n}b") + "⟩"


def extract_measurement_results(dev: DeviceState, measurement_qubits=None, shots: int = 1, seed: int = 0) -> dict:
    """result.clj:179-252 (ideal branch): `shots` full-register samples."""
    n = dev.num_qubits
    outcomes = [int(x) for x in dev.sample(shots, seed)]
    freq = dict(Counter(outcomes))
    return {
        "measurement-outcomes": outcomes,
        "measurement-probabilities": LazyVector(dev.size, dev.probabilities),
        "empirical-probabilities": {o: Fraction(c, shots) for o, c in freq.items()},
        "shot-count": shots,
        "measurement-qubits": list(measurement_qubits) if measurement_qubits is not None else list(range(n)),
        "frequencies": freq,
        "source": "ideal-simulation",
    }


def _state_vector(dev: DeviceState):
    """The state as a device vector of the :hip math backend: a view of a
    complex128 state, a complex128 copy of a complex64 one."""
    from . import la
    from ._abi import QSV_C128
    if dev.precision == QSV_C128:
        return la.DeviceMatrix.view_state(dev)
    return la.DeviceMatrix.from_numpy(dev.download().astype(np.complex128).reshape(-1, 1))


def _dense_expectation(dev: DeviceState, observable) -> complex:
    """obs/expectation-value (observables.clj:234-251) for an observable of the
    state's full dimension: <psi| (O psi) with the :hip backend's GEMV and inner
    product on the device, as the reference computes it through its LA facade."""
    from . import la
    hip = la.create_backend("hip")
    psi = _state_vector(dev)
    return hip.inner_product(psi, hip.matrix_vector_product(observable, psi))


def _embedded_expectation(dev: DeviceState, observable, target):
    obs = np.asarray(observable, dtype=np.complex128)
    if target is None:
        if obs.shape != (2 ** dev.num_qubits,) * 2:
            raise ValueError("observable dimension does not match the state")
        if dev.num_qubits != 1:
            return _dense_expectation(dev, obs)
        target = 0
    return dev.expect_1q(int(target), obs)


def extract_expectation_results(dev: DeviceState, observables, target_qubits=None) -> list:
    """result.clj:254-288: Re <psi| I..O..I |psi>."""
    targets = target_qubits if target_qubits is not None else [None] * len(observables)
    out = []
    for o, t in zip(observables, targets):
        v = _embedded_expectation(dev, o, t).real
        out.append({"expectation-value": v, "observable": o, "target-qubits": [t] if t is not None else None})
    return out


def extract_variance_results(dev: DeviceState, observables, target_qubits=None) -> list:
    """result.clj:290-325: <O^2> - <O>^2."""
    targets = target_qubits if target_qubits is not None else [None] * len(observables)
    out = []
    for o, t in zip(observables, targets):
        m = np.asarray(o, dtype=np.complex128)
        e = _embedded_expectation(dev, m, t).real
        if t is None and dev.num_qubits != 1 and m.shape == (2 ** dev.num_qubits,) * 2:
            # O^2 for a full-dimension observable: the device GEMM, then <psi|O^2 psi>
            from . import la
            e2 = _dense_expectation(dev, la.create_backend("hip").matrix_multiply(m, m)).real
        else:
            e2 = _embedded_expectation(dev, m @ m, t).real
        var = e2 - e * e
        out.append({"variance-value": var, "standard-deviation": float(np.sqrt(var)) if var >= 0 else float("nan"),
                    "observable": o, "target-qubits": [t] if t is not None else None})
    return out


def extract_hamiltonian_expectation(dev: DeviceState, h) -> dict:
    """result.clj:327-344."""
    if not ham.validate_hamiltonian(h):
        raise ValueError("invalid hamiltonian")
    return {"energy-expectation": ham.hamiltonian_expectation(h, dev),
            "hamiltonian": h,
            "measurement-groups": ham.group_commuting_terms(h),
            "measurement-bases": ham.group_pauli_terms_by_measurement_basis(h)}


def _bits_to_index(bits) -> int:
    n = len(bits)
    return sum(int(b) << (n - 1 - i) for i, b in enumerate(bits))


def extract_probability_results(dev: DeviceState, target_qubits=None, target_states=None) -> dict:
    """result.clj:346-380."""
    n = dev.num_qubits
    if target_states is not None:
        outcomes = {}
        for t in target_states:
            idx = _bits_to_index(t) if isinstance(t, (list, tuple)) else int(t)
            key = tuple(t) if isinstance(t, list) else t
            outcomes[key] = float(dev.probabilities(idx, 1)[0])
        return {"probability-outcomes": outcomes, "target-states": target_states, "target-qubits": target_qubits}
    allp = dev.probabilities()
    return {"probability-outcomes": dict(enumerate(allp.tolist())),
            "target-qubits": target_qubits if target_qubits is not None else list(range(n)),
            "all-probabilities": allp}


def extract_amplitude_results(dev: DeviceState, basis_states) -> dict:
    """result.clj:382-402."""
    vals = {int(b): complex(dev.download(int(b), 1)[0]) for b in basis_states}
    return {"amplitude-values": vals, "basis-states": list(basis_states)}


def extract_state_vector_result(dev: DeviceState) -> dict:
    """result.clj:404-418."""
    n = dev.num_qubits
    return {"state-vector": LazyVector(dev.size, dev.download), "num-qubits": n,
            "basis-labels": LazyVector(dev.size, lambda: [basis_label(i, n) for i in range(1 << n)])}


def extract_density_matrix_result(dev: DeviceState) -> dict:
    """result.clj:420-435: rho = |psi><psi| (host formatting; n <= 12)."""
    if dev.num_qubits > 12:
        raise ValueError("density matrix requested for more than 12 qubits")
    psi = dev.download()
    rho = np.outer(psi, psi.conj())
    return {"density-matrix": rho, "num-qubits": dev.num_qubits,
            "trace-valid": bool(abs(np.trace(rho).real - 1.0) < 1e-10)}


def _state_vector_of(ref) -> np.ndarray:
    if isinstance(ref, DeviceState):
        return ref.download()
    if isinstance(ref, FinalState):
        return ref.device_state.download()
    if isinstance(ref, dict):
        return np.asarray(ref["state-vector"], dtype=np.complex128)
    return np.asarray(ref, dtype=np.complex128)


def extract_fidelity_result(dev: DeviceState, reference_states) -> dict:
    """result.clj:437-457: |<psi|phi>| (the modulus, not squared)."""
    fids = {}
    for k, ref in enumerate(reference_states):
        if isinstance(ref, DeviceState):
            ov = dev.inner(ref)
        else:
            tmp = DeviceState(dev.num_qubits, dev.precision)
            try:
                tmp.upload(_state_vector_of(ref))
                ov = dev.inner(tmp)
            finally:
                tmp.destroy()
        fids[f"reference-{k}"] = abs(ov)
    return {"fidelities": fids, "reference-states": reference_states}


def _eigen_probabilities(dev: DeviceState, observable, target) -> list[tuple[float, float]]:
    """obs/measurement-probabilities (observables.clj:325-353): |<v_i|psi>|^2
    per eigenvector of the Hermitian observable, on the device. A 2x2
    observable with a target qubit is embedded as I..O..I: the probability of
    eigenvalue l is <psi| I..|v><v|..I |psi> (expect_1q). Eigenvalues that
    coincide are merged into one entry whose probability is that of the whole
    eigenspace (the reference keys a map by eigenvalue and keeps one
    eigenvector's overlap, which depends on its eigen-solver's basis choice)."""
    obs = np.asarray(observable, dtype=np.complex128)
    if obs.ndim != 2 or obs.shape[0] != obs.shape[1] or not np.allclose(obs, obs.conj().T, atol=1e-12):
        raise ValueError("observable must be a square Hermitian matrix")
    n = dev.num_qubits
    probs = []
    if obs.shape == (2, 2) and (target is not None or n == 1):
        w, v = np.linalg.eigh(obs)       # a 2x2 on the host, like a gate matrix
        q = 0 if target is None else int(target)
        for i in range(2):
            proj = np.outer(v[:, i], v[:, i].conj())
            probs.append(float(dev.expect_1q(q, proj).real))
    elif obs.shape == (1 << n, 1 << n):
        # the full-dimension eigensystem by the :hip backend's Jacobi solver
        from . import la
        e = la.create_backend("hip").eigen_hermitian(obs)
        w = np.array([x.real for x in e["eigenvalues"]])
        tmp = DeviceState(n, dev.precision)
        try:
            for vec in e["eigenvectors"]:
                tmp.upload(vec)
                probs.append(abs(tmp.inner(dev)) ** 2)
        finally:
            tmp.destroy()
    else:
        raise ValueError("observable dimension does not match the state (give a target qubit for a 2x2 observable)")
    merged: list[list[float]] = []
    for lam, p in zip(w, probs):
        if merged and abs(lam - merged[-1][0]) <= 1e-12 * max(1.0, abs(lam)):
            merged[-1][1] += p
        else:
            merged.append([float(lam), p])
    return [(lam, p) for lam, p in merged]


def extract_sample_results(dev: DeviceState, observables, shots: int = 1000, target_qubits=None,
                           seed: int = 0) -> list:
    """extract-sample-results (result.clj:459-494): per observable, `shots`
    eigenvalue draws; each draw walks the (eigenvalue, probability) entries in
    ascending eigenvalue order and takes the first whose running sum exceeds
    the uniform, else the last entry. Draw k of observable j uses
    qsv_uniform(seed ^ SAMPLE_SALT, j * 2^32 + k)."""
    from .engine import uniform
    targets = target_qubits if target_qubits is not None else [None] * len(observables)
    out = []
    for j, (o, t) in enumerate(zip(observables, targets)):
        entries = _eigen_probabilities(dev, o, t)
        outcomes = []
        for k in range(int(shots)):
            u = uniform(seed ^ SAMPLE_SALT, (j << 32) + k)
            acc = 0.0
            pick = entries[-1][0]
            for lam, p in entries:
                acc += p
                if u < acc:
                    pick = lam
                    break
            outcomes.append(pick)
        out.append({"sample-outcomes": outcomes, "observable": o, "shot-count": int(shots),
                    "target-qubits": target_qubits, "frequencies": dict(Counter(outcomes))})
    return out


SAMPLE_SALT = 0x53414D504C452121


def _truthy(v) -> bool:
    """Clojure truthiness: only nil and false are false (an empty map is true)."""
    return v is not None and v is not False


def extract_results(base_result: dict, result_specs: dict, dev: DeviceState, seed: int = 0) -> dict:
    """result.clj:535-639."""
    out = dict(base_result)
    rs = result_specs or {}
    if _truthy(rs.get("measurements")):
        spec = rs["measurements"]
        out["measurement-results"] = extract_measurement_results(
            dev, spec.get("qubits"), int(spec.get("shots") or 1), seed)
    if _truthy(rs.get("expectation")):
        spec = rs["expectation"]
        out["expectation-results"] = extract_expectation_results(dev, spec["observables"], spec.get("targets"))
    if _truthy(rs.get("variance")):
        spec = rs["variance"]
        out["variance-results"] = extract_variance_results(dev, spec["observables"], spec.get("targets"))
    if _truthy(rs.get("hamiltonian")):
        out["hamiltonian-result"] = extract_hamiltonian_expectation(dev, rs["hamiltonian"])
    if _truthy(rs.get("probabilities")):
        spec = rs["probabilities"]
        out["probability-results"] = extract_probability_results(dev, spec.get("qubits"), spec.get("targets"))
    if _truthy(rs.get("amplitudes")):
        out["amplitude-results"] = extract_amplitude_results(dev, rs["amplitudes"]["basis-states"])
    if _truthy(rs.get("state-vector")):
        out["state-vector-result"] = extract_state_vector_result(dev)
    if _truthy(rs.get("density-matrix")):
        out["density-matrix-result"] = extract_density_matrix_result(dev)
    if _truthy(rs.get("fidelity")):
        out["fidelity-results"] = extract_fidelity_result(dev, rs["fidelity"]["references"])
    if _truthy(rs.get("sample")):
        spec = rs["sample"]
        out["sample-results"] = extract_sample_results(dev, spec["observables"], int(spec.get("shots") or 1000),
                                                       spec.get("targets"), seed)
    return out

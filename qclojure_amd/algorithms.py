"""Callers of the path whose circuits the engine must reproduce exactly.

* MaxCut Hamiltonian, standard mixer, Hamiltonian-evolution circuit and the
  QAOA ansatz (src/org/soulspace/qclojure/application/algorithm/qaoa.clj:111-149,
  397-416, 453-564).
* The QFT circuit built from H + CRZ + SWAP
  (src/org/soulspace/qclojure/application/algorithm/quantum_fourier_transform.clj:34-62).

The builders collect operation maps in a list (the same maps the
circuit-builder functions append one by one) so building stays linear in the
number of operations.
"""
from __future__ import annotations

import math

from . import circuit as qc
from .hamiltonian import validate_hamiltonian


def _op(t: str, **params) -> dict:
    """The operation map add-gate appends (circuit.clj:88-109)."""
    return {"operation-type": t, "operation-params": {k.replace("_", "-"): v for k, v in params.items()}}


def max_cut_hamiltonian(graph, num_vertices: int) -> list[dict]:
    """qaoa.clj:111-149: per edge [i j w]: {w/2, I..I} and {-w/2, Z_i Z_j}."""
    terms = []
    for (i, j, w) in graph:
        coeff = w / 2.0
        ident = "I" * num_vertices
        zz = ["I"] * num_vertices
        zz[i] = "Z"
        zz[j] = "Z"
        terms.append({"coefficient": coeff, "pauli-string": ident})
        terms.append({"coefficient": -coeff, "pauli-string": "".join(zz)})
    return terms


def standard_mixer_hamiltonian(num_qubits: int) -> list[dict]:
    """qaoa.clj:397-416: sum_i X_i with coefficient 1.0."""
    out = []
    for i in range(num_qubits):
        s = ["I"] * num_qubits
        s[i] = "X"
        out.append({"coefficient": 1.0, "pauli-string": "".join(s)})
    return out


def _evolution_ops(hamiltonian, evolution_time: float) -> list[dict]:
    ops = []
    for term in hamiltonian:
        coeff, ps = term["coefficient"], term["pauli-string"]
        angle = 2.0 * evolution_time * coeff
        non_i = [(k, ch) for k, ch in enumerate(ps) if ch != "I"]
        if not non_i:
            continue                       # identity terms: global phase only
        if len(non_i) == 1:
            k, ch = non_i[0]
            if ch in "XYZ":
                ops.append(_op({"X": "rx", "Y": "ry", "Z": "rz"}[ch], target=k, angle=angle))
            continue
        if all(ch == "Z" for _, ch in non_i):
            idx = [k for k, _ in non_i]
            if len(idx) == 2:
                q1, q2 = idx
                ops += [_op("cnot", control=q1, target=q2), _op("rz", target=q2, angle=angle),
                        _op("cnot", control=q1, target=q2)]
            else:
                target, controls = idx[-1], idx[:-1]
                ops += [_op("cnot", control=c, target=target) for c in controls]
                ops.append(_op("rz", target=target, angle=angle))
                ops += [_op("cnot", control=c, target=target) for c in reversed(controls)]
        # X/Y multi-qubit terms are skipped, as in the reference
    return ops


def hamiltonian_evolution_circuit(hamiltonian, evolution_time: float, num_qubits: int) -> dict:
    """qaoa.clj:453-522."""
    if not validate_hamiltonian(hamiltonian):
        raise ValueError("invalid hamiltonian")
    c = qc.create_circuit(num_qubits, "Hamiltonian Evo")
    c["operations"] = _evolution_ops(hamiltonian, evolution_time)
    return c


def compose_circuits(c1: dict, c2: dict) -> dict:
    """circuit_composition.clj:87 with the identity qubit mapping."""
    out = dict(c1)
    out["num-qubits"] = max(c1["num-qubits"], c2["num-qubits"])
    out["operations"] = list(c1["operations"]) + list(c2["operations"])
    return out


def qaoa_ansatz_circuit(problem_h, mixer_h, parameters, num_qubits: int) -> dict:
    """qaoa.clj:524-564: H on every qubit, then per (gamma, beta) the problem
    evolution followed by the mixer evolution."""
    if len(parameters) % 2:
        raise ValueError("parameters must hold (gamma, beta) pairs")
    if not (validate_hamiltonian(problem_h) and validate_hamiltonian(mixer_h)):
        raise ValueError("invalid hamiltonian")
    c = qc.create_circuit(num_qubits, "QAOA Ansatz")
    ops = [_op("h", target=q) for q in range(num_qubits)]
    for k in range(0, len(parameters), 2):
        gamma, beta = parameters[k], parameters[k + 1]
        ops += _evolution_ops(problem_h, gamma)
        ops += _evolution_ops(mixer_h, beta)
    c["operations"] = ops
    return c


def quantum_fourier_transform_circuit(n: int) -> dict:
    """quantum_fourier_transform.clj:34-62: H(q), CRZ(q+k+1 -> q, pi/2^(k+1)),
    then SWAP(i, n-1-i)."""
    c = qc.create_circuit(n, "QFT", "Quantum Fourier Transform")
    ops = []
    for q in range(n):
        ops.append(_op("h", target=q))
        for k in range(n - q - 1):
            ctl = q + k + 1
            angle = math.pi / math.pow(2, k + 1)
            if ctl < n:
                ops.append(_op("crz", control=ctl, target=q, angle=angle))
    for i in range(n // 2):
        j = n - 1 - i
        if i < j:
            ops.append(_op("swap", qubit1=i, qubit2=j))
    c["operations"] = ops
    return c

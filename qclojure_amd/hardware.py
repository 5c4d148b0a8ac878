"""The hardware simulator's noisy execution on the GPU trajectory path.

Mirrors execute-circuit / execute-circuit-simulation-with-trajectories of
src/org/soulspace/qclojure/adapter/backend/hardware_simulator.clj:106-242:
every shot is one trajectory (clean operation, then the gate's noise channel,
noise.clj:53-103), measured once and passed through the readout-error model
(noise.clj:108-202). All trajectories run in one qsv_noisy_trajectories call
(one workgroup per trajectory with the state in LDS up to 13/14 qubits).

Result map keys follow the reference: :measurement-results {bitstring count},
:final-state, :trajectories (the first max-trajectories final states, when
collected), :trajectory-count, :density-matrix (trajectory average, n <= 10),
:density-matrix-trace, :trajectory-weights; the outer map carries :job-status,
:circuit, :circuit-metadata, :shots-executed and :execution-time-ms.
Differences: the reference's :final-state is the last shot's state, here it is
the last collected trajectory's (trajectories are independent, so the two have
the same distribution); shots are reproducible from :seed.
"""
from __future__ import annotations

import time

import numpy as np

from . import engine
from . import noise as nz
from .circuit import CircuitError, circuit_metadata, decode_operation, op_u, ops_from_tuples

READOUT_SALT = 0x52454144_4F555421  # separates the readout stream from the trajectory stream


def _density_matrix(states: np.ndarray) -> dict:
    """trajectory-to-density-matrix (state.clj:726-800) with equal weights:
    rho = (1/k) sum_t psi_t psi_t^H. With the k final states as the rows of T,
    T^H T = k conj(rho) = k rho^T (rho is Hermitian), so rho is one device GEMM
    (A^H B, the :hip backend's) and a transpose."""
    from . import la
    k = len(states)
    hip = la.create_backend("hip")
    T = la.DeviceMatrix.from_numpy(np.ascontiguousarray(states, dtype=np.complex128))
    rho = hip.transpose(hip.scale(hip._gemm(T, T, conja=1), 1.0 / k)).to_numpy()
    return {"density-matrix": rho, "trace": float(np.real(np.trace(rho))), "weights": [1.0 / k] * k}


def program_with_noise(circuit: dict, noise_model: dict | None):
    """Decoded records of the circuit and its noise sites, each site placed
    after the last record of its operation (an operation that decodes to no
    record, such as a dropped :controlled gate, still gets its noise)."""
    n = int(circuit["num-qubits"])
    rows, ends = [], []
    for op in circuit["operations"]:
        rows.extend(decode_operation(op, n))
        ends.append(len(rows))
    ops = ops_from_tuples(rows)
    sites = nz.noise_sites(circuit, noise_model or {})
    shift = 0
    if any(ends[st["after-op"]] == 0 for st in sites):
        # noise before the first record: anchor it on an identity record
        ops = np.concatenate([op_u(0, np.eye(2)), ops])
        shift = 1
    for st in sites:
        st["after-op"] = ends[st["after-op"]] - 1 + shift
    return ops, sites


def execute_noisy_circuit(circuit: dict, noise_model: dict | None, options: dict | None = None) -> dict:
    """hardware_simulator.clj:106-242 for one circuit on one device's noise model."""
    options = options or {}
    t0 = time.perf_counter()
    try:
        n = circuit["num-qubits"]
        shots = int(options.get("shots", 1024))
        seed = int(options.get("seed", 0)) & (2**64 - 1)
        result_specs = options.get("result-specs")
        needs_traj = bool(options.get("collect-trajectories") or (result_specs or {}).get("density-matrix")
                          or (result_specs or {}).get("hamiltonian") or (result_specs or {}).get("expectation"))
        max_traj = int(options.get("max-trajectories", 100))
        if not circuit["operations"]:
            raise CircuitError("circuit has no operations")
        ops, sites = program_with_noise(circuit, noise_model)
        n_keep = min(shots, max_traj) if needs_traj else min(shots, 1)
        precision = options.get("precision", "c128")
        outcomes, states = engine.noisy_trajectories(n, ops, sites, shots, seed, precision, n_keep)
        readout = (noise_model or {}).get("readout-error")
        counts: dict[str, int] = {}
        for t, idx in enumerate(outcomes):
            bits = nz.basis_string(int(idx), n)
            if readout:
                us = [engine.uniform(seed ^ READOUT_SALT, t * n + q) for q in range(n)]
                bits = nz.apply_readout_noise(bits, readout, us)
            counts[bits] = counts.get(bits, 0) + 1
        last = states[n_keep - 1] if n_keep else None
        res = {"measurement-results": counts,
               "final-state": {"state-vector": last, "num-qubits": n} if last is not None else None}
        if needs_traj and n_keep:
            res["trajectories"] = [{"state-vector": s, "num-qubits": n} for s in states]
            res["trajectory-count"] = n_keep
            if n <= 10:
                dm = _density_matrix(states)
                res["density-matrix"] = dm["density-matrix"]
                res["density-matrix-trace"] = dm["trace"]
                res["trajectory-weights"] = dm["weights"]
        return {"job-status": "completed", "circuit": circuit, "circuit-metadata": circuit_metadata(circuit),
                "shots-executed": shots, "execution-time-ms": int((time.perf_counter() - t0) * 1000),
                "results": res}
    except (CircuitError, nz.NoiseModelError, ValueError) as e:
        return {"job-status": "failed", "error-message": str(e), "exception-type": type(e).__name__}

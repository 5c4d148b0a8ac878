"""ctypes wrapper of the FP64 CPU oracle (qsv_ref.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker or the reported CPU
baseline. The product package never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "libqsv_ref.so")

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp, i32, u64, u32, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_double
        sig = {
            "ref_gate_matrix": (i32, [i32, dbl, vp]),
            "ref_apply_op": (i32, [vp, i32, vp, u32, i32, u64, vp]),
            "ref_apply_program": (i32, [vp, i32, vp, u64, u32, i32, u64]),
            "ref_probabilities": (None, [vp, i32, vp]),
            "ref_norm2": (dbl, [vp, i32]),
            "ref_measure_state": (u64, [vp, i32, dbl]),
            "ref_sample": (None, [vp, i32, u64, u64, vp]),
            "ref_sample_prefix": (None, [vp, i32, u64, u64, vp, vp]),
            "ref_measure_qubits": (i32, [vp, i32, u64, dbl, vp, vp]),
            "ref_pauli_expectation": (dbl, [vp, i32, ctypes.c_char_p]),
            "ref_pauli_expectation_literal": (dbl, [vp, i32, ctypes.c_char_p]),
            "ref_inner": (None, [vp, vp, i32, vp]),
            "ref_expect_1q": (None, [vp, i32, i32, vp, vp]),
            "ref_uniform": (dbl, [u64, u64]),
            "ref_zero_state": (None, [vp, i32]),
            "ref_literal_max_qubits": (i32, []),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def zero_state(n: int) -> np.ndarray:
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1.0
    return psi


def gate_matrix(kind: int, angle: float = 0.0) -> np.ndarray:
    out = np.zeros(8)
    if lib().ref_gate_matrix(kind, angle, _p(out)) != 0:
        raise ValueError(f"no matrix for kind {kind}")
    return out.view(np.complex128).reshape(2, 2)


def apply_program(psi: np.ndarray, n: int, ops: np.ndarray, flags: int = 0,
                  literal: bool = False, seed: int = 0) -> np.ndarray:
    """Apply decoded ops (qsv_op records) to a copy of psi (complex128)."""
    out = np.ascontiguousarray(psi, dtype=np.complex128).copy()
    ops = np.ascontiguousarray(ops)
    rc = lib().ref_apply_program(_p(out), n, _p(ops), len(ops), flags, 1 if literal else 0, seed)
    if rc != 0:
        raise RuntimeError(f"oracle rejected program (code {rc})")
    return out


def run(n: int, ops: np.ndarray, flags: int = 0, literal: bool = False, seed: int = 0) -> np.ndarray:
    return apply_program(zero_state(n), n, ops, flags, literal, seed)


def probabilities(psi: np.ndarray, n: int) -> np.ndarray:
    out = np.zeros(1 << n)
    lib().ref_probabilities(_p(np.ascontiguousarray(psi)), n, _p(out))
    return out


def norm2(psi: np.ndarray, n: int) -> float:
    return lib().ref_norm2(_p(np.ascontiguousarray(psi)), n)


def normalize(psi: np.ndarray, n: int, tolerance: float = 1e-12) -> np.ndarray:
    """normalize-state (state.clj:524-551): divide by norm = sqrt(sum |a|^2)
    when norm > 0 and norm > the LA tolerance (1e-12 by default,
    fastmath/backend.clj:23), else return the amplitudes unchanged. The scale
    is 1/norm applied as a multiplication (fc/scale a (/ 1.0 norm))."""
    a = np.ascontiguousarray(psi, dtype=np.complex128)
    nrm = float(np.sqrt(norm2(a, n)))
    if nrm > 0 and nrm > tolerance:
        return a * (1.0 / nrm)
    return a.copy()


def sample(psi: np.ndarray, n: int, shots: int, seed: int) -> np.ndarray:
    out = np.zeros(shots, dtype=np.uint64)
    lib().ref_sample(_p(np.ascontiguousarray(psi)), n, shots, seed, _p(out))
    return out


def sample_prefix(psi: np.ndarray, n: int, shots: int, seed: int) -> np.ndarray:
    """ref_sample's shots with the running sums formed once (large states)."""
    out = np.zeros(shots, dtype=np.uint64)
    prefix = np.empty(1 << n)
    lib().ref_sample_prefix(_p(np.ascontiguousarray(psi)), n, shots, seed, _p(prefix), _p(out))
    return out


def measure_qubits(psi: np.ndarray, n: int, mask: int, u: float):
    out = np.ascontiguousarray(psi, dtype=np.complex128).copy()
    m = bin(mask).count("1")
    probs = np.zeros(1 << m)
    oc = np.zeros(1, dtype=np.uint64)
    rc = lib().ref_measure_qubits(_p(out), n, mask, u, _p(oc), _p(probs))
    if rc != 0:
        raise RuntimeError(f"oracle measure failed ({rc})")
    return int(oc[0]), probs, out


def pauli_expectation(psi: np.ndarray, n: int, s: str, literal: bool = False) -> float:
    f = lib().ref_pauli_expectation_literal if literal else lib().ref_pauli_expectation
    return f(_p(np.ascontiguousarray(psi)), n, s.encode())


def inner(a: np.ndarray, b: np.ndarray, n: int) -> complex:
    out = np.zeros(2)
    lib().ref_inner(_p(np.ascontiguousarray(a)), _p(np.ascontiguousarray(b)), n, _p(out))
    return complex(out[0], out[1])


def expect_1q(psi: np.ndarray, n: int, q: int, m: np.ndarray) -> complex:
    mm = np.ascontiguousarray(np.asarray(m, dtype=np.complex128).reshape(4)).view(np.float64)
    out = np.zeros(2)
    lib().ref_expect_1q(_p(np.ascontiguousarray(psi)), n, q, _p(mm), _p(out))
    return complex(out[0], out[1])


def uniform(seed: int, k: int) -> float:
    return lib().ref_uniform(seed, k)


TRAJ_SLOTS = 1 << 20
_PHI = 0x9E3779B97F4A7C15


def trajectories(n: int, ops: np.ndarray, sites, n_traj: int, seed: int, make_u):
    """Noisy shots as the hardware simulator takes them
    (src/org/soulspace/qclojure/adapter/backend/hardware_simulator.clj:80-104):
    per trajectory, every op, then each site whose after-op is that op: the
    operator chosen by apply-quantum-channel's rule (channel.clj:185-222:
    weights max |k_ij|^2, first running sum above u, else the last) applied
    UNSCALED and the state renormalised (apply-single-qubit-kraus-operator,
    channel.clj:135-183); finally one measure-state draw. Site k of
    trajectory t draws uniform(seed, t*2^20 + k), the measurement
    uniform(seed, t*2^20 + 2^20 - 1). make_u(qubit, m) builds the explicit
    2x2 op record. Returns (outcomes, normalised final states)."""
    outs = np.zeros(n_traj, dtype=np.uint64)
    finals = np.zeros((n_traj, 1 << n), dtype=np.complex128)
    for t in range(n_traj):
        psi = zero_state(n)
        k = 0
        for i in range(len(ops)):
            psi = apply_program(psi, n, ops[i:i + 1])
            while k < len(sites) and sites[k]["after-op"] == i:
                kr = sites[k]["kraus"]
                j = 0
                if len(kr) > 1:
                    u = uniform(seed, t * TRAJ_SLOTS + k)
                    acc = 0.0
                    for j in range(len(kr)):
                        acc += float(np.max(np.abs(np.asarray(kr[j])) ** 2))
                        if u < acc or j == len(kr) - 1:
                            break
                psi = apply_program(psi, n, make_u(sites[k]["qubit"], np.asarray(kr[j], dtype=np.complex128)))
                psi = normalize(psi, n)
                k += 1
        seed_m = (seed + (t * TRAJ_SLOTS + TRAJ_SLOTS - 1) * _PHI) % (1 << 64)
        outs[t] = sample(psi, n, 1, seed_m)[0]
        finals[t] = psi / np.sqrt(norm2(psi, n))
    return outs, finals

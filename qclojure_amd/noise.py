"""Noise models of QClojure's hardware simulator, restated for the GPU
trajectory path.

* Kraus operators: depolarizing, amplitude damping, phase damping, coherent
  error (src/org/soulspace/qclojure/domain/channel.clj:39-133) and the
  decoherence parameters from T1/T2/gate time (:224-242).
* Which channel follows which operation (noise.clj:53-103
  apply-gate-noise): looked up by the operation's own :operation-type (no
  alias resolution), on its :target qubit (0 when absent), noise-strength
  0.01 by default; unknown noise types apply nothing.
* Operator choice (channel.clj:185-222 apply-quantum-channel): weight_i =
  max |k_ij|^2 over the entries of operator i, the first operator whose
  running sum exceeds u, else the last one; a single operator is applied
  directly. The choice does not depend on the state. The chosen operator is
  applied and the state renormalised (:135-183), which is the same as applying
  any positive multiple of it, so the device applies K / max|k_ij|.
* Readout errors (noise.clj:108-202): qubits in order, flip with
  prob-0-to-1 / prob-1-to-0 times the product of correlation factors from
  qubits that already flipped, clamped to [0, 1].

The uniforms are explicit (qsv_uniform streams) so that runs are
reproducible; the reference draws them from an unseeded `rand`.
"""
from __future__ import annotations

import math

import numpy as np

from .circuit import CircuitError

_I = np.array([[1, 0], [0, 1]], dtype=np.complex128)
_X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)


class NoiseModelError(ValueError):
    """A noise configuration the reference rejects with a failed precondition."""


def depolarizing_kraus_operators(p: float) -> list[np.ndarray]:
    """channel.clj:39-63: sqrt(1-p) I, sqrt(p/3) X, sqrt(p/3) Y, sqrt(p/3) Z."""
    if not (0.0 <= p <= 1.0) or p > 0.75:
        raise NoiseModelError(f"depolarizing probability {p} outside [0, 0.75]")
    a, b = math.sqrt(1.0 - p), math.sqrt(p / 3.0)
    return [a * _I, b * _X, b * _Y, b * _Z]


def amplitude_damping_kraus_operators(gamma: float) -> list[np.ndarray]:
    """channel.clj:65-81: [[1,0],[0,sqrt(1-g)]], [[0,sqrt(g)],[0,0]]."""
    if not (0.0 <= gamma <= 1.0):
        raise NoiseModelError(f"damping parameter {gamma} outside [0, 1]")
    return [np.array([[1, 0], [0, math.sqrt(1.0 - gamma)]], dtype=np.complex128),
            np.array([[0, math.sqrt(gamma)], [0, 0]], dtype=np.complex128)]


def phase_damping_kraus_operators(gamma: float) -> list[np.ndarray]:
    """channel.clj:83-99: [[1,0],[0,sqrt(1-g)]], [[0,0],[0,sqrt(g)]]."""
    if not (0.0 <= gamma <= 1.0):
        raise NoiseModelError(f"dephasing parameter {gamma} outside [0, 1]")
    return [np.array([[1, 0], [0, math.sqrt(1.0 - gamma)]], dtype=np.complex128),
            np.array([[0, 0], [0, math.sqrt(gamma)]], dtype=np.complex128)]


def coherent_error_kraus_operator(angle: float, axis: str) -> np.ndarray:
    """channel.clj:101-122. :x -> [[c,-s],[s,c]], :y -> [[c,s],[-s,c]] with
    c, s = cos, sin of angle/2; :z -> diag(cos angle, cos(-angle)), which is
    a multiple of the identity and so no change after renormalisation."""
    c, s = math.cos(angle / 2.0), math.sin(angle / 2.0)
    if axis == "x":
        return np.array([[c, -s], [s, c]], dtype=np.complex128)
    if axis == "y":
        return np.array([[c, s], [-s, c]], dtype=np.complex128)
    if axis == "z":
        return np.array([[math.cos(angle), 0], [0, math.cos(-angle)]], dtype=np.complex128)
    raise NoiseModelError(f"rotation axis {axis!r} is not one of x, y, z")


def calculate_decoherence_params(t1: float, t2: float, gate_time: float) -> dict:
    """channel.clj:224-242: gamma = 1 - exp(-(gate_time / 1000) / T), gate
    time in ns, T1/T2 in microseconds; all three must be positive."""
    if not (t1 > 0 and t2 > 0 and gate_time > 0):
        raise NoiseModelError("calculate-decoherence-params needs positive t1, t2 and gate-time")
    us = gate_time / 1000.0
    return {"gamma-1": 1.0 - math.exp(-(us / t1)), "gamma-2": 1.0 - math.exp(-(us / t2))}


def _present(cfg: dict, key: str) -> bool:
    # Clojure truthiness of a map lookup: nil and false are falsy, 0.0 is not
    v = cfg.get(key)
    return v is not None and v is not False


def channel_for(noise_config: dict) -> list[np.ndarray] | None:
    """The Kraus operators apply-gate-noise uses for one gate's configuration
    (noise.clj:73-101), or None for no noise."""
    kind = noise_config.get("noise-type")
    strength = noise_config.get("noise-strength", 0.01)
    if kind == "depolarizing":
        return depolarizing_kraus_operators(strength)
    if kind == "amplitude-damping":
        if _present(noise_config, "t1-time") and _present(noise_config, "gate-time"):
            t1 = noise_config["t1-time"]
            t2 = noise_config["t2-time"] if _present(noise_config, "t2-time") else t1
            g = calculate_decoherence_params(t1, t2, noise_config["gate-time"])["gamma-1"]
        else:
            g = strength
        return amplitude_damping_kraus_operators(g)
    if kind == "phase-damping":
        if _present(noise_config, "t2-time") and _present(noise_config, "gate-time"):
            t2 = noise_config["t2-time"]
            t1 = noise_config["t1-time"] if _present(noise_config, "t1-time") else t2
            g = calculate_decoherence_params(t1, t2, noise_config["gate-time"])["gamma-2"]
        else:
            g = strength
        return phase_damping_kraus_operators(g)
    if kind == "coherent":
        ce = noise_config.get("coherent-error", {"rotation-angle": 0.01, "rotation-axis": "z"})
        return [coherent_error_kraus_operator(ce["rotation-angle"], ce["rotation-axis"])]
    return None


def selection_weights(kraus: list[np.ndarray]) -> np.ndarray:
    """channel.clj:203-212: max |k_ij|^2 of each operator."""
    return np.array([float(np.max(np.abs(k) ** 2)) for k in kraus])


def choose_operator(kraus: list[np.ndarray], u: float) -> int:
    """channel.clj:213-222: the first index whose running weight sum exceeds u,
    else the last one. A single operator is always chosen."""
    if len(kraus) == 1:
        return 0
    acc = 0.0
    w = selection_weights(kraus)
    for i in range(len(kraus)):
        acc += w[i]
        if u < acc or i >= len(kraus) - 1:
            return i
    return len(kraus) - 1


def noise_sites(circuit: dict, noise_model: dict) -> list[dict]:
    """Every (operation index, target qubit, Kraus operators) the hardware
    simulator would apply after the clean operation (noise.clj:66-72)."""
    gate_noise = (noise_model or {}).get("gate-noise") or {}
    n = circuit["num-qubits"]
    sites = []
    for i, op in enumerate(circuit["operations"]):
        cfg = gate_noise.get(op["operation-type"])
        if not cfg:
            continue
        kraus = channel_for(cfg)
        if kraus is None:
            continue
        q = (op.get("operation-params") or {}).get("target", 0)
        if not (0 <= q < n):
            raise CircuitError(f"noise target qubit {q} out of range")
        sites.append({"after-op": i, "qubit": int(q), "kraus": kraus})
    return sites


def apply_readout_noise(bits: str, readout_config: dict | None, uniforms) -> str:
    """calculate-final-bitstring (noise.clj:108-155) with explicit uniforms
    (one per qubit, in qubit order)."""
    if not readout_config:
        return bits
    p01, p10 = readout_config["prob-0-to-1"], readout_config["prob-1-to-0"]
    corr = readout_config.get("correlated-errors") or {}
    cur = list(bits)
    history: list[int] = []
    for q in range(len(bits)):
        base = p01 if cur[q] == "0" else p10
        factor = 1.0
        if corr and history:
            for src in history:
                f = (corr.get(src) or {}).get(q)
                if f is not None:
                    factor *= f
        eff = min(1.0, max(0.0, base * factor))
        if uniforms[q] < eff:
            cur[q] = "1" if cur[q] == "0" else "0"
            history.append(q)
    return "".join(cur)


def basis_string(index: int, n: int) -> str:
    """state.clj basis-string: qubit 0 is the leftmost (most significant) bit."""
    return format(index, f"0{n}b") if n else ""

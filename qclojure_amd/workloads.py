"""Seeded synthetic workloads: random-layer circuits, QAOA MaxCut instances,
QFT + U3/CZ precision circuits and a Poisson request stream.

All generators use numpy's PCG64 seeded from the caller's seed, so the same
seed gives the same circuit in tests, the smoke check and the benchmark.
"""
from __future__ import annotations

import math

import numpy as np

from . import algorithms as alg
from . import circuit as qc
from ._abi import OP, OP_DTYPE

_ONEQ = ("h", "rx", "ry", "rz")


def random_layer_circuit(n: int, layers: int, seed: int) -> dict:
    """Per layer: H, RX, RY or RZ (angle ~ U[0, 2pi)) on every qubit, then CNOTs on a
    uniformly random perfect matching with random orientation (one qubit idles when n is odd)."""
    rng = np.random.default_rng(seed)
    c = qc.create_circuit(n, f"random-{n}q-{layers}l-s{seed}")
    ops = []
    for _ in range(layers):
        choice = rng.integers(0, 4, size=n)
        angles = rng.uniform(0.0, 2.0 * math.pi, size=n)
        for q in range(n):
            g = _ONEQ[choice[q]]
            if g == "h":
                ops.append({"operation-type": "h", "operation-params": {"target": q}})
            else:
                ops.append({"operation-type": g, "operation-params": {"target": q, "angle": float(angles[q])}})
        perm = rng.permutation(n)
        flips = rng.integers(0, 2, size=n // 2)
        for k in range(n // 2):
            a, b = int(perm[2 * k]), int(perm[2 * k + 1])
            if flips[k]:
                a, b = b, a
            ops.append({"operation-type": "cnot", "operation-params": {"control": a, "target": b}})
    c["operations"] = ops
    return c


def random_layer_ops(n: int, layers: int, seed: int) -> np.ndarray:
    """The same circuit as random_layer_circuit, written straight to qsv_op
    records with the same random draws (no per-op Python objects)."""
    rng = np.random.default_rng(seed)
    per = n + n // 2
    out = np.zeros(layers * per, dtype=OP_DTYPE)
    kinds = np.array([OP["H"], OP["RX"], OP["RY"], OP["RZ"]], dtype=np.int32)
    for L in range(layers):
        choice = rng.integers(0, 4, size=n)
        angles = rng.uniform(0.0, 2.0 * math.pi, size=n)
        blk = out[L * per:L * per + n]
        blk["kind"] = kinds[choice]
        blk["q"][:, 0] = np.arange(n)
        blk["angle"] = np.where(choice == 0, 0.0, angles)
        perm = rng.permutation(n)
        flips = rng.integers(0, 2, size=n // 2)
        first, second = perm[0:2 * (n // 2):2], perm[1:2 * (n // 2):2]
        sw = flips.astype(bool)
        cn = out[L * per + n:(L + 1) * per]
        cn["kind"] = OP["CNOT"]
        cn["q"][:, 0] = np.where(sw, second, first)
        cn["q"][:, 1] = np.where(sw, first, second)
    return out


def qaoa_ops(inst: dict) -> np.ndarray:
    """A QAOA instance's ansatz (qaoa.clj:524-564) straight to qsv_op records:
    H on every qubit; per (gamma, beta): for each edge term in Hamiltonian order
    CNOT(q1,q2) RZ(q2, 2 gamma c) CNOT(q1,q2), then RX(q, 2 beta) on every qubit."""
    n = inst["n"]
    rows = [(OP["H"], q, 0.0, 0) for q in range(n)]
    hp = inst["hamiltonian"]
    ps = inst["params"]
    for k in range(0, len(ps), 2):
        gamma, beta = ps[k], ps[k + 1]
        for t in hp:
            idx = [i for i, ch in enumerate(t["pauli-string"]) if ch != "I"]
            if not idx:
                continue
            q1, q2 = idx
            ang = 2.0 * gamma * t["coefficient"]
            rows += [(OP["CNOT"], q1, 0.0, q2), (OP["RZ"], q2, ang, 0), (OP["CNOT"], q1, 0.0, q2)]
        for q in range(n):
            rows.append((OP["RX"], q, 2.0 * beta * 1.0, 0))
    out = np.zeros(len(rows), dtype=OP_DTYPE)
    out["kind"] = [r[0] for r in rows]
    out["q"][:, 0] = [r[1] for r in rows]
    out["q"][:, 1] = [r[3] for r in rows]
    out["angle"] = [r[2] for r in rows]
    return out


def qaoa_terms(inst: dict) -> tuple[np.ndarray, np.ndarray]:
    """The instance's MaxCut Hamiltonian as (Z masks, coefficients), in term
    order; bit q of a mask is Z on qubit q."""
    hp = inst["hamiltonian"]
    masks = np.zeros(len(hp), dtype=np.uint64)
    coeffs = np.zeros(len(hp), dtype=np.float64)
    for k, t in enumerate(hp):
        m = 0
        for q, ch in enumerate(t["pauli-string"]):
            if ch == "Z":
                m |= 1 << q
        masks[k] = m
        coeffs[k] = t["coefficient"]
    return masks, coeffs


def layer_slices(n: int, layers: int) -> list[slice]:
    per = n + n // 2
    return [slice(k * per, (k + 1) * per) for k in range(layers)]


def random_3_regular_graph(n: int, rng: np.random.Generator, max_tries: int = 10000) -> list:
    """Pairing model with rejection: 3n half-edges matched uniformly; retry on
    self-loops or multi-edges. n must be even."""
    if n % 2 or n < 4:
        raise ValueError("3-regular graphs need an even n >= 4")
    for _ in range(max_tries):
        stubs = np.repeat(np.arange(n), 3)
        rng.shuffle(stubs)
        edges = set()
        ok = True
        for k in range(0, 3 * n, 2):
            a, b = int(stubs[k]), int(stubs[k + 1])
            if a == b:
                ok = False
                break
            e = (min(a, b), max(a, b))
            if e in edges:
                ok = False
                break
            edges.add(e)
        if ok:
            return [[a, b, 1.0] for (a, b) in sorted(edges)]
    raise RuntimeError("could not draw a simple 3-regular graph")


def qaoa_instance(seed: int, n_choices=(8, 10, 12, 14, 16, 18, 20), p_choices=(1, 2, 3, 4)) -> dict:
    """One MaxCut QAOA instance: n from the even sizes, p in 1..4, gamma/beta ~ U[0, pi)."""
    rng = np.random.default_rng(seed)
    n = int(rng.choice(n_choices))
    p = int(rng.choice(p_choices))
    graph = random_3_regular_graph(n, rng)
    params = []
    for _ in range(p):
        params += [float(rng.uniform(0, math.pi)), float(rng.uniform(0, math.pi))]
    hp = alg.max_cut_hamiltonian(graph, n)
    hm = alg.standard_mixer_hamiltonian(n)
    circ = alg.qaoa_ansatz_circuit(hp, hm, params, n)
    return {"n": n, "p": p, "graph": graph, "params": params, "hamiltonian": hp, "circuit": circ}


def u3_ops(q: int, theta: float, phi: float, lam: float) -> list[dict]:
    """U3(theta, phi, lambda) lowered to the reference gate set as
    RZ(lambda), RY(theta), RZ(phi) in circuit order (global phase dropped)."""
    return [{"operation-type": "rz", "operation-params": {"target": q, "angle": lam}},
            {"operation-type": "ry", "operation-params": {"target": q, "angle": theta}},
            {"operation-type": "rz", "operation-params": {"target": q, "angle": phi}}]


def precision_circuit(n: int, layers: int, seed: int) -> dict:
    """Seeded random product state (RY then RZ per qubit), the reference QFT,
    then `layers` of random U3 on every qubit plus CZ on a random perfect matching."""
    rng = np.random.default_rng(seed)
    c = qc.create_circuit(n, f"precision-{n}q-{layers}l-s{seed}")
    ops = []
    for q in range(n):
        ops.append({"operation-type": "ry", "operation-params": {"target": q, "angle": float(rng.uniform(0, math.pi))}})
        ops.append({"operation-type": "rz", "operation-params": {"target": q, "angle": float(rng.uniform(0, 2 * math.pi))}})
    ops += alg.quantum_fourier_transform_circuit(n)["operations"]
    for _ in range(layers):
        ang = rng.uniform(0, 2 * math.pi, size=(n, 3))
        for q in range(n):
            ops += u3_ops(q, float(ang[q, 0]), float(ang[q, 1]), float(ang[q, 2]))
        perm = rng.permutation(n)
        for k in range(n // 2):
            ops.append({"operation-type": "cz", "operation-params": {"control": int(perm[2 * k]),
                                                                      "target": int(perm[2 * k + 1])}})
    c["operations"] = ops
    return c


def serving_stream(num_requests: int, rate_hz: float, seed: int, n_min: int = 10, n_max: int = 30,
                   depth_min: int = 10, depth_max: int = 100, shots: int = 1000) -> list[dict]:
    """Poisson arrivals; n = floor(n_min * (n_max/n_min)^U) (log-uniform), depth ~ U{depth_min..depth_max}."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(num_requests):
        t += float(rng.exponential(1.0 / rate_hz))
        n = int(math.floor(n_min * (n_max / n_min) ** float(rng.uniform())))
        n = max(n_min, min(n_max, n))
        depth = int(rng.integers(depth_min, depth_max + 1))
        out.append({"arrival-s": t, "n": n, "depth": depth, "seed": int(rng.integers(0, 2**31)),
                    "shots": shots})
    return out


_SELF_INVERSE = ("X", "Y", "Z", "H", "CNOT", "CZ", "SWAP", "TOFFOLI", "FREDKIN", "RYDBERG_CZ")
_NEGATE = ("RX", "RY", "RZ", "PHASE", "CRX", "CRY", "CRZ", "RYDBERG_CPHASE", "RYDBERG_BLOCKADE",
           "GLOBAL_RX", "GLOBAL_RY", "GLOBAL_RZ")


def inverse_ops(ops: np.ndarray) -> np.ndarray:
    """U^-1 of a program of self-inverse gates and rotations: the ops reversed,
    rotation and phase angles negated. Raises for kinds outside those sets. The
    reference's inverse-gate (circuit.clj:1128-1173) is not followed: it
    negates only RX/RY/RZ and returns controlled rotations and phase gates
    unchanged."""
    allowed_self = {OP[k] for k in _SELF_INVERSE}
    allowed_neg = {OP[k] for k in _NEGATE}
    out = ops[::-1].copy()
    for r in out:
        k = int(r["kind"])
        if k in allowed_neg:
            r["angle"] = -r["angle"]
        elif k not in allowed_self:
            raise ValueError(f"no inverse rule for op kind {k}")
    return out

"""Circuit data, builders, metadata and the op decoder.

Circuits are plain maps shaped like the reference's EDN/JSON form
(src/org/soulspace/qclojure/domain/circuit.clj:16-37)::

    {"operations": [{"operation-type": "h", "operation-params": {"target": 0}}, ...],
     "num-qubits": 2, "name": "...", "description": "..."}

Operation types and parameter keys are the reference's keywords without the
leading colon ("cnot", "control", "target", "qubit1", "measurement-qubits" ...).
Host-side work here is exact graph/integer work; amplitudes are never touched.
"""
from __future__ import annotations

from typing import Any, Iterable

import numpy as np

from ._abi import OP, OP_DTYPE

# ---- operation registry (operation_registry.clj:51-422) -------------------
OPERATION_CATALOG = frozenset([
    "x", "y", "z", "h", "s", "s-dag", "t", "t-dag", "i", "rx", "ry", "rz", "phase",
    "cnot", "cz", "cy", "crx", "cry", "crz", "swap", "iswap", "toffoli", "fredkin",
    "rydberg-cz", "rydberg-cphase", "rydberg-blockade",
    "global-x", "global-y", "global-z", "global-h", "global-rx", "global-ry", "global-rz",
    "measure",
])

NATIVE_SIMULATOR_GATE_SET = frozenset([
    "i", "x", "y", "z", "h", "s", "s-dag", "t", "t-dag", "rx", "ry", "rz", "phase",
    "cnot", "cz", "cy", "swap", "crx", "cry", "crz", "toffoli", "fredkin",
])

GATE_ALIASES = {
    "not": "x", "bit-flip": "x", "phase-flip": "z", "id": "i", "cx": "cnot",
    "ccx": "toffoli", "ccnot": "toffoli", "cswap": "fredkin", "p": "phase",
    "u1": "phase", "sdg": "s-dag", "tdg": "t-dag", "phaseshift": "phase",
    "si": "s-dag", "ti": "t-dag",
}


def resolve_gate_alias(name: str) -> str:
    """operation_registry.clj:409-422."""
    return GATE_ALIASES.get(name, name)


class CircuitError(ValueError):
    """An operation the reference dispatcher would reject (ex-info)."""


# ---- builders (circuit.clj:47-870) ---------------------------------------
def create_circuit(num_qubits: int, name: str | None = None, description: str | None = None) -> dict:
    c: dict[str, Any] = {"operations": [], "num-qubits": int(num_qubits)}
    if name is not None:
        c["name"] = name
    if description is not None:
        c["description"] = description
    return c


def add_operation(circuit: dict, operation_type: str, params: dict | None) -> dict:
    op: dict[str, Any] = {"operation-type": operation_type}
    if params:
        op["operation-params"] = dict(params)
    out = dict(circuit)
    out["operations"] = list(circuit["operations"]) + [op]
    return out


def add_gate(circuit: dict, gate_type: str, **params) -> dict:
    """add-gate (circuit.clj:111-150): resolves aliases, rejects unknown types."""
    resolved = resolve_gate_alias(gate_type)
    if resolved not in OPERATION_CATALOG:
        raise CircuitError(f"Unknown gate type: {gate_type}")
    params = {k.replace("_", "-"): v for k, v in params.items() if v is not None}
    return add_operation(circuit, resolved, params)


def x_gate(c, target=0): return add_gate(c, "x", target=target)
def y_gate(c, target=0): return add_gate(c, "y", target=target)
def z_gate(c, target=0): return add_gate(c, "z", target=target)
def i_gate(c, target=0): return add_gate(c, "i", target=target)
def h_gate(c, target=0): return add_gate(c, "h", target=target)
def s_gate(c, target=0): return add_gate(c, "s", target=target)
def s_dag_gate(c, target=0): return add_gate(c, "s-dag", target=target)
def t_gate(c, target=0): return add_gate(c, "t", target=target)
def t_dag_gate(c, target=0): return add_gate(c, "t-dag", target=target)
def rx_gate(c, target, angle): return add_gate(c, "rx", target=target, angle=angle)
def ry_gate(c, target, angle): return add_gate(c, "ry", target=target, angle=angle)
def rz_gate(c, target, angle): return add_gate(c, "rz", target=target, angle=angle)
def phase_gate(c, target, angle): return add_gate(c, "phase", target=target, angle=angle)
def cnot_gate(c, control, target): return add_gate(c, "cnot", control=control, target=target)
def cz_gate(c, control, target): return add_gate(c, "cz", control=control, target=target)
def cy_gate(c, control, target): return add_gate(c, "cy", control=control, target=target)
def crx_gate(c, control, target, angle): return add_gate(c, "crx", control=control, target=target, angle=angle)
def cry_gate(c, control, target, angle): return add_gate(c, "cry", control=control, target=target, angle=angle)
def crz_gate(c, control, target, angle): return add_gate(c, "crz", control=control, target=target, angle=angle)
def swap_gate(c, qubit1, qubit2): return add_gate(c, "swap", qubit1=qubit1, qubit2=qubit2)
def iswap_gate(c, qubit1, qubit2): return add_gate(c, "iswap", qubit1=qubit1, qubit2=qubit2)
def toffoli_gate(c, control1, control2, target):
    return add_gate(c, "toffoli", control1=control1, control2=control2, target=target)
def fredkin_gate(c, control, target1, target2):
    return add_gate(c, "fredkin", control=control, target1=target1, target2=target2)
def rydberg_cz_gate(c, control, target): return add_gate(c, "rydberg-cz", control=control, target=target)
def rydberg_cphase_gate(c, control, target, phi):
    return add_gate(c, "rydberg-cphase", control=control, target=target, angle=phi)
def rydberg_blockade_gate(c, qubit_indices, phi):
    return add_gate(c, "rydberg-blockade", qubit_indices=list(qubit_indices), angle=phi)
def global_x_gate(c): return add_gate(c, "global-x")
def global_y_gate(c): return add_gate(c, "global-y")
def global_z_gate(c): return add_gate(c, "global-z")
def global_hadamard_gate(c): return add_gate(c, "global-h")
def global_rx_gate(c, angle): return add_gate(c, "global-rx", angle=angle)
def global_ry_gate(c, angle): return add_gate(c, "global-ry", angle=angle)
def global_rz_gate(c, angle): return add_gate(c, "global-rz", angle=angle)


def measure_operation(c, qubits: Iterable[int]):
    return add_operation(c, "measure", {"measurement-qubits": list(qubits)})


def measure_all_operation(c):
    return measure_operation(c, range(c["num-qubits"]))


def bell_state_circuit() -> dict:
    """circuit.clj:1651-1673."""
    c = create_circuit(2, "Bell State", "Prepares (|00⟩ + |11⟩)/√2")
    return cnot_gate(h_gate(c, 0), 0, 1)


def ghz_state_circuit(n: int) -> dict:
    """circuit.clj:1675-1703."""
    if n < 2:
        raise CircuitError("GHZ circuit needs n >= 2")
    c = create_circuit(n, "GHZ State", f"Prepares {n}-qubit GHZ state")
    c = h_gate(c, 0)
    for i in range(1, n):
        c = cnot_gate(c, 0, i)
    return c


# ---- metadata (circuit.clj:1241-1525) --------------------------------------
_SINGLE = {"x", "y", "z", "h", "s", "s-dag", "t", "t-dag", "rx", "ry", "rz", "phase"}
_CTRL2 = {"cnot", "cx", "cz", "cy", "crx", "cry", "crz", "rydberg-cz", "rydberg-cphase"}
_GLOBAL = {"global-x", "global-y", "global-z", "global-h", "global-rx", "global-ry", "global-rz"}


def operation_qubits_with_spans(op: dict) -> dict:
    """circuit.clj:1241-1335 (including the global-gate case, which reads a
    :qubits key the operation map does not carry and so yields no qubits)."""
    p = op.get("operation-params", {}) or {}
    t = op["operation-type"]
    if t == "measure":
        return {"qubits": list(p.get("measurement-qubits", []) or []), "spans": []}
    if t in _SINGLE:
        return {"qubits": [p.get("target")], "spans": []}
    if t in _CTRL2:
        c, tg = p.get("control"), p.get("target")
        return {"qubits": [c, tg], "spans": [[min(c, tg), max(c, tg)]]}
    if t == "rydberg-blockade":
        qi = list(p["qubit-indices"])
        return {"qubits": qi, "spans": [[min(qi), max(qi)]]}
    if t in _GLOBAL:
        return {"qubits": list(range(len(op.get("qubits", []) or []))), "spans": []}
    if t in ("swap", "iswap"):
        a, b = p.get("qubit1"), p.get("qubit2")
        return {"qubits": [a, b], "spans": [[min(a, b), max(a, b)]]}
    if t == "toffoli":
        q = [p.get("control1"), p.get("control2"), p.get("target")]
        return {"qubits": q, "spans": [[min(q), max(q)]]}
    if t == "fredkin":
        q = [p.get("control"), p.get("target1"), p.get("target2")]
        return {"qubits": q, "spans": [[min(q), max(q)]]}
    keys = ["target", "control", "qubit1", "qubit2", "control1", "control2", "target1", "target2"]
    return {"qubits": [p[k] for k in keys if p.get(k) is not None], "spans": []}


def circuit_depth(circuit: dict) -> int:
    """circuit.clj:1375-1449."""
    ops = circuit["operations"]
    if not ops:
        return 0
    layers = [0] * circuit["num-qubits"]
    span_layers: dict[tuple, int] = {}
    max_layer = 0
    for op in ops:
        info = operation_qubits_with_spans(op)
        qs = [q for q in info["qubits"] if q is not None]
        spans = [tuple(s) for s in info["spans"]]
        mq = max((layers[q] for q in qs), default=0)
        ms = max((lay for (s, lay) in span_layers.items() for o in spans
                  if not (s[1] < o[0] or o[1] < s[0])), default=0)
        if not spans:
            mc = max((lay for (s, lay) in span_layers.items() for q in qs
                      if s[0] <= q <= s[1]), default=0)
        else:
            mc = 0
        new = 1 + max(mq, ms, mc)
        for q in qs:
            layers[q] = new
        for s in spans:
            span_layers[s] = new
        max_layer = max(max_layer, new)
    return max_layer


def circuit_operation_count(circuit: dict) -> int:
    return len(circuit["operations"])


def circuit_gate_count(circuit: dict) -> int:
    return sum(1 for o in circuit["operations"] if o["operation-type"] != "measure")


def circuit_metadata(circuit: dict) -> dict:
    return {"circuit-depth": circuit_depth(circuit),
            "circuit-operation-count": circuit_operation_count(circuit),
            "circuit-gate-count": circuit_gate_count(circuit)}


# ---- decoder: apply-gate-to-state dispatch (circuit.clj:889-1111) -----------
_ONE_Q = {"x": "X", "y": "Y", "z": "Z", "h": "H", "s": "S", "s-dag": "SDG", "t": "T", "t-dag": "TDG",
          "rx": "RX", "ry": "RY", "rz": "RZ", "phase": "PHASE"}
_NEEDS_ANGLE = {"rx", "ry", "rz", "phase"}
_CTRL = {"cnot": ("CNOT", False), "cz": ("CZ", False), "crz": ("CRZ", True), "crx": ("CRX", True),
         "cry": ("CRY", True), "rydberg-cz": ("RYDBERG_CZ", False), "rydberg-cphase": ("RYDBERG_CPHASE", True)}
_GLOBALK = {"global-rx": "GLOBAL_RX", "global-ry": "GLOBAL_RY", "global-rz": "GLOBAL_RZ", "global-h": "GLOBAL_H",
            "global-x": "GLOBAL_X", "global-y": "GLOBAL_Y", "global-z": "GLOBAL_Z"}
_REQUIRE_MSG = {
    "cnot": "CNOT requires both control and target qubits",
    "cz": "CZ requires both control and target qubits",
    "crz": "CRZ requires control, target qubits and angle",
    "crx": "CRX requires control, target qubits and angle",
    "cry": "CRY requires control, target qubits and angle",
    "rydberg-cz": "Rydberg CZ requires both control and target qubits",
    "rydberg-cphase": "Rydberg CPhase requires control, target qubits and phase angle",
}


def _qubit(v, n: int, what: str) -> int:
    if v is None or isinstance(v, bool) or int(v) != v:
        raise CircuitError(f"{what} must be a qubit index, got {v!r}")
    v = int(v)
    if not 0 <= v < n:
        raise CircuitError(f"{what} {v} out of range for {n} qubits")
    return v


def _mask(qubits, n: int, what: str) -> int:
    qs = [_qubit(q, n, what) for q in qubits]
    if not qs or len(set(qs)) != len(qs):
        raise CircuitError(f"{what}: qubits must be a non-empty list of distinct indices")
    m = 0
    for q in qs:
        m |= 1 << q
    return m


def decode_operation(op: dict, n: int) -> list[tuple]:
    """Decode one operation map into zero or more (kind, q0, q1, q2, mask, angle) tuples."""
    otype = op.get("operation-type")
    p = op.get("operation-params", {}) or {}
    if otype == "measure":
        mq = p.get("measurement-qubits")
        if not mq:
            raise CircuitError("Measure requires measurement-qubits parameter")
        return [(OP["MEASURE"], 0, 0, 0, _mask(mq, n, "measurement qubit"), 0.0)]
    g = resolve_gate_alias(otype)
    angle = p.get("angle")
    if g in _ONE_Q:
        if g in _NEEDS_ANGLE and angle is None:
            raise CircuitError(f"{g} requires an angle")
        t = p.get("target")
        t = 0 if t is None else t
        return [(OP[_ONE_Q[g]], _qubit(t, n, "target"), 0, 0, 0, float(angle or 0.0))]
    if g in _CTRL:
        kind, needs_angle = _CTRL[g]
        c, t = p.get("control"), p.get("target")
        if c is None or t is None or (needs_angle and angle is None):
            raise CircuitError(_REQUIRE_MSG[g])
        c, t = _qubit(c, n, "control"), _qubit(t, n, "target")
        if c == t:
            raise CircuitError("control and target must differ")
        return [(OP[kind], c, t, 0, 0, float(angle or 0.0))]
    if g in ("swap", "iswap"):
        a, b = p.get("qubit1"), p.get("qubit2")
        if a is None or b is None:
            raise CircuitError(f"{'SWAP' if g == 'swap' else 'iSWAP'} requires both qubit1 and qubit2 parameters")
        a, b = _qubit(a, n, "qubit1"), _qubit(b, n, "qubit2")
        if a == b:
            raise CircuitError("qubit1 and qubit2 must differ")
        return [(OP["SWAP" if g == "swap" else "ISWAP"], a, b, 0, 0, 0.0)]
    if g == "toffoli":
        c1, c2, t = p.get("control1"), p.get("control2"), p.get("target")
        if c1 is None or c2 is None or t is None:
            raise CircuitError("Toffoli requires control1, control2, and target parameters")
        q = [_qubit(c1, n, "control1"), _qubit(c2, n, "control2"), _qubit(t, n, "target")]
        if len(set(q)) != 3:
            raise CircuitError("Toffoli qubits must be distinct")
        return [(OP["TOFFOLI"], *q, 0, 0.0)]
    if g == "fredkin":
        c, t1, t2 = p.get("control"), p.get("target1"), p.get("target2")
        if c is None or t1 is None or t2 is None:
            raise CircuitError("Fredkin requires control, target1, and target2 parameters")
        q = [_qubit(c, n, "control"), _qubit(t1, n, "target1"), _qubit(t2, n, "target2")]
        if len(set(q)) != 3:
            raise CircuitError("Fredkin qubits must be distinct")
        return [(OP["FREDKIN"], *q, 0, 0.0)]
    if g == "rydberg-blockade":
        qi = p.get("qubit-indices")
        if not qi or angle is None:
            raise CircuitError("Rydberg blockade requires qubit-indices vector and phase angle")
        return [(OP["RYDBERG_BLOCKADE"], 0, 0, 0, _mask(qi, n, "blockade qubit"), float(angle))]
    if g in _GLOBALK:
        if g in ("global-rx", "global-ry", "global-rz") and angle is None:
            raise CircuitError(f"Global {g[-2:].upper()} requires rotation angle")
        return [(OP[_GLOBALK[g]], 0, 0, 0, 0, float(angle or 0.0))]
    if g == "controlled":
        # apply-general-controlled-gate (circuit.clj:889-930)
        cq, tg = p.get("control-qubits"), p.get("target-gate")
        if not cq or tg is None:
            raise CircuitError("Controlled gate requires control-qubits and target-gate")
        targets = p.get("target-qubits") or [0]
        if len(cq) == 1 and len(targets) == 1:
            c, t = _qubit(cq[0], n, "control"), _qubit(targets[0], n, "target")
            if tg == "x":
                if c == t:
                    raise CircuitError("control and target must differ")
                return [(OP["CNOT"], c, t, 0, 0, 0.0)]
            if tg == "z":
                return [(OP["CZ"], c, t, 0, 0, 0.0)] if c != t else _raise("control and target must differ")
            if tg == "y":
                return [(OP["CY"], c, t, 0, 0, 0.0)] if c != t else _raise("control and target must differ")
            # :h :s :t are applied to the target unconditionally; others are no-ops
            if tg == "h":
                return [(OP["H"], t, 0, 0, 0, 0.0)]
            if tg == "s":
                return [(OP["S"], t, 0, 0, 0, 0.0)]
            if tg == "t":
                return [(OP["T"], t, 0, 0, 0, 0.0)]
            return []
        raise CircuitError("Multi-controlled gates with >1 control qubit not fully implemented")
    raise CircuitError(f"Unknown gate type: {otype}")


def _raise(msg):
    raise CircuitError(msg)


def decode_circuit(circuit: dict) -> np.ndarray:
    """Whole circuit -> array of qsv_op records (OP_DTYPE)."""
    n = int(circuit["num-qubits"])
    rows: list[tuple] = []
    for op in circuit["operations"]:
        rows.extend(decode_operation(op, n))
    return ops_from_tuples(rows)


def ops_from_tuples(rows: list[tuple]) -> np.ndarray:
    out = np.zeros(len(rows), dtype=OP_DTYPE)
    for i, (kind, a, b, c, mask, angle) in enumerate(rows):
        out[i]["kind"] = kind
        out[i]["q"] = (a, b, c)
        out[i]["mask"] = mask
        out[i]["angle"] = angle
    return out


def op_u(target: int, m: np.ndarray) -> np.ndarray:
    """An explicit 2x2 unitary op (QSV_OP_U) on `target` (extension kind)."""
    rec = np.zeros(1, dtype=OP_DTYPE)
    rec[0]["kind"] = OP["U"]
    rec[0]["q"] = (target, 0, 0)
    rec[0]["m"] = np.ascontiguousarray(np.asarray(m, dtype=np.complex128).reshape(4)).view(np.float64)
    return rec

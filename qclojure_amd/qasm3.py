"""OpenQASM 3.0 text <-> circuit maps: the line-oriented converter
(src/org/soulspace/qclojure/application/format/qasm3.clj).

The wire format a serving front-end accepts circuits in: ``qasm_to_circuit``
turns QASM 3 text (plus ``#pragma qclojure result ...`` lines) into the
circuit maps that ``backend.execute_circuit`` runs on the device, and
``circuit_to_qasm`` writes a circuit back. The text written and the maps read
are the reference's, quirks included:

* numbers print as Clojure's ``str`` does (``Double.toString``: "1.5707963267948966",
  "1.0E-5"; integers without a point);
* one gate per matching line, regexes tried in the reference's order, every
  other line ignored; ``c[i] = measure q[i];`` adds one single-qubit measure;
* angles accept numbers, ``pi``, ``pi/n``, ``n*pi``, ``pi*n`` and fractions only;
* result pragmas read values as integers (all digits), string lists (with a
  comma) or strings; a repeated expectation/variance/sample pragma appends its
  observable and target to ``observables``/``targets`` of the first one.

Host-side text work only; amplitudes are never touched here.
"""
from __future__ import annotations

import math
import re

from . import circuit as qc
from .jvmtext import clj_keyword_str, clj_str, java_div, java_parse_double, java_parse_int, \
    JvmNumberFormatError


class QasmError(ValueError):
    """What the reference throws while converting (ex-info or a JVM exception)."""

    def __init__(self, message: str, data: dict | None = None):
        super().__init__(message)
        self.data = data or {}


# Gate descriptions of the operation registry (operation_registry.clj:55-363), used by
# the emitter's fallback for operation types that resolve through an alias.
GATE_DESCRIPTIONS = {
    "x": "Pauli-X (NOT) gate - bit flip operation",
    "y": "Pauli-Y gate - bit flip + phase flip operation",
    "z": "Pauli-Z gate - phase flip operation",
    "i": "Identity gate - leaves qubit state unchanged",
    "h": "Hadamard gate - creates superposition",
    "s": "S gate (π/2 phase gate)",
    "s-dag": "S† gate (−π/2 phase gate)",
    "t": "T gate (π/4 phase gate)",
    "t-dag": "T† gate (−π/4 phase gate)",
    "rx": "Rotation around X-axis",
    "ry": "Rotation around Y-axis",
    "rz": "Rotation around Z-axis",
    "phase": "Arbitrary phase gate",
    "cnot": "Controlled-NOT gate",
    "cz": "Controlled-Z gate",
    "cy": "Controlled-Y gate",
    "swap": "SWAP gate - exchanges two qubits",
    "iswap": "iSWAP gate - swap with phase",
    "crx": "Controlled rotation around X-axis",
    "cry": "Controlled rotation around Y-axis",
    "crz": "Controlled rotation around Z-axis",
    "toffoli": "Toffoli gate (CCX) - doubly controlled NOT",
    "fredkin": "Fredkin gate (CSWAP) - controlled SWAP",
    "rydberg-cz": "Rydberg blockade-based controlled-Z gate",
    "rydberg-cphase": "Rydberg controlled phase gate with arbitrary phase",
    "rydberg-blockade": "Multi-qubit Rydberg blockade gate",
    "global-rx": "Global X rotation applied to all qubits",
    "global-ry": "Global Y rotation applied to all qubits",
    "global-rz": "Global Z rotation applied to all qubits",
    "global-h": "Global Hadamard gate applied to all qubits",
    "global-x": "Global X gate applied to all qubits",
    "global-y": "Global Y gate applied to all qubits",
    "global-z": "Global Z gate applied to all qubits",
    "measure": "Computational basis measurement (Z-basis) - collapses qubit to |0⟩ or |1⟩",
}

PRAGMA_TARGET = {"qclojure": "qclojure", "braket": "braket"}


# ---- small JVM/Clojure helpers ----------------------------------------------------
def java_split(s: str, pattern: str, limit: int = 0) -> list[str]:
    """``String.split(regex, limit)``: limit 0 drops trailing empty strings,
    limit n > 0 splits at most n−1 times and keeps them."""
    parts = re.split(pattern, s, maxsplit=max(limit - 1, 0), flags=re.ASCII) if limit > 0 \
        else re.split(pattern, s, flags=re.ASCII)
    if parts and parts[0] == "" and len(s) > 0:
        m = re.match(pattern, s, flags=re.ASCII)
        if m is not None and m.end() == 0:
            parts = parts[1:]
    if limit == 0:
        while parts and parts[-1] == "":
            parts.pop()
        if not parts:
            return [""] if s == "" else []
    return parts


def java_trim(s: str) -> str:
    """``String.trim``: strips characters <= U+0020 at both ends."""
    i, j = 0, len(s)
    while i < j and ord(s[i]) <= 32:
        i += 1
    while j > i and ord(s[j - 1]) <= 32:
        j -= 1
    return s[i:j]


def split_lines(text: str) -> list[str]:
    """``clojure.string/split-lines`` (split on \\r?\\n, trailing empties dropped)."""
    return java_split(text, r"\r?\n")


def clj_name(v) -> str:
    """``(name x)`` for a keyword or string (the mirror holds both as str)."""
    if isinstance(v, str):
        return v
    raise QasmError(f"class {type(v).__name__} cannot be cast to clojure.lang.Named")


def _truthy(v) -> bool:
    return v is not None and v is not False


def _seq(v) -> list:
    return [] if v is None or v is False else list(v)


def _join(sep: str, xs) -> str:
    return sep.join(clj_str(x) for x in _seq(xs))


# ---- emission ------------------------------------------------------------------------
def observable_to_qasm_comment(observable) -> str:
    """observable-to-qasm-comment (qasm3.clj:33-42)."""
    if observable == "pauli-x":
        return "// Observable: Pauli-X"
    if observable == "pauli-y":
        return "// Observable: Pauli-Y"
    if observable == "pauli-z":
        return "// Observable: Pauli-Z"
    if isinstance(observable, dict):
        from .qio import edn_str
        return "// Observable: " + edn_str(observable)
    return "// Observable: " + clj_str(observable)


def emit_qasm_pragmas(options: dict | None) -> str | None:
    """emit-qasm-pragmas (qasm3.clj:44-119): one ``#pragma <target> result``
    line per requested result, under a comment header; None when there is
    nothing to emit."""
    options = options or {}
    result_specs = options.get("result-specs")
    target = options.get("target", "qclojure")
    if not result_specs:
        return None
    head = "#pragma " + clj_str(PRAGMA_TARGET.get(target))
    lines: list[str] = []
    for rtype, spec in result_specs.items():
        spec = spec if isinstance(spec, dict) else {}
        if rtype == "measurements":
            s = head + " result measurement shots=" + clj_str(spec.get("shots", 1000))
            if _truthy(spec.get("qubits")):
                s += " qubits=" + _join(",", spec["qubits"])
            lines.append(s)
        elif rtype in ("expectation", "variance"):
            obs, tg = _seq(spec.get("observables", [])), _seq(spec.get("targets", []))
            for i, o in enumerate(obs):
                s = f"{head} result {rtype} observable=" + clj_name(o)
                if i < len(tg):
                    s += " target=" + clj_str(tg[i])
                lines.append(s)
        elif rtype == "probability":
            s = head + " result probability"
            if _truthy(spec.get("targets")):
                s += " targets=" + _join(",", spec["targets"])
            if _truthy(spec.get("states")):
                s += " states=" + _join(",", spec["states"])
            lines.append(s)
        elif rtype == "amplitude":
            lines.append(head + " result amplitude states=" + _join(",", spec.get("states", [])))
        elif rtype == "sample":
            obs, tg = _seq(spec.get("observables", [])), _seq(spec.get("targets", []))
            shots = spec.get("shots", 1000)
            for i, o in enumerate(obs):
                s = head + " result sample observable=" + clj_name(o) + " shots=" + clj_str(shots)
                if i < len(tg):
                    s += " target=" + clj_str(tg[i])
                lines.append(s)
        elif rtype == "state-vector":
            lines.append(head + " result state_vector // Simulation-only result")
        elif rtype == "density-matrix":
            lines.append(head + " result density_matrix // Simulation-only result")
        elif rtype == "fidelity":
            lines.append(head + " result fidelity // Simulation-only result")
        else:
            lines.append("// Unknown result type: " + clj_name(rtype))
    if not lines:
        return None
    return "\n// Result extraction specifications\n" + "\n".join(lines) + "\n"


def gate_to_qasm_fn(num_qubits: int, options: dict | None = None):
    """gate-to-qasm-fn (qasm3.clj:121-218): a function from one operation
    map to its QASM line(s)."""
    target = (options or {}).get("target", "qclojure")
    braket = target == "braket"

    def q(v) -> str:
        return "q[" + clj_str(v) + "]"

    def emit(op: dict) -> str:
        g = op.get("operation-type")
        p = op.get("operation-params") or {}
        t, c, a = p.get("target"), p.get("control"), clj_str(p.get("angle"))
        if g == "i":
            return f"id {q(t)};"
        if g in ("x", "y", "z", "h", "s", "t"):
            return f"{g} {q(t)};"
        if g == "s-dag":
            return f"{'si' if braket else 'sdg'} {q(t)};"
        if g == "t-dag":
            return f"{'ti' if braket else 'tdg'} {q(t)};"
        if g in ("rx", "ry", "rz"):
            return f"{g}({a}) {q(t)};"
        if g == "phase":
            return f"{'phaseshift' if braket else 'p'}({a}) {q(t)};"
        if g == "cnot":
            return f"{'cnot' if braket else 'cx'} {q(c)}, {q(t)};"
        if g in ("cx", "cz", "cy"):
            return f"{g} {q(c)}, {q(t)};"
        if g in ("crx", "cry", "crz"):
            return f"{g}({a}) {q(c)}, {q(t)};"
        if g == "controlled":
            gate = p.get("gate")
            gs = clj_keyword_str(gate) if isinstance(gate, str) else clj_str(gate)
            return f"ctrl @ {gs} {q(c)}, {q(t)};"
        if g in ("swap", "iswap"):
            return f"{g} {q(p.get('qubit1'))}, {q(p.get('qubit2'))};"
        if g == "toffoli":
            return f"{'ccnot' if braket else 'ccx'} {q(p.get('control1'))}, {q(p.get('control2'))}, {q(t)};"
        if g == "fredkin":
            return f"cswap {q(c)}, {q(p.get('target1'))}, {q(p.get('target2'))};"
        rng = range(num_qubits)
        if g in ("global-x", "global-y", "global-z"):
            k = g[-1]
            return f"// Global {k.upper()} gate - apply {k.upper()} to all qubits\n" + \
                "\n".join(f"{k} q[{i}];" for i in rng)
        if g == "global-h":
            return "// Global Hadamard gate - apply H to all qubits\n" + "\n".join(f"h q[{i}];" for i in rng)
        if g in ("global-rx", "global-ry", "global-rz"):
            k = g[-2:]
            return f"// Global {k.upper()}({a}) gate - apply {k.upper()} to all qubits\n" + \
                "\n".join(f"{k}({a}) q[{i}];" for i in rng)
        if g == "rydberg-cz":
            return "// Rydberg CZ gate - decomposed to standard CZ\n" + f"cz {q(c)}, {q(t)};"
        if g == "rydberg-cphase":
            return "// Rydberg controlled phase gate - decomposed to CRZ\n" + f"crz({a}) {q(c)}, {q(t)};"
        if g == "rydberg-blockade":
            return ("// Rydberg blockade gate - hardware specific\n"
                    "// Cannot be directly expressed in standard QASM\n"
                    "// Requires hardware-specific backend support")
        if g == "measure":
            qs = _seq(p.get("measurement-qubits"))
            return "\n".join(f"c[{clj_str(x)}] = measure q[{clj_str(x)}];" for x in qs)
        name = clj_name(g)
        info = GATE_DESCRIPTIONS.get(qc.resolve_gate_alias(name))
        if info is not None:
            return f"// QClojure gate: {info}\n// Gate type: {name} - requires decomposition"
        return "// Unknown gate: " + name

    return emit


def circuit_to_qasm(circuit: dict, options: dict | None = None) -> str:
    """circuit-to-qasm (qasm3.clj:221-259): header, result pragmas, one blank
    line, then the operations one per line. ``options``: "result-specs" (a
    result-spec map) and "target" ("braket" drops the include line and uses
    Braket gate names)."""
    options = options or {}
    target = options.get("target", "qclojure")
    n = circuit.get("num-qubits")
    header = "OPENQASM 3.0;\n" + ("" if target == "braket" else 'include "stdgates.inc";\n\n') + \
        f"qubit[{clj_str(n)}] q;\n" + f"bit[{clj_str(n)}] c;\n"
    pragmas = emit_qasm_pragmas(options) or ""
    emit = gate_to_qasm_fn(n, options)
    gates = "\n".join(emit(op) for op in circuit.get("operations") or [])
    return header + pragmas + "\n" + gates


# ---- parsing ---------------------------------------------------------------------------
_DIGITS = re.compile(r"\d+", re.ASCII)


def _parse_long(v: str) -> int:
    return java_parse_int(v, bits=64)


def parse_result_pragma(line: str):
    """parse-result-pragma (qasm3.clj:265-286): ``[result-type, params]`` or
    None. ``k=v`` values: all digits → int, containing a comma → list of
    trimmed strings, else the string."""
    if not line.startswith("#pragma qclojure result"):
        return None
    parts = java_split(line, r"\s+")
    if len(parts) <= 3:
        return None
    rtype = parts[3]
    params: dict = {}
    for pair in parts[4:]:
        kv = java_split(pair, "=", 2)
        k = kv[0]
        if len(kv) < 2:
            raise QasmError(f"NullPointerException: pragma parameter '{pair}' has no value")
        v = kv[1]
        if _DIGITS.fullmatch(v):
            params[k] = _parse_long(v)
        elif "," in v:
            params[k] = [java_trim(x) for x in java_split(v, ",")]
        else:
            params[k] = v
    return [rtype, params]


def collect_result_specs_from_qasm(lines) -> dict:
    """collect-result-specs-from-qasm (qasm3.clj:342-367)."""
    specs: dict = {}
    for line in lines:
        if not java_trim(line).startswith("#pragma qclojure result"):
            continue
        parsed = parse_result_pragma(line)
        if parsed is None:
            continue
        rtype, params = parsed
        existing = specs.get(rtype)
        if existing is not None:
            if rtype in ("expectation", "variance", "sample"):
                e = dict(existing)
                e["observables"] = list(e.get("observables") or []) + [params.get("observable")]
                e["targets"] = list(e.get("targets") or []) + [params.get("target")]
                specs[rtype] = e
            else:
                specs[rtype] = params
        else:
            specs[rtype] = params
    return specs


_NUM = re.compile(r"-?\d+(\.\d+)?", re.ASCII)


def parse_qasm_expression(expr_str: str) -> float:
    """parse-qasm-expression (qasm3.clj:288-340): numbers, ``pi``, ``-pi``,
    ``pi/n``, ``n*pi``, ``pi*n`` and fractions ``a/b``; anything else that
    ``Double.parseDouble`` rejects raises QasmError. Division follows IEEE
    doubles (``pi/0`` is Infinity)."""
    expr = java_trim(expr_str)
    if _NUM.fullmatch(expr):
        return java_parse_double(expr)
    if expr == "pi":
        return math.pi
    if expr == "-pi":
        return -math.pi
    if re.fullmatch(r"-?pi/\d+(\.\d+)?", expr, re.ASCII):
        m = re.match(r"(-?)pi/(.+)$", expr)
        d = java_parse_double(m.group(2))
        return java_div(-math.pi, d) if m.group(1) == "-" else java_div(math.pi, d)
    if re.fullmatch(r"-?\d+(\.\d+)?\*pi", expr, re.ASCII):
        m = re.match(r"(.+)\*pi$", expr)
        return java_parse_double(m.group(1)) * math.pi
    if re.fullmatch(r"-?pi\*\d+(\.\d+)?", expr, re.ASCII):
        m = re.match(r"pi\*(.+)$", expr)
        if m is None:  # "-pi*n": the reference's extraction regex misses the sign
            raise QasmError("NullPointerException: cannot parse multiplier of " + expr)
        return math.pi * java_parse_double(m.group(1))
    if re.fullmatch(r"-?\d+(\.\d+)?/\d+(\.\d+)?", expr, re.ASCII):
        num, den = java_split(expr, "/")
        return java_div(java_parse_double(num), java_parse_double(den))
    try:
        return java_parse_double(expr)
    except JvmNumberFormatError:
        raise QasmError(f"Unsupported QASM expression: {expr}. Supported: numbers, pi, pi/n, n*pi, pi*n, fractions",
                        {"expression": expr, "type": "unsupported-qasm-expression"}) from None


_Q = r"q\[(\d+)\]"
_RX_1Q = re.compile(r"(x|y|z|h|s|t|sdg|tdg|id)\s+" + _Q, re.ASCII)
_RX_C1 = re.compile(r"(cx|cz|cy)\s+" + _Q + r",\s*" + _Q, re.ASCII)
_RX_SWAP = re.compile(r"swap\s+" + _Q + r",\s*" + _Q, re.ASCII)
_RX_ISWAP = re.compile(r"iswap\s+" + _Q + r",\s*" + _Q, re.ASCII)
_RX_CCX = re.compile(r"ccx\s+" + _Q + r",\s*" + _Q + r",\s*" + _Q, re.ASCII)
_RX_CSWAP = re.compile(r"cswap\s+" + _Q + r",\s*" + _Q + r",\s*" + _Q, re.ASCII)
_RX_CROT = re.compile(r"cr([xyz])\((.+?)\)\s+" + _Q + r",\s*" + _Q, re.ASCII)
_RX_P = re.compile(r"p\((.+?)\)\s+" + _Q, re.ASCII)
_RX_ROT = re.compile(r"r([xyz])\((.+?)\)\s+" + _Q, re.ASCII)
_RX_MEAS = re.compile(r"c\[(\d+)\]\s*=\s*measure\s+" + _Q, re.ASCII)


def _idx(s: str) -> int:
    return java_parse_int(s)


def qasm_to_gate(circuit: dict, line: str) -> dict:
    """qasm-to-gate (qasm3.clj:369-464): the circuit with the gate of one QASM
    line appended (patterns tried in the reference's order), or unchanged."""
    line = java_trim(line)
    m = _RX_1Q.match(line)
    if m:
        return qc.add_gate(circuit, m.group(1), target=_idx(m.group(2)))
    m = _RX_C1.match(line)
    if m:
        return qc.add_gate(circuit, m.group(1), control=_idx(m.group(2)), target=_idx(m.group(3)))
    m = _RX_SWAP.match(line)
    if m:
        return qc.add_gate(circuit, "swap", qubit1=_idx(m.group(1)), qubit2=_idx(m.group(2)))
    m = _RX_ISWAP.match(line)
    if m:
        return qc.add_gate(circuit, "iswap", qubit1=_idx(m.group(1)), qubit2=_idx(m.group(2)))
    m = _RX_CCX.match(line)
    if m:
        return qc.add_gate(circuit, "toffoli", control1=_idx(m.group(1)), control2=_idx(m.group(2)),
                           target=_idx(m.group(3)))
    m = _RX_CSWAP.match(line)
    if m:
        return qc.add_gate(circuit, "fredkin", control=_idx(m.group(1)), target1=_idx(m.group(2)),
                           target2=_idx(m.group(3)))
    m = _RX_CROT.match(line)
    if m:
        angle = parse_qasm_expression(m.group(2))
        return qc.add_gate(circuit, "cr" + m.group(1), control=_idx(m.group(3)), target=_idx(m.group(4)),
                           angle=angle)
    m = _RX_P.match(line)
    if m:
        angle = parse_qasm_expression(m.group(1))
        return qc.add_gate(circuit, "phase", target=_idx(m.group(2)), angle=angle)
    m = _RX_ROT.match(line)
    if m:
        angle = parse_qasm_expression(m.group(2))
        return qc.add_gate(circuit, "r" + m.group(1), target=_idx(m.group(3)), angle=angle)
    m = _RX_MEAS.match(line)
    if m:
        return qc.measure_operation(circuit, [_idx(m.group(2))])
    return circuit


def qasm_to_circuit(qasm: str) -> dict:
    """qasm-to-circuit (qasm3.clj:466-490): the circuit map, named "Converted
    Circuit", with the parsed ``result-specs`` attached. The qubit count is
    the first number on the first line that starts with ``qubit[``."""
    lines = split_lines(qasm)
    decl = next((ln for ln in lines if java_trim(ln).startswith("qubit[")), None)
    if decl is None:
        raise QasmError("NullPointerException: no qubit[n] declaration")
    m = _DIGITS.search(decl)
    if m is None:
        raise QasmError("NumberFormatException: Cannot parse null string: null")
    c = qc.create_circuit(_idx(m.group(0)), "Converted Circuit")
    specs = collect_result_specs_from_qasm(lines)
    for ln in lines:
        c = qasm_to_gate(c, ln)
    c = dict(c)
    c["result-specs"] = specs
    return c

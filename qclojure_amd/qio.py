"""Portable state and circuit data: the serialisation maps, their JSON and
EDN text, and file import/export by format
(src/org/soulspace/qclojure/adapter/io.clj, adapter/io/{json,edn,qasm}.clj).

A serialised state is ``{"state-vector": [{"real": re, "imag": im}, ...],
"num-qubits": n, "metadata": {...}, "format-version": "1.0"}``
(io.clj:33-49). States may come from the device: a ``DeviceState`` (or a
job's final state) is downloaded once, and ``state_to_device`` uploads a
deserialised state into HBM, so a front-end can ship states in and out of the
engine in the reference's format.

Text follows the reference's writers: JSON as clojure.data.json 2.5.2 writes
it (compact, ``/`` escaped as ``\\/``, non-ASCII as ``\\uXXXX``, doubles as
``Double.toString``, NaN/Infinity rejected), EDN as ``pr-str`` prints it
(``{:a 1, :b [2 3]}``). Keyword values are ``Keyword`` (a str subclass) so
that EDN keeps them apart from strings; map keys are keyword names.
"""
from __future__ import annotations

import json
import math
import os
import re

import numpy as np

from .jvmtext import java_double_str


class QuantumIOError(ValueError):
    """The reference's ex-info for unsupported data or formats."""

    def __init__(self, message: str, data: dict | None = None):
        super().__init__(message)
        self.data = data or {}


class Keyword(str):
    """A keyword value (printed ``:name`` in EDN, ``"name"`` in JSON)."""

    def __repr__(self):
        return ":" + str(self)


# ---- serialisation maps (io.clj:10-117) ---------------------------------------------
def complex_to_map(z) -> dict:
    """complex-to-map (io.clj:10-20)."""
    z = complex(z)
    return {"real": z.real, "imag": z.imag}


def map_to_complex(m: dict) -> complex:
    """map-to-complex (io.clj:22-31)."""
    return complex(float(m.get("real")), float(m.get("imag")))


def _state_amplitudes(state) -> tuple[np.ndarray, int, dict | None]:
    dev = getattr(state, "device_state", None) or state
    if hasattr(dev, "download") and hasattr(dev, "num_qubits"):
        return np.asarray(dev.download(), dtype=np.complex128), int(dev.num_qubits), None
    sv = state["state-vector"]
    return np.asarray(sv, dtype=np.complex128).reshape(-1), state.get("num-qubits"), state.get("metadata")


def serialize_quantum_state(state) -> dict:
    """serialize-quantum-state (io.clj:33-49). ``state`` is a state map, a
    ``DeviceState`` or a job's final state (downloaded once)."""
    amps, n, meta = _state_amplitudes(state)
    re_, im_ = amps.real.tolist(), amps.imag.tolist()
    return {"state-vector": [{"real": r, "imag": i} for r, i in zip(re_, im_)],
            "num-qubits": n,
            "metadata": meta or {},
            "format-version": "1.0"}


def deserialize_quantum_state(data: dict) -> dict:
    """deserialize-quantum-state (io.clj:51-62): amplitudes as a complex128
    vector, ``num-qubits`` and ``metadata`` as given."""
    sv = data.get("state-vector") or []
    amps = np.array([map_to_complex(m) for m in sv], dtype=np.complex128)
    return {"state-vector": amps, "num-qubits": data.get("num-qubits"), "metadata": data.get("metadata")}


def state_to_device(state: dict, precision: str = "c128"):
    """Upload a (deserialised) state map into a new device state."""
    from .engine import DeviceState
    amps = np.asarray(state["state-vector"], dtype=np.complex128)
    n = int(state["num-qubits"])
    if amps.size != 1 << n:
        raise QuantumIOError(f"state vector has {amps.size} amplitudes, {1 << n} expected for {n} qubits")
    dev = DeviceState(n, precision)
    dev.upload(amps)
    return dev


def serialize_quantum_circuit(circuit: dict) -> dict:
    """serialize-quantum-circuit (io.clj:64-78)."""
    return {"operations": [_keywordize_op(op) for op in circuit.get("operations") or []],
            "num-qubits": circuit.get("num-qubits"),
            "name": circuit.get("name"),
            "description": circuit.get("description"),
            "metadata": circuit.get("metadata") or {},
            "format-version": "1.0"}


def deserialize_quantum_circuit(data: dict) -> dict:
    """deserialize-quantum-circuit (io.clj:80-89): the circuit keys present."""
    keep = ("operations", "num-qubits", "name", "description", "metadata")
    return {k: data[k] for k in keep if k in data}


def _is_state(d) -> bool:
    return isinstance(d, dict) and "state-vector" in d and isinstance(d.get("num-qubits"), int) \
        and not isinstance(d.get("num-qubits"), bool)


def _is_circuit(d) -> bool:
    return isinstance(d, dict) and isinstance(d.get("operations"), list) \
        and isinstance(d.get("num-qubits"), int) and not isinstance(d.get("num-qubits"), bool) \
        and d["num-qubits"] > 0


def serialize_quantum_data(data) -> dict:
    """serialize-quantum-data (io.clj:91-103)."""
    if hasattr(data, "download") or hasattr(data, "device_state") or _is_state(data):
        return serialize_quantum_state(data)
    if _is_circuit(data):
        return serialize_quantum_circuit(data)
    raise QuantumIOError("Unsupported quantum data type", {"data": data})


def deserialize_quantum_data(data: dict):
    """deserialize-quantum-data (io.clj:105-117)."""
    if isinstance(data, dict) and "state-vector" in data:
        return deserialize_quantum_state(data)
    if isinstance(data, dict) and "operations" in data:
        return deserialize_quantum_circuit(data)
    raise QuantumIOError("Unsupported quantum data format", {"data": data})


_KEYWORD_VALUED = ("operation-type", "gate", "target-gate")


def _keywordize_op(op: dict) -> dict:
    out = dict(op)
    if isinstance(out.get("operation-type"), str):
        out["operation-type"] = Keyword(out["operation-type"])
    p = out.get("operation-params")
    if isinstance(p, dict):
        p = dict(p)
        for k in _KEYWORD_VALUED[1:]:
            if isinstance(p.get(k), str):
                p[k] = Keyword(p[k])
        out["operation-params"] = p
    return out


# ---- JSON (clojure.data.json 2.5.2 write-str / read-str :key-fn keyword) --------------
def _json_string(s: str) -> str:
    return json.dumps(s, ensure_ascii=True).replace("/", "\\/")


def _json_write(v, out: list) -> None:
    if v is None:
        out.append("null")
    elif v is True:
        out.append("true")
    elif v is False:
        out.append("false")
    elif isinstance(v, str):
        out.append(_json_string(str(v)))
    elif isinstance(v, (int, np.integer)):
        out.append(str(int(v)))
    elif isinstance(v, (float, np.floating)):
        x = float(v)
        if math.isnan(x):
            raise QuantumIOError("JSON error: cannot write Double NaN")
        if math.isinf(x):
            raise QuantumIOError("JSON error: cannot write infinite Double")
        out.append(java_double_str(x))
    elif isinstance(v, dict):
        out.append("{")
        for i, (k, x) in enumerate(v.items()):
            if i:
                out.append(",")
            if k is None:
                raise QuantumIOError("JSON object properties may not be nil")
            out.append(_json_string(k if isinstance(k, str) else _edn_key_str(k)))
            out.append(":")
            _json_write(x, out)
        out.append("}")
    elif isinstance(v, (list, tuple, np.ndarray)):
        out.append("[")
        for i, x in enumerate(v):
            if i:
                out.append(",")
            _json_write(x, out)
        out.append("]")
    elif isinstance(v, complex):
        raise QuantumIOError(f"Don't know how to write JSON of class {type(v).__name__}")
    else:
        raise QuantumIOError(f"Don't know how to write JSON of class {type(v).__name__}")


def json_str(v) -> str:
    """``json/write-str`` of a value."""
    out: list[str] = []
    _json_write(v, out)
    return "".join(out)


def _reject_constant(name):
    raise QuantumIOError(f"JSON error (unexpected character): {name}")


def json_read(text: str):
    """``json/read-str`` with keyword keys: objects, arrays, integers, doubles."""
    return json.loads(text, parse_constant=_reject_constant)


# ---- EDN (pr-str / edn/read-string) ------------------------------------------------------
def _edn_key_str(k) -> str:
    return str(k)


_EDN_ESC = {'"': '\\"', "\\": "\\\\", "\n": "\\n", "\t": "\\t", "\r": "\\r", "\f": "\\f", "\b": "\\b"}


def _edn_write(v, out: list, as_key: bool = False) -> None:
    if v is None:
        out.append("nil")
    elif v is True:
        out.append("true")
    elif v is False:
        out.append("false")
    elif isinstance(v, Keyword) or (as_key and isinstance(v, str)):
        out.append(":" + str(v))
    elif isinstance(v, str):
        out.append('"' + "".join(_EDN_ESC.get(ch, ch) for ch in v) + '"')
    elif isinstance(v, (int, np.integer)):
        out.append(str(int(v)))
    elif isinstance(v, (float, np.floating)):
        x = float(v)
        if math.isnan(x):
            out.append("##NaN")
        elif math.isinf(x):
            out.append("##Inf" if x > 0 else "##-Inf")
        else:
            out.append(java_double_str(x))
    elif isinstance(v, dict):
        out.append("{")
        for i, (k, x) in enumerate(v.items()):
            if i:
                out.append(", ")
            _edn_write(k, out, as_key=True)
            out.append(" ")
            _edn_write(x, out)
        out.append("}")
    elif isinstance(v, (list, np.ndarray)):
        out.append("[")
        for i, x in enumerate(v):
            if i:
                out.append(" ")
            _edn_write(x, out)
        out.append("]")
    elif isinstance(v, tuple):
        out.append("(")
        for i, x in enumerate(v):
            if i:
                out.append(" ")
            _edn_write(x, out)
        out.append(")")
    elif isinstance(v, (set, frozenset)):
        out.append("#{")
        for i, x in enumerate(v):
            if i:
                out.append(" ")
            _edn_write(x, out)
        out.append("}")
    else:
        raise QuantumIOError(f"cannot print {type(v).__name__} as EDN")


def edn_str(v) -> str:
    """``pr-str`` of a value (map keys print as keywords)."""
    out: list[str] = []
    _edn_write(v, out)
    return "".join(out)


class _EdnReader:
    _DELIM = set(" \t\r\n,;()[]{}\"")
    _NUM = re.compile(r"[+-]?\d+(\.\d*)?([eE][+-]?\d+)?[MN]?\Z")

    def __init__(self, text: str):
        self.s, self.i = text, 0

    def err(self, msg):
        raise QuantumIOError(f"EDN read error at {self.i}: {msg}")

    def ws(self):
        s = self.s
        while self.i < len(s):
            c = s[self.i]
            if c in " \t\r\n,":
                self.i += 1
            elif c == ";":
                while self.i < len(s) and s[self.i] != "\n":
                    self.i += 1
            elif s.startswith("#_", self.i):
                self.i += 2
                self.read()
            else:
                break

    def token(self) -> str:
        j = self.i
        while j < len(self.s) and self.s[j] not in self._DELIM:
            j += 1
        t, self.i = self.s[self.i:j], j
        return t

    def seq(self, close: str) -> list:
        out = []
        while True:
            self.ws()
            if self.i >= len(self.s):
                self.err("EOF while reading")
            if self.s[self.i] == close:
                self.i += 1
                return out
            out.append(self.read())

    def read(self):
        self.ws()
        if self.i >= len(self.s):
            self.err("EOF while reading")
        c = self.s[self.i]
        if c == "{":
            self.i += 1
            items = self.seq("}")
            if len(items) % 2:
                self.err("map literal must contain an even number of forms")
            return {(str(items[k]) if isinstance(items[k], str) else items[k]): items[k + 1]
                    for k in range(0, len(items), 2)}
        if c == "[":
            self.i += 1
            return self.seq("]")
        if c == "(":
            self.i += 1
            return tuple(self.seq(")"))
        if c == '"':
            return self.string()
        if c == "#":
            if self.s.startswith("#{", self.i):
                self.i += 2
                return frozenset(self.seq("}"))
            if self.s.startswith("##", self.i):
                self.i += 2
                t = self.token()
                vals = {"NaN": math.nan, "Inf": math.inf, "-Inf": -math.inf}
                if t not in vals:
                    self.err("unknown symbolic value " + t)
                return vals[t]
            self.err("unsupported dispatch")
        if c == ":":
            self.i += 1
            return Keyword(self.token())
        if c == "\\":
            self.i += 1
            t = self.s[self.i] + self.token() if self.i < len(self.s) else self.err("EOF")
            names = {"newline": "\n", "space": " ", "tab": "\t", "return": "\r", "formfeed": "\f",
                     "backspace": "\b"}
            return names.get(t, t)
        if c in ")]}":
            self.err("unmatched delimiter " + c)
        t = self.token()
        if t == "nil":
            return None
        if t == "true":
            return True
        if t == "false":
            return False
        if self._NUM.match(t):
            body = t.rstrip("MN")
            if "." in body or "e" in body or "E" in body or t.endswith("M"):
                return float(body)
            return int(body)
        return t

    def string(self) -> str:
        self.i += 1
        out = []
        esc = {"n": "\n", "t": "\t", "r": "\r", "f": "\f", "b": "\b", '"': '"', "\\": "\\"}
        s = self.s
        while True:
            if self.i >= len(s):
                self.err("EOF while reading string")
            ch = s[self.i]
            if ch == '"':
                self.i += 1
                return "".join(out)
            if ch == "\\":
                nx = s[self.i + 1:self.i + 2]
                if nx == "u":
                    out.append(chr(int(s[self.i + 2:self.i + 6], 16)))
                    self.i += 6
                    continue
                if nx not in esc:
                    self.err("unsupported escape \\" + nx)
                out.append(esc[nx])
                self.i += 2
                continue
            out.append(ch)
            self.i += 1


def edn_read(text: str):
    """``edn/read-string`` of the first form (None for empty input)."""
    r = _EdnReader(text)
    r.ws()
    if r.i >= len(text):
        return None
    return r.read()


# ---- file import/export by format (adapter/io/{json,edn,qasm}.clj) --------------------------
def _save(filename: str, content: str) -> str:
    d = os.path.dirname(filename)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(filename, "w", encoding="utf-8") as f:
        f.write(content)
    return filename


def _slurp(filename: str) -> str:
    with open(filename, encoding="utf-8") as f:
        return f.read()


def _no_method(fn: str, fmt):
    raise QuantumIOError(f"No method in multimethod '{fn}' for dispatch value: {fmt!r}")


def export_quantum_state(fmt: str, state, filename: str):
    """export-quantum-state (json.clj:11-14, edn.clj:11-14)."""
    if fmt == "json":
        return _save(filename, json_str(serialize_quantum_state(state)))
    if fmt == "edn":
        _save(filename, edn_str(serialize_quantum_state(state)))
        return None
    _no_method("export-quantum-state", fmt)


def import_quantum_state(fmt: str, filename: str) -> dict:
    """import-quantum-state (json.clj:16-19, edn.clj:16-18)."""
    if fmt == "json":
        return deserialize_quantum_state(json_read(_slurp(filename)))
    if fmt == "edn":
        return deserialize_quantum_state(edn_read(_slurp(filename)))
    _no_method("import-quantum-state", fmt)


def export_quantum_circuit(fmt: str, circuit: dict, filename: str):
    """export-quantum-circuit (json.clj:21-24, edn.clj:20-22, qasm.clj:23-25)."""
    if fmt == "json":
        return _save(filename, json_str(serialize_quantum_circuit(circuit)))
    if fmt == "edn":
        return _save(filename, edn_str(serialize_quantum_circuit(circuit)))
    if fmt == "qasm3":
        from .qasm3 import circuit_to_qasm
        return _save(filename, circuit_to_qasm(circuit))
    _no_method("export-quantum-circuit", fmt)


def import_quantum_circuit(fmt: str, filename: str) -> dict:
    """import-quantum-circuit (json.clj:26-29, edn.clj:24-26 — the EDN map as
    read, qasm.clj:27-29)."""
    if fmt == "json":
        return deserialize_quantum_circuit(json_read(_slurp(filename)))
    if fmt == "edn":
        return edn_read(_slurp(filename))
    if fmt == "qasm3":
        from .qasm3 import qasm_to_circuit
        return qasm_to_circuit(_slurp(filename))
    _no_method("import-quantum-circuit", fmt)


def export_quantum_data(fmt: str, data, filename: str):
    """export-quantum-data (json.clj:31-34, edn.clj:28-30)."""
    if fmt == "json":
        return _save(filename, json_str(serialize_quantum_data(data)))
    if fmt == "edn":
        return _save(filename, edn_str(serialize_quantum_data(data)))
    _no_method("export-quantum-data", fmt)


def import_quantum_data(fmt: str, filename: str):
    """import-quantum-data (json.clj:36-39, edn.clj:32-34)."""
    if fmt == "json":
        return deserialize_quantum_data(json_read(_slurp(filename)))
    if fmt == "edn":
        return deserialize_quantum_data(edn_read(_slurp(filename)))
    _no_method("import-quantum-data", fmt)

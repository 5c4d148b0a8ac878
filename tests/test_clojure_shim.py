"""Static checks of the Clojure FFM shim (gpu_simulator.clj) against the C ABI.

There is no JVM here, so the shim cannot run; these checks pin what a typo
would break: every downcall names a symbol include/qsv.h declares, with the
argument layouts of that declaration; the op-kind numbers agree with enum
qsv_op_kind; and the dispatcher's rejections (:cy, :i as direct operation
types, src/org/soulspace/qclojure/domain/circuit.clj:964-1072) are kept."""
import ctypes
import os
import re

from qclojure_amd._abi import OP, SIGNATURES

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "qclojure_amd", "clj", "org", "soulspace", "qclojure", "adapter", "backend",
                    "gpu_simulator.clj")


def _src():
    with open(SHIM) as f:
        return f.read()


def _layout(ct):
    """The FFM value layout letter of a ctypes argument type."""
    if ct in (ctypes.c_int, ctypes.c_int32, ctypes.c_uint32):
        return "I"
    if ct in (ctypes.c_uint64, ctypes.c_int64, ctypes.c_long):
        return "J"
    if ct is ctypes.c_double:
        return "D"
    return "P"   # pointers, strings and structs by reference


def test_downcalls_match_header():
    downs = re.findall(r'\(defdown\s+\S+\s+"(qsv_\w+)"\s+([^)]*)\)', _src())
    assert len(downs) >= 20
    for sym, layouts in downs:
        assert sym in SIGNATURES, f"{sym} is not declared in include/qsv.h"
        res, args = SIGNATURES[sym]
        toks = layouts.split()
        ret, got = toks[0], toks[1:]
        want = [_layout(a) for a in args]
        assert got == want, f"{sym}: shim passes {got}, header takes {want}"
        exp_ret = "P" if res is ctypes.c_char_p else ("D" if res is ctypes.c_double else "I")
        assert ret == exp_ret, f"{sym}: return layout {ret}, header returns {exp_ret}"


def _kind_map(name):
    m = re.search(r"\(def \^:private " + name + r"\s*\{([^}]*)\}", _src())
    assert m, name
    return {k: int(v) for k, v in re.findall(r":([\w-]+)\s+\[?(\d+)", m.group(1))}


def test_kind_numbers_match_enum():
    one = _kind_map("one-qubit-kind")
    names = {"x": "X", "y": "Y", "z": "Z", "h": "H", "s": "S", "s-dag": "SDG", "t": "T", "t-dag": "TDG",
             "rx": "RX", "ry": "RY", "rz": "RZ", "phase": "PHASE"}
    assert one == {k: OP[v] for k, v in names.items()}
    ct = _kind_map("control-target")
    assert ct == {"cnot": OP["CNOT"], "cz": OP["CZ"], "crx": OP["CRX"], "cry": OP["CRY"], "crz": OP["CRZ"],
                  "rydberg-cz": OP["RYDBERG_CZ"], "rydberg-cphase": OP["RYDBERG_CPHASE"]}
    gl = _kind_map("global-kind")
    assert gl == {k: OP[k.upper().replace("-", "_")] for k in gl}
    assert len(gl) == 7


def test_dispatcher_rejections_kept():
    src = _src()
    # neither :cy nor :i may be decodable as a direct operation type
    for table in ("one-qubit-kind", "control-target", "global-kind"):
        assert "cy" not in _kind_map(table) and "i" not in _kind_map(table)
    assert "(throw (ex-info \"Unknown gate type\"" in src
    # :controlled lowers :y to CY (kind 15) and fails more than one control
    assert re.search(r":y \(do \(distinct! [^)]*\) \[\(rec 15 c t\)\]\)", src)
    assert "Multi-controlled gates with >1 control qubit not fully implemented" in src


def test_every_protocol_method_has_a_body():
    src = _src()
    for m in ("backend-info", "device", "available?", "submit-circuit", "job-status", "job-result",
              "cancel-job", "queue-status", "batch-submit", "batch-status", "batch-results"):
        assert re.search(r"\(" + re.escape(m) + r" \[", src), m
    assert "..." not in src.split("(defrecord GpuSimulator", 1)[1]

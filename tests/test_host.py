"""Host-side logic: op decoding (dispatcher semantics and quirks), circuit
metadata, the QAOA/QFT builders and the C ABI surface of libqsv (loaded, not
called, so this runs without a GPU)."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from qclojure_amd import algorithms as alg
from qclojure_amd import circuit as qc
from qclojure_amd import hamiltonian as ham
from qclojure_amd import workloads as wl
from qclojure_amd._abi import OP, OP_DTYPE, QsvOp, SIGNATURES, LIB_PATH

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_op_record_layout():
    assert ctypes.sizeof(QsvOp) == OP_DTYPE.itemsize == 96
    for name in ("kind", "q", "mask", "angle", "m"):
        assert OP_DTYPE.fields[name][1] == getattr(QsvOp, name).offset


def test_header_enum_matches_python():
    src = open(os.path.join(ROOT, "include", "qsv.h")).read()
    body = re.search(r"enum qsv_op_kind \{(.*?)\};", src, re.S).group(1)
    names = re.findall(r"QSV_OP_([A-Z0-9_]+)", body)
    names = [n for n in names if n != "KIND_END"]
    assert names[0] == "X" and {n: i + 1 for i, n in enumerate(names)} == OP


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB_PATH), "libqsv.so is not built"
    hdr = open(os.path.join(ROOT, "include", "qsv.h")).read()
    declared = set(re.findall(r"^\s*(?:int|double|const char\*)\s+(qsv_[a-z0-9_]+)\s*\(", hdr, re.M))
    assert declared, "no declarations parsed"
    lib = ctypes.CDLL(LIB_PATH)
    for name in sorted(declared):
        assert hasattr(lib, name), f"libqsv.so does not export {name}"
    assert declared == set(SIGNATURES), "python bindings and header disagree"


def test_aliases_resolve():
    # operation_registry.clj:387-422
    assert qc.resolve_gate_alias("cx") == "cnot"
    assert qc.resolve_gate_alias("ccx") == "toffoli"
    assert qc.resolve_gate_alias("cswap") == "fredkin"
    assert qc.resolve_gate_alias("u1") == "phase"
    assert qc.resolve_gate_alias("sdg") == "s-dag"
    assert qc.resolve_gate_alias("h") == "h"
    ops = qc.decode_circuit({"num-qubits": 3, "operations": [
        {"operation-type": "cx", "operation-params": {"control": 0, "target": 1}},
        {"operation-type": "not", "operation-params": {"target": 2}},
        {"operation-type": "p", "operation-params": {"target": 0, "angle": 0.5}}]})
    assert [int(k) for k in ops["kind"]] == [OP["CNOT"], OP["X"], OP["PHASE"]]


def test_dispatcher_rejections():
    def dec(op, n=3):
        return qc.decode_circuit({"num-qubits": n, "operations": [op]})
    # :cy and :i are catalogued but the dispatcher has no branch (circuit.clj:964-1072)
    for t in ("cy", "i", "bogus"):
        with pytest.raises(qc.CircuitError, match="Unknown gate type"):
            dec({"operation-type": t, "operation-params": {"control": 0, "target": 1}})
    with pytest.raises(qc.CircuitError, match="CNOT requires"):
        dec({"operation-type": "cnot", "operation-params": {"target": 1}})
    with pytest.raises(qc.CircuitError, match="CRZ requires"):
        dec({"operation-type": "crz", "operation-params": {"control": 0, "target": 1}})
    with pytest.raises(qc.CircuitError, match="measurement-qubits"):
        dec({"operation-type": "measure", "operation-params": {}})
    with pytest.raises(qc.CircuitError, match="Multi-controlled"):
        dec({"operation-type": "controlled",
             "operation-params": {"control-qubits": [0, 1], "target-gate": "x", "target-qubits": [2]}})
    with pytest.raises(qc.CircuitError, match="out of range"):
        dec({"operation-type": "h", "operation-params": {"target": 3}})


def test_controlled_quirks():
    """apply-general-controlled-gate (circuit.clj:889-930): :x/:z/:y are
    controlled; :h/:s/:t act on the target unconditionally; others vanish."""
    def dec(tg):
        return qc.decode_circuit({"num-qubits": 2, "operations": [
            {"operation-type": "controlled",
             "operation-params": {"control-qubits": [0], "target-gate": tg, "target-qubits": [1]}}]})
    assert int(dec("x")[0]["kind"]) == OP["CNOT"]
    assert int(dec("z")[0]["kind"]) == OP["CZ"]
    assert int(dec("y")[0]["kind"]) == OP["CY"]
    h = dec("h")[0]
    assert int(h["kind"]) == OP["H"] and int(h["q"][0]) == 1
    assert len(dec("rx")) == 0
    # target-qubits defaults to [0]
    d = qc.decode_circuit({"num-qubits": 2, "operations": [
        {"operation-type": "controlled", "operation-params": {"control-qubits": [1], "target-gate": "x"}}]})
    assert list(d[0]["q"][:2]) == [1, 0]


def test_default_targets():
    # 1-qubit gates without :target act on qubit 0 ((or target 0), circuit.clj:966-984)
    d = qc.decode_circuit({"num-qubits": 2, "operations": [{"operation-type": "h"},
                                                           {"operation-type": "s"}]})
    assert list(d["q"][:, 0]) == [0, 0]


def test_circuit_metadata():
    bell = qc.bell_state_circuit()
    assert qc.circuit_depth(bell) == 2
    assert qc.circuit_operation_count(bell) == 2 and qc.circuit_gate_count(bell) == 2
    c = qc.measure_all_operation(qc.ghz_state_circuit(3))
    assert qc.circuit_gate_count(c) == 3 and qc.circuit_operation_count(c) == 4
    # parallel single-qubit gates share a layer; a CNOT(0,2) span blocks qubit 1
    c = qc.create_circuit(3)
    c = qc.h_gate(qc.h_gate(qc.h_gate(c, 0), 1), 2)
    assert qc.circuit_depth(c) == 1
    c = qc.h_gate(qc.cnot_gate(c, 0, 2), 1)
    assert qc.circuit_depth(c) == 3


def test_builders_reject_unknown():
    with pytest.raises(qc.CircuitError):
        qc.add_gate(qc.create_circuit(2), "controlled", control=0, target=1)


def test_maxcut_kats(kats):
    for case in kats["maxcut-cases"]:
        h = alg.max_cut_hamiltonian(case["graph"], case["n"])
        assert len(h) == case["num-terms"]
        assert ham.validate_hamiltonian(h)
        if "coefficient-set" in case:
            assert {t["coefficient"] for t in h} == set(case["coefficient-set"])


def test_qaoa_circuit_shape():
    inst = wl.qaoa_instance(5)
    n, p = inst["n"], inst["p"]
    ops = inst["circuit"]["operations"]
    edges = len(inst["graph"])
    assert edges == 3 * n // 2
    assert len(ops) == n + p * (3 * edges + n)
    # ZZ terms use the sorted qubit pair: CNOT(q1,q2) RZ(q2, -gamma w) CNOT(q1,q2)
    first = ops[n:n + 3]
    (i, j, w) = inst["graph"][0]
    assert first[0]["operation-params"] == {"control": min(i, j), "target": max(i, j)}
    assert first[1]["operation-params"]["target"] == max(i, j)
    assert first[1]["operation-params"]["angle"] == 2.0 * inst["params"][0] * (-(w / 2.0))
    # mixer: RX(q, 2 beta)
    mix = ops[n + 3 * edges]
    assert mix["operation-type"] == "rx" and mix["operation-params"]["angle"] == 2.0 * inst["params"][1] * 1.0


def test_qft_circuit_shape():
    c = alg.quantum_fourier_transform_circuit(32)
    kinds = [o["operation-type"] for o in c["operations"]]
    assert kinds.count("h") == 32 and kinds.count("crz") == 496 and kinds.count("swap") == 16
    crz = [o for o in c["operations"] if o["operation-type"] == "crz"][0]
    assert crz["operation-params"] == {"control": 1, "target": 0, "angle": math.pi / 2}


def test_hamiltonian_grouping():
    h = ham.hamiltonian([[1.0, "XX"], [1.0, "ZZ"], [0.5, "XI"], [0.5, "ZI"], [2.0, "II"]])
    groups = ham.group_commuting_terms(h)
    assert sum(len(g) for g in groups) == len(h)
    bases = ham.group_pauli_terms_by_measurement_basis(h)
    assert set(bases) == {"x", "z", "identity"}


def test_fast_generators_match_circuit_maps():
    """The benchmark's direct-to-record generators reproduce the circuit-map path."""
    for (n, layers, seed) in [(12, 20, 1), (30, 3, 9), (7, 5, 4), (1, 2, 3)]:
        assert np.array_equal(wl.random_layer_ops(n, layers, seed),
                              qc.decode_circuit(wl.random_layer_circuit(n, layers, seed)))
    for seed in range(4):
        inst = wl.qaoa_instance(seed)
        assert np.array_equal(wl.qaoa_ops(inst), qc.decode_circuit(inst["circuit"]))


def test_precision_circuit_shape():
    c = wl.precision_circuit(8, 3, seed=2)
    kinds = [o["operation-type"] for o in c["operations"]]
    assert kinds.count("crz") == 28 and kinds.count("swap") == 4
    assert kinds.count("cz") == 3 * 4 and len(kinds) == 2 * 8 + (8 + 28 + 4) + 3 * (3 * 8 + 4)


def test_workloads_deterministic():
    a = wl.random_layer_ops(30, 3, seed=9)
    b = wl.random_layer_ops(30, 3, seed=9)
    assert np.array_equal(a, b) and len(a) == 3 * 45
    s = wl.serving_stream(50, 100.0, seed=1)
    assert all(10 <= r["n"] <= 30 and 10 <= r["depth"] <= 100 for r in s)
    assert all(s[i]["arrival-s"] < s[i + 1]["arrival-s"] for i in range(len(s) - 1))

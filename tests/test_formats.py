"""Wire formats: QASM 3 (line converter and grammar parser), JSON/EDN state
and circuit data, and the JVM number text they use. Cases follow the
reference's test/org/soulspace/qclojure/application/format/qasm3_test.clj,
qasm3_parser_test.clj and adapter/io/qasm_test.clj; number text is pinned
against the JDK's documented Double.toString / parseDouble results."""
import math
import time

import numpy as np
import pytest

from qclojure_amd import circuit as qc
from qclojure_amd import qasm3
from qclojure_amd import qasm3_parser as q3p
from qclojure_amd import qio
from qclojure_amd.jvmtext import java_double_str, java_parse_double, JvmNumberFormatError

PI = math.pi


def lines(text):
    return text.split("\n")


def bell_measured():
    return qc.measure_operation(qc.cnot_gate(qc.h_gate(qc.create_circuit(2, "Bell State"), 0), 0, 1), [0, 1])


def complex_circuit():
    c = qc.create_circuit(3, "Complex Circuit")
    c = qc.h_gate(c, 0); c = qc.x_gate(c, 1); c = qc.y_gate(c, 2); c = qc.z_gate(c, 0)
    c = qc.s_gate(c, 1); c = qc.t_gate(c, 2); c = qc.s_dag_gate(c, 0); c = qc.t_dag_gate(c, 1)
    c = qc.cnot_gate(c, 0, 1); c = qc.cz_gate(c, 1, 2); c = qc.cy_gate(c, 0, 2)
    c = qc.swap_gate(c, 0, 2); c = qc.phase_gate(c, 1, PI / 6)
    return c


def advanced_circuit():
    c = qc.create_circuit(4, "Advanced Gates")
    c = qc.iswap_gate(c, 0, 1); c = qc.toffoli_gate(c, 0, 1, 2); c = qc.fredkin_gate(c, 3, 0, 2)
    c = qc.crx_gate(c, 1, 3, PI / 4); c = qc.cry_gate(c, 2, 0, PI / 3); c = qc.crz_gate(c, 0, 3, PI / 8)
    return c


# ---- Double.toString / parseDouble (java/lang/Double.java specification, JDK 19+) ----
@pytest.mark.parametrize("x,text", [
    (1.0, "1.0"), (0.1, "0.1"), (100.0, "100.0"), (1e7, "1.0E7"), (9999999.0, "9999999.0"),
    (1e-3, "0.001"), (9.999e-4, "9.999E-4"), (1e-5, "1.0E-5"), (1.0e21, "1.0E21"), (1e23, "1.0E23"),
    (123456789.0, "1.23456789E8"), (-0.0, "-0.0"), (0.0, "0.0"), (PI / 2, "1.5707963267948966"),
    (5e-324, "4.9E-324"), (1.7976931348623157e308, "1.7976931348623157E308"),
    (2.2250738585072014e-308, "2.2250738585072014E-308"), (2.0 ** 63, "9.223372036854776E18"),
    (math.inf, "Infinity"), (-math.inf, "-Infinity"), (math.nan, "NaN"), (-1.5, "-1.5"),
])
def test_java_double_str(x, text):
    assert java_double_str(x) == text
    if math.isfinite(x):
        assert java_parse_double(text) == x


def test_java_double_str_round_trips_random_doubles():
    rng = np.random.default_rng(7)
    bits = rng.integers(0, 2 ** 63 - 2 ** 52, size=2000, dtype=np.int64)   # finite positive doubles
    for b in bits:
        x = float(np.int64(b).view(np.float64))
        t = java_double_str(x)
        assert java_parse_double(t) == x
        mant = t.split("E")[0].replace("-", "").replace(".", "").strip("0")
        assert len(mant) <= 17


@pytest.mark.parametrize("text,value", [
    ("1.5", 1.5), ("  2 ", 2.0), ("1e3", 1000.0), ("-.5", -0.5), ("1.", 1.0), ("1.5f", 1.5), ("2d", 2.0),
    ("0x1.8p1", 3.0), ("+7", 7.0), ("-Infinity", -math.inf),
])
def test_java_parse_double(text, value):
    assert java_parse_double(text) == value


@pytest.mark.parametrize("text", ["1_0", "pi", "", "1e", "inf", "nan", "0x1.8"])
def test_java_parse_double_rejects(text):
    with pytest.raises(JvmNumberFormatError):
        java_parse_double(text)


# ---- qasm3.clj (qasm3_test.clj) --------------------------------------------------------------
def test_header():
    text = qasm3.circuit_to_qasm(qc.create_circuit(3, "Test Circuit"))
    assert text.startswith("OPENQASM 3.0;")
    assert "qubit[3] q;" in lines(text) and "bit[3] c;" in lines(text)
    assert "OPENQASM 2.0;" not in text and "qreg q[3];" not in text and "creg c[3];" not in text


def test_basic_and_phase_gates():
    c = qc.create_circuit(2)
    for g, t in (("x", 0), ("y", 1), ("z", 0), ("h", 1), ("s", 0), ("t", 1), ("s-dag", 0), ("t-dag", 1)):
        c = qc.add_gate(c, g, target=t)
    ls = lines(qasm3.circuit_to_qasm(c))
    for want in ("x q[0];", "y q[1];", "z q[0];", "h q[1];", "s q[0];", "t q[1];", "sdg q[0];", "tdg q[1];"):
        assert want in ls


def test_controlled_rotation_swap_three_qubit_gates():
    c = qc.create_circuit(4)
    c = qc.cnot_gate(c, 0, 1); c = qc.cz_gate(c, 1, 2); c = qc.cy_gate(c, 0, 2)
    c = qc.rx_gate(c, 0, PI / 2); c = qc.ry_gate(c, 1, PI / 4); c = qc.rz_gate(c, 0, PI / 3)
    c = qc.phase_gate(c, 0, PI / 4)
    c = qc.crx_gate(c, 0, 1, PI / 4); c = qc.cry_gate(c, 1, 0, PI / 3); c = qc.crz_gate(c, 0, 1, PI / 6)
    c = qc.swap_gate(c, 0, 1); c = qc.iswap_gate(c, 1, 2)
    c = qc.toffoli_gate(c, 0, 1, 2); c = qc.fredkin_gate(c, 3, 0, 2)
    ls = lines(qasm3.circuit_to_qasm(c))
    for want in ("cx q[0], q[1];", "cz q[1], q[2];", "cy q[0], q[2];",
                 "rx(1.5707963267948966) q[0];", "ry(0.7853981633974483) q[1];", "rz(1.0471975511965976) q[0];",
                 "p(0.7853981633974483) q[0];", "crx(0.7853981633974483) q[0], q[1];",
                 "cry(1.0471975511965976) q[1], q[0];", "crz(0.5235987755982988) q[0], q[1];",
                 "swap q[0], q[1];", "iswap q[1], q[2];", "ccx q[0], q[1], q[2];", "cswap q[3], q[0], q[2];"):
        assert want in ls, want


def test_measurements_and_braket_target():
    text = qasm3.circuit_to_qasm(bell_measured())
    assert "c[0] = measure q[0];" in lines(text) and "c[1] = measure q[1];" in lines(text)
    assert "measure q -> c;" not in text
    bk = qasm3.circuit_to_qasm(complex_circuit(), {"target": "braket"})
    assert 'include "stdgates.inc";' not in bk
    assert "si q[0];" in lines(bk) and "ti q[1];" in lines(bk) and "cnot q[0], q[1];" in lines(bk)
    assert "phaseshift(0.5235987755982988) q[1];" in lines(bk)


def test_exact_text_of_bell_with_pragmas():
    opts = {"result-specs": {"measurements": {"shots": 1000, "qubits": [0, 1]},
                             "expectation": {"observables": ["pauli-z", "pauli-x"], "targets": [0]},
                             "probability": {"targets": [0, 1], "states": ["00", "11"]},
                             "amplitude": {"states": ["00", "11"]},
                             "state-vector": True}}
    text = qasm3.circuit_to_qasm(bell_measured(), opts)
    assert text == (
        'OPENQASM 3.0;\ninclude "stdgates.inc";\n\nqubit[2] q;\nbit[2] c;\n'
        "\n// Result extraction specifications\n"
        "#pragma qclojure result measurement shots=1000 qubits=0,1\n"
        "#pragma qclojure result expectation observable=pauli-z target=0\n"
        "#pragma qclojure result expectation observable=pauli-x\n"
        "#pragma qclojure result probability targets=0,1 states=00,11\n"
        "#pragma qclojure result amplitude states=00,11\n"
        "#pragma qclojure result state_vector // Simulation-only result\n"
        "\nh q[0];\ncx q[0], q[1];\nc[0] = measure q[0];\nc[1] = measure q[1];")


def test_global_rydberg_unknown_and_controlled_emission():
    c = qc.create_circuit(3)
    c = qc.global_x_gate(c); c = qc.global_rx_gate(c, 0.5)
    c = qc.rydberg_cz_gate(c, 0, 1); c = qc.rydberg_cphase_gate(c, 0, 2, 0.25)
    c = qc.add_operation(c, "controlled", {"gate": "x", "control": 0, "target": 1})
    text = qasm3.circuit_to_qasm(c)
    assert "// Global X gate - apply X to all qubits\nx q[0];\nx q[1];\nx q[2];" in text
    assert "// Global RX(0.5) gate - apply RX to all qubits\nrx(0.5) q[0];" in text
    assert "// Rydberg CZ gate - decomposed to standard CZ\ncz q[0], q[1];" in text
    assert "// Rydberg controlled phase gate - decomposed to CRZ\ncrz(0.25) q[0], q[2];" in text
    assert "ctrl @ :x q[0], q[1];" in lines(text)
    unknown = qasm3.circuit_to_qasm({"operations": [{"operation-type": "unknown-gate",
                                                     "operation-params": {"target": 0}}], "num-qubits": 1})
    assert "// Unknown gate: unknown-gate" in unknown
    alias = qasm3.circuit_to_qasm({"operations": [{"operation-type": "sdg"}], "num-qubits": 1})
    assert "// QClojure gate: S† gate (−π/2 phase gate)\n// Gate type: sdg - requires decomposition" in alias


def test_qasm_to_circuit_basic():
    c = qasm3.qasm_to_circuit("OPENQASM 3.0;\n\nqubit[2] q;\nbit[2] c;\n\nh q[0];\ncx q[0], q[1];\n"
                              "c[0] = measure q[0];\nc[1] = measure q[1];")
    assert c["num-qubits"] == 2 and len(c["operations"]) == 4
    ops = c["operations"]
    assert ops[0] == {"operation-type": "h", "operation-params": {"target": 0}}
    assert ops[1] == {"operation-type": "cnot", "operation-params": {"control": 0, "target": 1}}
    assert ops[2]["operation-type"] == "measure" and ops[2]["operation-params"]["measurement-qubits"] == [0]
    assert c["name"] == "Converted Circuit" and c["result-specs"] == {}


def test_qasm_to_circuit_single_parametric_advanced():
    c = qasm3.qasm_to_circuit("OPENQASM 3.0;\n\nqubit[3] q;\nbit[3] c;\n\nx q[0];\ny q[1];\nz q[2];\nh q[0];\n"
                              "s q[1];\nt q[2];\nsdg q[0];\ntdg q[1];")
    assert [o["operation-type"] for o in c["operations"]] == ["x", "y", "z", "h", "s", "t", "s-dag", "t-dag"]
    c = qasm3.qasm_to_circuit("OPENQASM 3.0;\n\nqubit[2] q;\nbit[2] c;\n\nrx(1.5708) q[0];\nry(0.7854) q[1];\n"
                              "rz(1.0472) q[0];\np(0.5236) q[1];")
    assert [o["operation-type"] for o in c["operations"]] == ["rx", "ry", "rz", "phase"]
    assert abs(c["operations"][0]["operation-params"]["angle"] - 1.5708) < 1e-6
    c = qasm3.qasm_to_circuit("OPENQASM 3.0;\n\nqubit[4] q;\nbit[4] c;\n\nswap q[0], q[1];\niswap q[1], q[2];\n"
                              "ccx q[0], q[1], q[2];\ncswap q[3], q[0], q[2];")
    assert [o["operation-type"] for o in c["operations"]] == ["swap", "iswap", "toffoli", "fredkin"]


def test_round_trips():
    orig = bell_measured()
    back = qasm3.qasm_to_circuit(qasm3.circuit_to_qasm(orig))
    assert [o["operation-type"] for o in back["operations"] if o["operation-type"] != "measure"] == ["h", "cnot"]
    for orig in (complex_circuit(), advanced_circuit()):
        back = qasm3.qasm_to_circuit(qasm3.circuit_to_qasm(orig))
        assert back["num-qubits"] == orig["num-qubits"]
        assert back["operations"] == orig["operations"]     # exact, angles included
    c = qc.rx_gate(qc.create_circuit(1), 0, 1.2345678901234567)
    back = qasm3.qasm_to_circuit(qasm3.circuit_to_qasm(c))
    assert back["operations"][0]["operation-params"]["angle"] == 1.2345678901234567


def test_edge_cases_empty_and_large():
    t = qasm3.circuit_to_qasm(qc.create_circuit(1, "Empty"))
    assert "OPENQASM 3.0;" in t and "qubit[1] q;" in t
    t = qasm3.circuit_to_qasm(qc.create_circuit(10, "Large"))
    assert "qubit[10] q;" in t and "bit[10] c;" in t


@pytest.mark.parametrize("expr,value", [
    ("pi", PI), ("-pi", -PI), ("pi/2", PI / 2), ("-pi/4", -PI / 4), ("pi/2.5", PI / 2.5), ("2*pi", 2 * PI),
    ("-0.5*pi", -0.5 * PI), ("pi*2", PI * 2), ("1/2", 0.5), ("-3/4", -0.75), ("1.5", 1.5), ("3", 3.0),
    (" 0.25 ", 0.25), ("1e-3", 0.001), ("pi/0", math.inf), ("-pi/0", -math.inf),
])
def test_parse_qasm_expression(expr, value):
    assert qasm3.parse_qasm_expression(expr) == value


@pytest.mark.parametrize("expr", ["2*pi/3", "theta", "pi+1", "-pi*2"])
def test_parse_qasm_expression_rejects(expr):
    with pytest.raises(qasm3.QasmError):
        qasm3.parse_qasm_expression(expr)


def test_result_pragmas_parse_and_merge():
    text = ("OPENQASM 3.0;\nqubit[2] q;\n"
            "#pragma qclojure result measurement shots=1000 qubits=0,1\n"
            "#pragma qclojure result expectation observable=pauli-z target=0\n"
            "#pragma qclojure result expectation observable=pauli-x target=1\n"
            "#pragma qclojure result probability targets=0,1 states=00\n"
            "  #pragma qclojure result variance observable=pauli-y target=0\n")
    specs = qasm3.qasm_to_circuit(text)["result-specs"]
    assert specs == {"measurement": {"shots": 1000, "qubits": ["0", "1"]},
                     "expectation": {"observable": "pauli-z", "target": 0,
                                     "observables": ["pauli-x"], "targets": [1]},
                     "probability": {"targets": ["0", "1"], "states": 0}}
    # the grammar parser reads digit list elements as integers
    specs2 = q3p.collect_result_specs_from_qasm(text.split("\n"))
    assert specs2["measurement"] == {"shots": 1000, "qubits": [0, 1]}
    with pytest.raises(qasm3.QasmError):   # "// Simulation-only" has no '=': the reference throws
        qasm3.parse_result_pragma("#pragma qclojure result state_vector // Simulation-only result")


def test_qubit_count_from_first_declaration():
    c = qasm3.qasm_to_circuit("qubit[12] q;\nqubit[3] r;\nh q[11];")
    assert c["num-qubits"] == 12 and c["operations"][0]["operation-params"]["target"] == 11
    with pytest.raises(qasm3.QasmError):
        qasm3.qasm_to_circuit("h q[0];")


# ---- qasm3_parser.clj (qasm3_parser_test.clj) --------------------------------------------------
def _parse_emit(text):
    c = q3p.qasm_to_circuit(text)
    return c, q3p.circuit_to_qasm(c, c["result-specs"])


def test_parser_gates_and_emission():
    c, e = _parse_emit('OPENQASM 3.0;\ninclude "stdgates.inc";\nqubit[2] q;\nh q[0];\nh q[1];')
    assert len(c["operations"]) == 2 and "h q[0];" in e
    c, e = _parse_emit('OPENQASM 3.0;\ninclude "stdgates.inc";\nqubit[4] q;\ns q[0];\nt q[1];\nsdg q[2];\ntdg q[3];')
    assert len(c["operations"]) == 4
    for want in ("s q[0];", "t q[1];", "sdg q[2];", "tdg q[3];"):
        assert want in e
    c, e = _parse_emit('OPENQASM 3.0;\ninclude "stdgates.inc";\nqubit[3] q;\n'
                       'rx(1.5708) q[0];\nry(0.7854) q[1];\nrz(3.1416) q[2];')
    assert [o["operation-params"] for o in c["operations"]] == [
        {"target": 0, "angle": 1.5708}, {"target": 1, "angle": 0.7854}, {"target": 2, "angle": 3.1416}]
    assert "rx(1.5708) q[0];" in e and "rz(3.1416) q[2];" in e
    c, e = _parse_emit('OPENQASM 3.0;\nqubit[1] q;\np(1.2345) q[0];')
    assert c["operations"][0] == {"operation-type": "phase", "operation-params": {"target": 0, "angle": 1.2345}}
    assert "p(1.2345) q[0];" in e
    c, e = _parse_emit('OPENQASM 3.0;\nqubit[4] q;\ncx q[0], q[1];\ncy q[0], q[1];\ncz q[2], q[3];\n'
                       'swap q[0], q[3];\niswap q[1], q[2];\nccx q[0], q[1], q[2];\ncswap q[1], q[2], q[3];')
    assert [(o["operation-type"], o["operation-params"]) for o in c["operations"]] == [
        ("cnot", {"control": 0, "target": 1}), ("cy", {"control": 0, "target": 1}),
        ("cz", {"control": 2, "target": 3}), ("swap", {"qubit1": 0, "qubit2": 3}),
        ("iswap", {"qubit1": 1, "qubit2": 2}), ("toffoli", {"control1": 0, "control2": 1, "target": 2}),
        ("fredkin", {"control": 1, "target1": 2, "target2": 3})]
    for want in ("cx q[0], q[1];", "cy q[0], q[1];", "cz q[2], q[3];", "swap q[0], q[3];", "iswap q[1], q[2];",
                 "ccx q[0], q[1], q[2];", "cswap q[1], q[2], q[3];"):
        assert want in e


def test_parser_comprehensive_round_trip():
    text = ('OPENQASM 3.0;\ninclude "stdgates.inc";\nqubit[5] q;\nh q[0];\nx q[1];\ny q[2];\nz q[3];\ns q[4];\n'
            't q[0];\nsdg q[1];\ntdg q[2];\nrx(pi/2) q[3];\nry(0.7854) q[4];\nrz(pi) q[0];\np(0.5236) q[1];\n'
            'cx q[0], q[1];\ncy q[1], q[2];\ncz q[2], q[3];\nswap q[0], q[4];\niswap q[1], q[3];\n'
            'ccx q[0], q[1], q[2];\ncswap q[2], q[3], q[4];')
    c = q3p.qasm_to_circuit(text)
    e = q3p.circuit_to_qasm(c)
    c2 = q3p.qasm_to_circuit(e)
    assert len(c["operations"]) == 19 and c2["operations"] == c["operations"]
    assert c["name"] == "Parsed QASM3 Circuit"
    for want in ("swap q[0], q[4];", "iswap q[1], q[3];", "ccx q[0], q[1], q[2];", "cswap q[2], q[3], q[4];"):
        assert want in e


def test_parser_empty_invalid_and_large():
    c, e = _parse_emit('OPENQASM 3.0;\ninclude "stdgates.inc";\nqubit[1] q;')
    assert c["operations"] == [] and "OPENQASM 3.0;" in e and "qubit[1] q;" in e
    for bad in ("INVALID QASM CONTENT", "OPENQASM 3.0;\nqreg q[2];", "qubit[2] q;\nh q[0]",
                "qubit[2] index;", "qubit[2] q;\nrx(π) q[0];"):
        with pytest.raises(q3p.QasmParseError):
            q3p.qasm_to_circuit(bad)
    big = 'OPENQASM 3.0;\ninclude "stdgates.inc";\nqubit[20] q;\n' + "\n".join(f"h q[{i % 20}];" for i in range(50))
    t0 = time.perf_counter()
    c = q3p.qasm_to_circuit(big)
    t1 = time.perf_counter()
    e = q3p.circuit_to_qasm(c)
    t2 = time.perf_counter()
    assert len(c["operations"]) == 50 and "h q[0];" in e
    assert t1 - t0 < 0.75 and t2 - t1 < 0.1


def test_parser_comments():
    c = q3p.qasm_to_circuit('OPENQASM 3.0;\n// This is a line comment\ninclude "stdgates.inc";\nqubit[2] q;\n'
                            '// Another comment\nh q[0]; // End-of-line comment\ncx q[0], q[1];')
    assert [o["operation-type"] for o in c["operations"]] == ["h", "cnot"]
    c = q3p.qasm_to_circuit('OPENQASM 3.0;\n/* This is a block comment */\ninclude "stdgates.inc";\nqubit[2] q;\n'
                            '/* Multi-line\n   block comment\n   with multiple lines */\nh q[0];\n'
                            '/* Another block comment */ cx q[0], q[1]; /* Final comment */')
    assert [o["operation-type"] for o in c["operations"]] == ["h", "cnot"]
    c = q3p.qasm_to_circuit('OPENQASM 3.0;\n// Line comment at start\n/* Block comment */ include "stdgates.inc"; '
                            '// End-of-line comment\nqubit[3] q;\n// Before gate operations\nh q[0]; /* Inline block */ '
                            'y q[1]; // Another line comment\n/* Multi-line block\n   comment here */\n'
                            'cx q[0], q[1]; // Final comment')
    assert [o["operation-type"] for o in c["operations"]] == ["h", "y", "cnot"]


def test_parser_keeps_only_plain_top_level_gate_calls():
    text = """OPENQASM 3.0;
include "stdgates.inc";
qubit[3] q;
bit[3] c;
const float theta = pi / 4;
gate mygate a, b { h a; cx a, b; }
ctrl @ x q[0], q[1];
inv @ s q[2];
gphase(0.1);
u3(0.1, 0.2, 0.3) q[0];
if (1) { x q[0]; }
for int i in [0:2] { h q[1]; }
barrier q[0], q[1];
reset q[2];
measure q[0] -> c[0];
c[1] = measure q[1];
id q[2];
u1(0.5) q[1];
cnot q[0], q[2];
ccnot q[0], q[1], q[2];
h $0;
rz(2 * pi / 8 + 1) q[1];
"""
    c = q3p.qasm_to_circuit(text)
    assert [(o["operation-type"], o.get("operation-params")) for o in c["operations"]] == [
        ("i", {"target": 2}), ("phase", {"target": 1, "angle": 0.5}), ("cnot", {"control": 0, "target": 2}),
        ("toffoli", {"control1": 0, "control2": 1, "target": 2}), ("h", {"target": None}),
        ("rz", {"target": 1, "angle": 2 * PI / 8 + 1})]


@pytest.mark.parametrize("expr,value", [
    ("1+2*3", 7), ("(1+2)*3", 9), ("7/2", 3.5), ("7 % 3", 1), ("-7 % 3", 2.0), ("7 % -3", -2.0),
    ("2**3", 8.0), ("-2**2", 4.0), ("2**3**2", 512.0), ("-1", -1.0), ("pi/2", PI / 2), ("tau", 2 * PI),
    ("euler", math.e), ("sin(pi/2)", 1.0), ("cos(0)", 1.0), ("arctan(1)", PI / 4), ("sqrt(2)", math.sqrt(2)),
    ("log(8, 2)", math.log(8) / math.log(2)), ("exp(1)", math.e), ("floor(2.7)", 2.0), ("ceiling(2.1)", 3.0),
    ("mod(7, 3)", 1), ("popcount(255)", 8), ("rotl(1, 65)", 2), ("rotr(1, 1)", -(2 ** 63)), ("~0", -1),
    ("5 & 3", 1), ("5 | 3", 7), ("5 ^ 3", 6), ("3 > 2 && 1", True), ("0 || 0", False), ("1 == 1.0", True),
    ("2 != 2", False), ("!true", False), ("10/0", math.inf), ("sqrt(-1)", math.nan), ("log(0)", -math.inf),
    ("SIN(0)", 0.0), ("1.5e1", 15.0), ("3 <= 3", True),
])
def test_parser_expressions(expr, value):
    v = q3p.evaluate_expression(q3p.parse_expression(expr))
    if isinstance(value, float) and math.isnan(value):
        assert math.isnan(v)
    else:
        assert v == value and type(v) is type(value)


@pytest.mark.parametrize("expr", ["theta", "9223372036854775807 + 1", "1 % 0", "2im", "$0", "unknownfn(1)",
                                  "3000000000 * 3000000000 * 3000000000"])
def test_parser_expression_errors(expr):
    with pytest.raises(q3p.QasmError):
        q3p.evaluate_expression(q3p.parse_expression(expr))


def test_parser_emitter_registry_shapes():
    c = {"num-qubits": 3, "operations": [
        {"operation-type": "crx", "operation-params": {"control": 0, "target": 1, "angle": 0.5}},
        {"operation-type": "i", "operation-params": {"target": 2}},
        {"operation-type": "rydberg-cz", "operation-params": {"control": 0, "target": 1}},
        {"operation-type": "global-x"},
        {"operation-type": "global-ry", "operation-params": {"angle": 0.25}},
        {"operation-type": "rydberg-blockade", "operation-params": {"qubit-indices": [0, 1], "angle": 0.1}},
        {"operation-type": "measure", "operation-params": {"measurement-qubits": [0]}},
        {"operation-type": "nope"}]}
    e = q3p.circuit_to_qasm(c, {"result-specs": {"result-specs": {"measurement": {"shots": 10}}}})
    assert e.split("\n")[6:] == [
        "crx(0.5) q[1];", "i q[2];", "rydberg cz q[0], q[1];", "// Global gate: global-x",
        "// Global parametric gate: global-ry(0.25)", "// Multi-qubit gate: rydberg-blockade", "measure q[];",
        "// Unknown gate: nope", "#pragma qclojure result measurement shots=10", ""]


# ---- io.clj (JSON / EDN) ---------------------------------------------------------------------
def test_state_serialisation_json_edn_exact():
    s = 1 / math.sqrt(2)
    state = {"state-vector": np.array([s, 0, 0, s], dtype=complex), "num-qubits": 2}
    data = qio.serialize_quantum_state(state)
    assert data["format-version"] == "1.0" and data["metadata"] == {}
    j = qio.json_str(data)
    assert j == ('{"state-vector":[{"real":0.7071067811865475,"imag":0.0},{"real":0.0,"imag":0.0},'
                 '{"real":0.0,"imag":0.0},{"real":0.7071067811865475,"imag":0.0}],"num-qubits":2,'
                 '"metadata":{},"format-version":"1.0"}')
    e = qio.edn_str(data)
    assert e == ('{:state-vector [{:real 0.7071067811865475, :imag 0.0} {:real 0.0, :imag 0.0} '
                 '{:real 0.0, :imag 0.0} {:real 0.7071067811865475, :imag 0.0}], :num-qubits 2, '
                 ':metadata {}, :format-version "1.0"}')
    for back in (qio.deserialize_quantum_state(qio.json_read(j)), qio.deserialize_quantum_state(qio.edn_read(e))):
        assert back["num-qubits"] == 2 and np.array_equal(back["state-vector"], state["state-vector"])


def test_circuit_serialisation_and_files(tmp_path):
    c = complex_circuit()
    data = qio.serialize_quantum_circuit(c)
    e = qio.edn_str(data)
    assert e.startswith('{:operations [{:operation-type :h, :operation-params {:target 0}} ')
    assert ':name "Complex Circuit", :description nil, :metadata {}, :format-version "1.0"}' in e
    j = qio.json_str(data)
    assert j.startswith('{"operations":[{"operation-type":"h","operation-params":{"target":0}},')
    for fmt in ("json", "edn", "qasm3"):
        path = str(tmp_path / f"c.{fmt}")
        qio.export_quantum_circuit(fmt, c, path)
        back = qio.import_quantum_circuit(fmt, path)
        assert back["num-qubits"] == 3
        assert [dict(o) for o in back["operations"]] == c["operations"]
    state = {"state-vector": np.exp(1j * np.arange(8)) / np.sqrt(8), "num-qubits": 3}
    for fmt in ("json", "edn"):
        path = str(tmp_path / f"s.{fmt}")
        qio.export_quantum_state(fmt, state, path)
        back = qio.import_quantum_state(fmt, path)
        assert np.array_equal(back["state-vector"], state["state-vector"])
        qio.export_quantum_data(fmt, c, path)
        assert qio.import_quantum_data(fmt, path)["num-qubits"] == 3
    with pytest.raises(qio.QuantumIOError):
        qio.serialize_quantum_data({"foo": 1})
    with pytest.raises(qio.QuantumIOError):
        qio.deserialize_quantum_data({"foo": 1})
    with pytest.raises(qio.QuantumIOError):
        qio.json_str({"x": math.nan})


def test_json_string_escapes_like_data_json():
    assert qio.json_str({"name": 'a/b "q" é\n'}) == '{"name":"a\\/b \\"q\\" \\u00e9\\n"}'
    assert qio.edn_str({"name": 'a "q"\n', "k": qio.Keyword("x"), "v": [1, 2.5, None, True]}) == \
        '{:name "a \\"q\\"\\n", :k :x, :v [1 2.5 nil true]}'
    assert qio.edn_read('{:a [1 2.5 "s" :kw nil] ; comment\n :b #_ 3 4, :c ##Inf}') == \
        {"a": [1, 2.5, "s", "kw", None], "b": 4, "c": math.inf}

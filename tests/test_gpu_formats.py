"""Wire formats into and out of the device: circuits arriving as OpenQASM 3
text run on the GPU and match the oracle on the same circuit; states leave
and re-enter HBM through the reference's JSON/EDN state maps
(adapter/io.clj:33-62) bit for bit."""
import math
import sys

import numpy as np
import pytest

from qclojure_amd import backend as be
from qclojure_amd import circuit as qc
from qclojure_amd import qasm3
from qclojure_amd import qasm3_parser as q3p
from qclojure_amd import qio
from qclojure_amd import engine
from qclojure_amd.engine import DeviceState

pytestmark = pytest.mark.gpu
TOL_C128 = 1e-12


def note(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def eng():
    engine.init(0)
    return engine


def _qft_text(n):
    lines = ["OPENQASM 3.0;", 'include "stdgates.inc";', f"qubit[{n}] q;", f"bit[{n}] c;"]
    for k in range(n):
        lines.append(f"x q[{k}];" if k % 3 == 0 else f"ry({0.1 * (k + 1)}) q[{k}];")
    for j in range(n):
        lines.append(f"h q[{j}];")
        for k in range(j + 1, n):
            lines.append(f"crz(pi/{2 ** (k - j)}) q[{k}], q[{j}];")
    for j in range(n // 2):
        lines.append(f"swap q[{j}], q[{n - 1 - j}];")
    lines += ["cx q[0], q[1];", "ccx q[0], q[1], q[2];", "cswap q[2], q[0], q[1];", "iswap q[1], q[2];",
              "p(-pi/8) q[3];", "sdg q[0];", "tdg q[1];", "cz q[0], q[3];"]
    return "\n".join(lines)


@pytest.mark.parametrize("parser", ["lines", "grammar"])
@pytest.mark.parametrize("n", [4, 9, 14])
def test_qasm_circuits_run_on_device_like_the_oracle(eng, oracle, parser, n):
    note(f"qasm {parser} n={n}")
    text = _qft_text(n)
    c = (qasm3 if parser == "lines" else q3p).qasm_to_circuit(text)
    assert c["num-qubits"] == n
    ops = qc.decode_circuit(c)
    with DeviceState(n) as st:
        st.init_basis(0)
        st.apply(ops)
        got = st.download()
    exp = oracle.run(n, ops)
    assert np.max(np.abs(got - exp)) < TOL_C128


def test_qasm_job_through_backend_with_pragma_results():
    note("qasm backend job")
    text = ("OPENQASM 3.0;\nqubit[3] q;\nbit[3] c;\n"
            "#pragma qclojure result measurement shots=500 qubits=0,1,2\n"
            "h q[0];\ncx q[0], q[1];\ncx q[1], q[2];")
    c = qasm3.qasm_to_circuit(text)
    shots = c["result-specs"]["measurement"]["shots"]
    be.reset_simulator_state()
    sim = be.create_simulator({"seed": 11})
    try:
        r = be.execute_circuit(sim, c, {"result-specs": {"measurements": {"shots": shots}, "state-vector": True}})
    finally:
        sim.close()
        be.reset_simulator_state()
    assert r["job-status"] == "completed"
    m = r["results"]["measurement-results"]
    assert m["shot-count"] == 500 and set(m["frequencies"]) <= {0, 7}
    sv = np.asarray(r["results"]["final-state"]["state-vector"])
    assert np.allclose(sv[[0, 7]], [1 / math.sqrt(2)] * 2, atol=1e-12)


@pytest.mark.parametrize("fmt", ["json", "edn"])
@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_state_leaves_and_reenters_hbm(eng, oracle, fmt, prec, tmp_path):
    n = 10
    note(f"state io {fmt} {prec}")
    text = _qft_text(n)
    ops = qc.decode_circuit(qasm3.qasm_to_circuit(text))
    with DeviceState(n, prec) as st:
        st.init_basis(0)
        st.apply(ops)
        path = str(tmp_path / f"state.{fmt}")
        qio.export_quantum_state(fmt, st, path)
        amps = st.download()
    back = qio.import_quantum_state(fmt, path)
    assert back["num-qubits"] == n
    assert np.array_equal(back["state-vector"], amps.astype(np.complex128))
    with qio.state_to_device(back, prec) as st2:
        assert np.array_equal(st2.download(), amps)
        # continue the computation on the re-uploaded state: H on every qubit, compare to the oracle
        more = qc.ops_from_tuples([(qc.OP["H"], q, 0, 0, 0, 0.0) for q in range(n)])
        st2.apply(more)
        got = st2.download()
    exp = oracle.apply_program(back["state-vector"], n, more)
    tol = TOL_C128 if prec == "c128" else 1e-6
    assert np.max(np.abs(got - exp)) < tol


def test_qasm_cy_is_rejected_like_the_reference_dispatcher():
    # qasm3.clj parses "cy" into a :cy operation; the ideal simulator's dispatcher has no :cy
    c = qasm3.qasm_to_circuit("qubit[2] q;\ncy q[0], q[1];")
    assert c["operations"] == [{"operation-type": "cy", "operation-params": {"control": 0, "target": 1}}]
    with pytest.raises(qc.CircuitError):
        qc.decode_circuit(c)


def test_state_to_device_rejects_wrong_length():
    with pytest.raises(qio.QuantumIOError):
        qio.state_to_device({"state-vector": np.zeros(6, complex), "num-qubits": 3})

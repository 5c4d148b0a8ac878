"""The QuantumBackend protocol over libqsv, following the reference's
ideal_simulator_test.clj and result_test.clj."""
import time

import numpy as np
import pytest

from qclojure_amd import algorithms as alg
from qclojure_amd import backend as be
from qclojure_amd import circuit as qc
from qclojure_amd import workloads as wl

pytestmark = pytest.mark.gpu


@pytest.fixture
def sim():
    be.reset_simulator_state()   # reset-simulator-fixture (ideal_simulator_test.clj:12-20)
    s = be.create_simulator({"seed": 2024})
    yield s
    s.close()
    be.reset_simulator_state()


def test_backend_info_and_availability(sim):
    info = sim.backend_info()
    assert info["backend-type"] == "simulator"
    assert "quantum-backend" in info["capabilities"]
    assert sim.is_available() is True
    assert sim.device()["technology"] == "simulator"


def test_submission_and_tracking(sim):
    job = sim.submit_circuit(qc.bell_state_circuit(), {"result-specs": {"measurements": {"shots": 100}}})
    assert isinstance(job, str)
    res = be.wait_for_job(sim, job, 30000)
    assert sim.job_status(job) == "completed"
    r = sim.job_result(job)
    assert r["job-id"] == job and r["job-status"] == "completed"
    m = r["results"]["measurement-results"]
    assert set(m["frequencies"]) == {0, 3} and m["shot-count"] == 100
    assert len(m["measurement-outcomes"]) == 100
    assert np.allclose(np.asarray(m["measurement-probabilities"]), [0.5, 0, 0, 0.5], atol=1e-15)
    assert res["job-status"] == "completed"


def test_execute_circuit_result_shape(sim):
    r = be.execute_circuit(sim, qc.bell_state_circuit(),
                           {"result-specs": {"measurements": {"shots": 50}, "state-vector": True,
                                             "probabilities": {}, "amplitudes": {"basis-states": [0, 3]}}})
    assert r["job-status"] == "completed"
    res = r["results"]
    assert res["circuit-metadata"] == {"circuit-depth": 2, "circuit-operation-count": 2, "circuit-gate-count": 2}
    assert res["result-types"] == {"measurements", "state-vector", "probabilities", "amplitudes"}
    sv = np.asarray(res["final-state"]["state-vector"])
    assert np.allclose(sv, [2 ** -0.5, 0, 0, 2 ** -0.5], atol=1e-12)
    assert list(res["state-vector-result"]["basis-labels"]) == ["|00⟩", "|01⟩", "|10⟩", "|11⟩"]
    assert abs(res["amplitude-results"]["amplitude-values"][3] - 2 ** -0.5) < 1e-12
    assert abs(res["probability-results"]["probability-outcomes"][0] - 0.5) < 1e-12
    assert "execution-time-ms" in r


def test_hamiltonian_result(sim):
    inst = wl.qaoa_instance(11, n_choices=(8,))
    r = be.execute_circuit(sim, inst["circuit"], {"result-specs": {"hamiltonian": inst["hamiltonian"]}})
    hr = r["results"]["hamiltonian-result"]
    assert 0.0 <= hr["energy-expectation"] <= len(inst["graph"])
    assert sum(len(g) for g in hr["measurement-groups"]) == len(inst["hamiltonian"])


def test_expectation_variance_fidelity(sim):
    Z = [[1, 0], [0, -1]]
    X = [[0, 1], [1, 0]]
    plus = {"num-qubits": 1, "state-vector": [2 ** -0.5, 2 ** -0.5]}
    c = qc.h_gate(qc.create_circuit(1), 0)
    r = be.execute_circuit(sim, c, {"result-specs": {
        "expectation": {"observables": [Z, X], "targets": [0, 0]},
        "variance": {"observables": [Z], "targets": [0]},
        "fidelity": {"references": [plus]}}})
    res = r["results"]
    assert abs(res["expectation-results"][0]["expectation-value"]) < 1e-12
    assert abs(res["expectation-results"][1]["expectation-value"] - 1.0) < 1e-12
    assert abs(res["variance-results"][0]["variance-value"] - 1.0) < 1e-12
    assert abs(res["fidelity-results"]["fidelities"]["reference-0"] - 1.0) < 1e-12


def test_initial_state_option(sim):
    init = {"num-qubits": 2, "state-vector": [0, 0, 1, 0]}   # |10>
    r = be.execute_circuit(sim, qc.cnot_gate(qc.create_circuit(2), 0, 1),
                           {"initial-state": init, "result-specs": {"state-vector": True}})
    assert np.allclose(np.asarray(r["results"]["final-state"]["state-vector"]), [0, 0, 0, 1])


def test_failed_jobs_never_hang(sim):
    bad = {"num-qubits": 2, "operations": [{"operation-type": "cy", "operation-params": {"control": 0, "target": 1}}]}
    r = be.execute_circuit(sim, bad, {})
    assert r["job-status"] == "failed" and "Unknown gate type" in r["error-message"]
    oob = {"num-qubits": 2, "operations": [{"operation-type": "h", "operation-params": {"target": 7}}]}
    job = sim.submit_circuit(oob, {})
    t0 = time.monotonic()
    while sim.job_status(job) in ("queued", "running") and time.monotonic() - t0 < 30:
        time.sleep(0.005)
    assert sim.job_status(job) == "failed"
    assert sim.job_result(job)["exception-type"] == "CircuitError"


def test_cleanup_and_keep_final_state_release_hbm(sim):
    """Completed jobs hold their final states in HBM until cleanup-completed-jobs
    (ideal_simulator.clj:243-263) drops them; with keep-final-state false a job
    copies its results to host memory and frees the device state at once."""
    from qclojure_amd.engine import device_info
    n, jobs = 27, 4                       # 2 GiB per complex128 state
    circ = qc.ghz_state_circuit(n)
    base = device_info()["hbm-free-bytes"]
    ids = [sim.submit_circuit(circ, {"result-specs": {"probabilities": {"targets": [0]}}}) for _ in range(jobs)]
    for j in ids:
        assert be.wait_for_job(sim, j, 60000)["job-status"] == "completed"
    held = base - device_info()["hbm-free-bytes"]
    assert held >= jobs * (16 << n) * 0.99, held
    assert be.cleanup_completed_jobs(0) == jobs
    back = base - device_info()["hbm-free-bytes"]
    assert back < (16 << n) * 0.5, back
    assert sim.job_status(ids[0]) == "not-found"
    host = be.create_simulator({"seed": 1, "keep-final-state": False})
    try:
        r = be.execute_circuit(host, circ, {"result-specs": {"measurements": {"shots": 10}}})
        assert r["job-status"] == "completed"
        assert base - device_info()["hbm-free-bytes"] < (16 << n) * 0.5
        sv = np.asarray(r["results"]["final-state"]["state-vector"])
        assert abs(sv[0] - 2 ** -0.5) < 1e-12 and abs(sv[-1] - 2 ** -0.5) < 1e-12
        p = np.asarray(r["results"]["measurement-results"]["measurement-probabilities"])
        assert abs(p[0] - 0.5) < 1e-12 and set(r["results"]["measurement-results"]["frequencies"]) <= {0, (1 << n) - 1}
    finally:
        host.close()


def test_not_found_and_cancel(sim):
    r = sim.job_result("non-existent-id")
    assert r["job-status"] == "not-found" and r["job-id"] == "non-existent-id" and "error-message" in r
    assert sim.cancel_job("non-existent-id") == "not-found"
    job = sim.submit_circuit(qc.ghz_state_circuit(5), {"shots": 5000})
    assert sim.cancel_job(job) in ("cancelled", "cannot-cancel")


def test_queue_status_and_stats(sim):
    sim.submit_circuit(qc.bell_state_circuit(), {})
    sim.submit_circuit(qc.ghz_state_circuit(3), {})
    q = sim.queue_status()
    assert q["total-jobs"] >= 2 and {"queued", "running", "completed"} <= set(q)
    be.execute_circuit(sim, qc.bell_state_circuit(), {"shots": 100})
    st = be.get_simulator_stats()
    assert {"total-jobs", "completed-jobs", "average-execution-time", "job-counter"} <= set(st)


def test_batch_job_backend(sim):
    circs = [qc.ghz_state_circuit(n) for n in (2, 3, 4, 5)]
    bid = sim.batch_submit(circs, {"result-specs": {"measurements": {"shots": 64}}})
    t0 = time.monotonic()
    while sim.batch_status(bid)["status"] != "completed" and time.monotonic() - t0 < 30:
        time.sleep(0.005)
    res = sim.batch_results(bid)
    assert len(res) == 4
    for (jid, r), n in zip(res.items(), (2, 3, 4, 5)):
        assert set(r["results"]["measurement-results"]["frequencies"]) <= {0, (1 << n) - 1}


def test_results_same_alone_and_concurrent(sim):
    """Jobs running together on the device give the results they give alone."""
    circs = [wl.random_layer_circuit(n, 6, seed=n) for n in (10, 12, 14, 16)]
    spec = {"result-specs": {"state-vector": True}}
    alone = [np.asarray(be.execute_circuit(sim, c, spec)["results"]["final-state"]["state-vector"]) for c in circs]
    ids = [sim.submit_circuit(c, spec) for c in circs]
    mixed = [np.asarray(be.wait_for_job(sim, j)["results"]["final-state"]["state-vector"]) for j in ids]
    for a, m in zip(alone, mixed):
        assert np.array_equal(a, m)


def test_sample_result_type(sim):
    """extract-sample-results (result.clj:459-494): eigenvalue draws from
    |<v|psi>|^2, one uniform per draw walking the entries in ascending
    eigenvalue order."""
    from qclojure_amd import engine
    from qclojure_amd.result import SAMPLE_SALT
    theta = 1.1
    Z = [[1, 0], [0, -1]]
    X = [[0, 1], [1, 0]]
    c = qc.add_gate(qc.create_circuit(1), "ry", target=0, angle=theta)
    r = sim.execute(c, {"result-specs": {"sample": {"observables": [Z, X], "shots": 400}}}, 17)
    s = r["results"]["sample-results"]
    rs = 17 ^ 0xA5A5A5A5A5A5A5A5 ^ SAMPLE_SALT   # the result seed execute derives from the job seed
    p_minus = np.sin(theta / 2) ** 2           # Z = -1 on |1>
    exp = [-1.0 if engine.uniform(rs, k) < p_minus else 1.0 for k in range(400)]
    assert s[0]["sample-outcomes"] == exp and s[0]["shot-count"] == 400
    px_minus = (1 - np.sin(theta)) / 2          # <X> = sin(theta)
    exp = [-1.0 if engine.uniform(rs, (1 << 32) + k) < px_minus else 1.0 for k in range(400)]
    assert s[1]["sample-outcomes"] == exp
    # a full-register observable with a degenerate spectrum (Z x Z) on a Bell state: always +1
    ZZ = np.kron(Z, Z)
    r = be.execute_circuit(sim, qc.bell_state_circuit(), {"result-specs": {"sample": {"observables": [ZZ], "shots": 50}}})
    assert r["results"]["sample-results"][0]["frequencies"] == {1.0: 50}
    # a 2x2 observable on a target qubit of a larger state: X on qubit 1 after H(1) -> always +1
    c3 = qc.h_gate(qc.x_gate(qc.create_circuit(3), 0), 1)
    r = be.execute_circuit(sim, c3, {"result-specs": {"sample": {"observables": [X], "targets": [1], "shots": 30}}})
    assert r["results"]["sample-results"][0]["frequencies"] == {1.0: 30}


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_full_dimension_observables(oracle, prec):
    """result.clj:279-281: an observable given without a target is used as-is on
    the whole state; <psi|O psi> through the :hip GEMV and inner product."""
    from qclojure_amd import engine
    from qclojure_amd import result as rs
    engine.init(0)
    n = 4
    ops = wl.random_layer_ops(n, 3, 11)
    psi = oracle.run(n, ops)
    rng = np.random.default_rng(3)
    M = rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))
    O = (M + M.conj().T) / 2
    exp = float(np.real(np.vdot(psi, O @ psi)))
    var = float(np.real(np.vdot(psi, O @ (O @ psi)))) - exp * exp
    tol = 1e-11 if prec == "c128" else 1e-5
    with engine.DeviceState(n, prec) as s:
        s.apply(ops)
        e = rs.extract_expectation_results(s, [O])[0]
        v = rs.extract_variance_results(s, [O])[0]
        assert abs(e["expectation-value"] - exp) < tol and e["target-qubits"] is None
        assert abs(v["variance-value"] - var) < 10 * tol
        with pytest.raises(ValueError):
            rs.extract_expectation_results(s, [np.eye(8)])


def test_precision_report_on_c64_results(oracle):
    """Every complex64 job states what it is worth: the FP64 norm drift of
    its state over the program and, below the shadow size, the overlap
    fidelity |<psi32|psi64>|^2 and max |d amplitude| against a complex128 run
    of the same program; checked on the 12-qubit, 20-layer random circuit,
    with the complex128 shadow itself against the oracle."""
    be.reset_simulator_state()
    circ = wl.random_layer_circuit(12, 20, seed=1)
    sim64 = be.create_simulator({"seed": 7, "precision": "c64"})
    sim128 = be.create_simulator({"seed": 7})
    try:
        r = be.execute_circuit(sim64, circ, {"result-specs": {"measurements": {"shots": 1024}}})
        assert r["job-status"] == "completed", r
        rep = r["results"]["precision-report"]
        assert rep["precision"] == "c64" and rep["reference"] == "complex128-device-shadow"
        assert rep["fidelity-vs-c128"] >= 1 - 1e-6
        assert rep["max-abs-amplitude-error"] < 1e-6
        assert rep["norm-drift"] < 1e-6
        # the complex128 shadow is the oracle's state: the reported figures
        # equal those of the c64 state against the oracle
        psi = oracle.run(12, qc.decode_circuit(circ))
        got = np.asarray(r["results"]["final-state"]["state-vector"])
        assert abs(abs(np.vdot(psi, got)) ** 2 - rep["fidelity-vs-c128"]) < 1e-9
        assert abs(np.max(np.abs(got - psi)) - rep["max-abs-amplitude-error"]) < 1e-9
        r2 = be.execute_circuit(sim128, circ, {"result-specs": {}})
        assert r2["results"]["precision-report"] == {"precision": "c128"}
        # above the shadow size only the norm drift is reported
        big = be.create_simulator({"precision": "c64", "shadow-max-qubits": 10})
        try:
            r3 = be.execute_circuit(big, circ, {"result-specs": {}})
            assert set(r3["results"]["precision-report"]) == {"precision", "norm-drift"}
        finally:
            big.close()
    finally:
        sim64.close()
        sim128.close()
        be.reset_simulator_state()

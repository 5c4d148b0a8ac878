"""Noisy trajectories (qsv_noisy_trajectories) against the oracle's
restatement of the hardware simulator: same uniforms, so every trajectory's
outcome is compared exactly and its final state to 1e-12 (complex128)."""
import math
import sys

import numpy as np
import pytest

from qclojure_amd import circuit as qc
from qclojure_amd import noise as nz
from qclojure_amd import workloads as wl
from qclojure_amd.hardware import execute_noisy_circuit, program_with_noise

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def eng():
    from qclojure_amd import engine
    engine.init(0)
    return engine


MODEL = {"gate-noise": {"h": {"noise-type": "depolarizing", "noise-strength": 0.05},
                        "rx": {"noise-type": "amplitude-damping", "noise-strength": 0.1},
                        "ry": {"noise-type": "phase-damping", "t1-time": 80.0, "t2-time": 40.0, "gate-time": 500.0},
                        "rz": {"noise-type": "coherent", "coherent-error": {"rotation-angle": 0.2,
                                                                          "rotation-axis": "x"}},
                        "cnot": {"noise-type": "depolarizing", "noise-strength": 0.1}},
         "readout-error": {"prob-0-to-1": 0.02, "prob-1-to-0": 0.05}}


def _case(n, layers, seed):
    c = wl.random_layer_circuit(n, layers, seed)
    return program_with_noise(c, MODEL)


@pytest.mark.parametrize("n,T", [(3, 200), (7, 96), (12, 16)])
def test_trajectories_match_oracle_c128(eng, oracle, n, T):
    ops, sites = _case(n, 3, 40 + n)
    note(f"trajectories c128 n={n} T={T} ops={len(ops)} sites={len(sites)}")
    got, states = eng.noisy_trajectories(n, ops, sites, T, 99, "c128", n_keep=T)
    exp, fin = oracle.trajectories(n, ops, sites, T, 99, qc.op_u)
    assert np.array_equal(got, exp)
    assert np.max(np.abs(states - fin)) < 1e-12


def test_trajectories_large_state_path(eng, oracle):
    """Above the LDS size the trajectories run one after another on a state of
    their own through fused passes."""
    n, T = 15, 3
    ops, sites = _case(n, 2, 5)
    note(f"trajectories c128 n={n} T={T}")
    got, states = eng.noisy_trajectories(n, ops, sites, T, 3, "c128", n_keep=T)
    exp, fin = oracle.trajectories(n, ops, sites, T, 3, qc.op_u)
    assert np.array_equal(got, exp)
    assert np.max(np.abs(states - fin)) < 1e-12


def test_trajectories_c64(eng, oracle):
    n, T = 8, 128
    ops, sites = _case(n, 3, 8)
    note(f"trajectories c64 n={n} T={T}")
    got, states = eng.noisy_trajectories(n, ops, sites, T, 21, "c64", n_keep=T)
    exp, fin = oracle.trajectories(n, ops, sites, T, 21, qc.op_u)
    assert np.mean(got == exp) > 0.97
    ov = np.abs(np.sum(states.conj() * fin, axis=1)) ** 2
    assert np.min(ov) > 1 - 1e-6


@pytest.mark.parametrize("prec", ["c64", "c128"])
@pytest.mark.parametrize("n", [4, 15])
def test_deep_damping_stays_normalised(eng, oracle, prec, n):
    """800 amplitude-damping sites (gamma = 0.5), each after an H: the chosen
    operator K0 shrinks the norm^2 by about 3/4 per site, 1e-100 over the
    program, which underflows complex64 (and nears FP64's range in the large
    path) unless the state is renormalised as the reference does after every
    Kraus operator (channel.clj:148-199). Outcomes and final states against
    the oracle, in the LDS batch path (n = 4) and the fused path (n = 15)."""
    sites = 800 if n == 4 else 200
    c = qc.create_circuit(n)
    for k in range(sites):
        c = qc.h_gate(c, k % n)
    model = {"gate-noise": {"h": {"noise-type": "amplitude-damping", "noise-strength": 0.5}}}
    ops, st = program_with_noise(c, model)
    T = 16 if n == 4 else 2
    note(f"deep damping {prec} n={n} sites={len(st)} T={T}")
    got, states = eng.noisy_trajectories(n, ops, st, T, 77, prec, n_keep=T)
    exp, fin = oracle.trajectories(n, ops, st, T, 77, qc.op_u)
    assert np.all(np.isfinite(states))
    ov = np.abs(np.sum(states.conj() * fin, axis=1)) ** 2
    if prec == "c128":
        assert np.array_equal(got, exp)
        assert np.max(np.abs(states - fin)) < 1e-12
    else:
        assert np.mean(got == exp) > 0.9
        assert np.min(ov) > 1 - 1e-5, ov


def test_depolarizing_statistics(eng):
    """X then depolarizing(p) on 1 qubit: P(1) = 1 - 2p/3."""
    p, T = 0.3, 20000
    c = qc.x_gate(qc.create_circuit(1), 0)
    ops, sites = program_with_noise(c, {"gate-noise": {"x": {"noise-type": "depolarizing", "noise-strength": p}}})
    got, _ = eng.noisy_trajectories(1, ops, sites, T, 1234, "c128")
    assert abs(float(np.mean(got)) - (1 - 2 * p / 3)) < 4 * math.sqrt(0.8 * 0.2 / T)


def test_execute_noisy_circuit_result_map(eng):
    """hardware_simulator_test.clj:61-80: a noisy GHZ keeps most shots in
    000/111 and shows some noise; counts sum to the shot count."""
    kyoto_like = {"gate-noise": {"h": {"noise-type": "depolarizing", "noise-strength": 0.0003},
                                 "cnot": {"noise-type": "depolarizing", "noise-strength": 0.0035}},
                  "readout-error": {"prob-0-to-1": 0.008, "prob-1-to-0": 0.018}}
    res = execute_noisy_circuit(qc.ghz_state_circuit(3), kyoto_like, {"shots": 1000, "seed": 5})
    assert res["job-status"] == "completed" and res["shots-executed"] == 1000
    m = res["results"]["measurement-results"]
    assert sum(m.values()) == 1000 and 2 <= len(m) <= 8
    ideal = (m.get("000", 0) + m.get("111", 0)) / 1000
    assert 0.85 < ideal < 1.0
    # density matrix from the collected trajectories: trace 1, Hermitian
    res = execute_noisy_circuit(qc.ghz_state_circuit(3), kyoto_like,
                                {"shots": 200, "seed": 6, "result-specs": {"density-matrix": True}})
    r = res["results"]
    assert r["trajectory-count"] == 100 and abs(r["density-matrix-trace"] - 1) < 1e-12
    assert np.allclose(r["density-matrix"], r["density-matrix"].conj().T)
    # an empty circuit fails (hardware_simulator_test.clj:133-141)
    assert execute_noisy_circuit(qc.create_circuit(1), kyoto_like, {"shots": 10})["job-status"] == "failed"


def test_density_matrix_equals_trajectory_average(eng):
    """rho = (1/k) sum_t psi_t psi_t^H (state.clj:726-800) over the returned
    trajectories, on complex amplitudes (RY then S: rho differs from rho^T
    and from conj(rho), so a dropped or extra transpose or conjugation fails)."""
    circ = qc.create_circuit(2)
    circ = qc.ry_gate(circ, 0, 0.7)
    circ = qc.s_gate(circ, 0)
    circ = qc.cnot_gate(circ, 0, 1)
    circ = qc.ry_gate(circ, 1, 1.9)
    circ = qc.t_gate(circ, 1)
    model = {"gate-noise": {"cnot": {"noise-type": "depolarizing", "noise-strength": 0.2},
                            "ry": {"noise-type": "amplitude-damping", "noise-strength": 0.3}}}
    res = execute_noisy_circuit(circ, model, {"shots": 64, "seed": 8, "result-specs": {"density-matrix": True}})
    r = res["results"]
    states = np.array([np.asarray(t["state-vector"]) for t in r["trajectories"]])
    k = len(states)
    exp = np.einsum("ti,tj->ij", states, states.conj()) / k
    rho = np.asarray(r["density-matrix"])
    assert np.max(np.abs(rho - exp)) < 1e-13
    assert np.max(np.abs(exp - exp.T)) > 1e-3          # the case tells rho from its transpose


def test_trajectory_errors(eng):
    ops = qc.ops_from_tuples([(qc.OP["H"], 0, 0, 0, 0, 0.0), (qc.OP["MEASURE"], 0, 0, 0, 1, 0.0)])
    with pytest.raises(Exception, match="mid-circuit"):
        eng.noisy_trajectories(1, ops, [], 4, 0)
    ops = qc.ops_from_tuples([(qc.OP["H"], 0, 0, 0, 0, 0.0)])
    with pytest.raises(Exception, match="zero"):
        eng.noisy_trajectories(1, ops, [{"after-op": 0, "qubit": 0, "kraus": [np.zeros((2, 2))]}], 4, 0)
    with pytest.raises(Exception, match="after_op"):
        eng.noisy_trajectories(1, ops, [{"after-op": 3, "qubit": 0, "kraus": [np.eye(2)]}], 4, 0)

"""Batched circuits in one launch (qsv_batch_expect_z): per-instance <H_C> of
QAOA MaxCut instances against the FP64 oracle, and against the single-state
path."""
import sys

import numpy as np
import pytest

from qclojure_amd import circuit as qc
from qclojure_amd import workloads as wl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from qclojure_amd import engine
    engine.init(0)
    return engine


def _oracle_energy(oracle, inst):
    n = inst["n"]
    psi = oracle.run(n, qc.decode_circuit(inst["circuit"]))
    e = None
    for t in inst["hamiltonian"]:
        v = t["coefficient"] * oracle.pauli_expectation(psi, n, t["pauli-string"])
        e = v if e is None else e + v
    return e


def _batch(insts):
    return [(i["n"], qc.decode_circuit(i["circuit"]), [(t["coefficient"], t["pauli-string"]) for t in i["hamiltonian"]])
            for i in insts]


@pytest.mark.parametrize("prec,tol", [("c128", 1e-10), ("c64", 1e-5)])
def test_batch_qaoa_vs_oracle(eng, oracle, prec, tol):
    insts = [wl.qaoa_instance(1000 + k) for k in range(36)]
    # make sure every size class is present, including the LDS-resident and pooled paths
    insts += [wl.qaoa_instance(77, n_choices=(n,)) for n in (8, 12, 14, 16, 20)]
    print(f"sizes {[i['n'] for i in insts]}", file=sys.stderr, flush=True)
    got = eng.batch_expect_z(_batch(insts), prec)
    for inst, g in zip(insts, got):
        exp = _oracle_energy(oracle, inst)
        assert abs(g - exp) < tol * max(1.0, abs(exp)), (inst["n"], inst["p"], g, exp)


def test_batch_matches_single_state_path(eng):
    from qclojure_amd.hamiltonian import hamiltonian_expectation
    insts = [wl.qaoa_instance(2000 + k, n_choices=(10, 16)) for k in range(6)]
    got = eng.batch_expect_z(_batch(insts), "c128")
    for inst, g in zip(insts, got):
        with eng.DeviceState(inst["n"], "c128") as s:
            s.apply(qc.decode_circuit(inst["circuit"]))
            e = hamiltonian_expectation(inst["hamiltonian"], s)
        assert abs(g - e) < 1e-10


def test_batch_general_gates(eng, oracle):
    """Every gate kind through the batched interpreter (small and pooled sizes)."""
    from test_gpu_parity import _random_program
    rng = np.random.default_rng(5)
    insts = []
    for n in (3, 6, 9, 13, 15):
        ops = _random_program(rng, n, 80)
        terms = [(float(rng.normal()), "".join(rng.choice(["I", "Z"], size=n))) for _ in range(5)]
        insts.append((n, ops, terms))
    got = eng.batch_expect_z(insts, "c128")
    for (n, ops, terms), g in zip(insts, got):
        psi = oracle.run(n, ops)
        exp = sum(c * oracle.pauli_expectation(psi, n, s) for c, s in terms)
        assert abs(g - exp) < 1e-10, n


def test_batch_rejects_measurements(eng):
    from qclojure_amd._abi import OP, QsvError
    ops = qc.ops_from_tuples([(OP["H"], 0, 0, 0, 0, 0.0), (OP["MEASURE"], 0, 0, 0, 1, 0.0)])
    with pytest.raises(QsvError, match="not batched"):
        eng.batch_expect_z([(2, ops, [(1.0, "ZI")])], "c64")


@pytest.mark.parametrize("prec,tol", [("c128", 1e-10), ("c64", 1e-5)])
def test_grouped_passes_uneven_programs(eng, oracle, prec, tol):
    """Several instances per qubit count with programs of different lengths:
    the grouped launches run pass rounds where some instances have no pass."""
    from test_gpu_parity import _random_program
    rng = np.random.default_rng(77)
    insts = []
    for n in (15, 15, 15, 16, 16, 17):
        ops = _random_program(rng, n, int(rng.integers(5, 160)))
        terms = [(float(rng.normal()), "".join(rng.choice(["I", "Z"], size=n))) for _ in range(4)]
        insts.append((n, ops, terms))
    print(f"grouped {prec} lengths {[len(i[1]) for i in insts]}", file=sys.stderr, flush=True)
    got = eng.batch_expect_z(insts, prec)
    for (n, ops, terms), g in zip(insts, got):
        psi = oracle.run(n, ops)
        exp = sum(c * oracle.pauli_expectation(psi, n, s) for c, s in terms)
        assert abs(g - exp) < tol * max(1.0, sum(abs(c) for c, _ in terms)), (n, len(ops), g, exp)


def test_batch_rejects_decreasing_offsets(eng):
    """An instance's op or term range [off[i], off[i+1]) must not run backwards:
    the call fails before reading past the arrays, and the next call works."""
    import ctypes
    from qclojure_amd import _abi
    nq = np.array([2, 2], dtype=np.int32)
    ops = qc.ops_from_tuples([(_abi.OP["H"], 0, 0, 0, 0, 0.0), (_abi.OP["CNOT"], 0, 1, 0, 0, 0.0)])
    zm = np.array([1, 2], dtype=np.uint64)
    co = np.array([1.0, 1.0])
    out = np.zeros(2)
    for opoff, toff in (([0, 2, 1], [0, 1, 2]), ([0, 1, 2], [0, 2, 1])):
        opo = np.array(opoff, dtype=np.uint64)
        to = np.array(toff, dtype=np.uint64)
        desc = _abi.QsvBatchDesc(2, nq.ctypes.data, opo.ctypes.data, ops.ctypes.data, to.ctypes.data,
                                 zm.ctypes.data, co.ctypes.data, _abi.QSV_C128)
        with pytest.raises(_abi.QsvError, match="non-decreasing"):
            _abi.check(eng.lib().qsv_batch_expect_z(ctypes.byref(desc), out.ctypes.data))
    e = eng.batch_expect_z([(2, ops, [(1.0, "ZI"), (1.0, "IZ")])], "c128")
    assert abs(e[0]) < 1e-12


def _graph_case(prec, shift=0.0):
    """Grouped instances (15-17 qubits, one with no gates so that the arena is
    filled) plus LDS-resident ones; `shift` moves every angle but keeps the
    launch structure."""
    from test_gpu_parity import _random_program
    rng = np.random.default_rng(4242)
    insts = []
    for n, g in ((15, 120), (15, 0), (16, 90), (17, 60), (10, 40)):
        ops = _random_program(rng, n, g) if g else qc.ops_from_tuples([])
        ops = np.array(ops, copy=True)
        if len(ops):
            ops["angle"] += shift
        terms = [(float(rng.normal()), "".join(rng.choice(["I", "Z"], size=n))) for _ in range(4)]
        insts.append((n, ops, terms))
    return insts


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_group_graph_replays_match_eager(eng, oracle, prec, monkeypatch):
    """The grouped launches of a batch replayed as one captured graph: three
    replays with other device work and host waits between them, then a replay
    with new angles, each equal to the kernels issued one by one."""
    insts = _graph_case(prec)
    moved = _graph_case(prec, shift=0.37)
    monkeypatch.setenv("QSV_BATCH_GRAPH", "0")
    eager = eng.batch_expect_z(insts, prec)
    eager_moved = eng.batch_expect_z(moved, prec)
    monkeypatch.setenv("QSV_BATCH_GRAPH", "1")
    other = [(16, qc.ops_from_tuples([(qc.OP["H"], q, 0, 0, 0, 0.0) for q in range(16)]), [(1.0, "Z" * 16)])]
    for rep in range(3):
        print(f"graph {prec} replay {rep}", file=sys.stderr, flush=True)
        got = eng.batch_expect_z(insts, prec)
        assert np.array_equal(got, eager), (rep, np.max(np.abs(got - eager)))
        # device work of another shape between the replays
        eng.batch_expect_z(other, prec)
        with eng.DeviceState(12, prec) as s:
            s.apply(qc.ops_from_tuples([(qc.OP["H"], 0, 0, 0, 0, 0.0)]))
            s.probabilities()
    got = eng.batch_expect_z(moved, prec)
    assert np.array_equal(got, eager_moved), np.max(np.abs(got - eager_moved))
    tol = 1e-10 if prec == "c128" else 1e-5
    for (n, ops, terms), g in zip(moved, got):
        psi = oracle.run(n, ops)
        exp = sum(c * oracle.pauli_expectation(psi, n, s) for c, s in terms)
        assert abs(g - exp) < tol * max(1.0, sum(abs(c) for c, _ in terms)), (n, g, exp)

"""Pins the FP64 oracle (oracle/qsv_ref.c) before anything is checked against it:
the reference test suite's known answers, textbook closed forms, brute force on
tiny inputs, and the quirks read from the reference source."""
import math

import numpy as np
import pytest

from qclojure_amd import algorithms as alg
from qclojure_amd import circuit as qc
from qclojure_amd import workloads as wl
from qclojure_amd._abi import OP
from kat_checks import as_vec, check_expectation, check_shots, run_checks

I2 = np.eye(2, dtype=complex)
X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.diag([1, -1]).astype(complex)


def kron_embed(g, n, q):
    """I ⊗ .. ⊗ g (factor q) ⊗ .. ⊗ I with numpy's kron (qubit 0 outermost)."""
    m = np.array([[1.0 + 0j]])
    for k in range(n):
        m = np.kron(m, g if k == q else I2)
    return m


def rand_state(n, rng):
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


# ---- the reference's own known answers -------------------------------------
@pytest.mark.parametrize("literal", [True, False])
def test_gate_kats(kats, oracle, literal):
    for case in kats["gate-cases"]:
        init = as_vec(case["init"])
        circ = {"num-qubits": case["n"], "operations": case["ops"]}
        ops = qc.decode_circuit(circ)
        psi = oracle.apply_program(init, case["n"], ops, literal=literal)
        run_checks(case, psi, init)


@pytest.mark.parametrize("literal", [True, False])
def test_expectation_kats(kats, oracle, literal):
    for case in kats["expectation-cases"]:
        psi = as_vec(case["init"])
        total = None
        for c, s in case["hamiltonian"]:
            v = c * oracle.pauli_expectation(psi, case["n"], s, literal=literal)
            total = v if total is None else total + v
        check_expectation(case, total)


def test_shot_windows(kats, oracle):
    for case in kats["shot-cases"]:
        psi = oracle.run(case["n"], qc.decode_circuit({"num-qubits": case["n"], "operations": case["ops"]}))
        outs = oracle.sample(psi, case["n"], case["shots"], seed=1234)
        check_shots(case, outs)


def test_probability_kats(kats, oracle):
    """qs/probability on |0>, |1>, |+>, |00> (state_test.clj:145-166); exact
    where the reference compares with =."""
    for case in kats["probability-cases"]:
        p = oracle.probabilities(as_vec(case["init"]), case["n"])
        exp = np.asarray(case["probs"])
        if case["tol"] == "exact":
            assert np.array_equal(p, exp), (case["id"], p)
        else:
            assert np.max(np.abs(p - exp)) < case["tol"], (case["id"], p)


def test_normalize_kats(kats, oracle):
    """normalize-state on {3, 4} and |+> (state_test.clj:114-140)."""
    for case in kats["normalize-cases"]:
        out = oracle.normalize(as_vec(case["init"]), case["n"])
        assert abs(np.sum(np.abs(out) ** 2) - 1.0) < case["norm2-tol"], case["id"]
        assert np.max(np.abs(out - as_vec(case["expect"]))) < case["tol"], (case["id"], out)


def test_probabilities_norm_and_normalize_closed_form(oracle):
    """Born rule |a_i|^2 at index i (numpy's abs, so a dropped imaginary part,
    a missing square or a bit-reversed index fails), sum = squared 2-norm, and
    normalize = psi / ||psi|| except at or below the 1e-12 tolerance."""
    rng = np.random.default_rng(17)
    for n in range(1, 9):
        v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        v *= rng.uniform(0.1, 10.0)
        p = oracle.probabilities(v, n)
        assert np.allclose(p, np.abs(v) ** 2, rtol=1e-14, atol=0)
        assert abs(oracle.norm2(v, n) - np.linalg.norm(v) ** 2) < 1e-12 * np.linalg.norm(v) ** 2
        out = oracle.normalize(v, n)
        assert np.allclose(out, v / np.linalg.norm(v), rtol=1e-14, atol=1e-16)
    # asymmetric single amplitude: only index 5 carries probability
    v = np.zeros(8, complex)
    v[5] = 0.6 + 0.8j
    assert np.array_equal(oracle.probabilities(v, 3) > 0, np.arange(8) == 5)
    # at or below the tolerance (and the zero vector) the amplitudes are kept
    tiny = np.array([1e-13, 0.0], complex)
    assert np.array_equal(oracle.normalize(tiny, 1), tiny)
    assert np.array_equal(oracle.normalize(np.zeros(4, complex), 2), np.zeros(4, complex))


def test_bits_to_index(kats):
    from qclojure_amd.result import _bits_to_index
    for c in kats["bits-to-index"]:
        assert _bits_to_index(c["bits"]) == c["index"]


# ---- textbook closed forms ---------------------------------------------------
def test_rotation_matrices_closed_form(oracle):
    for th in np.linspace(-7, 7, 29):
        c, s = math.cos(th / 2), math.sin(th / 2)
        assert np.allclose(oracle.gate_matrix(OP["RX"], th), c * I2 - 1j * s * X, atol=1e-15)
        assert np.allclose(oracle.gate_matrix(OP["RY"], th), c * I2 - 1j * s * Y, atol=1e-15)
        assert np.allclose(oracle.gate_matrix(OP["RZ"], th), c * I2 - 1j * s * Z, atol=1e-15)
        assert np.allclose(oracle.gate_matrix(OP["PHASE"], th), np.diag([1, np.exp(1j * th)]), atol=1e-15)


@pytest.mark.parametrize("literal", [True, False])
def test_one_qubit_application_matches_kron(oracle, literal):
    rng = np.random.default_rng(7)
    for n in range(1, 7):
        for q in range(n):
            for kind in ("H", "RX", "RY", "RZ", "Y", "T"):
                ang = float(rng.uniform(0, 2 * math.pi))
                psi = rand_state(n, rng)
                ops = qc.ops_from_tuples([(OP[kind], q, 0, 0, 0, ang)])
                got = oracle.apply_program(psi, n, ops, literal=literal)
                exp = kron_embed(oracle.gate_matrix(OP[kind], ang), n, q) @ psi
                assert np.max(np.abs(got - exp)) < 1e-13, (n, q, kind)


def test_ghz_closed_form(oracle):
    for n in range(2, 13):
        psi = oracle.run(n, qc.decode_circuit(qc.ghz_state_circuit(n)))
        exp = np.zeros(1 << n, complex)
        exp[0] = exp[-1] = 1 / math.sqrt(2)
        assert np.max(np.abs(psi - exp)) < 1e-15


def test_hadamard_wall_uniform(oracle):
    for n in range(1, 13):
        c = qc.create_circuit(n)
        for q in range(n):
            c = qc.h_gate(c, q)
        psi = oracle.run(n, qc.decode_circuit(c))
        assert np.max(np.abs(psi - 2 ** (-n / 2))) < 1e-14


def _ref_qft_closed_form(n, x):
    """The reference QFT (H + CRZ(pi/2^(k+1)) + SWAPs) on |x>: CRZ = CP times a
    phase e^{-i theta c/2} on the control bit, and every control is still a
    basis bit when its CRZ acts, so the result is the textbook QFT times the
    global phase exp(-i/2 * sum_q sum_k theta_k x_{q+k+1})."""
    N = 1 << n
    bits = [(x >> (n - 1 - q)) & 1 for q in range(n)]
    phi = 0.0
    for q in range(n):
        for k in range(n - q - 1):
            phi -= (math.pi / 2 ** (k + 1)) * bits[q + k + 1] / 2
    y = np.arange(N)
    return np.exp(1j * phi) * np.exp(2j * math.pi * x * y / N) / math.sqrt(N)


@pytest.mark.parametrize("literal", [True, False])
def test_qft_closed_form(oracle, literal):
    rng = np.random.default_rng(3)
    for n in range(1, 9):
        ops = qc.decode_circuit(alg.quantum_fourier_transform_circuit(n))
        for x in rng.integers(0, 1 << n, size=4):
            psi = np.zeros(1 << n, complex)
            psi[x] = 1
            got = oracle.apply_program(psi, n, ops, literal=literal)
            assert np.max(np.abs(got - _ref_qft_closed_form(n, int(x)))) < 1e-12, (n, x)


def test_global_x_phase(oracle):
    # global-x = RX(pi) on every qubit (gate.clj:1209-1219): |0..0> -> (-i)^n |1..1>
    for n in range(1, 8):
        psi = oracle.run(n, qc.decode_circuit(qc.global_x_gate(qc.create_circuit(n))))
        assert abs(psi[-1] - (-1j) ** n) < 1e-14
        assert np.sum(np.abs(psi[:-1])) < 1e-14


@pytest.mark.parametrize("literal", [True, False])
def test_global_gates_match_kron(oracle, literal):
    """global-rx/ry/rz/hadamard apply the gate to qubits 0..n-1 in turn, and
    global-x/y/z are RX/RY/RZ(pi) on every qubit (gate.clj:1098-1253): the
    n-fold numpy Kronecker power of the textbook matrix, on a complex state.
    (A global-y taken as Pauli Y would differ by the phase (-i)^n.)"""
    rng = np.random.default_rng(23)
    H = np.array([[1, 1], [1, -1]], complex) / math.sqrt(2)

    def rot(P, th):
        return math.cos(th / 2) * I2 - 1j * math.sin(th / 2) * P

    for n in (1, 2, 4, 6):
        psi = rand_state(n, rng)
        th = float(rng.uniform(0, 2 * math.pi))
        cases = {"GLOBAL_RX": rot(X, th), "GLOBAL_RY": rot(Y, th), "GLOBAL_RZ": rot(Z, th), "GLOBAL_H": H,
                 "GLOBAL_X": rot(X, math.pi), "GLOBAL_Y": rot(Y, math.pi), "GLOBAL_Z": rot(Z, math.pi)}
        for kind, g in cases.items():
            m = np.array([[1.0 + 0j]])
            for _ in range(n):
                m = np.kron(m, g)
            ops = qc.ops_from_tuples([(OP[kind], 0, 0, 0, 0, th)])
            got = oracle.apply_program(psi, n, ops, literal=literal)
            assert np.max(np.abs(got - m @ psi)) < 1e-13, (n, kind)


def test_pauli_expectation_matches_dense(oracle):
    rng = np.random.default_rng(11)
    P = {"I": I2, "X": X, "Y": Y, "Z": Z}
    for n in range(1, 7):
        for _ in range(6):
            psi = rand_state(n, rng)
            s = "".join(rng.choice(list("IXYZ"), size=n))
            m = np.array([[1.0 + 0j]])
            for ch in s:
                m = np.kron(m, P[ch])
            exp = np.vdot(psi, m @ psi).real
            assert abs(oracle.pauli_expectation(psi, n, s) - exp) < 1e-13
            assert abs(oracle.pauli_expectation(psi, n, s, literal=True) - exp) < 1e-13
    # <Y> on |+i> = (|0> + i|1>)/sqrt2 is +1
    assert abs(oracle.pauli_expectation(np.array([1, 1j]) / math.sqrt(2), 1, "Y") - 1.0) < 1e-15


def test_inner_and_expect_1q(oracle):
    rng = np.random.default_rng(5)
    for n in range(1, 7):
        a, b = rand_state(n, rng), rand_state(n, rng)
        assert abs(oracle.inner(a, b, n) - np.vdot(a, b)) < 1e-14
        o = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
        for q in range(n):
            assert abs(oracle.expect_1q(a, n, q, o) - np.vdot(a, kron_embed(o, n, q) @ a)) < 1e-13


# ---- quirks read from the reference source ----------------------------------
def test_controlled_gates_apply_the_transpose(oracle):
    """apply-controlled-gate accumulates a0*g01 into the flipped index
    (gate.clj:461-486): on control=1 the target sees g^T. So CRY(theta)|10> =
    cos(theta/2)|10> - sin(theta/2)|11> and CY|10> = -i|11>."""
    th = 0.7
    psi = oracle.apply_program(np.array([0, 0, 1, 0], complex), 2,
                               qc.ops_from_tuples([(OP["CRY"], 0, 1, 0, 0, th)]), literal=True)
    assert np.allclose(psi, [0, 0, math.cos(th / 2), -math.sin(th / 2)], atol=1e-15)
    psi = oracle.apply_program(np.array([0, 0, 1, 0], complex), 2,
                               qc.ops_from_tuples([(OP["CY"], 0, 1, 0, 0, 0.0)]), literal=True)
    assert np.allclose(psi, [0, 0, 0, -1j], atol=1e-15)
    # pair mode agrees with the literal loop
    rng = np.random.default_rng(2)
    for kind in ("CNOT", "CZ", "CY", "CRX", "CRY", "CRZ", "RYDBERG_CZ"):
        for (c, t) in [(0, 3), (3, 0), (1, 2), (4, 1)]:
            psi0 = rand_state(5, rng)
            ops = qc.ops_from_tuples([(OP[kind], c, t, 0, 0, float(rng.uniform(0, 6)))])
            a = oracle.apply_program(psi0, 5, ops, literal=True)
            b = oracle.apply_program(psi0, 5, ops, literal=False)
            assert np.max(np.abs(a - b)) < 1e-15


def test_swap_uses_lsb_positions(oracle):
    """swap-gate shifts by the qubit number itself (gate.clj:768-778): for
    n = 3, SWAP(0, 1) exchanges index bits 0 and 1, i.e. reference qubits 2 and 1."""
    ops = qc.ops_from_tuples([(OP["SWAP"], 0, 1, 0, 0, 0.0)])
    for x in range(8):
        psi = np.zeros(8, complex)
        psi[x] = 1
        out = oracle.apply_program(psi, 3, ops)
        b0, b1 = x & 1, (x >> 1) & 1
        y = (x & ~3) | (b0 << 1) | b1
        assert out[y] == 1
    # the MSB flag swaps reference qubits 0 and 1 (index bits 2 and 1)
    for x in range(8):
        psi = np.zeros(8, complex)
        psi[x] = 1
        out = oracle.apply_program(psi, 3, ops, flags=1)
        b2, b1 = (x >> 2) & 1, (x >> 1) & 1
        y = (x & 1) | (b1 << 2) | (b2 << 1)
        assert out[y] == 1


def test_measure_collapse_closed_form(oracle):
    n = 3
    ghz = oracle.run(n, qc.decode_circuit(qc.ghz_state_circuit(n)))
    for u, exp_o, exp_idx in [(0.1, 0, 0), (0.49, 0, 0), (0.51, 1, 7), (0.99, 1, 7)]:
        o, probs, st = oracle.measure_qubits(ghz, n, 1 << 0, u)
        assert o == exp_o and np.allclose(probs, [0.5, 0.5], atol=1e-15)
        exp = np.zeros(8, complex)
        exp[exp_idx] = 1
        assert np.max(np.abs(st - exp)) < 1e-15
    # two measured qubits of H^3: four outcomes of 1/4; outcome bit j = j-th qubit ascending
    c = qc.create_circuit(3)
    for q in range(3):
        c = qc.h_gate(c, q)
    psi = oracle.run(3, qc.decode_circuit(c))
    o, probs, st = oracle.measure_qubits(psi, 3, (1 << 0) | (1 << 2), 0.6)   # cum 0.25,0.5,0.75 -> outcome 2
    assert o == 2 and np.allclose(probs, 0.25)
    # outcome 2 = qubit0 -> 0, qubit2 -> 1: kept indices have bit2 (q0) = 0 and bit0 (q2) = 1
    kept = [i for i in range(8) if abs(st[i]) > 0]
    assert kept == [1, 3] and np.allclose(np.abs(st[kept]), 1 / math.sqrt(2))


def test_measure_state_inverse_cdf(oracle):
    rng = np.random.default_rng(9)
    n = 5
    psi = rand_state(n, rng)
    p = np.abs(psi) ** 2
    cum = np.cumsum(p)
    for _ in range(200):
        u = float(rng.uniform())
        exp = min(int(np.sum(cum < u * cum[-1])), (1 << n) - 1)
        assert oracle.lib().ref_measure_state(oracle._p(psi), n, u) == exp


def test_sample_prefix_equals_sample(oracle):
    """The binary-search sampler for large states gives ref_sample's shots."""
    rng = np.random.default_rng(19)
    for n in (1, 3, 8, 12):
        psi = rand_state(n, rng)
        psi[rng.integers(0, 1 << n, size=(1 << n) // 3)] = 0     # zero runs: equal running sums
        assert np.array_equal(oracle.sample_prefix(psi, n, 3000, 5 + n), oracle.sample(psi, n, 3000, 5 + n))


def test_qaoa_p1_closed_form(oracle):
    """p = 1 MaxCut on triangle-free 3-regular graphs: <C_uv> = 1/2 +
    1/2 sin(4 beta) sin(gamma) cos^2(gamma) per edge (Wang, Hadfield, Jiang,
    Rieffel, PRA 97, 022304, 2018). The reference circuit is e^{-i gamma C}
    e^{-i beta B} up to a global phase (CNOT RZ(-gamma w) CNOT per edge, RX(2 beta))."""
    cube = [[a, b, 1.0] for a in range(8) for b in range(a + 1, 8) if bin(a ^ b).count("1") == 1]
    k33 = [[a, b, 1.0] for a in range(3) for b in range(3, 6)]
    for graph, n in ((cube, 8), (k33, 6)):
        for gamma, beta in ((0.3, 0.2), (1.1, 2.5), (2.9, 0.4)):
            hp = alg.max_cut_hamiltonian(graph, n)
            circ = alg.qaoa_ansatz_circuit(hp, alg.standard_mixer_hamiltonian(n), [gamma, beta], n)
            psi = oracle.run(n, qc.decode_circuit(circ))
            e = None
            for t in hp:
                v = t["coefficient"] * oracle.pauli_expectation(psi, n, t["pauli-string"])
                e = v if e is None else e + v
            m = len(graph)
            exp = m * (0.5 + 0.5 * math.sin(4 * beta) * math.sin(gamma) * math.cos(gamma) ** 2)
            assert abs(e - exp) < 1e-12, (n, gamma, beta, e, exp)


def test_literal_equals_pair_on_random_programs(oracle):
    rng = np.random.default_rng(21)
    kinds1 = ["X", "Y", "Z", "H", "S", "SDG", "T", "TDG", "RX", "RY", "RZ", "PHASE"]
    kinds2 = ["CNOT", "CZ", "CY", "CRX", "CRY", "CRZ", "RYDBERG_CZ", "RYDBERG_CPHASE", "SWAP", "ISWAP"]
    for n in (3, 5, 7):
        rows = []
        for _ in range(60):
            r = rng.uniform()
            ang = float(rng.uniform(0, 2 * math.pi))
            if r < 0.5:
                rows.append((OP[rng.choice(kinds1)], int(rng.integers(n)), 0, 0, 0, ang))
            elif r < 0.85:
                a, b = rng.choice(n, 2, replace=False)
                rows.append((OP[rng.choice(kinds2)], int(a), int(b), 0, 0, ang))
            elif r < 0.95:
                a, b, c = rng.choice(n, 3, replace=False)
                rows.append((OP[rng.choice(["TOFFOLI", "FREDKIN"])], int(a), int(b), int(c), 0, 0.0))
            else:
                m = int(rng.integers(1, 1 << n))
                rows.append((OP["RYDBERG_BLOCKADE"], 0, 0, 0, m, ang))
        ops = qc.ops_from_tuples(rows)
        a = oracle.run(n, ops, literal=True)
        b = oracle.run(n, ops, literal=False)
        assert np.max(np.abs(a - b)) < 1e-13
        assert abs(np.linalg.norm(a) - 1) < 1e-12


def test_c1_config_literal_vs_pair(oracle):
    """The 12-qubit, 20-layer random circuit: the literal dense-Kronecker path
    and the pair path agree."""
    ops = wl.random_layer_ops(12, 20, seed=1)
    a = oracle.run(12, ops, literal=True)
    b = oracle.run(12, ops, literal=False)
    assert np.max(np.abs(a - b)) < 1e-13

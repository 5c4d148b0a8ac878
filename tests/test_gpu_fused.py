"""Parity of the fused tiled-pass path (QSV_FLAG_FUSE) with the oracle and with
the per-gate path."""
import math
import sys

import numpy as np
import pytest

from qclojure_amd import circuit as qc
from qclojure_amd import workloads as wl
from qclojure_amd._abi import OP, QSV_FLAG_FUSE, QSV_FLAG_MSB_SWAP

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def eng():
    from qclojure_amd import engine
    engine.init(0)
    return engine


def _program(rng, n, count):
    kinds1 = ["X", "Y", "Z", "H", "S", "SDG", "T", "TDG", "RX", "RY", "RZ", "PHASE"]
    kinds2 = ["CNOT", "CZ", "CY", "CRX", "CRY", "CRZ", "RYDBERG_CZ", "RYDBERG_CPHASE", "SWAP", "ISWAP"]
    rows = []
    for _ in range(count):
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
            qs = rng.choice(n, 3, replace=False)
            rows.append((OP["RYDBERG_BLOCKADE"], 0, 0, 0, int(sum(1 << int(q) for q in qs)), ang))
    return qc.ops_from_tuples(rows)


def _run(eng, n, ops, prec="c128", flags=0, seed=0, init=None):
    with eng.DeviceState(n, prec) as s:
        if init is not None:
            s.upload(init)
        s.apply(ops, flags, seed)
        return s.download()


@pytest.mark.parametrize("n", [12, 13, 14, 16])
def test_fused_random_programs_c128(eng, oracle, n):
    rng = np.random.default_rng(500 + n)
    init = oracle.run(n, _program(rng, n, 20))
    ops = _program(rng, n, 300)
    note(f"fused c128 n={n}")
    exp = oracle.apply_program(init, n, ops)
    got = _run(eng, n, ops, flags=QSV_FLAG_FUSE, init=init)
    assert np.max(np.abs(got - exp)) < 1e-12
    got = _run(eng, n, ops, flags=QSV_FLAG_FUSE | QSV_FLAG_MSB_SWAP, init=init)
    assert np.max(np.abs(got - oracle.apply_program(init, n, ops, flags=QSV_FLAG_MSB_SWAP))) < 1e-12


@pytest.mark.parametrize("n", [13, 15, 17])
def test_fused_random_programs_c64(eng, oracle, n):
    rng = np.random.default_rng(900 + n)
    init = oracle.run(n, _program(rng, n, 20))
    ops = _program(rng, n, 300)
    note(f"fused c64 n={n}")
    exp = oracle.apply_program(init, n, ops)
    got = _run(eng, n, ops, prec="c64", flags=QSV_FLAG_FUSE, init=init)
    assert abs(np.vdot(exp, got)) ** 2 > 1 - 1e-6   # the complex64 fidelity floor
    assert np.max(np.abs(got - exp)) < 1e-5


def test_fused_every_position(eng, oracle):
    """Dense gates on every bit, controls on every bit, swaps on low/high mixes."""
    n = 16
    rng = np.random.default_rng(3)
    rows = []
    for q in range(n):
        rows += [(OP["RY"], q, 0, 0, 0, 0.3 + q), (OP["H"], q, 0, 0, 0, 0.0), (OP["RZ"], q, 0, 0, 0, 0.1 * q)]
        c = (q + 5) % n
        rows += [(OP["CRX"], c, q, 0, 0, 0.2 + q), (OP["CNOT"], q, c, 0, 0, 0.0), (OP["T"], q, 0, 0, 0, 0.0)]
        rows += [(OP["SWAP"], q, (q + 1) % n, 0, 0, 0.0), (OP["ISWAP"], q, (q + 9) % n, 0, 0, 0.0)]
    ops = qc.ops_from_tuples(rows)
    init = oracle.run(n, _program(rng, n, 30))
    exp = oracle.apply_program(init, n, ops)
    assert np.max(np.abs(_run(eng, n, ops, flags=QSV_FLAG_FUSE, init=init) - exp)) < 1e-12


def test_fused_measure_barrier(eng, oracle):
    n = 14
    rows = [(OP["H"], q, 0, 0, 0, 0.0) for q in range(n)] + [(OP["MEASURE"], 0, 0, 0, 0b1011, 0.0)]
    rows += [(OP["RY"], q, 0, 0, 0, 0.3 * q) for q in range(n)] + [(OP["CNOT"], 0, 13, 0, 0, 0.0)]
    rows += [(OP["MEASURE"], 0, 0, 0, 1 << 13, 0.0)] + [(OP["RX"], q, 0, 0, 0, 0.2) for q in range(n)]
    ops = qc.ops_from_tuples(rows)
    exp = oracle.run(n, ops, seed=77)
    assert np.max(np.abs(_run(eng, n, ops, flags=QSV_FLAG_FUSE, seed=77) - exp)) < 1e-12


def test_fused_equals_per_gate_layers(eng):
    """The 22-qubit random-layer circuit: fused and per-gate paths agree."""
    n = 22
    ops = wl.random_layer_ops(n, 10, seed=4)
    a = _run(eng, n, ops)
    b = _run(eng, n, ops, flags=QSV_FLAG_FUSE)
    assert np.max(np.abs(a - b)) < 1e-12


def test_small_states_fall_back(eng, oracle):
    n = 8
    rng = np.random.default_rng(1)
    ops = _program(rng, n, 100)
    assert np.max(np.abs(_run(eng, n, ops, flags=QSV_FLAG_FUSE) - oracle.run(n, ops))) < 1e-12


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_tuner_picks_a_candidate_and_results_hold(eng, prec):
    """qsv_tune_fused times every candidate width and keeps the narrowest
    within 1 % of the fastest; the fused path under the tuned width still
    equals the per-gate path."""
    note(f"tune_fused prec={prec} n=20")
    res = eng.tune_fused(prec, 20)
    times = res["ms_per_layer"]
    assert len(times) >= 2 and res["low_bits"] in times
    assert all(t > 0 for t in times.values())
    assert times[res["low_bits"]] <= 1.01 * min(times.values())
    assert all(times[w] > 1.01 * min(times.values()) for w in times if w < res["low_bits"])
    n = 21
    ops = wl.random_layer_ops(n, 6, seed=12)
    a = _run(eng, n, ops, prec)
    b = _run(eng, n, ops, prec, flags=QSV_FLAG_FUSE)
    tol = 1e-12 if prec == "c128" else 1e-5
    assert np.max(np.abs(a - b)) < tol
    with pytest.raises(Exception, match="16..30"):
        eng.tune_fused(prec, 40)


def test_c64_fp32_passes_renormalise(eng):
    """complex64 passes computed in FP32 keep the norm (the per-pass
    renormalisation removes the drift of the rounded matrices) and stay within
    the stated fidelity of the FP64 passes and of complex128."""
    from qclojure_amd._abi import QSV_FLAG_FP64_PASSES
    n = 22
    ops = wl.random_layer_ops(n, 60, seed=31)
    note(f"c64 fp32 passes n={n} gates={len(ops)}")
    with eng.DeviceState(n, "c128") as ref, eng.DeviceState(n, "c64") as f32, \
            eng.DeviceState(n, "c64") as f64:
        ref.apply(ops, QSV_FLAG_FUSE)
        f32.apply(ops, QSV_FLAG_FUSE)
        f64.apply(ops, QSV_FLAG_FUSE | QSV_FLAG_FP64_PASSES)
        assert abs(f32.norm2() - 1.0) < 2e-6
        for s in (f32, f64):
            c = s.compare(ref)
            assert c["fidelity"] > 1 - 1e-6 and c["max_abs_diff"] < 1e-5, c


def _block_program(rng, n, layers):
    """Layers of U3 (as RZ RY RZ, folded into one dense general 2x2) on every
    qubit, then a random matching of CZ, CNOT, CRX/CRY/CRZ, SWAP, Toffoli and
    two-atom blockades: long runs of register-bit gates between gates with
    thread-bit parts and layout barriers."""
    rows = []
    for _ in range(layers):
        for q in range(n):
            for k in ("RZ", "RY", "RZ"):
                rows.append((OP[k], q, 0, 0, 0, float(rng.uniform(0, 2 * math.pi))))
        perm = rng.permutation(n)
        for j in range(0, n - 1, 2):
            a, b = int(perm[j]), int(perm[j + 1])
            r = rng.uniform()
            ang = float(rng.uniform(0, 2 * math.pi))
            if r < 0.35:
                rows.append((OP["CZ"], a, b, 0, 0, 0.0))
            elif r < 0.6:
                rows.append((OP["CNOT"], a, b, 0, 0, 0.0))
            elif r < 0.75:
                rows.append((OP[rng.choice(["CRX", "CRY", "CRZ"])], a, b, 0, 0, ang))
            elif r < 0.85:
                rows.append((OP["SWAP"], a, b, 0, 0, 0.0))
            elif r < 0.95:
                c = int(rng.integers(n))
                if c not in (a, b):
                    rows.append((OP["TOFFOLI"], a, c, b, 0, 0.0))
            else:
                rows.append((OP["RYDBERG_BLOCKADE"], 0, 0, 0, (1 << a) | (1 << b), ang))
    return qc.ops_from_tuples(rows)


@pytest.mark.parametrize("n", [13, 17, 22])
def test_c64_matrix_core_blocks(eng, oracle, n, monkeypatch):
    """complex64 FP32 passes fold a stage's long register-bit runs into dense
    16x16 blocks on the matrix cores (moving them past thread-bit gates on
    other bits); QSV_PASS_BLOCKS=1 makes every run a block, and
    QSV_FLAG_VALU_PASSES keeps every gate on the vector units. Each must reach
    the oracle's state (complex128 fused passes at 22 qubits) at the complex64
    bound, on random programs (controls on thread and register bits,
    blockades, swaps, Toffolis), random layers and U3/CZ-style layers."""
    from qclojure_amd._abi import QSV_FLAG_VALU_PASSES
    rng = np.random.default_rng(1300 + n)
    progs = (("random", _program(rng, n, 300)), ("layers", wl.random_layer_ops(n, 8, seed=n)),
             ("u3", _block_program(rng, n, 12)))

    def check(got, exp):
        assert abs(np.vdot(exp, got)) ** 2 > 1 - 1e-6
        assert np.max(np.abs(got - exp)) < 1e-5
        assert abs(float(np.vdot(got, got).real) - 1.0) < 2e-6

    for name, ops in progs:
        note(f"c64 blocks n={n} {name} gates={len(ops)}")
        exp = oracle.run(n, ops) if n <= 17 else _run(eng, n, ops, "c128", QSV_FLAG_FUSE)
        check(_run(eng, n, ops, "c64", QSV_FLAG_FUSE | QSV_FLAG_VALU_PASSES), exp)
        for cost in ("1", None):
            if cost:
                monkeypatch.setenv("QSV_PASS_BLOCKS", cost)
            else:
                monkeypatch.delenv("QSV_PASS_BLOCKS", raising=False)
            check(_run(eng, n, ops, "c64", QSV_FLAG_FUSE), exp)
        monkeypatch.delenv("QSV_PASS_BLOCKS", raising=False)


def _fold_heavy_program(rng, n, layers):
    """Programs full of what the host folds and of what must stop a fold:
    RZ RY RZ chains, RZ on a CNOT control and RX on its target (both commute
    through it), CNOT RZ CNOT and CNOT PHASE CNOT triples, CZ between
    diagonals, and H on a CNOT target or a SWAP in between (no fold)."""
    rows = []
    for _ in range(layers):
        for q in range(n):
            for k in ("RZ", "RY", "RZ"):
                rows.append((OP[k], q, 0, 0, 0, float(rng.uniform(0, 2 * math.pi))))
        perm = rng.permutation(n)
        for j in range(0, n - 1, 2):
            c, t = int(perm[j]), int(perm[j + 1])
            r = rng.uniform()
            ang = float(rng.uniform(0, 2 * math.pi))
            if r < 0.3:
                rows += [(OP["CNOT"], c, t, 0, 0, 0.0), (OP["RZ"], t, 0, 0, 0, ang), (OP["CNOT"], c, t, 0, 0, 0.0)]
            elif r < 0.4:
                rows += [(OP["CNOT"], c, t, 0, 0, 0.0), (OP["PHASE"], t, 0, 0, 0, ang), (OP["CNOT"], c, t, 0, 0, 0.0)]
            elif r < 0.6:
                rows += [(OP["CNOT"], c, t, 0, 0, 0.0), (OP["RZ"], c, 0, 0, 0, ang), (OP["RX"], t, 0, 0, 0, ang)]
            elif r < 0.75:
                rows += [(OP["CZ"], c, t, 0, 0, 0.0), (OP["T"], c, 0, 0, 0, 0.0), (OP["S"], t, 0, 0, 0, 0.0)]
            elif r < 0.9:
                rows += [(OP["CNOT"], c, t, 0, 0, 0.0), (OP["H"], t, 0, 0, 0, 0.0), (OP["H"], t, 0, 0, 0, 0.0)]
            else:
                rows += [(OP["SWAP"], c, t, 0, 0, 0.0), (OP["RZ"], c, 0, 0, 0, ang)]
    return qc.ops_from_tuples(rows)


@pytest.mark.parametrize("n", [12, 14, 17])
def test_folded_programs_vs_oracle(eng, oracle, n):
    """The fused path folds gates on the host; the result must still be the
    oracle's state: 1e-12 for complex128, complex64 at its fidelity bound."""
    rng = np.random.default_rng(900 + n)
    ops = _fold_heavy_program(rng, n, 6)
    exp = oracle.run(n, ops)
    got = _run(eng, n, ops, "c128", QSV_FLAG_FUSE)
    err = float(np.max(np.abs(got - exp)))
    note(f"folded n={n} ops={len(ops)} c128 max|d|={err:.2e}")
    assert err < 1e-12
    got64 = _run(eng, n, ops, "c64", QSV_FLAG_FUSE)
    fid = abs(np.vdot(exp, got64)) ** 2
    note(f"folded n={n} c64 1-F={1 - fid:.2e}")
    assert abs(1 - fid) < 1e-6

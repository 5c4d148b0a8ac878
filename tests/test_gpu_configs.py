"""The benchmark configurations at their full sizes on the MI355X:

* the 30-qubit random-layer circuit (100 layers), complex128 and complex64,
  per gate and through fused passes: agreement of the two paths, and the
  round trip U then U^-1 back to |0..0>;
* 32 qubits: the reference QFT on a basis state against its closed form, then
  the precision circuit (product state, QFT, 50 layers of U3 + CZ) with the
  complex64 state's fidelity against complex128 asserted every 10 layers;
* the serving mix with 25-30-qubit requests and 1000 shots each, one checked
  shot for shot against the oracle and one against the known outcome of a
  30-qubit U U^-1 X circuit.

Sizes independent of the oracle are checked through properties (inverse round
trips, closed forms); the oracle checks the 25-qubit request directly."""
import math
import sys

import numpy as np
import pytest

from qclojure_amd import algorithms as alg
from qclojure_amd import circuit as qc
from qclojure_amd import workloads as wl
from qclojure_amd._abi import OP, QSV_FLAG_FUSE

pytestmark = pytest.mark.gpu

TOL_C128 = 1e-12
TOL_C64 = 1e-6
FID_C64 = 1.0 - 1e-6


def note(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def eng():
    from qclojure_amd import engine
    engine.init(0)
    return engine


def _free_gib(eng):
    return eng.device_info()["hbm-free-bytes"] / 2**30


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_c2_30q_per_gate_vs_fused_and_round_trip(eng, prec):
    n = 30
    per = n + n // 2
    ops = wl.random_layer_ops(n, 100, seed=2026)
    two = ops[:2 * per]
    tol = TOL_C128 if prec == "c128" else TOL_C64
    note(f"C2 {prec}: 30 qubits, free HBM {_free_gib(eng):.0f} GiB")
    with eng.DeviceState(n, prec) as a, eng.DeviceState(n, prec) as b, eng.DeviceState(n, prec) as e0:
        e0.init_basis(0)
        note(f"C2 {prec}: two layers per gate and fused")
        a.apply(two)
        b.apply(two, QSV_FLAG_FUSE)
        c = a.compare(b)
        note(f"C2 {prec}: per-gate vs fused {c}")
        if prec == "c128":
            assert c["max_abs_diff"] < TOL_C128
        else:
            assert c["fidelity"] / (c["norm2_self"] * c["norm2_other"]) > FID_C64
            assert c["max_abs_diff"] < TOL_C64
        assert abs(c["norm2_self"] - 1) < (1e-12 if prec == "c128" else 1e-6)
        assert abs(c["norm2_other"] - 1) < (1e-12 if prec == "c128" else 1e-6)
        inv2 = wl.inverse_ops(two)
        a.apply(inv2)
        b.apply(inv2, QSV_FLAG_FUSE)
        for name, s in (("per-gate", a), ("fused", b)):
            r = s.compare(e0)
            note(f"C2 {prec}: two layers and inverse, {name}: {r}")
            assert r["max_abs_diff"] < tol, name
            assert r["fidelity"] > (1 - 1e-12 if prec == "c128" else FID_C64), name
        # the whole 100-layer circuit and its inverse through fused passes
        note(f"C2 {prec}: 100 layers + inverse fused")
        b.init_basis(0)
        b.apply(ops, QSV_FLAG_FUSE)
        nb = b.norm2()
        assert abs(nb - 1) < (1e-12 if prec == "c128" else 1e-6), nb
        b.apply(wl.inverse_ops(ops), QSV_FLAG_FUSE)
        r = b.compare(e0)
        note(f"C2 {prec}: 100 layers and inverse, fused: {r}")
        assert r["max_abs_diff"] < tol
        assert abs(r["norm2_self"] - 1) < (1e-12 if prec == "c128" else 1e-6)


def _qft_amplitudes(n, x, ys):
    """Closed form of the reference QFT (H + CRZ + SWAPs) on |x> at indices
    ys: the textbook QFT exp(2 pi i x y / 2^n) / 2^(n/2) times the global
    phase exp(-i/2 sum_q sum_k theta_k x_{q+k+1}) that the CRZs leave (each
    CRZ is a controlled phase times e^{-i theta/2} on its control, which is
    still a basis bit when it acts). x*y is reduced modulo 2^n in exact
    integer arithmetic."""
    bits = [(x >> (n - 1 - q)) & 1 for q in range(n)]
    phi = 0.0
    for q in range(n):
        for k in range(n - q - 1):
            phi -= (math.pi / 2 ** (k + 1)) * bits[q + k + 1] / 2
    N = 1 << n
    frac = np.array([((x * int(y)) % N) / N for y in ys])
    return np.exp(1j * phi) * np.exp(2j * math.pi * frac) / math.sqrt(N)


def test_c4_32q_qft_closed_form(eng):
    n = 32
    rng = np.random.default_rng(2032)
    x = int(rng.integers(0, 1 << n))
    ops = qc.decode_circuit(alg.quantum_fourier_transform_circuit(n))
    offsets = [0, (1 << n) - 512] + [int(v) for v in rng.integers(0, (1 << n) - 512, size=14)]
    for prec, flags, tol in (("c128", QSV_FLAG_FUSE, TOL_C128), ("c64", QSV_FLAG_FUSE, 1e-9),
                             ("c128", 0, TOL_C128)):
        note(f"C4 QFT 32q {prec} flags={flags} x={x}, free HBM {_free_gib(eng):.0f} GiB")
        with eng.DeviceState(n, prec) as s:
            s.init_basis(x)
            s.apply(ops, flags)
            worst = 0.0
            for off in offsets:
                got = s.download(off, 512)
                exp = _qft_amplitudes(n, x, range(off, off + 512))
                worst = max(worst, float(np.max(np.abs(got - exp))))
            norm = s.norm2()
        note(f"C4 QFT {prec} flags={flags}: max|d| over {len(offsets) * 512} amplitudes {worst:.2e}, "
             f"norm^2-1 {norm - 1:.2e}")
        assert worst < tol
        assert abs(norm - 1) < (1e-12 if prec == "c128" else 1e-6)


def test_c4_32q_precision_study(eng):
    """complex64 fused passes vs complex128 on the precision circuit: overlap
    fidelity |<psi64|psi128>|^2 >= 1 - 1e-6, max |d| and norm drift, every 10
    layers."""
    n, layers = 32, 50
    circ = wl.precision_circuit(n, layers, seed=2026)
    ops = qc.decode_circuit(circ)
    prologue = 2 * n + len(alg.quantum_fourier_transform_circuit(n)["operations"])
    per_layer = 3 * n + n // 2
    assert len(ops) == prologue + layers * per_layer
    with eng.DeviceState(n, "c128") as a, eng.DeviceState(n, "c64") as b:
        a.apply(ops[:prologue], QSV_FLAG_FUSE)
        b.apply(ops[:prologue], QSV_FLAG_FUSE)
        for L in range(0, layers, 10):
            chunk = ops[prologue + L * per_layer:prologue + (L + 10) * per_layer]
            a.apply(chunk, QSV_FLAG_FUSE)
            b.apply(chunk, QSV_FLAG_FUSE)
            c = b.compare(a)
            note(f"C4 precision after {L + 10} layers: {c}")
            assert c["fidelity"] > FID_C64, (L + 10, c)
            assert c["max_abs_diff"] < 1e-8
            assert abs(c["norm2_self"] - 1) < 1e-6 and abs(c["norm2_other"] - 1) < 1e-12


def test_c5_serving_large_requests(eng, oracle):
    """A serving mix with 25-30-qubit requests and 1000 shots each: the
    25-qubit request against the oracle shot for shot (inverse-CDF rule on
    the exact FP64 state), the 30-qubit U U^-1 X(0,5,29) request against its
    one possible outcome, the others against the same circuit run alone."""
    def uux(n, depth, seed, flips):
        ops = wl.random_layer_ops(n, depth, seed)
        x = qc.ops_from_tuples([(OP["X"], q, 0, 0, 0, 0.0) for q in flips])
        return np.concatenate([ops, wl.inverse_ops(ops), x])

    reqs = [
        ("oracle", 25, wl.random_layer_ops(25, 10, seed=525), 1000, 9001),
        ("basis", 30, uux(30, 10, 530, [0, 5, 29]), 1000, 9002),
        ("alone", 28, wl.random_layer_ops(28, 12, seed=528), 1000, 9003),
        ("alone", 27, wl.random_layer_ops(27, 30, seed=527), 1000, 9004),
        ("basis", 26, uux(26, 40, 526, [3]), 1000, 9005),
        ("alone", 12, wl.random_layer_ops(12, 60, seed=512), 1000, 9006),
        ("alone", 14, wl.random_layer_ops(14, 20, seed=514), 1000, 9007),
        ("alone", 20, wl.random_layer_ops(20, 80, seed=520), 1000, 9008),
    ]
    note(f"C5: serving {len(reqs)} requests, free HBM {_free_gib(eng):.0f} GiB")
    with eng.Server("c64", flags=QSV_FLAG_FUSE) as srv:
        ids = [srv.submit(n, ops, shots, seed) for (_, n, ops, shots, seed) in reqs]
        got = [srv.wait(i, r[3]) for i, r in zip(ids, reqs)]
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    for (kind, n, ops, shots, seed), (info, outs) in zip(reqs, got):
        note(f"C5: {kind} n={n} {info['status']} latency {info['latency-ms']:.1f} ms "
             f"norm drift {info.get('norm-drift')}")
        assert info["status"] == "completed" and len(outs) == shots
        assert info["norm-drift"] < 1e-6
        if kind == "oracle":
            psi = oracle.run(n, ops)
            ref = oracle.sample_prefix(psi, n, shots, seed)
            agree = float(np.mean(outs == ref))
            note(f"C5: 25q shots vs oracle agreement {agree:.4f}")
            assert agree > 0.99
            assert np.all(np.abs(psi[outs.astype(np.int64)]) ** 2 > 0)
        elif kind == "basis":
            flips = [0, 5, 29] if n == 30 else [3]
            idx = sum(1 << (n - 1 - q) for q in flips)
            assert np.all(outs == idx), (n, np.unique(outs)[:5])
        else:
            with eng.DeviceState(n, "c64") as s:
                s.apply(ops, QSV_FLAG_FUSE)
                ref = s.sample(shots, seed)
            assert np.mean(outs == ref) > 0.99, n


def test_c3_4096_qaoa_batch_vs_oracle_stratified(eng, oracle):
    """The batched QAOA workload at its full size: 4096 MaxCut instances
    (n in {8..20} even, p in 1..4, gamma/beta uniform) in one complex64
    batch call, as the benchmark runs it; every (n, p) cell is checked on
    two instances against the FP64 oracle (sum_k c_k <P_k> in term order),
    and every energy lies in [min cut term sum, max] bounds of the
    Hamiltonian (|<H>| <= sum |c_k|)."""
    insts = [wl.qaoa_instance(2026 * 100003 + k) for k in range(4096)]
    batch = [(i["n"], wl.qaoa_ops(i), wl.qaoa_terms(i)) for i in insts]
    note("C3: 4096-instance complex64 batch")
    e64 = np.asarray(eng.PackedBatch(batch, "c64").run())
    assert e64.shape == (4096,) and np.all(np.isfinite(e64))
    bound = np.array([sum(abs(t["coefficient"]) for t in i["hamiltonian"]) for i in insts])
    assert np.all(np.abs(e64) <= bound + 1e-6)
    cells, picked = {}, []
    for k, inst in enumerate(insts):
        key = (inst["n"], inst["p"])
        if cells.get(key, 0) < 2:
            cells[key] = cells.get(key, 0) + 1
            picked.append(k)
    assert len(cells) == 28
    worst = 0.0
    for k in picked:
        inst = insts[k]
        psi = oracle.run(inst["n"], wl.qaoa_ops(inst))
        ref = 0.0
        for t in inst["hamiltonian"]:
            ref += t["coefficient"] * oracle.pauli_expectation(psi, inst["n"], t["pauli-string"])
        worst = max(worst, abs(ref - e64[k]))
    note(f"C3: {len(picked)} instances over {len(cells)} cells, max |dE| vs FP64 oracle {worst:.2e}")
    assert worst < 1e-5

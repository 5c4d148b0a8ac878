"""Benchmark: HBM throughput of the per-gate state-vector path at 30 qubits.

Workload: the 30-qubit random-layer circuit (each layer: H/RX/RY/RZ on every
qubit, then CNOTs on a random perfect matching; 100 layers, seed 2026) on a
complex128 state of 2^30 amplitudes (16 GiB) resident in HBM. One step = one
gate layer (30 one-qubit gates + 15 CNOTs), one kernel per gate.

value = algorithmic bytes of the timed layers / wall time (GB/s), where a gate
that touches t amplitudes moves 2 * t * bytes_per_amplitude (read + write).

    python bench.py [--steps K] [--warmup W] [--prec c128|c64] [--qubits 30]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "HBM GB/s (% peak) per gate layer at 30q; served circuits/s + p99 vs plain"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def _host_info() -> dict:
    mem = None
    try:
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemTotal:"):
                    mem = round(int(line.split()[1]) / 2**20, 1)
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"cpu_model": cpu_model(), "cpus_available": avail, "ram_gib": mem}


_QAOA_SAMPLE: dict = {}


def cpu_baseline(n: int, seed: int) -> dict:
    """The CPU leg: the oracle (an FP64 C restatement of the reference's
    algorithms, oracle/qsv_ref.c) timed on this host, and the oracle as the
    checker of the batched QAOA energies.

    value: the pair-update mode (the same mathematics as in-place O(2^n)
    updates, OpenMP on the host cores) on one n-qubit layer of the headline
    construction, in the headline's unit. Beside it: the literal mode (the
    reference's own algorithm, gate.clj:323-395: every one-qubit gate as a
    dense 2^n x 2^n Kronecker product times the state, one thread) on two
    layers of the 12-qubit, 20-layer random circuit."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from qclojure_amd import workloads as wl
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ops = wl.random_layer_ops(n, 1, seed)
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1.0
    S = 1 << n
    t0 = time.perf_counter()
    rc = oracle.lib().ref_apply_program(oracle._p(psi), n, oracle._p(ops), len(ops), 0, 0, 0)
    dt = time.perf_counter() - t0
    if rc != 0:
        raise RuntimeError("oracle failed")
    del psi
    nbytes = sum((16 if int(r["kind"]) == 13 else 32) * S for r in ops)  # CNOT touches S/2, a 1-qubit gate S
    out = {"value": round(nbytes / dt / 1e9, 2), "unit": "GB/s", "cores": threads, "kind": "port",
           "sample": f"one {n}-qubit complex128 layer of the same construction ({len(ops)} gates, "
                     f"{nbytes / 1e9:.0f} GB algorithmic) by the oracle's OpenMP pair-update mode in {dt:.2f} s",
           "host": _host_info()}
    log(f"bench: cpu pair mode {out['value']} GB/s on {threads} threads")
    # the reference's algorithm, literally, on the 12-qubit random circuit
    n1, layers1 = 12, 2
    ops1 = wl.random_layer_ops(n1, 20, 1)[:layers1 * (n1 + n1 // 2)]
    t0 = time.perf_counter()
    psi1 = oracle.run(n1, ops1, literal=True)
    dt1 = time.perf_counter() - t0
    b1 = sum((16 if int(r["kind"]) == 13 else 32) * (1 << n1) for r in ops1)
    out["literal_12q"] = {"seconds_per_layer": round(dt1 / layers1, 3), "GB_s": round(b1 / dt1 / 1e9, 6),
                          "cores": 1, "layers_timed": layers1,
                          "projected_20_layer_circuit_s": round(dt1 / layers1 * 20, 1),
                          "norm2_minus_1": float(np.sum(np.abs(psi1) ** 2) - 1.0)}
    log(f"bench: cpu literal 12q {out['literal_12q']}")
    # the batched complex64 QAOA energies against the FP64 oracle on a
    # stratified sample: up to two instances per (qubits, p) cell
    if _QAOA_SAMPLE:
        insts, e64 = _QAOA_SAMPLE["insts"], _QAOA_SAMPLE["e64"]
        picked, cells = [], {}
        for k, inst in enumerate(insts):
            key = (inst["n"], inst["p"])
            if cells.get(key, 0) < 2:
                cells[key] = cells.get(key, 0) + 1
                picked.append(k)
        t0 = time.perf_counter()
        errs = []
        for k in picked:
            inst = insts[k]
            st = oracle.run(inst["n"], wl.qaoa_ops(inst))
            ref = 0.0
            for t in inst["hamiltonian"]:
                ref += t["coefficient"] * oracle.pauli_expectation(st, inst["n"], t["pauli-string"])
            errs.append(abs(ref - e64[k]))
        out["qaoa_oracle_check"] = {"instances": len(picked), "cells": len(cells),
                                    "max_abs_energy_err_c64_vs_fp64_oracle": float(max(errs)),
                                    "mean_abs_energy_err": float(np.mean(errs)),
                                    "oracle_seconds": round(time.perf_counter() - t0, 2)}
        log(f"bench: qaoa oracle check {out['qaoa_oracle_check']}")
    return out


def position_sweep(engine, n: int, prec: str, reps: int = 3) -> dict:
    """GB/s of each per-gate kernel class at every target bit position
    (bit b = n-1-q for reference qubit q), timed with HIP events."""
    from qclojure_amd import circuit as qc
    from qclojure_amd._abi import OP
    dev = engine.DeviceState(n, prec)
    table = {}
    gates = {"H": lambda q: [(OP["H"], q, 0, 0, 0, 0.0)],
             "RZ": lambda q: [(OP["RZ"], q, 0, 0, 0, 0.37)],
             "T": lambda q: [(OP["T"], q, 0, 0, 0, 0.0)],
             "CNOT(c=q-1)": lambda q: [(OP["CNOT"], (q - 1) % n, q, 0, 0, 0.0)],
             "CZ(c=q-1)": lambda q: [(OP["CZ"], (q - 1) % n, q, 0, 0, 0.0)],
             # SWAP's qubit numbers are LSB bit positions (gate.clj:768-778): this swaps index bits b and b-1
             "SWAP(b,b-1)": lambda q: [(OP["SWAP"], n - 1 - q, n - 2 - q, 0, 0, 0.0)]}
    for name, mk in gates.items():
        row = {}
        for q in range(n - 1 if name.startswith("SWAP") else n):
            ops = qc.ops_from_tuples(mk(q) * reps)
            dev.apply(ops[:1])
            dev.sync()
            dev.timing(True)
            dev.timing_read(reset=True)
            dev.apply(ops)
            st = dev.timing_read(reset=True)
            dev.timing(False)
            ms = sum(v["total_ms"] for v in st.values())
            by = sum(v["bytes"] for v in st.values())
            row[f"b{n - 1 - q}"] = round(by / (ms / 1e3) / 1e9, 1)
        table[name] = row
        log(f"sweep {name}: " + " ".join(f"{k}={v:.0f}" for k, v in row.items()))
    dev.destroy()
    return table


def _touched(ops: "np.ndarray", n: int) -> float:
    """Amplitude updates of a program: S per 1-qubit gate, S/2 per controlled gate."""
    import numpy as np
    from qclojure_amd._abi import OP
    k = ops["kind"]
    ctl = np.isin(k, [OP["CNOT"], OP["CZ"], OP["CRX"], OP["CRY"], OP["CRZ"], OP["CY"]])
    return float((len(ops) - ctl.sum()) * (1 << n) + ctl.sum() * (1 << (n - 1)))


def bench_batched_qaoa(engine, count: int, seed: int, plain_every: int) -> dict:
    """4096 MaxCut QAOA instances (n in {8..20} even, p in 1..4), complex64:
    per-instance <H_C> from one batched call vs the plain path (one circuit at a
    time: per-gate kernels, then a Pauli-expectation pass per 8 terms)."""
    import numpy as np
    from qclojure_amd import workloads as wl
    insts = [wl.qaoa_instance(seed * 100003 + k) for k in range(count)]
    batch = []
    upd = 0.0
    for inst in insts:
        ops = wl.qaoa_ops(inst)
        batch.append((inst["n"], ops, wl.qaoa_terms(inst)))
        upd += _touched(ops, inst["n"])
    engine.batch_expect_z(batch[:64], "c64")               # warm-up: kernels, pooled states
    t0 = time.perf_counter()
    packed = engine.PackedBatch(batch, "c64")
    t_pack = time.perf_counter() - t0
    t0 = time.perf_counter()
    e64 = packed.run()
    t_call = time.perf_counter() - t0
    t_batch = t_pack + t_call
    # the same batch again: its grouped launches replay the graph captured by
    # the first call (an optimiser loop's steady state), then issued kernel by
    # kernel for comparison
    t0 = time.perf_counter()
    e64_replay = packed.run()
    t_replay = time.perf_counter() - t0
    os.environ["QSV_BATCH_GRAPH"] = "0"
    try:
        t0 = time.perf_counter()
        e64_eager = packed.run()
        t_eager = time.perf_counter() - t0
    finally:
        del os.environ["QSV_BATCH_GRAPH"]
    graph_diff = float(max(np.max(np.abs(e64_replay - e64)), np.max(np.abs(e64_eager - e64))))
    _QAOA_SAMPLE.update(insts=insts, e64=e64)
    e128 = engine.batch_expect_z(batch, "c128")
    err = float(max(abs(a - b) for a, b in zip(e64, e128)))
    # plain path on every plain_every-th instance
    sample = [(n, ops, [(t["coefficient"], t["pauli-string"]) for t in inst["hamiltonian"]])
              for (n, ops, _), inst in zip(batch[::plain_every], insts[::plain_every])]
    s_upd = sum(_touched(ops, n) for (n, ops, _) in sample)
    states = {}
    t0 = time.perf_counter()
    e_plain = []
    for (n, ops, terms) in sample:
        if n not in states:
            states[n] = engine.DeviceState(n, "c64")
        s = states[n]
        s.init_basis(0)
        s.apply(ops)
        vals = s.expect_pauli([p for _, p in terms])
        e_plain.append(sum(c * v for (c, _), v in zip(terms, vals)))
    t_plain = time.perf_counter() - t0
    for s in states.values():
        s.destroy()
    plain_err = float(max(abs(a - b) for a, b in zip(e_plain, e64[::plain_every])))
    sizes = sorted({b[0] for b in batch})
    return {"instances": count, "qubits": sizes, "gates": int(sum(len(b[1]) for b in batch)),
            "batched_s": round(t_batch, 4), "pack_s": round(t_pack, 4), "library_call_s": round(t_call, 4),
            "library_call_graph_replay_s": round(t_replay, 4), "library_call_eager_launches_s": round(t_eager, 4),
            "max_abs_diff_replay_eager_vs_first": graph_diff,
            "batched_amp_updates_per_s": round(upd / t_batch, 1),
            "plain_sample_instances": len(sample), "plain_s": round(t_plain, 4),
            "plain_amp_updates_per_s": round(s_upd / t_plain, 1),
            "speedup": round((upd / t_batch) / (s_upd / t_plain), 2),
            "max_abs_energy_err_c64_vs_c128": err, "max_abs_energy_diff_plain_vs_batched": plain_err,
            "note": "plain = one circuit at a time, per-gate kernels + Pauli reductions; timed on a stratified sample"}


def _sequential_real(engine, reqs, arr, flags) -> dict:
    """One request at a time in arrival order, with real arrivals: a request
    that arrives while another runs waits in a FIFO queue; latency = results
    on the host - arrival. flags 0: per-gate kernels; QSV_FLAG_FUSE: fused
    passes. Then the shots, as the served path takes them."""
    import numpy as np
    states = {}
    lat = []
    t_start = time.perf_counter()
    for (n, ops, shots, sd), a in zip(reqs, arr):
        now = time.perf_counter() - t_start
        if now < a:
            time.sleep(a - now)
        if n not in states:
            states[n] = engine.DeviceState(n, "c64")
        s = states[n]
        s.init_basis(0)
        s.apply(ops, flags)
        s.sample(shots, sd)
        lat.append(time.perf_counter() - t_start - a)
    t_total = time.perf_counter() - t_start
    for s in states.values():
        s.destroy()
    return {"circuits_per_s": round(len(reqs) / t_total, 2),
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 1),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 1)}


def bench_serving(engine, rates, n_req: int, n_plain: int, seed: int) -> dict:
    """Poisson request stream (n log-uniform in [10, 30], depth 10..100 layers,
    1000 shots, complex64): the native serving loop against two one-at-a-time
    baselines with the same real arrivals: "plain" (per-gate kernels, the
    plain GPU path) and "sequential_fused" (fused passes, no batching or
    overlap), so that the ratio served/sequential_fused is what continuous
    batching earns on top of fusion. The request contents are fixed; each rate
    rescales the same unit-rate arrival times."""
    import threading
    import numpy as np
    from qclojure_amd import workloads as wl
    from qclojure_amd._abi import QSV_FLAG_FUSE
    stream = wl.serving_stream(n_req, 1.0, seed)
    reqs = [(r["n"], wl.random_layer_ops(r["n"], r["depth"], r["seed"]), r["shots"], r["seed"]) for r in stream]
    unit_arrivals = [r["arrival-s"] for r in stream]
    gates = sum(len(r[1]) for r in reqs)
    n_plain = n_plain if 0 < n_plain < n_req else n_req
    # warm every state size and kernel once, untimed
    _sequential_real(engine, reqs[:4], [0.0] * 4, QSV_FLAG_FUSE)
    out = {"requests": n_req, "gates": gates, "plain_requests": n_plain, "rates": []}
    for lam in rates:
        arr = [a / lam for a in unit_arrivals]
        plain = _sequential_real(engine, reqs[:n_plain], arr[:n_plain], 0)
        seqf = _sequential_real(engine, reqs, arr, QSV_FLAG_FUSE)
        # served: real arrivals into the native loop
        with engine.Server("c64", flags=QSV_FLAG_FUSE) as srv:
            ids = [None] * n_req
            t_start = time.perf_counter()

            def submitter():
                for k, (n, ops, shots, sd) in enumerate(reqs):
                    delay = arr[k] - (time.perf_counter() - t_start)
                    if delay > 0:
                        time.sleep(delay)
                    ids[k] = srv.submit(n, ops, shots, sd)

            th = threading.Thread(target=submitter)
            th.start()
            lats, drift = [], 0.0
            for k in range(n_req):
                while ids[k] is None:
                    time.sleep(0.0005)
                info, _ = srv.wait(ids[k], reqs[k][2])
                lats.append(info["latency-ms"])
                drift = max(drift, info.get("norm-drift", 0.0))
            th.join()
            t_total = time.perf_counter() - t_start
            st = srv.stats()
        served = {"circuits_per_s": round(n_req / t_total, 2),
                  "p50_ms": round(float(np.percentile(lats, 50)), 1),
                  "p99_ms": round(float(np.percentile(lats, 99)), 1),
                  "small_batched": int(st["small_batched_requests"]), "batches": int(st["small_batches"]),
                  "max_queue_depth": int(st["max_queue_depth"]), "max_norm_drift": drift}
        out["rates"].append({"lambda_per_s": lam, "served": served, "plain": plain, "sequential_fused": seqf,
                             "throughput_ratio": round(served["circuits_per_s"] / plain["circuits_per_s"], 2),
                             "batching_ratio": round(served["circuits_per_s"] / seqf["circuits_per_s"], 2)})
        log(f"bench: serving lambda={lam}: served {served} plain {plain} sequential_fused {seqf}")
    return out


def bench_precision(engine, n: int, layers: int, seed: int) -> dict:
    """Seeded random product state, the reference QFT, then layers of random U3
    (as RZ RY RZ) and CZ on a random matching: complex64 vs complex128, both
    through fused passes, compared every 10 layers. A third state runs the
    complex64 program with every gate on the vector units
    (QSV_FLAG_VALU_PASSES), so the section also times the matrix-core blocks
    of the FP32 passes against the vector units on the same chunks."""
    from qclojure_amd import algorithms as alg
    from qclojure_amd import circuit as qc
    from qclojure_amd import workloads as wl
    from qclojure_amd._abi import QSV_FLAG_FUSE, QSV_FLAG_VALU_PASSES
    circ = wl.precision_circuit(n, layers, seed)
    ops = qc.decode_circuit(circ)
    prologue = 2 * n + len(alg.quantum_fourier_transform_circuit(n)["operations"])
    per_layer = 3 * n + n // 2
    a = engine.DeviceState(n, "c128")
    b = engine.DeviceState(n, "c64")
    v = engine.DeviceState(n, "c64")
    rows = []
    t_blocks = t_valu = 0.0

    def timed_apply(s, part, flags):
        s.sync()
        t = time.perf_counter()
        s.apply(part, flags)
        s.sync()
        return time.perf_counter() - t

    t0 = time.perf_counter()
    a.apply(ops[:prologue], QSV_FLAG_FUSE)
    a.sync()   # the complex64 timings must not overlap the complex128 kernels
    t_blocks += timed_apply(b, ops[:prologue], QSV_FLAG_FUSE)
    t_valu += timed_apply(v, ops[:prologue], QSV_FLAG_FUSE | QSV_FLAG_VALU_PASSES)
    for L in range(0, layers, 10):
        chunk = ops[prologue + L * per_layer:prologue + min(layers, L + 10) * per_layer]
        a.apply(chunk, QSV_FLAG_FUSE)
        a.sync()
        t_blocks += timed_apply(b, chunk, QSV_FLAG_FUSE)
        t_valu += timed_apply(v, chunk, QSV_FLAG_FUSE | QSV_FLAG_VALU_PASSES)
        c = b.compare(a)
        cv = v.compare(a)
        rows.append({"layers": min(layers, L + 10), "fidelity": c["fidelity"],
                     "fidelity_normalized": c["fidelity"] / (c["norm2_self"] * c["norm2_other"]),
                     "max_abs_diff": c["max_abs_diff"], "norm_drift_c64": abs(c["norm2_self"] - 1.0),
                     "norm_drift_c128": abs(c["norm2_other"] - 1.0), "fidelity_valu_passes": cv["fidelity"]})
        log(f"bench: precision {rows[-1]}")
    dt = time.perf_counter() - t0
    a.destroy()
    b.destroy()
    v.destroy()
    return {"qubits": n, "layers": layers, "gates": int(len(ops)), "seconds": round(dt, 2), "per_10_layers": rows,
            "min_fidelity": min(r["fidelity"] for r in rows),
            "min_fidelity_normalized": min(r["fidelity_normalized"] for r in rows),
            "c64_fused_s_matrix_core_blocks": round(t_blocks, 4), "c64_fused_s_vector_units_only": round(t_valu, 4),
            "c64_blocks_speedup": round(t_valu / t_blocks, 4)}


def bench_la(engine, n: int, seed: int) -> dict:
    """The :hip math backend: complex128 GEMM on FP64 MFMA (8 n^3 flops), GEMV
    (HBM-bound), the Jacobi eigensolver, and the reference's gate application
    (a 2^q x 2^q Kronecker expansion times the state) dense vs kept as factors."""
    import numpy as np
    from qclojure_amd import la
    engine.init(0)
    hip = la.create_backend("hip")
    rng = np.random.default_rng(seed)
    out = {}
    A = la.DeviceMatrix.from_numpy(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    B = la.DeviceMatrix.from_numpy(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    hip.matrix_multiply(A, B).close()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        hip.matrix_multiply(A, B).close()
    ms = (time.perf_counter() - t0) / reps * 1e3
    tf = 8.0 * n ** 3 / (ms * 1e-3) / 1e12
    out["zgemm"] = {"n": n, "ms": round(ms, 3), "TFLOP_s": round(tf, 2), "peak_fp64_TFLOP_s": 78.6,
                    "frac": round(tf / 78.6, 4), "flops": 8 * n ** 3}
    log(f"bench: la zgemm {out['zgemm']}")
    A.close()
    B.close()
    m = 8192
    G = la.DeviceMatrix.from_numpy(rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m)))
    x = la.DeviceMatrix.from_numpy(rng.standard_normal(m) + 1j * rng.standard_normal(m))
    hip.matrix_vector_product(G, x).close()
    t0 = time.perf_counter()
    for _ in range(reps):
        hip.matrix_vector_product(G, x).close()
    ms = (time.perf_counter() - t0) / reps * 1e3
    gbs = (16.0 * m * m + 32.0 * m) / (ms * 1e-3) / 1e9
    out["zgemv"] = {"n": m, "ms": round(ms, 3), "GB_s": round(gbs, 1), "frac_of_8TB_s": round(gbs / 8000.0, 4)}
    G.close()
    x.close()
    for k in (16, 64, 256):
        H = rng.standard_normal((k, k)) + 1j * rng.standard_normal((k, k))
        H = H + H.conj().T
        Hd = la.DeviceMatrix.from_numpy(H)
        t0 = time.perf_counter()
        e = hip.eigen_hermitian(Hd)
        ms = (time.perf_counter() - t0) * 1e3
        w = np.array([v.real for v in e["eigenvalues"]])
        out[f"eigh_{k}"] = {"ms": round(ms, 2), "sweeps": e["sweeps"],
                            "max_abs_eigenvalue_err_vs_lapack": float(np.max(np.abs(w - np.linalg.eigvalsh(H))))}
        Hd.close()
    # the reference's apply-single-qubit-gate at q qubits: (reduce kronecker-product [I .. G .. I]) * psi
    q = 12
    Gm = np.array([[0.6, 0.8j], [0.8j, 0.6]], complex)
    psi = la.DeviceMatrix.from_numpy(rng.standard_normal(1 << q) + 1j * rng.standard_normal(1 << q))
    E = np.eye(2, dtype=complex)
    for j in range(1, q):
        E = hip.kronecker_product(E, Gm if j == q // 2 else np.eye(2, dtype=complex))
    hip.matrix_vector_product(E, psi).close()
    t0 = time.perf_counter()
    for _ in range(reps):
        hip.matrix_vector_product(E, psi).close()
    lazy_ms = (time.perf_counter() - t0) / reps * 1e3
    D = E.materialize()
    t0 = time.perf_counter()
    for _ in range(reps):
        hip.matrix_vector_product(D, psi).close()
    dense_ms = (time.perf_counter() - t0) / reps * 1e3
    Dh = D.to_numpy()
    ph = psi.to_numpy()[:, 0]
    t0 = time.perf_counter()
    Dh @ ph
    host_ms = (time.perf_counter() - t0) * 1e3
    out["gate_expansion_12q"] = {"kron_factors_ms": round(lazy_ms, 4), "dense_gemv_ms": round(dense_ms, 4),
                                 "host_dense_matvec_ms": round(host_ms, 3),
                                 "note": "one 1-qubit gate applied as the reference does (Kronecker expansion "
                                         "times the state): kept as factors on the device, as a dense 4096x4096 "
                                         "GEMV on the device, and the dense product on the host (numpy)"}
    D.close()
    psi.close()
    # decompositions and matrix functions at n = 128: wall clock of one call, and
    # the residual of the identity each one satisfies
    k = 128
    M = rng.standard_normal((k, k)) + 1j * rng.standard_normal((k, k))
    Hp = M @ M.conj().T / k + np.eye(k)
    Md = la.DeviceMatrix.from_numpy(M)
    Hd = la.DeviceMatrix.from_numpy(Hp)
    dec = {}

    def timed(name, fn, resid):
        t0 = time.perf_counter()
        r = fn()
        ms = (time.perf_counter() - t0) * 1e3
        dec[name] = {"ms": round(ms, 2), "max_abs_residual": float(resid(r))}

    timed("lu", lambda: hip.lu_decomposition(Md),
          lambda r: np.max(np.abs(np.asarray(r["P"]) @ M - np.asarray(r["L"]) @ np.asarray(r["U"]))))
    timed("qr", lambda: hip.qr_decomposition(Md),
          lambda r: np.max(np.abs(np.asarray(r["Q"]) @ np.asarray(r["R"]) - M)))
    timed("cholesky", lambda: hip.cholesky_decomposition(Hd),
          lambda r: np.max(np.abs(np.asarray(r["L"]) @ np.asarray(r["L"]).conj().T - Hp)))
    timed("svd", lambda: hip.svd(Md),
          lambda r: np.max(np.abs(np.asarray(r["U"]) @ np.diag([s.real for s in r["S"]]) @ np.asarray(r["V†"]) - M)))
    timed("matrix_sqrt_hermitian", lambda: hip.matrix_sqrt(Hd),
          lambda r: np.max(np.abs(np.asarray(r) @ np.asarray(r) - Hp)))
    timed("matrix_exp_general", lambda: hip.matrix_exp(hip.scale(Md, 1.0 / k)),
          lambda r: np.max(np.abs(np.asarray(r) @ np.asarray(hip.matrix_exp(hip.scale(Md, -1.0 / k))) - np.eye(k))))
    out["decompositions_128"] = dec
    Md.close()
    Hd.close()
    return out


def measure_layers(engine, n: int, prec: str, ops, args, tuned=None) -> dict:
    """The random-layer circuit at n qubits and one precision: `steps` layers
    through the per-gate kernels (one kernel per gate, HIP-event timing per
    kernel class), then the same layers as one fused program."""
    import numpy as np
    layers_total = 100
    per = n + n // 2
    layer = lambda k: ops[(k % layers_total) * per:((k % layers_total) + 1) * per]
    dev = engine.DeviceState(n, prec)
    log(f"bench: state {n} qubits {prec} ({dev.nbytes / 2**30:.1f} GiB) allocated")
    for w in range(args.warmup):
        dev.apply(layer(w))
    dev.sync()
    dev.timing(True)
    dev.timing_read(reset=True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        dev.apply(layer(args.warmup + k))
    dev.sync()
    elapsed = time.perf_counter() - t0
    stats = dev.timing_read(reset=True)
    dev.timing(False)
    norm = dev.norm2()
    log(f"bench: {prec} {args.steps} layers in {elapsed:.3f} s, norm^2-1 = {norm - 1:.2e}")
    total_bytes = sum(v["bytes"] for v in stats.values())
    kern = {}
    for name, v in stats.items():
        if v["launches"]:
            kern[name] = {"launches": v["launches"], "avg_ms": round(v["total_ms"] / v["launches"], 4),
                          "GB_s": round(v["bytes"] / (v["total_ms"] / 1e3) / 1e9, 1)}
    dom = max(stats, key=lambda k: stats[k]["total_ms"])
    d = stats[dom]
    if args.calibrate_copy:
        other = engine.DeviceState(n, prec)
        other.copy_from(dev)
        other.destroy()
        log(f"bench: calibration copy of {dev.nbytes} B done")
    fused = None
    if not args.no_fused:
        from qclojure_amd._abi import QSV_FLAG_FUSE
        if tuned is None and "QSV_TILE_LOWBITS" not in os.environ:
            # tuned on a scratch state of the workload's size (the tuner takes 16..30 qubits)
            tuned = engine.tune_fused(prec, num_qubits=max(16, min(n, 30)))
            tuned["qubits"] = max(16, min(n, 30))
            log(f"bench: tuned tile low bits {tuned}")
        # the fused path takes the timed layers as one program, as a circuit is
        # submitted, so passes may span layer boundaries
        if args.warmup:
            dev.apply(np.concatenate([layer(w) for w in range(args.warmup)]), QSV_FLAG_FUSE)
        program = np.concatenate([layer(args.warmup + k) for k in range(args.steps)])
        dev.sync()
        dev.timing(True)
        dev.timing_read(reset=True)
        t0 = time.perf_counter()
        dev.apply(program, QSV_FLAG_FUSE)
        dev.sync()
        f_elapsed = time.perf_counter() - t0
        fst = dev.timing_read(reset=True)
        dev.timing(False)
        blk = fst["block"]
        fused = {"ms_per_layer": round(f_elapsed / args.steps * 1e3, 3),
                 "speedup_vs_per_gate": round(elapsed / f_elapsed, 2),
                 "passes_per_layer": round(blk["launches"] / args.steps, 2),
                 "tile_kernel_avg_ms": round(blk["total_ms"] / max(1, blk["launches"]), 4),
                 "tile_kernel_GB_s": round(blk["bytes"] / (blk["total_ms"] / 1e3) / 1e9, 1) if blk["launches"] else None,
                 "tile_kernel_frac_of_peak": round(blk["bytes"] / (blk["total_ms"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
                 if blk["launches"] else None,
                 "effective_GB_s": round(total_bytes / args.steps / (f_elapsed / args.steps) / 1e9, 1),
                 "tile_low_bits": tuned["low_bits"] if tuned else int(os.environ["QSV_TILE_LOWBITS"]),
                 "tuner_qubits": tuned.get("qubits") if tuned else None,
                 "tuner_ms_per_layer": {str(k): round(v, 3) for k, v in tuned["ms_per_layer"].items()} if tuned else None,
                 "note": "effective_GB_s = per-gate algorithmic bytes of the layer / fused layer time"}
        log(f"bench: {prec} fused {fused}")
    dev.destroy()
    achieved = d["bytes"] / (d["total_ms"] / 1e3) / 1e9
    summary = {"precision": prec, "value_GB_s": round(total_bytes / elapsed / 1e9, 1),
               "ms_per_step": round(elapsed / args.steps * 1e3, 3),
               "hbm_frac_of_peak": round(total_bytes / elapsed / 1e9 / HBM_PEAK_GBS, 4),
               "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(achieved / HBM_PEAK_GBS, 4), "kernel": f"k_{dom}",
                            "avg_launch_ms": round(d["total_ms"] / d["launches"], 4)},
               "kernels": kern, "fused": fused}
    return {"elapsed": elapsed, "stats": stats, "total_bytes": total_bytes, "dom": dom, "kernels": kern,
            "fused": fused, "tuned": tuned, "summary": summary}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--prec", default="c128", choices=["c128", "c64"])
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fused", action="store_true", help="skip the fused tiled-pass measurement")
    ap.add_argument("--cpu-qubits", type=int, default=30)
    ap.add_argument("--sweep", metavar="PATH", help="also write GB/s per gate type x bit position to PATH")
    ap.add_argument("--qaoa", type=int, default=4096, help="batched QAOA instances (0 to skip)")
    ap.add_argument("--serve", type=int, default=120, help="served requests per arrival rate (0 to skip)")
    ap.add_argument("--serve-plain", type=int, default=0,
                    help="requests timed on the plain path (0: the same requests as served)")
    ap.add_argument("--rates", default="50,200,1000,2000")
    ap.add_argument("--precision-qubits", type=int, default=32, help="0 to skip the precision study")
    ap.add_argument("--precision-layers", type=int, default=50)
    ap.add_argument("--la", type=int, default=4096, help="GEMM size of the linear-algebra section (0 to skip)")
    ap.add_argument("--no-other-precision", action="store_true",
                    help="skip the same measurement at the other precision (c64 beside c128)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also copy the state once with the library's copy kernel (a kernel of known bytes for "
                         "scaling hardware byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the state-vector path "
              "here runs on a single MI355X; multi-GPU sharding is not built)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from qclojure_amd import engine, workloads as wl
    engine.init(0)
    info = engine.device_info()
    log(f"bench: device {info['name']} {info['arch']} CUs={info['compute-units']} "
        f"HBM={info['hbm-bytes'] / 2**30:.0f} GiB")

    n, layers_total = args.qubits, 100
    ops = wl.random_layer_ops(n, layers_total, args.seed)
    main_run = measure_layers(engine, n, args.prec, ops, args)
    other = "c64" if args.prec == "c128" else "c128"
    other_run = None
    if not args.no_other_precision:
        other_run = measure_layers(engine, n, other, ops, args)
        other_run.pop("tuned", None)
    elapsed, stats, total_bytes = main_run["elapsed"], main_run["stats"], main_run["total_bytes"]
    value = total_bytes / elapsed / 1e9
    dom = main_run["dom"]
    d = stats[dom]
    achieved = d["bytes"] / (d["total_ms"] / 1e3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        try:
            with open(tpath) as f:
                tr = json.load(f)
            traffic = tr.get(f"{dom}:{args.prec}:{n}")
        except (OSError, ValueError):
            traffic = None

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.prec == "c128" else "f32",
        "data": "synthetic",
        "config": {"workload": f"{n}-qubit random-layer circuit, {args.prec}, per-gate kernels "
                               f"(one step = one layer: {n} one-qubit gates + {n // 2} CNOTs)",
                   "qubits": n, "precision": args.prec, "layers_in_circuit": layers_total,
                   "seed": args.seed, "state_bytes": (1 << n) * (16 if args.prec == "c128" else 8)},
        "hbm_frac_of_peak": round(value / HBM_PEAK_GBS, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": f"k_{dom}", "avg_launch_ms": round(d["total_ms"] / d["launches"], 4)},
        "kernels": main_run["kernels"],
        "fused": main_run["fused"],
    }
    if other_run is not None:
        out[other] = other_run["summary"]
    def section(name, fn):
        try:
            log(f"bench: {name} ...")
            out[name] = fn()
        except Exception as e:  # a section that cannot run on this box is reported, not fatal
            out[name] = {"error": f"{type(e).__name__}: {e}"}
            log(f"bench: {name} failed: {e}")

    if args.qaoa:
        section("batched_qaoa", lambda: bench_batched_qaoa(engine, args.qaoa, args.seed, plain_every=16))
    if args.serve:
        rates = [float(r) for r in args.rates.split(",") if r]
        section("serving", lambda: bench_serving(engine, rates, args.serve, args.serve_plain, args.seed))
    if args.precision_qubits:
        section("precision", lambda: bench_precision(engine, args.precision_qubits, args.precision_layers, args.seed))
    if args.la:
        section("linear_algebra", lambda: bench_la(engine, args.la, args.seed))
    if args.sweep:
        table = position_sweep(engine, n, args.prec)
        with open(args.sweep, "w") as f:
            json.dump({"qubits": n, "precision": args.prec, "unit": "GB/s", "table": table}, f, indent=1)
    if not args.no_cpu_baseline:
        log(f"bench: CPU baseline at {args.cpu_qubits} qubits ...")
        out["cpu_baseline"] = cpu_baseline(args.cpu_qubits, args.seed)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

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


def cpu_baseline(n: int, seed: int) -> dict:
    """The oracle's pair-update mode (OpenMP, FP64) on one layer of the same
    construction at n qubits, timed on this host."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from qclojure_amd import workloads as wl
    ops = wl.random_layer_ops(n, 1, seed)
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1.0
    S = 1 << n
    t0 = time.perf_counter()
    rc = oracle.lib().ref_apply_program(oracle._p(psi), n, oracle._p(ops), len(ops), 0, 0, 0)
    dt = time.perf_counter() - t0
    if rc != 0:
        raise RuntimeError("oracle failed")
    nbytes = 0
    for r in ops:
        nbytes += (16 if int(r["kind"]) == 13 else 32) * S   # CNOT touches S/2, a 1-qubit gate S
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(nbytes / dt / 1e9, 2), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"one {n}-qubit complex128 layer of the same construction ({len(ops)} gates, "
                      f"{nbytes / 1e9:.0f} GB algorithmic) by the oracle's OpenMP pair-update mode: "
                      f"{dt:.2f} s on {cpu_model()}"}


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
             "SWAP(q,q-1)": lambda q: [(OP["SWAP"], q, (q - 1) % n, 0, 0, 0.0)]}
    for name, mk in gates.items():
        row = {}
        for q in range(n):
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
    ap.add_argument("--cpu-qubits", type=int, default=28)
    ap.add_argument("--sweep", metavar="PATH", help="also write GB/s per gate type x bit position to PATH")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the state-vector path "
              "here runs on a single MI355X; sharding is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from qclojure_amd import engine, workloads as wl
    engine.init(0)
    info = engine.device_info()
    log(f"bench: device {info['name']} {info['arch']} CUs={info['compute-units']} "
        f"HBM={info['hbm-bytes'] / 2**30:.0f} GiB")

    n, layers_total = args.qubits, 100
    ops = wl.random_layer_ops(n, layers_total, args.seed)
    per = n + n // 2
    layer = lambda k: ops[(k % layers_total) * per:((k % layers_total) + 1) * per]

    dev = engine.DeviceState(n, args.prec)
    log(f"bench: state {n} qubits {args.prec} ({dev.nbytes / 2**30:.1f} GiB) allocated")
    for w in range(args.warmup):
        dev.apply(layer(w))
    dev.sync()
    log(f"bench: {args.warmup} warmup layers done")

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
    log(f"bench: {args.steps} layers in {elapsed:.3f} s, norm^2-1 = {norm - 1:.2e}")

    total_bytes = sum(v["bytes"] for v in stats.values())
    value = total_bytes / elapsed / 1e9
    kern = {}
    for name, v in stats.items():
        if v["launches"]:
            kern[name] = {"launches": v["launches"], "avg_ms": round(v["total_ms"] / v["launches"], 4),
                          "GB_s": round(v["bytes"] / (v["total_ms"] / 1e3) / 1e9, 1)}
    dom = max(stats, key=lambda k: stats[k]["total_ms"])
    d = stats[dom]
    achieved = d["bytes"] / (d["total_ms"] / 1e3) / 1e9
    fused = None
    if not args.no_fused:
        from qclojure_amd._abi import QSV_FLAG_FUSE
        for w in range(args.warmup):
            dev.apply(layer(w), QSV_FLAG_FUSE)
        dev.sync()
        dev.timing(True)
        dev.timing_read(reset=True)
        t0 = time.perf_counter()
        for k in range(args.steps):
            dev.apply(layer(args.warmup + k), QSV_FLAG_FUSE)
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
                 "effective_GB_s": round(total_bytes / args.steps / (f_elapsed / args.steps) / 1e9, 1),
                 "note": "effective_GB_s = per-gate algorithmic bytes of the layer / fused layer time"}
        log(f"bench: fused {fused}")
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        try:
            with open(tpath) as f:
                tr = json.load(f)
            key = f"{dom}:{args.prec}:{n}"
            traffic = tr.get(key)
        except (OSError, ValueError):
            traffic = None
    dev.destroy()

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
        "kernels": kern,
        "fused": fused,
    }
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

"""The drop-in boundary: QClojure's backend protocols over libqsv.

Mirrors ``QuantumBackend`` / ``BatchJobBackend``
(src/org/soulspace/qclojure/application/backend.clj:72-165) and the job
lifecycle and result map of ``LocalQuantumSimulator``
(src/org/soulspace/qclojure/adapter/backend/ideal_simulator.clj:56-195), with
the gate path and every result reduction running on the GPU.

Status keywords are the reference's, as strings: "queued", "running",
"completed", "failed", "cancelled", "not-found".
"""
from __future__ import annotations

import abc
import itertools
import os
import threading
import time
import traceback
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import _abi
from .circuit import CircuitError, circuit_metadata, decode_circuit, NATIVE_SIMULATOR_GATE_SET
from .engine import DeviceState, device_info
from .result import FinalState, detach_results, extract_results, release_results


class QuantumBackend(abc.ABC):
    """application/backend.clj:72-112."""

    @abc.abstractmethod
    def backend_info(self) -> dict: ...

    @abc.abstractmethod
    def device(self) -> dict: ...

    @abc.abstractmethod
    def is_available(self) -> bool: ...

    @abc.abstractmethod
    def submit_circuit(self, circuit: dict, options: dict | None = None) -> str: ...

    @abc.abstractmethod
    def _precision_report(self, dev: DeviceState, norm_in, ops, flags: int, seed: int, init) -> dict:
        """What a complex64 result is worth: the drift of its FP64 norm over
        the program, and (up to shadow-max-qubits) the overlap fidelity
        |<psi32|psi64>|^2 (state-fidelity of state.clj:1164-1185, squared) and
        max |a32 - a64| against the same program run on a complex128 state."""
        if dev.precision != _abi.QSV_C64:
            return {"precision": "c128"}
        rep = {"precision": "c64", "norm-drift": abs(dev.norm2() - norm_in)}
        if dev.num_qubits <= int(self.config.get("shadow-max-qubits", 24)):
            with DeviceState(dev.num_qubits, "c128") as sh:
                if init is not None:
                    sh.upload(np.asarray(init["state-vector"], dtype=np.complex128))
                sh.apply(ops, flags, seed)
                c = dev.compare(sh)
            rep.update({"fidelity-vs-c128": c["fidelity"], "max-abs-amplitude-error": c["max_abs_diff"],
                        "reference": "complex128-device-shadow"})
        return rep

    def job_status(self, job_id: str) -> str: ...

    @abc.abstractmethod
    def job_result(self, job_id: str) -> dict: ...

    @abc.abstractmethod
    def cancel_job(self, job_id: str) -> str: ...

    @abc.abstractmethod
    def queue_status(self) -> dict: ...


class BatchJobBackend(abc.ABC):
    """application/backend.clj:136-165 (declared, never implemented, in the reference)."""

    @abc.abstractmethod
    def batch_submit(self, circuits, options: dict | None = None) -> str: ...

    @abc.abstractmethod
    def batch_status(self, batch_id: str) -> dict: ...

    @abc.abstractmethod
    def batch_results(self, batch_id: str) -> dict: ...


DEVICE = {
    "id": "mi355x-simulator",
    "name": "MI355X State-Vector Simulator",
    "provider": "QClojure",
    "platform": "Local",
    "technology": "simulator",
    "num-qubits": 34,
    "topology": "all-to-all",
    "connectivity": "full",
    "native-gates": NATIVE_SIMULATOR_GATE_SET,
    "virtual-gates": frozenset(),
    "supported-operations": NATIVE_SIMULATOR_GATE_SET,
    "measurement-basis": "any",
    "noise-model": {},
    "performance": {"gate-fidelity": 1.0, "readout-fidelity": 1.0, "gate-times": {"all": 0},
                    "simulation-time": {"exponential-scaling": True}},
}

# One job table shared by every simulator instance, as the reference's
# namespace-level atom (ideal_simulator.clj:56-57).
_STATE_LOCK = threading.Lock()
_SIM_STATE = {"job-counter": 0, "active-jobs": {}}


def reset_simulator_state() -> None:
    """ideal_simulator.clj:200-206."""
    with _STATE_LOCK:
        _SIM_STATE["job-counter"] = 0
        _SIM_STATE["active-jobs"] = {}


def _now_ms() -> int:
    return int(time.time() * 1000)


def _generate_job_id(prefix: str = "gpu_job") -> str:
    with _STATE_LOCK:
        _SIM_STATE["job-counter"] += 1
        k = _SIM_STATE["job-counter"]
    return f"{prefix}_{k}_{_now_ms()}"


class GpuSimulator(QuantumBackend, BatchJobBackend):
    """A QuantumBackend whose gate path runs on the MI355X.

    config keys: "max-qubits" (default 34), "precision" ("c128" default, or
    "c64"), "seed" (base seed for shots and mid-circuit measurement; each job
    uses seed + job counter), "workers" (concurrent jobs, default 4),
    "msb-swap" (True = SWAP/iSWAP on MSB bit positions instead of the
    reference's LSB positions), "keep-final-state" (default True: the result's
    :final-state and lazily read vectors stay in HBM until read or until
    cleanup_completed_jobs drops the job; False: they are copied to host memory
    when the job completes and the device state is freed at once),
    "fuse" (default True: fused LDS-tiled passes), "shadow-max-qubits"
    (default 24: complex64 jobs up to this size also run the program on a
    complex128 state and report |<psi32|psi64>|^2 and max |d amplitude|
    against it in the result's "precision-report"; 0 turns the shadow off).
    """

    def __init__(self, config: dict | None = None):
        self.config = dict(config or {})
        self._pool = ThreadPoolExecutor(max_workers=int(self.config.get("workers", 4)),
                                        thread_name_prefix="qsv-job")
        self._seed_counter = itertools.count()

    # -- QuantumBackend ------------------------------------------------------
    def backend_info(self) -> dict:
        return {"backend-type": "simulator",
                "backend-name": "MI355X GPU State-Vector Simulator",
                "description": "State-vector simulator on hand-written CDNA4 HIP kernels (libqsv)",
                "backend-config": self.config,
                "max-qubits": self.config.get("max-qubits", 34),
                "capabilities": {"quantum-backend", "batch-job-backend"},
                "device": DEVICE,
                "version": _abi.load_library().qsv_version().decode()}

    def device(self) -> dict:
        return DEVICE

    def is_available(self) -> bool:
        try:
            device_info()
            return True
        except Exception:
            return False

    def submit_circuit(self, circuit: dict, options: dict | None = None) -> str:
        options = dict(options or {})
        job_id = _generate_job_id()
        job = {"job-id": job_id, "circuit": circuit, "options": options, "status": "queued",
               "result": None, "created-at": _now_ms(), "completed-at": None}
        with _STATE_LOCK:
            _SIM_STATE["active-jobs"][job_id] = job
        seed = (int(self.config.get("seed", 0x5EED)) + next(self._seed_counter)) & (2**64 - 1)
        self._pool.submit(self._run_job, job_id, circuit, options, seed)
        return job_id

    def _run_job(self, job_id: str, circuit: dict, options: dict, seed: int) -> None:
        with _STATE_LOCK:
            job = _SIM_STATE["active-jobs"].get(job_id)
            if job is None or job["status"] != "queued":
                return  # cancelled before it started
            job["status"] = "running"
        result = self.execute(circuit, options, seed)
        with _STATE_LOCK:
            job = _SIM_STATE["active-jobs"].get(job_id)
            if job is not None and job["status"] == "running":
                job["status"] = result["job-status"]
                job["result"] = result
                job["completed-at"] = _now_ms()

    def execute(self, circuit: dict, options: dict, seed: int) -> dict:
        """The private execute-circuit of ideal_simulator.clj:70-96."""
        start = time.perf_counter()
        try:
            n = int(circuit["num-qubits"])
            if n > int(self.config.get("max-qubits", 34)):
                raise CircuitError(f"circuit has {n} qubits, backend max-qubits is {self.config.get('max-qubits', 34)}")
            ops = decode_circuit(circuit)
            dev = DeviceState(n, self.config.get("precision", "c128"))
            init = options.get("initial-state")
            if init is not None:
                if int(init["num-qubits"]) != n:
                    raise CircuitError("initial state and circuit have different numbers of qubits")
                dev.upload(np.asarray(init["state-vector"], dtype=np.complex128))
            flags = _abi.QSV_FLAG_MSB_SWAP if self.config.get("msb-swap") else 0
            if self.config.get("fuse", True):
                flags |= _abi.QSV_FLAG_FUSE
            norm_in = dev.norm2() if dev.precision == _abi.QSV_C64 else None
            dev.apply(ops, flags, seed)
            specs = options.get("result-specs") or {}
            base = {"final-state": FinalState(dev),
                    "result-types": set(specs.keys()),
                    "circuit": circuit,
                    "circuit-metadata": circuit_metadata(circuit),
                    "precision-report": self._precision_report(dev, norm_in, ops, flags, seed, init)}
            results = extract_results(base, specs, dev, seed ^ 0xA5A5A5A5A5A5A5A5) if specs else base
            dev.sync()
            if not self.config.get("keep-final-state", True):
                detach_results(results)
            return {"job-status": "completed", "results": results,
                    "execution-time-ms": (time.perf_counter() - start) * 1000.0}
        except Exception as e:  # every failure becomes a :failed job, never a hang
            return {"job-status": "failed", "error-message": str(e),
                    "exception-type": type(e).__name__,
                    "trace": traceback.format_exc(limit=3) if os.environ.get("QSV_DEBUG") else None}

    def _precision_report(self, dev: DeviceState, norm_in, ops, flags: int, seed: int, init) -> dict:
        """What a complex64 result is worth: the drift of its FP64 norm over
        the program, and (up to shadow-max-qubits) the overlap fidelity
        |<psi32|psi64>|^2 (state-fidelity of state.clj:1164-1185, squared) and
        max |a32 - a64| against the same program run on a complex128 state."""
        if dev.precision != _abi.QSV_C64:
            return {"precision": "c128"}
        rep = {"precision": "c64", "norm-drift": abs(dev.norm2() - norm_in)}
        if dev.num_qubits <= int(self.config.get("shadow-max-qubits", 24)):
            with DeviceState(dev.num_qubits, "c128") as sh:
                if init is not None:
                    sh.upload(np.asarray(init["state-vector"], dtype=np.complex128))
                sh.apply(ops, flags, seed)
                c = dev.compare(sh)
            rep.update({"fidelity-vs-c128": c["fidelity"], "max-abs-amplitude-error": c["max_abs_diff"],
                        "reference": "complex128-device-shadow"})
        return rep

    def job_status(self, job_id: str) -> str:
        with _STATE_LOCK:
            job = _SIM_STATE["active-jobs"].get(job_id)
            return job["status"] if job is not None else "not-found"

    def job_result(self, job_id: str) -> dict:
        with _STATE_LOCK:
            job = _SIM_STATE["active-jobs"].get(job_id)
            if job is None:
                return {"job-id": job_id, "job-status": "not-found", "error-message": "Job not found"}
            if job["status"] == "completed":
                r = dict(job["result"])
                r["job-id"] = job_id
                return r
            if job["status"] == "failed" and job["result"] is not None:
                r = dict(job["result"])
                r["job-id"] = job_id
                return r
            return {"job-id": job_id, "job-status": job["status"], "error-message": "Job not completed"}

    def cancel_job(self, job_id: str) -> str:
        with _STATE_LOCK:
            job = _SIM_STATE["active-jobs"].get(job_id)
            if job is None:
                return "not-found"
            if job["status"] in ("queued", "running"):
                job["status"] = "cancelled"
                job["completed-at"] = _now_ms()
                return "cancelled"
            return "cannot-cancel"

    def queue_status(self) -> dict:
        with _STATE_LOCK:
            jobs = list(_SIM_STATE["active-jobs"].values())
        count = lambda s: sum(1 for j in jobs if j["status"] == s)
        return {"total-jobs": len(jobs), "queued": count("queued"), "running": count("running"),
                "completed": count("completed"), "backend-load": 0.0, "estimated-wait-time": 0}

    # -- BatchJobBackend -------------------------------------------------------
    def batch_submit(self, circuits, options: dict | None = None) -> str:
        batch_id = _generate_job_id("gpu_batch")
        ids = [self.submit_circuit(c, options) for c in circuits]
        with _STATE_LOCK:
            _SIM_STATE.setdefault("batches", {})[batch_id] = ids
        return batch_id

    def batch_status(self, batch_id: str) -> dict:
        with _STATE_LOCK:
            ids = _SIM_STATE.get("batches", {}).get(batch_id)
        if ids is None:
            return {"batch-id": batch_id, "status": "not-found"}
        st = {j: self.job_status(j) for j in ids}
        done = all(s in ("completed", "failed", "cancelled") for s in st.values())
        return {"batch-id": batch_id, "status": "completed" if done else "running", "job-statuses": st}

    def batch_results(self, batch_id: str) -> dict:
        with _STATE_LOCK:
            ids = _SIM_STATE.get("batches", {}).get(batch_id)
        if ids is None:
            return {}
        return {j: self.job_result(j) for j in ids}

    def close(self) -> None:
        self._pool.shutdown(wait=True)


def create_simulator(config: dict | None = None) -> GpuSimulator:
    return GpuSimulator(config or {})


def cleanup_completed_jobs(max_age_ms: int = 60 * 60 * 1000) -> int:
    """cleanup-completed-jobs! (ideal_simulator.clj:243-263): drop completed
    jobs that completed more than max_age_ms ago and return how many were
    dropped. The device memory of a dropped job's final state is freed."""
    cutoff = _now_ms() - max_age_ms
    with _STATE_LOCK:
        jobs = _SIM_STATE["active-jobs"]
        drop = [jid for jid, j in jobs.items()
                if j["status"] == "completed" and (j["completed-at"] or 0) <= cutoff]
        dropped = [jobs.pop(jid) for jid in drop]
    for j in dropped:
        if j.get("result"):
            release_results(j["result"].get("results"))
    return len(dropped)


def get_simulator_stats() -> dict:
    """ideal_simulator.clj:208-241."""
    with _STATE_LOCK:
        jobs = list(_SIM_STATE["active-jobs"].values())
        counter = _SIM_STATE["job-counter"]
    completed = [j for j in jobs if j["status"] == "completed"]
    times = [j["completed-at"] - j["created-at"] for j in completed if j["completed-at"] is not None]
    by_status: dict = {}
    for j in jobs:
        by_status[j["status"]] = by_status.get(j["status"], 0) + 1
    return {"total-jobs": len(jobs), "completed-jobs": len(completed),
            "average-execution-time": (sum(times) / len(times)) if times else 0,
            "job-counter": counter, "jobs-by-status": by_status,
            "active-jobs-count": sum(1 for j in jobs if j["status"] in ("queued", "running")),
            "oldest-job": min((j["created-at"] for j in jobs), default=None),
            "newest-job": max((j["created-at"] for j in jobs), default=None)}


# ---- high-level helpers (application/backend.clj:209-315) --------------------
def execute_circuit(backend: QuantumBackend, circuit: dict, options: dict | None = None,
                    timeout_s: float = 60.0) -> dict:
    """Submit and wait. The reference polls every 100 ms up to 600 times; this
    polls every millisecond up to the same 60 s."""
    if options is None:
        options = {"shots": 128}
    if not backend.is_available():
        return {"job-status": "failed", "error-message": "Backend is not available"}
    job_id = backend.submit_circuit(circuit, options)
    deadline = time.monotonic() + timeout_s
    while time.monotonic() < deadline:
        st = backend.job_status(job_id)
        if st == "completed":
            return backend.job_result(job_id)
        if st == "failed":
            r = backend.job_result(job_id)
            return {"job-status": "failed", "job-id": job_id,
                    "error-message": r.get("error-message", "Job execution failed")}
        if st == "cancelled":
            return {"job-status": "cancelled", "job-id": job_id}
        time.sleep(0.001)
    return {"job-status": "failed", "error-message": "Job execution timeout"}


def execute_circuit_async(backend: QuantumBackend, circuit: dict, options: dict | None = None) -> str:
    return backend.submit_circuit(circuit, options if options is not None else {"shots": 1024})


def wait_for_job(backend: QuantumBackend, job_id: str, timeout_ms: int = 30000) -> dict:
    deadline = time.monotonic() + timeout_ms / 1000.0
    while time.monotonic() < deadline:
        st = backend.job_status(job_id)
        if st == "completed":
            return backend.job_result(job_id)
        if st == "failed":
            return {"job-status": "failed", "job-id": job_id, "error-message": "Job execution failed"}
        if st == "cancelled":
            return {"job-status": "cancelled", "job-id": job_id}
        time.sleep(0.001)
    return {"job-status": "failed", "error-message": "Job wait timeout"}


def analyze_measurement_results(results: dict) -> dict:
    """application/backend.clj:320-343."""
    import math
    from fractions import Fraction
    total = sum(results.values())
    probs = {o: Fraction(c, total) for o, c in results.items()}
    most = max(results.items(), key=lambda kv: kv[1]) if results else None
    return {"total-shots": total, "outcomes": results, "probabilities": probs, "num-outcomes": len(results),
            "most-frequent": most,
            "entropy": -sum(float(p) * math.log(float(p)) for p in probs.values() if p != 0)}

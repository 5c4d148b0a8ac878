"""Device-resident state vectors over libqsv's C ABI.

A :class:`DeviceState` owns one ``qsv_state`` handle: 2^n amplitudes in HBM on
their own HIP stream. Every method is one C-ABI call; nothing here computes
amplitudes on the host.
"""
from __future__ import annotations

import ctypes
import threading

import numpy as np

from . import _abi
from ._abi import QSV_C128, QSV_C64, check

_init_lock = threading.Lock()
_initialised = False


def lib():
    return _abi.load_library()


def init(device: int = 0) -> None:
    global _initialised
    with _init_lock:
        if not _initialised:
            check(lib().qsv_init(device))
            _initialised = True


def device_info() -> dict:
    init()
    info = _abi.QsvDeviceInfo()
    check(lib().qsv_device_query(ctypes.byref(info)))
    return {"name": info.name.decode(), "arch": info.arch.decode(), "compute-units": info.compute_units,
            "wavefront-size": info.wavefront_size, "hbm-bytes": info.hbm_bytes,
            "hbm-free-bytes": info.hbm_free_bytes}


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return a.ctypes.data_as(ctypes.c_void_p)


PRECISIONS = {"c128": QSV_C128, "complex128": QSV_C128, "c64": QSV_C64, "complex64": QSV_C64,
              QSV_C128: QSV_C128, QSV_C64: QSV_C64}


class DeviceState:
    """One state vector in HBM (qsv_state_create / qsv_state_destroy)."""

    def __init__(self, num_qubits: int, precision="c128"):
        init()
        self.num_qubits = int(num_qubits)
        self.precision = PRECISIONS[precision]
        h = ctypes.c_void_p()
        check(lib().qsv_state_create(self.num_qubits, self.precision, ctypes.byref(h)))
        self._h = h

    # -- lifecycle ------------------------------------------------------------
    @property
    def handle(self) -> ctypes.c_void_p:
        if self._h is None:
            raise RuntimeError("state already destroyed")
        return self._h

    def destroy(self) -> None:
        if getattr(self, "_h", None) is not None:
            lib().qsv_state_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()

    @property
    def size(self) -> int:
        return 1 << self.num_qubits

    @property
    def nbytes(self) -> int:
        return self.size * (16 if self.precision == QSV_C128 else 8)

    # -- data movement ---------------------------------------------------------
    def init_basis(self, index: int = 0) -> "DeviceState":
        check(lib().qsv_state_init_basis(self.handle, int(index)))
        return self

    def upload(self, amps, offset: int = 0) -> "DeviceState":
        a = np.ascontiguousarray(np.asarray(amps, dtype=np.complex128))
        check(lib().qsv_state_upload(self.handle, _ptr(a), offset, a.size))
        return self

    def download(self, offset: int = 0, count: int | None = None) -> np.ndarray:
        count = self.size - offset if count is None else count
        out = np.empty(count, dtype=np.complex128)
        check(lib().qsv_state_download(self.handle, _ptr(out), offset, count))
        return out

    def copy_from(self, other: "DeviceState") -> "DeviceState":
        check(lib().qsv_state_copy(self.handle, other.handle))
        return self

    def sync(self) -> None:
        check(lib().qsv_state_sync(self.handle))

    # -- gates -------------------------------------------------------------------
    def apply(self, ops: np.ndarray, flags: int = 0, seed: int = 0) -> "DeviceState":
        ops = np.ascontiguousarray(ops, dtype=_abi.OP_DTYPE)
        check(lib().qsv_apply_program(self.handle, _ptr(ops), len(ops), flags, seed & (2**64 - 1)))
        return self

    # -- results -----------------------------------------------------------------
    def norm2(self) -> float:
        v = ctypes.c_double()
        check(lib().qsv_norm2(self.handle, ctypes.byref(v)))
        return v.value

    def normalize(self) -> "DeviceState":
        check(lib().qsv_normalize(self.handle))
        return self

    def probabilities(self, offset: int = 0, count: int | None = None) -> np.ndarray:
        count = self.size - offset if count is None else count
        out = np.empty(count, dtype=np.float64)
        check(lib().qsv_probabilities(self.handle, offset, count, _ptr(out)))
        return out

    def sample(self, shots: int, seed: int) -> np.ndarray:
        out = np.empty(int(shots), dtype=np.uint64)
        check(lib().qsv_sample(self.handle, int(shots), seed & (2**64 - 1), _ptr(out)))
        return out

    def measure_qubits(self, qubits, u: float):
        mask = 0
        for q in qubits:
            mask |= 1 << int(q)
        m = bin(mask).count("1")
        probs = np.empty(1 << m, dtype=np.float64)
        oc = ctypes.c_uint64()
        check(lib().qsv_measure_qubits(self.handle, mask, float(u), ctypes.byref(oc), _ptr(probs)))
        return int(oc.value), probs

    def expect_pauli(self, pauli_strings) -> np.ndarray:
        strs = list(pauli_strings)
        out = np.empty(len(strs), dtype=np.float64)
        if not strs:
            return out
        for s in strs:
            if len(s) != self.num_qubits:
                raise ValueError(f"pauli string {s!r} has length {len(s)}, state has {self.num_qubits} qubits")
        blob = "".join(strs).encode()
        check(lib().qsv_expect_pauli(self.handle, blob, len(strs), _ptr(out)))
        return out

    def expect_1q(self, qubit: int, matrix) -> complex:
        m = np.ascontiguousarray(np.asarray(matrix, dtype=np.complex128).reshape(4)).view(np.float64)
        out = np.empty(2, dtype=np.float64)
        check(lib().qsv_expect_1q(self.handle, int(qubit), _ptr(m), _ptr(out)))
        return complex(out[0], out[1])

    def inner(self, other: "DeviceState") -> complex:
        out = np.empty(2, dtype=np.float64)
        check(lib().qsv_inner(self.handle, other.handle, _ptr(out)))
        return complex(out[0], out[1])

    def compare(self, other: "DeviceState") -> dict:
        """One device pass over both states (any precisions): overlap fidelity
        |<self|other>|^2, max |a_i - b_i| and both norms."""
        out = np.empty(5, dtype=np.float64)
        check(lib().qsv_state_compare(self.handle, other.handle, _ptr(out)))
        ov = complex(out[0], out[1])
        return {"overlap": ov, "fidelity": abs(ov) ** 2, "norm2_self": out[2], "norm2_other": out[3],
                "max_abs_diff": out[4]}

    # -- timing ------------------------------------------------------------------
    def timing(self, enable: bool) -> None:
        check(lib().qsv_timing_enable(self.handle, 1 if enable else 0))

    def timing_read(self, reset: bool = True) -> dict:
        arr = (_abi.QsvKernelStat * 5)()
        check(lib().qsv_timing_read(self.handle, arr, 1 if reset else 0))
        return {name: {"launches": arr[i].launches, "total_ms": arr[i].total_ms, "bytes": arr[i].bytes}
                for i, name in enumerate(_abi.KCLASS_NAMES)}


class PackedBatch:
    """Instances flattened into the arrays of a qsv_batch_desc. An instance's
    terms are either a list of (coefficient, I/Z string) or a pair of arrays
    (Z masks with bit q = Z on qubit q, coefficients)."""

    def __init__(self, instances, precision="c64"):
        nq, opoff, toff, zm, co, all_ops = [], [0], [0], [], [], []
        zparts, cparts = [], []
        masks: dict[str, int] = {}
        for (n, ops, terms) in instances:
            nq.append(int(n))
            ops = np.ascontiguousarray(ops, dtype=_abi.OP_DTYPE)
            all_ops.append(ops)
            opoff.append(opoff[-1] + len(ops))
            if isinstance(terms, tuple) and len(terms) == 2 and isinstance(terms[0], np.ndarray):
                zmask = np.asarray(terms[0], dtype=np.uint64)
                if zm:
                    zparts.append(np.asarray(zm, dtype=np.uint64))
                    cparts.append(np.asarray(co, dtype=np.float64))
                    zm, co = [], []
                zparts.append(zmask)
                cparts.append(np.asarray(terms[1], dtype=np.float64))
                toff.append(toff[-1] + len(zmask))
                continue
            start = len(zm)
            for (c, s) in terms:
                m = masks.get(s)
                if m is None:
                    if len(s) != n or set(s) - {"I", "Z"}:
                        raise ValueError(f"batched Hamiltonians take I/Z strings of length {n}: {s!r}")
                    m = masks[s] = sum(1 << q for q, ch in enumerate(s) if ch == "Z")
                elif len(s) != n:
                    raise ValueError(f"batched Hamiltonians take I/Z strings of length {n}: {s!r}")
                zm.append(m)
                co.append(float(c))
            toff.append(toff[-1] + len(zm) - start)
        if zm:
            zparts.append(np.asarray(zm, dtype=np.uint64))
            cparts.append(np.asarray(co, dtype=np.float64))
        self.n = len(nq)
        self.precision = PRECISIONS[precision]
        self._nq = np.asarray(nq, dtype=np.int32)
        self._opoff = np.asarray(opoff, dtype=np.uint64)
        self._toff = np.asarray(toff, dtype=np.uint64)
        self._zm = np.concatenate(zparts) if zparts else np.zeros(1, dtype=np.uint64)
        self._co = np.concatenate(cparts) if cparts else np.zeros(1, dtype=np.float64)
        # one byte copy per instance into the flat record array
        flat = np.zeros(max(1, opoff[-1]), dtype=_abi.OP_DTYPE)
        raw = flat.view(np.uint8)
        rec = _abi.OP_DTYPE.itemsize
        for k, ops in enumerate(all_ops):
            if len(ops):
                raw[opoff[k] * rec:opoff[k + 1] * rec] = ops.view(np.uint8)
        self._ops = flat

    def run(self) -> np.ndarray:
        """qsv_batch_expect_z: <H> of every instance."""
        init()
        desc = _abi.QsvBatchDesc(self.n, self._nq.ctypes.data, self._opoff.ctypes.data, self._ops.ctypes.data,
                                 self._toff.ctypes.data, self._zm.ctypes.data, self._co.ctypes.data, self.precision)
        out = np.zeros(self.n, dtype=np.float64)
        check(lib().qsv_batch_expect_z(ctypes.byref(desc), _ptr(out)))
        return out


def batch_expect_z(instances, precision="c64") -> np.ndarray:
    """Run many independent circuits from |0..0> and return each one's
    <H> = sum_k c_k <Z-string_k> (qsv_batch_expect_z).

    instances: iterable of (num_qubits, ops (qsv_op records), terms) where terms
    is a list of (coefficient, pauli_string) with only I/Z characters.
    """
    return PackedBatch(instances, precision).run()


class Server:
    """The native request loop (qsv_server_*): continuous batching with
    admission by state bytes against the HBM budget."""

    def __init__(self, precision="c64", hbm_budget_bytes: int = 0, max_inflight: int = 0,
                 flags: int = _abi.QSV_FLAG_FUSE, batch_small: bool = True):
        init()
        cfg = _abi.QsvServerConfig(int(hbm_budget_bytes), PRECISIONS[precision], int(max_inflight),
                                   int(flags), 1 if batch_small else 0)
        h = ctypes.c_void_p()
        check(lib().qsv_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h

    def submit(self, num_qubits: int, ops: np.ndarray, shots: int, seed: int) -> int:
        ops = np.ascontiguousarray(ops, dtype=_abi.OP_DTYPE)
        rid = ctypes.c_uint64()
        check(lib().qsv_server_submit(self._h, int(num_qubits), _ptr(ops), len(ops), int(shots),
                                      seed & (2**64 - 1), ctypes.byref(rid)))
        return rid.value

    @staticmethod
    def _info(info) -> dict:
        out = {"status": _abi.REQ_STATUS[info.status], "num-qubits": info.num_qubits, "shots": info.shots,
               "queue-ms": info.queue_ms, "latency-ms": info.latency_ms, "error": info.error.decode()}
        if info.status == 2:
            # the served state's FP64 sum |a|^2 before sampling: its drift from 1
            out["norm2"] = info.norm2
            out["norm-drift"] = abs(info.norm2 - 1.0)
        return out

    def status(self, rid: int) -> dict:
        info = _abi.QsvRequestInfo()
        check(lib().qsv_server_status(self._h, rid, ctypes.byref(info)))
        return self._info(info)

    def wait(self, rid: int, shots: int, timeout_s: float = 600.0):
        info = _abi.QsvRequestInfo()
        out = np.zeros(max(1, int(shots)), dtype=np.uint64)
        check(lib().qsv_server_wait(self._h, rid, float(timeout_s), ctypes.byref(info), _ptr(out)))
        return self._info(info), out[:int(shots)]

    def cancel(self, rid: int) -> str:
        """cancel-job: "cancelled", "cannot-cancel" or "not-found"."""
        out = ctypes.c_int()
        rc = lib().qsv_server_cancel(self._h, rid, ctypes.byref(out))
        if rc == _abi.QSV_E_NOTFOUND:
            return "not-found"
        check(rc)
        return _abi.CANCEL_OUTCOME[out.value]

    def stats(self) -> dict:
        st = _abi.QsvServerStats()
        check(lib().qsv_server_stats_read(self._h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in _abi.QsvServerStats._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            lib().qsv_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tune_fused(precision: str = "c128", num_qubits: int = 26) -> dict:
    """Time a fused four-layer random program for every candidate tile low-run width and
    make the fastest the default for `precision` (qsv_tune_fused)."""
    init()
    res = _abi.QsvTuneResult()
    check(lib().qsv_tune_fused(PRECISIONS[precision], num_qubits, ctypes.byref(res)))
    return {"low_bits": res.low_bits,
            "ms_per_layer": {int(res.candidate[i]): float(res.ms[i]) for i in range(res.n_candidates)}}


def noisy_trajectories(num_qubits: int, ops: np.ndarray, sites, n_traj: int, seed: int,
                       precision: str = "c128", n_keep: int = 0):
    """qsv_noisy_trajectories: n_traj independent noisy trajectories of the
    program `ops`, one measured outcome each, plus the first n_keep final
    states (normalised). sites: list of {"after-op", "qubit", "kraus": [2x2]}
    ordered by after-op."""
    init()
    ops = np.ascontiguousarray(ops, dtype=_abi.OP_DTYPE)
    arr = (_abi.QsvKrausSite * max(1, len(sites)))()
    for i, st in enumerate(sites):
        kr = st["kraus"]
        if not 1 <= len(kr) <= _abi.QSV_MAX_KRAUS:
            raise ValueError(f"site {i}: 1..{_abi.QSV_MAX_KRAUS} Kraus operators expected")
        arr[i].after_op = int(st["after-op"])
        arr[i].qubit = int(st["qubit"])
        arr[i].n_kraus = len(kr)
        for j, k in enumerate(kr):
            flat = np.asarray(k, dtype=np.complex128).reshape(4)
            for e in range(4):
                arr[i].m[j][2 * e] = flat[e].real
                arr[i].m[j][2 * e + 1] = flat[e].imag
    out = np.empty(int(n_traj), dtype=np.uint64)
    states = np.empty((int(n_keep), 1 << num_qubits), dtype=np.complex128)
    check(lib().qsv_noisy_trajectories(int(num_qubits), PRECISIONS[precision], _ptr(ops), len(ops),
                                       ctypes.byref(arr), len(sites), int(n_traj), seed & (2**64 - 1),
                                       _ptr(out), int(n_keep), _ptr(states) if n_keep else None))
    return out, states


def uniform(seed: int, k: int) -> float:
    """Counter-based uniform shared with the device sampler (qsv_uniform)."""
    return lib().qsv_uniform(seed & (2**64 - 1), k)

"""ctypes binding of libqsv's C ABI (include/qsv.h).

The shared library is built in-tree (``qclojure_amd/libqsv.so``). Loading it
never falls back to anything: if the library is missing the import of the
engine raises, naming the build step.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqsv.so")

# ---- constants mirrored from include/qsv.h --------------------------------
QSV_OK = 0
QSV_E_INVALID = -1
QSV_E_UNSUPPORTED = -2
QSV_E_NOMEM = -3
QSV_E_DEVICE = -4
QSV_E_STATE = -5
QSV_E_NOTFOUND = -6

QSV_C128 = 0
QSV_C64 = 1

QSV_FLAG_MSB_SWAP = 0x1
QSV_FLAG_FUSE = 0x2
QSV_FLAG_FP64_PASSES = 0x4
QSV_FLAG_VALU_PASSES = 0x8

_KIND_NAMES = [
    "X", "Y", "Z", "H", "S", "SDG", "T", "TDG", "RX", "RY", "RZ", "PHASE",
    "CNOT", "CZ", "CY", "CRX", "CRY", "CRZ",
    "SWAP", "ISWAP", "TOFFOLI", "FREDKIN",
    "RYDBERG_CZ", "RYDBERG_CPHASE", "RYDBERG_BLOCKADE",
    "GLOBAL_RX", "GLOBAL_RY", "GLOBAL_RZ", "GLOBAL_H",
    "GLOBAL_X", "GLOBAL_Y", "GLOBAL_Z",
    "MEASURE", "U", "CU",
]
OP = {name: i + 1 for i, name in enumerate(_KIND_NAMES)}
OP_NAME = {v: k for k, v in OP.items()}

KCLASS_NAMES = ["pair", "diag", "perm", "block", "other"]

# qsv_op: int32 kind; int32 q[3]; uint64 mask; double angle; double m[8]  (96 bytes)
OP_DTYPE = np.dtype([
    ("kind", "<i4"), ("q", "<i4", (3,)), ("mask", "<u8"),
    ("angle", "<f8"), ("m", "<f8", (8,)),
], align=True)
assert OP_DTYPE.itemsize == 96


class QsvOp(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("q", ctypes.c_int32 * 3),
                ("mask", ctypes.c_uint64), ("angle", ctypes.c_double),
                ("m", ctypes.c_double * 8)]


class QsvDeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 64), ("arch", ctypes.c_char * 32),
                ("compute_units", ctypes.c_int32), ("wavefront_size", ctypes.c_int32),
                ("hbm_bytes", ctypes.c_uint64), ("hbm_free_bytes", ctypes.c_uint64)]


class QsvKernelStat(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_uint64), ("total_ms", ctypes.c_double),
                ("bytes", ctypes.c_double)]


QSV_TUNE_MAX = 8


class QsvTuneResult(ctypes.Structure):
    _fields_ = [("low_bits", ctypes.c_int32), ("n_candidates", ctypes.c_int32),
                ("candidate", ctypes.c_int32 * QSV_TUNE_MAX), ("ms", ctypes.c_double * QSV_TUNE_MAX)]


QSV_MAX_KRAUS = 4
QSV_TRAJ_SLOTS = 1 << 20
QSV_NOALLOC_MAX_QUBITS = 30
QSV_NOALLOC_MAX_SHOTS = 1 << 19


class QsvKrausSite(ctypes.Structure):
    _fields_ = [("after_op", ctypes.c_uint64), ("qubit", ctypes.c_int32), ("n_kraus", ctypes.c_int32),
                ("m", (ctypes.c_double * 8) * QSV_MAX_KRAUS)]


class QsvBatchDesc(ctypes.Structure):
    _fields_ = [("n_instances", ctypes.c_uint64),
                ("num_qubits", ctypes.c_void_p),
                ("op_offsets", ctypes.c_void_p),
                ("ops", ctypes.c_void_p),
                ("term_offsets", ctypes.c_void_p),
                ("term_zmask", ctypes.c_void_p),
                ("term_coeff", ctypes.c_void_p),
                ("precision", ctypes.c_int32)]


class QsvServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("precision", ctypes.c_int32),
                ("max_inflight", ctypes.c_int32), ("flags", ctypes.c_uint32), ("batch_small", ctypes.c_int32)]


class QsvRequestInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("num_qubits", ctypes.c_int32), ("shots", ctypes.c_uint64),
                ("queue_ms", ctypes.c_double), ("latency_ms", ctypes.c_double), ("error", ctypes.c_char * 128),
                ("norm2", ctypes.c_double)]


class QsvServerStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in (
        "submitted", "completed", "failed", "queued_now", "inflight_now", "max_queue_depth",
        "small_batches", "small_batched_requests", "hbm_budget_bytes", "hbm_in_use_bytes", "cancelled")]


REQ_STATUS = {0: "queued", 1: "running", 2: "completed", 3: "failed", 4: "cancelled"}
CANCEL_OUTCOME = {0: "cancelled", 1: "cannot-cancel"}

# Every symbol include/qsv.h declares, with (restype, argtypes).
_vp, _u64, _i32, _u32, _dbl = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint32, ctypes.c_double
_pd = ctypes.POINTER(ctypes.c_double)
_pu64 = ctypes.POINTER(ctypes.c_uint64)
SIGNATURES = {
    "qsv_init": (_i32, [_i32]),
    "qsv_shutdown": (_i32, []),
    "qsv_device_query": (_i32, [ctypes.POINTER(QsvDeviceInfo)]),
    "qsv_last_error": (ctypes.c_char_p, []),
    "qsv_version": (ctypes.c_char_p, []),
    "qsv_state_create": (_i32, [_i32, _i32, ctypes.POINTER(_vp)]),
    "qsv_state_destroy": (_i32, [_vp]),
    "qsv_state_num_qubits": (_i32, [_vp]),
    "qsv_state_precision": (_i32, [_vp]),
    "qsv_state_init_basis": (_i32, [_vp, _u64]),
    "qsv_state_upload": (_i32, [_vp, _vp, _u64, _u64]),
    "qsv_state_download": (_i32, [_vp, _vp, _u64, _u64]),
    "qsv_state_copy": (_i32, [_vp, _vp]),
    "qsv_state_sync": (_i32, [_vp]),
    "qsv_apply_program": (_i32, [_vp, _vp, _u64, _u32, _u64]),
    "qsv_norm2": (_i32, [_vp, _pd]),
    "qsv_normalize": (_i32, [_vp]),
    "qsv_probabilities": (_i32, [_vp, _u64, _u64, _vp]),
    "qsv_sample": (_i32, [_vp, _u64, _u64, _vp]),
    "qsv_sample_noalloc": (_i32, [_vp, _u64, _u64, _vp]),
    "qsv_measure_qubits": (_i32, [_vp, _u64, _dbl, _pu64, _vp]),
    "qsv_expect_pauli": (_i32, [_vp, ctypes.c_char_p, _u64, _vp]),
    "qsv_expect_1q": (_i32, [_vp, _i32, _vp, _vp]),
    "qsv_inner": (_i32, [_vp, _vp, _vp]),
    "qsv_state_compare": (_i32, [_vp, _vp, _vp]),
    "qsv_uniform": (_dbl, [_u64, _u64]),
    "qsv_batch_expect_z": (_i32, [ctypes.POINTER(QsvBatchDesc), _vp]),
    "qsv_server_create": (_i32, [ctypes.POINTER(QsvServerConfig), ctypes.POINTER(_vp)]),
    "qsv_server_submit": (_i32, [_vp, _i32, _vp, _u64, _u64, _u64, _pu64]),
    "qsv_server_status": (_i32, [_vp, _u64, ctypes.POINTER(QsvRequestInfo)]),
    "qsv_server_wait": (_i32, [_vp, _u64, _dbl, ctypes.POINTER(QsvRequestInfo), _vp]),
    "qsv_server_cancel": (_i32, [_vp, _u64, ctypes.POINTER(ctypes.c_int)]),
    "qsv_server_stats_read": (_i32, [_vp, ctypes.POINTER(QsvServerStats)]),
    "qsv_server_destroy": (_i32, [_vp]),
    "qsv_timing_enable": (_i32, [_vp, _i32]),
    "qsv_timing_read": (_i32, [_vp, _vp, _i32]),
    "qsv_tune_fused": (_i32, [_i32, _i32, _vp]),
    "qsv_noisy_trajectories": (_i32, [_i32, _i32, _vp, _u64, _vp, _u64, _u64, _u64, _vp, _u64, _vp]),
    "qsv_mat_create": (_i32, [ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(_vp)]),
    "qsv_mat_destroy": (_i32, [_vp]),
    "qsv_mat_upload": (_i32, [_vp, _vp]),
    "qsv_mat_download": (_i32, [_vp, _vp]),
    "qsv_mat_copy": (_i32, [_vp, _vp]),
    "qsv_mat_shape": (_i32, [_vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "qsv_mat_view_state": (_i32, [_vp, ctypes.POINTER(_vp)]),
    "qsv_la_axpby": (_i32, [_dbl, _dbl, _vp, _dbl, _dbl, _vp, _vp]),
    "qsv_la_hadamard": (_i32, [_vp, _vp, _vp]),
    "qsv_la_gemm": (_i32, [_i32, _vp, _vp, _vp]),
    "qsv_la_gemv": (_i32, [_vp, _vp, _vp]),
    "qsv_la_dot": (_i32, [_i32, _vp, _vp, _pd]),
    "qsv_la_norm2": (_i32, [_vp, _pd]),
    "qsv_la_outer": (_i32, [_vp, _vp, _vp]),
    "qsv_la_kron": (_i32, [_vp, _vp, _vp]),
    "qsv_la_transpose": (_i32, [_i32, _vp, _vp]),
    "qsv_la_trace": (_i32, [_vp, _pd]),
    "qsv_la_deviation": (_i32, [_i32, _vp, _pd]),
    "qsv_la_solve": (_i32, [_vp, _vp, _vp]),
    "qsv_la_eigh": (_i32, [_vp, _pd, _vp, ctypes.POINTER(_i32)]),
    "qsv_la_eig": (_i32, [_vp, _pd, _vp]),
    "qsv_la_kron_apply": (_i32, [ctypes.POINTER(_pd), _i32, _vp]),
    "qsv_la_cholesky": (_i32, [_vp, _dbl, _vp]),
    "qsv_la_lu": (_i32, [_vp, _vp, _vp, ctypes.POINTER(ctypes.c_int64)]),
    "qsv_la_qr": (_i32, [_vp, _vp, _vp]),
}

QSV_LA_DEV_HERMITIAN, QSV_LA_DEV_OFFDIAG, QSV_LA_DEV_IDENTITY = 0, 1, 2


class QsvError(RuntimeError):
    """A libqsv call returned a non-zero status."""

    def __init__(self, code: int, message: str):
        super().__init__(f"libqsv error {code}: {message}")
        self.code = code
        self.message = message


_lib = None


def load_library(path: str | None = None) -> ctypes.CDLL:
    """Load libqsv.so (once) and attach the C signatures.

    Raises ``ImportError`` when the library has not been built: there is no
    fallback engine.
    """
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(
            f"libqsv.so not found at {p}: build it first with "
            "`python -c 'import __graft_entry__ as g; g.build()'` (hipcc, gfx950)")
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(code: int) -> None:
    if code != QSV_OK:
        msg = _lib.qsv_last_error().decode() if _lib is not None else ""
        raise QsvError(code, msg)


def ops_array(n: int) -> np.ndarray:
    return np.zeros(n, dtype=OP_DTYPE)

"""qsv_sample_noalloc (the allocation-free sampler) against qsv_sample.

Both run the same kernels on the same state; the no-alloc entry point keeps
its chunk prefix and device outcomes in the library's fixed per-GPU scratch
array. Outcomes are integers, so they must agree exactly. Two calls in a row,
the second on a smaller state with fewer shots, show that the scratch left by
one call does not leak into the next; the first call fills the scratch to its
limits (30 qubits, 2^19 shots)."""
import ctypes
import sys

import numpy as np
import pytest

from qclojure_amd import workloads as wl
from qclojure_amd._abi import (QSV_E_INVALID, QSV_FLAG_FUSE, QSV_NOALLOC_MAX_QUBITS, QSV_NOALLOC_MAX_SHOTS,
                               check)

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def eng():
    from qclojure_amd import engine
    engine.init(0)
    return engine


def _noalloc(eng, s, shots, seed):
    out = np.zeros(max(shots, 1), dtype=np.uint64)
    rc = eng.lib().qsv_sample_noalloc(s.handle, int(shots), seed, out.ctypes.data_as(ctypes.c_void_p))
    return rc, out[:shots]


def test_noalloc_matches_sample_two_calls(eng):
    cases = [
        (QSV_NOALLOC_MAX_QUBITS, "c64", 6, QSV_NOALLOC_MAX_SHOTS, 71),
        (12, "c128", 20, 1000, 72),
    ]
    for n, prec, depth, shots, seed in cases:
        note(f"noalloc: n={n} {prec} depth={depth} shots={shots} seed={seed}")
        ops = wl.random_layer_ops(n, depth, seed=seed)
        with eng.DeviceState(n, prec) as s:
            s.apply(ops, QSV_FLAG_FUSE)
            ref = s.sample(shots, seed)
            rc, got = _noalloc(eng, s, shots, seed)
            check(rc)
        assert got.shape == ref.shape
        mism = int(np.sum(got != ref))
        note(f"noalloc: n={n} mismatches {mism} of {shots}")
        assert mism == 0


def test_noalloc_rejects_beyond_limits(eng):
    with eng.DeviceState(12, "c64") as s:
        s.init_basis(5)
        rc, _ = _noalloc(eng, s, QSV_NOALLOC_MAX_SHOTS + 1, 1)
        assert rc == QSV_E_INVALID
        assert "scratch limit" in eng.lib().qsv_last_error().decode()
        rc, got = _noalloc(eng, s, 64, 1)
        check(rc)
        assert np.all(got == 5)
    with eng.DeviceState(QSV_NOALLOC_MAX_QUBITS + 1, "c64") as s:
        s.init_basis(0)
        rc, _ = _noalloc(eng, s, 16, 1)
        assert rc == QSV_E_INVALID
        assert "scratch limit" in eng.lib().qsv_last_error().decode()

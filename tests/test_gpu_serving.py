"""The native serving loop: results are the same alone and in the mix,
admission respects the HBM budget, failures are reported, latencies are
recorded."""
import threading
import time

import numpy as np
import pytest

from qclojure_amd import workloads as wl
from qclojure_amd._abi import OP, QSV_FLAG_FUSE, QsvError
from qclojure_amd import circuit as qc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from qclojure_amd import engine
    engine.init(0)
    return engine


def _alone(eng, n, ops, shots, seed, prec):
    with eng.DeviceState(n, prec) as s:
        s.apply(ops, QSV_FLAG_FUSE)
        return s.sample(shots, seed), s.download()


@pytest.mark.parametrize("prec", ["c64", "c128"])
def test_same_alone_and_in_the_mix(eng, prec):
    """Each request's shots equal the shots of the same circuit run alone
    (same seed): batched-LDS, fused-large and per-gate paths all reproduce it."""
    reqs = [(n, wl.random_layer_ops(n, d, seed=n * 7 + d), 200, 1000 + n) for n, d in
            [(10, 12), (11, 30), (12, 8), (13, 20), (14, 15), (15, 10), (16, 25), (18, 12), (20, 6), (22, 4)]]
    with eng.Server(prec) as srv:
        ids = [srv.submit(n, ops, shots, seed) for (n, ops, shots, seed) in reqs]
        got = [srv.wait(i, r[2]) for i, r in zip(ids, reqs)]
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    lds_n = 13 if prec == "c128" else 14   # largest state one workgroup's LDS holds
    assert st["small_batched_requests"] == sum(1 for r in reqs if r[0] <= lds_n)
    for (n, ops, shots, seed), (info, outs) in zip(reqs, got):
        assert info["status"] == "completed" and info["latency-ms"] > 0
        # every served result states its FP64 norm drift
        assert info["norm-drift"] < (1e-12 if prec == "c128" else 1e-6), (n, info)
        ref, psi = _alone(eng, n, ops, shots, seed, prec)
        # the batched sampler sums chunks in another order: allow rare boundary flips
        assert np.mean(outs == ref) > 0.99, (n, np.mean(outs == ref))
        p = np.abs(psi) ** 2
        assert np.all(p[outs.astype(np.int64)] > 0)


def test_admission_by_budget(eng):
    """With a budget of two 20-qubit c64 states, a burst of six is admitted two at a time."""
    budget = 2 * (1 << 20) * 8
    ops = wl.random_layer_ops(20, 3, seed=1)
    with eng.Server("c64", hbm_budget_bytes=budget) as srv:
        ids = [srv.submit(20, ops, 10, k) for k in range(6)]
        infos = [srv.wait(i, 10)[0] for i in ids]
        st = srv.stats()
    assert st["completed"] == 6
    assert st["max_queue_depth"] >= 3
    assert max(i["queue-ms"] for i in infos) > 0


def test_concurrent_submitters(eng):
    results = {}
    with eng.Server("c64") as srv:
        def worker(k):
            n = 10 + (k % 8)
            ops = wl.random_layer_ops(n, 5, seed=k)
            rid = srv.submit(n, ops, 50, k)
            results[k] = srv.wait(rid, 50)
        th = [threading.Thread(target=worker, args=(k,)) for k in range(24)]
        for t in th:
            t.start()
        for t in th:
            t.join()
    assert len(results) == 24 and all(v[0]["status"] == "completed" for v in results.values())


def test_bad_requests_fail_fast(eng):
    with eng.Server("c64") as srv:
        with pytest.raises(QsvError, match="out of range"):
            srv.submit(4, qc.ops_from_tuples([(OP["H"], 9, 0, 0, 0, 0.0)]), 10, 0)
        with pytest.raises(QsvError, match="not served"):
            srv.submit(4, qc.ops_from_tuples([(OP["MEASURE"], 0, 0, 0, 1, 0.0)]), 10, 0)
        with pytest.raises(QsvError):
            srv.status(12345)


def test_cancel(eng):
    """cancel-job (ideal_simulator.clj:154-163) on the serving loop: a queued
    request leaves the queue, an admitted one finishes as cancelled (or, if it
    already finished, cannot be cancelled), unknown ids are not found."""
    n = 22
    budget = (1 << n) * 8                       # one 22-qubit complex64 state at a time
    ops = wl.random_layer_ops(n, 100, seed=5)
    with eng.Server("c64", hbm_budget_bytes=budget) as srv:
        ids = [srv.submit(n, ops, 100, k) for k in range(4)]
        assert srv.cancel(ids[3]) == "cancelled"      # behind three others of the same size: still queued
        first = srv.cancel(ids[0])
        assert first in ("cancelled", "cannot-cancel")
        infos = [srv.wait(i, 100)[0] for i in ids[:3]]
        info3, outs3 = srv.wait(ids[3], 100)
        assert info3["status"] == "cancelled"
        assert infos[0]["status"] == ("cancelled" if first == "cancelled" else "completed")
        assert infos[1]["status"] == infos[2]["status"] == "completed"
        assert srv.cancel(ids[1]) == "not-found"      # collected by wait
        done = srv.submit(10, wl.random_layer_ops(10, 2, seed=1), 10, 1)
        while srv.status(done)["status"] != "completed":
            time.sleep(0.001)
        assert srv.cancel(done) == "cannot-cancel"
        assert srv.cancel(987654321) == "not-found"
        st = srv.stats()
        assert st["cancelled"] == 1 + (first == "cancelled")
        assert st["hbm_in_use_bytes"] == 0

"""Evaluates the reference test suite's checks (tests/golden/reference_kats.json)
against a final state vector, whoever produced it."""
import numpy as np


def as_vec(v):
    a = np.asarray(v, dtype=np.float64)
    return a[:, 0] + 1j * a[:, 1]


def run_checks(case, psi, init):
    for chk in case["checks"]:
        kind = chk["check"]
        tol = chk.get("tol", 0.0)
        where = f"{case['id']} ({case['source']}) {kind}"
        if kind == "state":
            exp = as_vec(chk["value"])
            if tol == 0.0:
                assert np.array_equal(psi, exp), f"{where}: {psi} != {exp}"
            else:
                assert np.max(np.abs(psi - exp)) < tol, f"{where}: {psi} vs {exp}"
        elif kind == "amplitude":
            exp = complex(*chk["value"])
            d = abs(psi[chk["index"]] - exp)
            assert (d == 0.0) if tol == 0.0 else (d < tol), f"{where}: {psi[chk['index']]} vs {exp}"
        elif kind == "abs":
            assert abs(abs(psi[chk["index"]]) - chk["value"]) < tol, f"{where}: {psi}"
        elif kind == "abs-greater":
            assert abs(psi[chk["index"]]) > chk["value"], f"{where}: {psi}"
        elif kind == "imag":
            d = abs(psi[chk["index"]].imag - chk["value"])
            assert (d == 0.0) if tol == 0.0 else (d < tol), f"{where}: {psi}"
        elif kind == "real":
            d = abs(psi[chk["index"]].real - chk["value"])
            assert (d == 0.0) if tol == 0.0 else (d < tol), f"{where}: {psi}"
        elif kind == "state-changed":
            assert not np.array_equal(psi, init), f"{where}: state unchanged"
        else:
            raise AssertionError(f"unknown check {kind}")


def check_expectation(case, value):
    if case["tol"] == "exact":
        assert value == case["value"], f"{case['id']} ({case['source']}): {value!r} != {case['value']!r}"
    else:
        assert abs(value - case["value"]) < case["tol"], f"{case['id']}: {value} vs {case['value']}"


def check_shots(case, outcomes):
    from collections import Counter
    freq = Counter(int(x) for x in outcomes)
    assert set(freq) <= set(case["allowed"]), f"{case['id']}: outcomes {dict(freq)}"
    if "distinct" in case:
        assert len(freq) == case["distinct"], f"{case['id']}: {dict(freq)}"
    for o in case.get("must-contain", []):
        assert o in freq, f"{case['id']}: missing {o} in {dict(freq)}"
    if "ratio-of" in case:
        a, b = case["ratio-of"]
        r = freq.get(a, 0) / (freq.get(a, 0) + freq.get(b, 0))
        lo, hi = case["ratio-window"]
        assert lo < r < hi, f"{case['id']}: ratio {r}"
    if "count" in case:
        o, c = case["count"]
        assert freq.get(o, 0) == c, f"{case['id']}: {dict(freq)}"

"""Pauli-term Hamiltonians (src/org/soulspace/qclojure/domain/hamiltonian.clj).

A Hamiltonian is a list of ``{"coefficient": c, "pauli-string": "IZZX"}``;
character k acts on qubit k (qubit 0 is the most significant index bit).
Expectation values are evaluated on the device (one FP64 reduction pass per
8 terms); the host sums c_k <P_k> sequentially in term order, as the
reference's ``(reduce + ...)`` does (hamiltonian.clj:96-114).
"""
from __future__ import annotations

from collections import OrderedDict

_PAULI = set("IXYZ")


def pauli_term(coefficient: float, pauli_string: str) -> dict:
    """hamiltonian.clj:33-45."""
    if not isinstance(pauli_string, str) or not set(pauli_string) <= _PAULI:
        raise ValueError(f"invalid pauli string {pauli_string!r}")
    return {"coefficient": coefficient, "pauli-string": pauli_string}


def hamiltonian(pairs) -> list[dict]:
    return [pauli_term(c, s) for (c, s) in pairs]


def validate_hamiltonian(h) -> bool:
    """hamiltonian.clj:61-73."""
    try:
        lengths = set()
        for t in h:
            c, s = t["coefficient"], t["pauli-string"]
            if not isinstance(c, (int, float)) or not isinstance(s, str):
                return False
            lengths.add(len(s))
        return len(lengths) <= 1
    except (TypeError, KeyError):
        return False


def hamiltonian_expectation(h, state) -> float:
    """sum_k c_k <psi|P_k|psi> on a device state (hamiltonian.clj:96-114)."""
    if not validate_hamiltonian(h):
        raise ValueError("invalid hamiltonian")
    if not h:
        return 0
    vals = state.expect_pauli([t["pauli-string"] for t in h])
    total = None
    for t, v in zip(h, vals):
        x = t["coefficient"] * float(v)
        total = x if total is None else total + x
    return total


def _key(t):
    return (t["coefficient"], t["pauli-string"])


def group_commuting_terms(h) -> list[list[dict]]:
    """Greedy grouping (hamiltonian.clj:172-204)."""
    ungrouped = list(h)
    groups = []
    while ungrouped:
        cur = ungrouped[0]["pauli-string"]
        commuting = [t for t in ungrouped
                     if sum(1 for a, b in zip(cur, t["pauli-string"])
                            if a != "I" and b != "I" and a != b) % 2 == 0]
        keys = {_key(t) for t in commuting}
        ungrouped = [t for t in ungrouped if _key(t) not in keys]
        groups.append(commuting)
    return groups


def group_pauli_terms_by_measurement_basis(h) -> dict:
    """hamiltonian.clj:206-232 (group-by keeps first-seen key order)."""
    out: "OrderedDict[str, list]" = OrderedDict()
    for t in h:
        u = set(t["pauli-string"]) - {"I"}
        key = ("identity" if not u else "z" if u == {"Z"} else "x" if u == {"X"}
               else "y" if u == {"Y"} else "mixed")
        out.setdefault(key, []).append(t)
    return dict(out)

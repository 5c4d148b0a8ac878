"""qclojure_amd — an MI355X state-vector engine behind QClojure's backend protocol.

The gate-application path and every result reduction run in libqsv
(hand-written HIP kernels for gfx950, C ABI in include/qsv.h). This package is
the host-side mirror of the reference's backend protocol, circuit data and
result extraction.
"""
from . import _abi, circuit, algorithms, hamiltonian, workloads  # noqa: F401
from .circuit import decode_circuit  # noqa: F401

__version__ = "0.1.0"


def load_engine():
    """Import the device engine (loads libqsv.so; raises if it is not built)."""
    from . import engine as _engine
    _engine.lib()
    return _engine

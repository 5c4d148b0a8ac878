"""Text conversions with the JVM's semantics, for the wire formats.

The reference writes numbers into QASM, JSON and EDN text with Clojure's
``str``/``pr-str``, which print a double with ``Double.toString`` (JDK 19+:
the shortest decimal that reads back to the same double, with at least two
significant digits considered when one would do), and reads them back with
``Double.parseDouble``. This module restates both, plus the few other
Clojure printing rules the formats rely on, so that text written here is the
text the reference writes for the same circuit.
"""
from __future__ import annotations

import math
import re
from decimal import Decimal


class JvmNumberFormatError(ValueError):
    """What ``Double.parseDouble``/``Integer.parseInt`` reject (NumberFormatException)."""


def _digits_exponent(text: str) -> tuple[str, int]:
    """Significant digits and decimal exponent E (value = d.ddd x 10^E) of a
    positive decimal string, trailing zeros removed."""
    sign, digits, exp = Decimal(text).as_tuple()
    ds = "".join(map(str, digits)).lstrip("0")
    stripped = ds.rstrip("0")
    exp += len(ds) - len(stripped)
    return stripped, exp + len(stripped) - 1


def java_double_str(x: float) -> str:
    """``Double.toString`` (java/lang/Double.java, JDK 19+ specification).

    Finite non-zero values in [1e-3, 1e7) print as plain decimals with at
    least one fraction digit ("0.5", "100.0"); others in computerized
    scientific notation ("1.0E-5", "1.2345E10"). The digits are the shortest
    that read back to x; where that is a single digit, the closest two-digit
    decimal is taken instead ("4.9E-324" for the smallest subnormal)."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Infinity" if x > 0 else "-Infinity"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    sign = "-" if x < 0 else ""
    a = abs(x)
    digits, e = _digits_exponent(repr(a))
    if len(digits) == 1:
        digits, e = _digits_exponent(format(a, ".1e"))
    if 1e-3 <= a < 1e7:
        if e >= 0:
            ip = digits[: e + 1].ljust(e + 1, "0")
            fp = digits[e + 1:] or "0"
        else:
            ip = "0"
            fp = "0" * (-e - 1) + digits
        return f"{sign}{ip}.{fp}"
    return f"{sign}{digits[0]}.{digits[1:] or '0'}E{e}"


def clj_str(v) -> str:
    """Clojure ``str`` of one value as the host mirror represents it: None is
    "", booleans "true"/"false", integers in decimal, floats as
    ``Double.toString``, strings as themselves."""
    if v is None:
        return ""
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, int):
        return str(v)
    if isinstance(v, float):
        return java_double_str(v)
    try:
        import numpy as np
        if isinstance(v, np.integer):
            return str(int(v))
        if isinstance(v, np.floating):
            return java_double_str(float(v))
    except ImportError:  # pragma: no cover
        pass
    return str(v)


def clj_keyword_str(v) -> str:
    """``(str kw)`` of a keyword held as its name: ":name"."""
    return ":" + str(v)


_JAVA_DEC = re.compile(r"([+-]?)(?:(NaN)|(Infinity)|((?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)[fFdD]?)\Z",
                       re.ASCII)
_JAVA_HEX = re.compile(r"([+-]?)0[xX]([0-9a-fA-F]+\.?|[0-9a-fA-F]*\.[0-9a-fA-F]+)[pP]([+-]?\d+)[fFdD]?\Z",
                       re.ASCII)


def java_parse_double(s) -> float:
    """``Double.parseDouble``: leading/trailing characters <= ' ' are
    trimmed; accepts an optional sign, "NaN", "Infinity", decimal digits with
    optional point and exponent, hexadecimal floating literals, and a type
    suffix f/F/d/D. Anything else raises JvmNumberFormatError."""
    if s is None:
        raise JvmNumberFormatError("null")
    t = str(s)
    i, j = 0, len(t)
    while i < j and ord(t[i]) <= 32:
        i += 1
    while j > i and ord(t[j - 1]) <= 32:
        j -= 1
    t = t[i:j]
    m = _JAVA_DEC.match(t)
    if m:
        sign = -1.0 if m.group(1) == "-" else 1.0
        if m.group(2):
            return math.nan
        if m.group(3):
            return sign * math.inf
        return sign * float(m.group(4))
    m = _JAVA_HEX.match(t)
    if m:
        mant = m.group(2)
        v = float.fromhex("0x" + (mant if "." in mant else mant + ".") + "p" + m.group(3))
        return -v if m.group(1) == "-" else v
    raise JvmNumberFormatError(f'For input string: "{t}"')


_JAVA_INT = re.compile(r"[+-]?\d+\Z", re.ASCII)


def java_parse_int(s, bits: int = 32) -> int:
    """``Integer.parseInt`` (bits=32) / ``Long.parseLong`` (bits=64): an
    optional sign and ASCII digits, in range, else JvmNumberFormatError."""
    t = "" if s is None else str(s)
    if not _JAVA_INT.match(t):
        raise JvmNumberFormatError(f'For input string: "{t}"')
    v = int(t)
    if not -(1 << (bits - 1)) <= v < (1 << (bits - 1)):
        raise JvmNumberFormatError(f'For input string: "{t}"')
    return v


def java_div(a: float, b: float) -> float:
    """IEEE division as the JVM does it for doubles (x/0 is ±Infinity or NaN)."""
    a, b = float(a), float(b)
    if b == 0.0:
        if a == 0.0 or math.isnan(a):
            return math.nan
        neg = (math.copysign(1.0, a) < 0) != (math.copysign(1.0, b) < 0)
        return -math.inf if neg else math.inf
    return a / b
